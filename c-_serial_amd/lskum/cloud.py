"""Point clouds: the reference's partGrid content (main.cpp:30-58) as numpy
arrays, the legacy text reader/writer and the synthetic NACA generator, all
through liblskum.so."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class Cloud:
    """0-based cloud: left/right are -1 for interior points; neighbours of
    point i are conn[conn_off[i]:conn_off[i+1]] in file order."""
    x: np.ndarray
    y: np.ndarray
    left: np.ndarray
    right: np.ndarray
    flag_1: np.ndarray
    flag_2: np.ndarray
    min_dist: np.ndarray
    conn_off: np.ndarray
    conn: np.ndarray

    @property
    def n(self):
        return int(self.x.shape[0])

    def nbhs(self):
        return np.diff(self.conn_off)

    def normalized(self):
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        return Cloud(f64(self.x), f64(self.y), i32(self.left), i32(self.right),
                     i32(self.flag_1), i32(self.flag_2), f64(self.min_dist),
                     i32(self.conn_off), i32(self.conn))

    def view(self):
        """CloudView pointing into this cloud's arrays (keep self alive)."""
        c = self.normalized()
        self.__dict__.update(c.__dict__)
        v = L.CloudView()
        v.n_points = c.n
        v.x, v.y = L.dptr(c.x), L.dptr(c.y)
        v.left, v.right = L.iptr(c.left), L.iptr(c.right)
        v.flag_1, v.flag_2 = L.iptr(c.flag_1), L.iptr(c.flag_2)
        v.min_dist = L.dptr(c.min_dist)
        v.conn_off, v.conn = L.iptr(c.conn_off), L.iptr(c.conn)
        return v

    # ---------------------------------------------------------------- I/O
    @staticmethod
    def _from_handle(h):
        v = L.CloudView()
        L.check(L.lib().lskum_cloud_view(h, C.byref(v)))
        n = v.n_points
        nnz = v.conn_off[n]
        arr = lambda p, k, dt: np.ctypeslib.as_array(p, shape=(k,)).astype(dt, copy=True)
        return Cloud(arr(v.x, n, np.float64), arr(v.y, n, np.float64),
                     arr(v.left, n, np.int32), arr(v.right, n, np.int32),
                     arr(v.flag_1, n, np.int32), arr(v.flag_2, n, np.int32),
                     arr(v.min_dist, n, np.float64), arr(v.conn_off, n + 1, np.int32),
                     arr(v.conn, nnz, np.int32) if nnz else np.zeros(0, np.int32))

    @classmethod
    def read_partgrid(cls, path):
        h = C.c_void_p()
        L.check(L.lib().lskum_cloud_read_partgrid(str(path).encode(), C.byref(h)))
        try:
            return cls._from_handle(h)
        finally:
            L.lib().lskum_cloud_free(h)

    def _handle(self):
        v = self.view()
        h = C.c_void_p()
        L.check(L.lib().lskum_cloud_from_arrays(C.byref(v), C.byref(h)))
        return h

    def write_partgrid(self, path):
        h = self._handle()
        try:
            L.check(L.lib().lskum_cloud_write_partgrid(h, str(path).encode()))
        finally:
            L.lib().lskum_cloud_free(h)

    def connectivity(self):
        """(nx, ny, mask): host set-up as lskum_create does it; mask bits per
        conn entry 1 xpos, 2 xneg, 4 ypos, 8 yneg."""
        h = self._handle()
        try:
            nx = np.zeros(self.n)
            ny = np.zeros(self.n)
            mask = np.zeros(int(self.conn_off[-1]), np.uint8)
            L.check(L.lib().lskum_cloud_connectivity(
                h, L.dptr(nx), L.dptr(ny), mask.ctypes.data_as(C.POINTER(C.c_uint8))))
            return nx, ny, mask
        finally:
            L.lib().lskum_cloud_free(h)

    def rcm(self):
        """Reverse Cuthill-McKee order, perm[new] = old."""
        h = self._handle()
        try:
            perm = np.zeros(self.n, np.int32)
            L.check(L.lib().lskum_cloud_rcm(h, L.iptr(perm)))
            return perm
        finally:
            L.lib().lskum_cloud_free(h)

    def stats(self):
        """(min nbhs, max nbhs, min relative sub-stencil det per point type)."""
        h = self._handle()
        try:
            lo, hi = C.c_int32(), C.c_int32()
            dets = (C.c_double * 3)()
            L.check(L.lib().lskum_cloud_stats(h, C.byref(lo), C.byref(hi), dets))
            return lo.value, hi.value, list(dets)
        finally:
            L.lib().lskum_cloud_free(h)

    @classmethod
    def generate(cls, naca="0012", n_wall=160, n_radial=64, n_circ=150, r_far=20.0,
                 stretch=1.08, jitter=0.1, k=20, radius_factor=3.0, seed=1, wall_blend=0.0):
        """Synthetic NACA 4-digit cloud (see lskum_gen_params in lskum.h)."""
        digits = str(naca)
        if len(digits) != 4 or not digits.isdigit():
            raise ValueError("naca must be a 4-digit string such as '0012'")
        p = L.GenParams(int(digits[0]), int(digits[1]), int(digits[2:]), n_wall, n_radial,
                        n_circ, r_far, stretch, jitter, k, radius_factor, seed, wall_blend)
        h = C.c_void_p()
        L.check(L.lib().lskum_cloud_generate(C.byref(p), C.byref(h)))
        try:
            return cls._from_handle(h)
        finally:
            L.lib().lskum_cloud_free(h)


# Named workloads: the clouds the benchmark and the tests use.
WORKLOADS = {
    # NACA0012, 160 wall, 64 x 150 O-grid, R = 20, stretch 1.08, k = 20
    "naca0012_9760_k20": dict(naca="0012", n_wall=160, n_radial=64, n_circ=150,
                              stretch=1.08, k=20, seed=1),
    # NACA0012, 2,000 wall, 500 x 5,000, k = 24 (stretch matched to spacing)
    "naca0012_2p5m_k24": dict(naca="0012", n_wall=2000, n_radial=500, n_circ=5000,
                              stretch=0.0, k=24, seed=2),
    # NACA0012, 1,500 wall, 806 x 1,240 near-isotropic O-grid, k = 24 (the
    # shape-gradient cloud; a 266 x 3,750 layout of the same size, far
    # cells 14 times longer radially than around, loses stability within
    # a few hundred iterations at M 0.8).  The wall points' spacing is
    # blended half way from cosine to uniform arc length: cosine spacing
    # alone puts the trailing-edge points 4.4e-6 chords apart against
    # 1.6e-3 to the first ring, and the iteration then ran away at the
    # trailing edge after about 9,000 iterations even at CFL 0.1; blended
    # (wall spacing 6.8e-4 to 1.7e-3) it ran 120,000 iterations at CFL 0.2
    "naca0012_1m_k24": dict(naca="0012", n_wall=1500, n_radial=806, n_circ=1240,
                            stretch=0.0, k=24, seed=5, wall_blend=0.5),
    # NACA0012, 4,000 wall, 1,000 x 10,000, k = 24
    "naca0012_10m_k24": dict(naca="0012", n_wall=4000, n_radial=1000, n_circ=10000,
                             stretch=0.0, k=24, seed=3),
    # small clouds for parity tests
    "naca0012_1480_k20": dict(naca="0012", n_wall=40, n_radial=24, n_circ=60,
                              stretch=1.2, k=20, seed=11),
    "naca0012_3620_k24": dict(naca="0012", n_wall=80, n_radial=30, n_circ=118,
                              stretch=0.0, k=24, seed=12),
}


def workload(name):
    return Cloud.generate(**WORKLOADS[name])
