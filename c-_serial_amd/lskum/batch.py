"""Many flow conditions over several clouds in one context.

``Batch`` wraps lskum_create_batch: every stage launch and every iteration
advance all cases together, and each case's arithmetic is that of a
``Solver`` of its own, so per-case results are bit-identical to individual
runs.  ``batch_workload`` defines the seeded many-case workload: 8 NACA
4-digit airfoil clouds of uneven sizes and 64 flow conditions.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L
from .cloud import Cloud


class Batch:
    def __init__(self, clouds, cases, cfl=0.2, vl_const=5.0, power=0.0, inner_iterations=0,
                 precision=L.FP64, reorder=L.REORDER_HILBERT, lds_tiles=L.TILES_AUTO,
                 tile_block=0, points_per_thread=0, graph_iters=0):
        """clouds: list of Cloud; cases: list of (cloud index, mach, aoa_deg)."""
        self.clouds = list(clouds)
        self.cases = [(int(c), float(m), float(a)) for c, m, a in cases]
        self.params = L.Params(self.cases[0][1], self.cases[0][2], cfl, vl_const, power,
                               inner_iterations, precision, reorder, lds_tiles, tile_block,
                               points_per_thread)
        self._views = [c.view() for c in self.clouds]
        views = (L.CloudView * len(self._views))(*self._views)
        cs = (L.Case * len(self.cases))(*[L.Case(c, m, a) for c, m, a in self.cases])
        ctx = C.c_void_p()
        rc = L.lib().lskum_create_batch(views, len(self._views), cs, len(self.cases),
                                        C.byref(self.params), C.byref(ctx))
        self._ctx = ctx
        if rc != L.LSKUM_OK:
            msg = L.lib().lskum_last_error(ctx) if ctx else b""
            if ctx:
                L.lib().lskum_destroy(ctx)
                self._ctx = None
            raise L.LskumError(rc, (msg or b"").decode())
        self.n_cases = len(self.cases)
        self.points = np.array([self.clouds[c].n for c, _, _ in self.cases], np.int64)
        if graph_iters:
            self.set_graph(graph_iters)

    def close(self):
        if getattr(self, "_ctx", None):
            L.lib().lskum_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        L.check(rc, self._ctx)

    def set_graph(self, iters_per_graph):
        self._chk(L.lib().lskum_set_graph(self._ctx, iters_per_graph))

    def q_lskum(self, n_iter):
        """n_iter iterations of every case; (res_new, residue) as [n_cases, n_iter]."""
        rn = np.zeros((self.n_cases, max(n_iter, 1)))
        rs = np.zeros_like(rn)
        self._chk(L.lib().lskum_iterate(self._ctx, n_iter, L.dptr(rn), L.dptr(rs)))
        return rn[:, :n_iter], rs[:, :n_iter]

    def failed_at(self):
        """Per case, the first iteration with a non-finite residual (0: none)."""
        out = np.zeros(self.n_cases, np.int32)
        self._chk(L.lib().lskum_case_status(self._ctx, L.iptr(out)))
        return out

    def compute_cl_cd_cm(self):
        """[n_cases, 3] of Cl, Cd, Cm."""
        cl, cd, cm = (np.zeros(self.n_cases) for _ in range(3))
        self._chk(L.lib().lskum_forces(self._ctx, L.dptr(cl), L.dptr(cd), L.dptr(cm)))
        return np.stack([cl, cd, cm], axis=1)

    def get(self, which, ncomp):
        """Per-case list of [ncomp, n_case] arrays."""
        out = np.zeros(int(ncomp * self.points.sum()))
        self._chk(L.lib().lskum_get_array(self._ctx, which, L.dptr(out)))
        parts, o = [], 0
        for n in self.points:
            parts.append(out[o:o + ncomp * n].reshape(ncomp, n))
            o += ncomp * n
        return parts

    def prim(self):
        return self.get(L.A_PRIM, 4)

    def diagnostics(self, which):
        """Per case: "entropy" -> ([entropy per point], totals), "enstrophy" ->
        ([vorticity per point], totals), "stagnation" -> ([n_cases, 4] values,
        [n_cases, 2] points), "J" -> [n_cases] stagnation-pressure losses
        (compute_entropy_mod.h, compute_enstrophy_mod.h, stagnation_values_mod.h)."""
        nc, tot = len(self.cases), int(self.points.sum())

        def split(a):
            return np.split(a, np.cumsum(self.points)[:-1])

        if which == "entropy":
            e, t = np.zeros(tot), np.zeros(nc)
            self._chk(L.lib().lskum_compute_entropy(self._ctx, L.dptr(e), L.dptr(t)))
            return split(e), t
        if which == "enstrophy":
            v, t = np.zeros(tot), np.zeros(nc)
            self._chk(L.lib().lskum_compute_enstrophy(self._ctx, L.dptr(v), None, L.dptr(t)))
            return split(v), t
        if which == "stagnation":
            vals, idx = np.zeros((nc, 4)), np.zeros((nc, 2), np.int32)
            self._chk(L.lib().lskum_stagnation_pressure(self._ctx, L.dptr(vals), L.iptr(idx)))
            return vals, idx
        if which == "J":
            J = np.zeros(nc)
            self._chk(L.lib().lskum_objective_J(self._ctx, L.dptr(J)))
            return J
        raise ValueError(which)

    def set_timing(self, enable):
        self._chk(L.lib().lskum_set_timing(self._ctx, 1 if enable else 0))

    def kernel_time(self, kernel):
        ms, cnt = C.c_double(), C.c_int32()
        self._chk(L.lib().lskum_kernel_time(self._ctx, kernel, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


def autotune(clouds, cases, iters=5, cfl=0.2, vl_const=5.0, power=0.0, inner_iterations=0,
             precision=L.FP64, reorder=L.REORDER_HILBERT):
    """lskum_autotune over the given cases: returns (best, trials) as dicts
    with tile_block, lds_tiles, points_per_thread and ms_per_iter."""
    cases = [(int(c), float(m), float(a)) for c, m, a in cases]
    views = [c.view() for c in clouds]
    va = (L.CloudView * len(views))(*views)
    cs = (L.Case * len(cases))(*[L.Case(c, m, a) for c, m, a in cases])
    p = L.Params(cases[0][1], cases[0][2], cfl, vl_const, power, inner_iterations, precision,
                 reorder, L.TILES_AUTO, 0, 0)
    best = L.TuneResult()
    trials = (L.TuneResult * 16)()
    n = C.c_int32()
    L.check(L.lib().lskum_autotune(va, len(views), cs, len(cases), C.byref(p), iters,
                                   C.byref(best), trials, 16, C.byref(n)))
    as_dict = lambda r: {"tile_block": r.tile_block, "lds_tiles": r.lds_tiles,
                         "points_per_thread": r.points_per_thread,
                         "ms_per_iter": r.ms_per_iter}
    return as_dict(best), [as_dict(trials[i]) for i in range(n.value)]


def _grid_for(points):
    """(n_wall, n_radial, n_circ) of a cloud of about `points` points whose
    cells are near-isotropic from the wall to the far field: with the radial
    stretch matched to the wall's circumferential spacing (about 2.04/n_circ
    for a unit chord) and growing to the far-field circle's (2 pi 20/n_circ),
    the spacing grows 62-fold over 19.5 chords, which takes n_radial = 0.65
    n_circ geometric layers.  (Fewer layers leave far-field cells many times
    longer radially than circumferentially, and the explicit scheme then
    loses stability there after several hundred iterations.)
    n_wall = 0.4 n_circ; points = n_wall + n_radial*n_circ."""
    n_circ = max(60, int(round(math.sqrt(points / 0.65))))
    n_radial = max(12, int(round((points - 0.4 * n_circ) / n_circ)))
    n_wall = max(40, 2 * int(round(0.2 * n_circ)))
    return n_wall, n_radial, n_circ


def batch_workload(n_clouds=8, conditions_per_cloud=8, min_points=50_000, max_points=800_000,
                   k=20, seed=4):
    """The seeded many-case workload: n_clouds NACA 4-digit airfoils
    (thickness 8-15 %, camber 0-4 % at 20-60 % chord; camber 0 gives a
    symmetric section) whose cloud sizes step geometrically from min_points
    to max_points, each with conditions_per_cloud flow conditions (Mach
    uniform in [0.5, 0.85], AoA uniform in [-2, 6] deg).  Returns (clouds,
    cases) with cases = [(cloud index, mach, aoa_deg)], cloud-major."""
    rng = np.random.Generator(np.random.PCG64(seed))
    clouds, cases = [], []
    for j in range(n_clouds):
        t = int(rng.integers(8, 16))
        m = int(rng.integers(0, 5))
        p = int(rng.integers(2, 7)) if m else 0
        frac = j / max(1, n_clouds - 1)
        pts = min_points * (max_points / min_points) ** frac
        n_wall, n_radial, n_circ = _grid_for(pts)
        clouds.append(Cloud.generate(naca=f"{m}{p}{t:02d}", n_wall=n_wall, n_radial=n_radial,
                                     n_circ=n_circ, stretch=0.0, k=k, seed=seed * 1000 + j))
        for _ in range(conditions_per_cloud):
            cases.append((j, float(rng.uniform(0.5, 0.85)), float(rng.uniform(-2.0, 6.0))))
    return clouds, cases
