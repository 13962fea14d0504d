"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, and only there as the checker or the
timed CPU baseline.  The product (c-_serial_amd/lskum) never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
        h = C.CDLL(LIB_PATH)
        h.orc_create.restype = C.c_void_p
        h.orc_create.argtypes = [C.c_int, _dp, _dp, _ip, _ip, _ip, _ip, _dp, _ip, _ip]
        h.orc_destroy.argtypes = [C.c_void_p]
        h.orc_set_params.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_int]
        h.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        for name in ["orc_compute_normals", "orc_generate_connectivity",
                     "orc_setup_case_parameters", "orc_initial_conditions", "orc_func_delta",
                     "orc_eval_q_variables", "orc_eval_q_derivatives", "orc_eval_q_inner_loop",
                     "orc_eval_update_innerloop", "orc_cal_flux_residual",
                     "orc_compute_cl_cd_cm"]:
            getattr(h, name).argtypes = [C.c_void_p]
            getattr(h, name).restype = None
        h.orc_state_update.argtypes = [C.c_void_p, C.c_int]
        h.orc_fpi_solver.argtypes = [C.c_void_p, C.c_int]
        h.orc_q_lskum.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        h.orc_qtilde_to_primitive.argtypes = [_dp, _dp, _dp, _dp, _dp]
        h.orc_flux.argtypes = [C.c_int] + [C.c_double] * 6 + [_dp]
        h.orc_array.restype = _dp
        h.orc_array.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        h.orc_sub_count.argtypes = [C.c_void_p, C.c_int, C.c_int]
        h.orc_sub_list.restype = _ip
        h.orc_sub_list.argtypes = [C.c_void_p, C.c_int, C.c_int]
        h.orc_scalars.argtypes = [C.c_void_p, _dp]
        h.orc_set_iteration.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        h.orc_compute_entropy.argtypes = [C.c_void_p, _dp]
        h.orc_compute_entropy.restype = C.c_double
        h.orc_compute_enstrophy.argtypes = [C.c_void_p, _dp, _dp]
        h.orc_compute_enstrophy.restype = C.c_double
        h.orc_stagnation_pressure.argtypes = [C.c_void_p, _dp, _ip]
        h.orc_objective_J.argtypes = [C.c_void_p]
        h.orc_objective_J.restype = C.c_double
        _lib = h
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Oracle:
    """The reference's solver state for one cloud, restated in C."""

    ARRAYS = {"prim": 0, "prim_old": 1, "q": 2, "dq": 3, "qm": 4, "delta": 5,
              "flux_res": 6, "nx": 7, "ny": 8, "res_pt": 9}

    def __init__(self, cloud, mach=0.63, aoa=2.0, cfl=0.1, vl_const=150.0, power=0.0,
                 inner_iterations=3, threads=1):
        L = lib()
        self._keep = [_d(cloud.x), _d(cloud.y), _i(cloud.left), _i(cloud.right),
                      _i(cloud.flag_1), _i(cloud.flag_2), _d(cloud.min_dist),
                      _i(cloud.conn_off), _i(cloud.conn)]
        k = self._keep
        ptrs = [k[0].ctypes.data_as(_dp), k[1].ctypes.data_as(_dp)]
        ptrs += [a.ctypes.data_as(_ip) for a in k[2:6]]
        ptrs += [k[6].ctypes.data_as(_dp), k[7].ctypes.data_as(_ip), k[8].ctypes.data_as(_ip)]
        self.n = cloud.n
        self.c = L.orc_create(self.n, *ptrs)
        L.orc_set_params(self.c, mach, aoa, cfl, vl_const, power, inner_iterations)
        L.orc_set_threads(self.c, threads)
        L.orc_compute_normals(self.c)
        L.orc_generate_connectivity(self.c)
        L.orc_initial_conditions(self.c)

    def __del__(self):
        if getattr(self, "c", None):
            lib().orc_destroy(self.c)
            self.c = None

    # reference stage functions
    def func_delta(self): lib().orc_func_delta(self.c)
    def eval_q_variables(self): lib().orc_eval_q_variables(self.c)
    def eval_q_derivatives(self): lib().orc_eval_q_derivatives(self.c)
    def eval_q_inner_loop(self): lib().orc_eval_q_inner_loop(self.c)
    def eval_update_innerloop(self): lib().orc_eval_update_innerloop(self.c)
    def cal_flux_residual(self): lib().orc_cal_flux_residual(self.c)
    def state_update(self, rk): lib().orc_state_update(self.c, rk)
    def compute_cl_cd_cm(self): lib().orc_compute_cl_cd_cm(self.c)
    def fpi_solver(self, it): lib().orc_fpi_solver(self.c, it)

    def snapshot(self):
        self.set("prim_old", self.get("prim"))

    # diagnostics of the current state
    def compute_entropy(self):
        """(entropy [n], total) -- compute_entropy_mod.h:5-24"""
        e = np.zeros(self.n)
        t = lib().orc_compute_entropy(self.c, e.ctypes.data_as(_dp))
        return e, t

    def compute_enstrophy(self):
        """(vorticity [n], vorticity_sqr [n], total) -- compute_enstrophy_mod.h:5-76"""
        v, v2 = np.zeros(self.n), np.zeros(self.n)
        t = lib().orc_compute_enstrophy(self.c, v.ctypes.data_as(_dp), v2.ctypes.data_as(_dp))
        return v, v2, t

    def stagnation_pressure(self):
        """(pMin, pMax, pMin/p0_inf, pMax/p0_inf), (indexMin, indexMax) 0-based --
        stagnation_values_mod.h:8-46"""
        out = np.zeros(4)
        idx = np.zeros(2, np.int32)
        lib().orc_stagnation_pressure(self.c, out.ctypes.data_as(_dp), idx.ctypes.data_as(_ip))
        return out, idx

    def objective_J(self):
        """stagnation-pressure loss, stagnation_values_mod.h:49-80 (the sum the
        reference forms, in place of its uninitialised total_p0)"""
        return lib().orc_objective_J(self.c)

    def q_lskum(self, n_iter):
        rn = np.zeros(n_iter)
        rs = np.zeros(n_iter)
        lib().orc_q_lskum(self.c, n_iter, rn.ctypes.data_as(_dp), rs.ctypes.data_as(_dp))
        return rn, rs

    def _view(self, name):
        ln = C.c_int()
        p = lib().orc_array(self.c, self.ARRAYS[name], C.byref(ln))
        return np.ctypeslib.as_array(p, shape=(ln.value,))

    def get(self, name):
        a = self._view(name).copy()
        if name in ("prim", "prim_old", "q", "flux_res"):
            return a.reshape(4, self.n)
        if name in ("dq", "qm"):
            return a.reshape(2, 4, self.n)
        return a

    def set(self, name, value):
        self._view(name)[:] = np.asarray(value, dtype=np.float64).ravel()

    def substencil(self, which):
        """(counts, members) of sub-stencil 0 xpos 1 xneg 2 ypos 3 yneg."""
        L = lib()
        counts = np.array([L.orc_sub_count(self.c, which, i) for i in range(self.n)], np.int32)
        members = []
        for i in range(self.n):
            if counts[i]:
                p = L.orc_sub_list(self.c, which, i)
                members.append(np.ctypeslib.as_array(p, shape=(counts[i],)).copy())
        return counts, (np.concatenate(members) if members else np.zeros(0, np.int32))

    def scalars(self):
        """dict of the globals the stage functions set."""
        out = np.zeros(11)
        lib().orc_scalars(self.c, out.ctypes.data_as(_dp))
        keys = ["sum_res_sqr", "max_res", "res_old", "res_new", "residue", "max_res_point",
                "it", "cl", "cd", "cm", "theta"]
        d = dict(zip(keys, out.tolist()))
        d["max_res_point"] = int(d["max_res_point"])
        d["it"] = int(d["it"])
        return d

    def set_iteration(self, it, res_old, restart=1):
        lib().orc_set_iteration(self.c, it, res_old, restart)


def flux(which, nx, ny, u1, u2, rho, pr):
    g = np.zeros(4)
    lib().orc_flux(which, nx, ny, u1, u2, rho, pr, g.ctypes.data_as(_dp))
    return g


def qtilde_to_primitive(qt):
    qt = _d(qt)
    out = [C.c_double() for _ in range(4)]
    lib().orc_qtilde_to_primitive(qt.ctypes.data_as(_dp), *[C.byref(o) for o in out])
    return np.array([o.value for o in out])
