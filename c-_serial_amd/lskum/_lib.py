"""ctypes binding of include/lskum.h (liblskum.so).

The library is loaded from this package directory only.  There is no
fallback: if liblskum.so is missing or fails to load, every entry point
raises, so a run can never silently compute on something else.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblskum.so")

LSKUM_OK = 0
LSKUM_EINVAL, LSKUM_EHIP, LSKUM_ENOMEM, LSKUM_EIO, LSKUM_ENONFINITE, LSKUM_ESTENCIL = (
    -1, -2, -3, -4, -5, -6)
ERRORS = {
    -1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EIO", -5: "ENONFINITE", -6: "ESTENCIL",
}

FP64, MIXED = 0, 1
REORDER_NONE, REORDER_RCM, REORDER_HILBERT = 0, 1, 2
TILES_AUTO = -1
TILE_QDERIV, TILE_QINNER, TILE_DELTA, TILE_FLUX = 1, 2, 4, 8
A_PRIM, A_PRIM_OLD, A_Q, A_DQ, A_QM, A_DELTA, A_FLUX_RES, A_NORMALS = range(8)
K_FLUX, K_QDERIV, K_QINNER, K_STATE, K_DELTA, K_QVAR = range(6)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class LskumError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"lskum error {code} ({ERRORS.get(code, '?')}): {msg}")


class CloudView(C.Structure):
    _fields_ = [
        ("n_points", C.c_int32),
        ("x", _dp), ("y", _dp),
        ("left", _ip), ("right", _ip),
        ("flag_1", _ip), ("flag_2", _ip),
        ("min_dist", _dp),
        ("conn_off", _ip), ("conn", _ip),
    ]


class GenParams(C.Structure):
    _fields_ = [
        ("naca_m", C.c_int32), ("naca_p", C.c_int32), ("naca_t", C.c_int32),
        ("n_wall", C.c_int32), ("n_radial", C.c_int32), ("n_circ", C.c_int32),
        ("r_far", C.c_double), ("stretch", C.c_double), ("jitter", C.c_double),
        ("k", C.c_int32), ("radius_factor", C.c_double), ("seed", C.c_uint64),
        ("wall_blend", C.c_double),
    ]


class Params(C.Structure):
    _fields_ = [
        ("mach", C.c_double), ("aoa_deg", C.c_double), ("cfl", C.c_double),
        ("vl_const", C.c_double), ("power", C.c_double),
        ("inner_iterations", C.c_int32), ("precision", C.c_int32),
        ("reorder", C.c_int32), ("lds_tiles", C.c_int32), ("tile_block", C.c_int32),
        ("points_per_thread", C.c_int32),
    ]


class Case(C.Structure):
    _fields_ = [("cloud", C.c_int32), ("mach", C.c_double), ("aoa_deg", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("tile_block", C.c_int32), ("lds_tiles", C.c_int32),
                ("points_per_thread", C.c_int32), ("ms_per_iter", C.c_double)]


# (name, restype, argtypes)
_SIGS = [
    ("lskum_cloud_generate", C.c_int, [C.POINTER(GenParams), C.POINTER(C.c_void_p)]),
    ("lskum_cloud_read_partgrid", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("lskum_cloud_write_partgrid", C.c_int, [C.c_void_p, C.c_char_p]),
    ("lskum_cloud_from_arrays", C.c_int, [C.POINTER(CloudView), C.POINTER(C.c_void_p)]),
    ("lskum_cloud_view", C.c_int, [C.c_void_p, C.POINTER(CloudView)]),
    ("lskum_cloud_stats", C.c_int, [C.c_void_p, _ip, _ip, _dp]),
    ("lskum_cloud_connectivity", C.c_int, [C.c_void_p, _dp, _dp, C.POINTER(C.c_uint8)]),
    ("lskum_cloud_rcm", C.c_int, [C.c_void_p, _ip]),
    ("lskum_cloud_free", None, [C.c_void_p]),
    ("lskum_create", C.c_int, [C.POINTER(CloudView), C.POINTER(Params), C.POINTER(C.c_void_p)]),
    ("lskum_create_batch", C.c_int, [C.POINTER(CloudView), C.c_int32, C.POINTER(Case),
                                     C.c_int32, C.POINTER(Params), C.POINTER(C.c_void_p)]),
    ("lskum_case_count", C.c_int, [C.c_void_p, _ip, _ip]),
    ("lskum_case_status", C.c_int, [C.c_void_p, _ip]),
    ("lskum_set_graph", C.c_int, [C.c_void_p, C.c_int32]),
    ("lskum_autotune", C.c_int, [C.POINTER(CloudView), C.c_int32, C.POINTER(Case), C.c_int32,
                                 C.POINTER(Params), C.c_int32, C.POINTER(TuneResult),
                                 C.POINTER(TuneResult), C.c_int32, _ip]),
    ("lskum_destroy", None, [C.c_void_p]),
    ("lskum_last_error", C.c_char_p, [C.c_void_p]),
    ("lskum_set_state", C.c_int, [C.c_void_p, _dp]),
    ("lskum_get_state", C.c_int, [C.c_void_p, _dp]),
    ("lskum_snapshot", C.c_int, [C.c_void_p]),
    ("lskum_func_delta", C.c_int, [C.c_void_p]),
    ("lskum_q_variables", C.c_int, [C.c_void_p]),
    ("lskum_q_derivatives", C.c_int, [C.c_void_p]),
    ("lskum_q_inner", C.c_int, [C.c_void_p, C.c_int32]),
    ("lskum_flux_residual", C.c_int, [C.c_void_p]),
    ("lskum_state_update", C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _ip]),
    ("lskum_forces", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("lskum_iterate", C.c_int, [C.c_void_p, C.c_int32, _dp, _dp]),
    ("lskum_iterate_noalloc", C.c_int, [C.c_void_p, C.c_int32, _dp, _dp]),
    ("lskum_get_iteration", C.c_int, [C.c_void_p, _ip, _dp]),
    ("lskum_set_iteration", C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_int32]),
    ("lskum_restart_write", C.c_int, [C.c_char_p, C.POINTER(CloudView), C.c_int32, C.c_double,
                                      _dp]),
    ("lskum_restart_read", C.c_int, [C.c_char_p, C.c_int32, _ip, _dp, _dp]),
    ("lskum_solution_write", C.c_int, [C.c_char_p, C.POINTER(CloudView), C.c_int32,
                                       C.c_double, _dp, _dp]),
    ("lskum_cp_write", C.c_int, [C.c_char_p, C.POINTER(CloudView), _dp, C.c_double]),
    ("lskum_get_array", C.c_int, [C.c_void_p, C.c_int32, _dp]),
    ("lskum_get_substencil", C.c_int, [C.c_void_p, C.c_int32, _ip, _ip]),
    ("lskum_stream", C.c_void_p, [C.c_void_p]),
    ("lskum_layout", C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    ("lskum_flux_tile", C.c_int, [C.c_void_p, _ip]),
    ("lskum_set_timing", C.c_int, [C.c_void_p, C.c_int32]),
    ("lskum_kernel_time", C.c_int, [C.c_void_p, C.c_int32, _dp, _ip]),
    ("lskum_flux_vjp", C.c_int, [C.c_void_p, _dp, _dp]),
    ("lskum_flux_jvp", C.c_int, [C.c_void_p, _dp, _dp]),
    ("lskum_forces_grad", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("lskum_compute_entropy", C.c_int, [C.c_void_p, _dp, _dp]),
    ("lskum_compute_enstrophy", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("lskum_stagnation_pressure", C.c_int, [C.c_void_p, _dp, _ip]),
    ("lskum_objective_J", C.c_int, [C.c_void_p, _dp]),
    ("lskum_set_coordinates", C.c_int, [C.c_void_p, _dp, _dp]),
    ("lskum_shape_vjp", C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp, _dp]),
    ("lskum_shape_jvp", C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("lskum_hicks_henne_fields", C.c_int, [C.POINTER(CloudView), C.c_int32, C.c_double,
                                           C.c_double, C.c_int32, C.c_double, _dp]),
    ("lskum_kat_fluxes", C.c_int, [C.c_int32, _dp, _dp, C.c_int32]),
    ("lskum_kat_qtilde", C.c_int, [C.c_int32, _dp, _dp]),
]

_lib = None


def lib():
    """Loads liblskum.so once; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() "
                "(or python -m lskum.build in c-_serial_amd/)")
        handle = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols():
    return [s[0] for s in _SIGS]


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def check(code, ctx=None):
    if code != LSKUM_OK:
        msg = ""
        if ctx:
            raw = lib().lskum_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise LskumError(code, msg)
