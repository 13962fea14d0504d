"""Python face of the C ABI, with the reference's stage names.

``Solver`` owns one lskum_ctx.  Its methods are the reference's hot-path
functions (fpi_solver_mod.h:32-61 call order), each one C-ABI call that runs
HIP kernels on the context's stream:

    func_delta            state_update_mod.h:448
    eval_q_variables      q_variables_mod.h:9
    eval_q_derivatives    q_variables_mod.h:36
    eval_q_inner_loop     q_variables_mod.h:120 (+ eval_update_innerloop :215)
    cal_flux_residual     flux_residual_mod.h:10
    state_update(rk)      state_update_mod.h:92
    compute_cl_cd_cm      compute_force_coeffs_mod.h:7
    fpi_solver(it)        fpi_solver_mod.h:14
    q_lskum(n)            q_lskum_mod.h:13

Arrays come back in the caller's point order, 0-based, component-major.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .cloud import Cloud

_SHAPES = {L.A_PRIM: (4,), L.A_PRIM_OLD: (4,), L.A_Q: (4,), L.A_DQ: (2, 4),
           L.A_QM: (2, 4), L.A_DELTA: (), L.A_FLUX_RES: (4,), L.A_NORMALS: (2,)}


class Solver:
    def __init__(self, cloud: Cloud, mach=0.63, aoa=2.0, cfl=0.1, vl_const=150.0, power=0.0,
                 inner_iterations=3, precision=L.FP64, reorder=L.REORDER_HILBERT,
                 lds_tiles=L.TILES_AUTO, tile_block=0, points_per_thread=0):
        self.cloud = cloud
        self.n = cloud.n
        self.params = L.Params(mach, aoa, cfl, vl_const, power, inner_iterations, precision,
                               reorder, lds_tiles, tile_block, points_per_thread)
        self.inner_iterations = inner_iterations
        view = cloud.view()
        ctx = C.c_void_p()
        rc = L.lib().lskum_create(C.byref(view), C.byref(self.params), C.byref(ctx))
        self._ctx = ctx
        if rc != L.LSKUM_OK:
            msg = L.lib().lskum_last_error(ctx) if ctx else b""
            if ctx:
                L.lib().lskum_destroy(ctx)
                self._ctx = None
            raise L.LskumError(rc, (msg or b"").decode())
        self.sum_res_sqr = 0.0
        self.max_res = 0.0
        self.max_res_point = 0
        self.res_new = 0.0
        self.res_old = 0.0
        self.residue = 0.0
        self.Cl = self.Cd = self.Cm = 0.0

    def close(self):
        if getattr(self, "_ctx", None):
            L.lib().lskum_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        L.check(rc, self._ctx)

    # ---------------------------------------------------------- state I/O
    def set_state(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(4, self.n)
        self._chk(L.lib().lskum_set_state(self._ctx, L.dptr(prim)))

    def get(self, which):
        shape = _SHAPES[which] + (self.n,)
        out = np.empty(shape, dtype=np.float64)
        self._chk(L.lib().lskum_get_array(self._ctx, which, L.dptr(out)))
        return out

    @property
    def prim(self):
        return self.get(L.A_PRIM)

    @prim.setter
    def prim(self, value):
        self.set_state(value)

    @property
    def q(self):
        return self.get(L.A_Q)

    @property
    def dq(self):
        return self.get(L.A_DQ)

    @property
    def qm(self):
        return self.get(L.A_QM)

    @property
    def delta(self):
        return self.get(L.A_DELTA)

    @property
    def flux_res(self):
        return self.get(L.A_FLUX_RES)

    @property
    def normals(self):
        return self.get(L.A_NORMALS)

    def substencil(self, which):
        """(counts, members) of sub-stencil 0 xpos, 1 xneg, 2 ypos, 3 yneg."""
        counts = np.zeros(self.n, np.int32)
        self._chk(L.lib().lskum_get_substencil(self._ctx, which, L.iptr(counts), None))
        members = np.zeros(int(counts.sum()), np.int32)
        if members.size:
            self._chk(L.lib().lskum_get_substencil(self._ctx, which, L.iptr(counts),
                                                   L.iptr(members)))
        return counts, members

    # ------------------------------------------------- reference stages
    def snapshot(self):
        self._chk(L.lib().lskum_snapshot(self._ctx))

    def func_delta(self):
        self._chk(L.lib().lskum_func_delta(self._ctx))

    def eval_q_variables(self):
        self._chk(L.lib().lskum_q_variables(self._ctx))

    def eval_q_derivatives(self):
        self._chk(L.lib().lskum_q_derivatives(self._ctx))

    def eval_q_inner_loop(self, passes=1):
        """One (or more) Jacobi passes; the reference's temp-then-copy
        (eval_update_innerloop) is the kernels' buffer swap."""
        self._chk(L.lib().lskum_q_inner(self._ctx, passes))

    def eval_update_innerloop(self):
        """No-op: eval_q_inner_loop already published the new derivatives."""

    def cal_flux_residual(self):
        self._chk(L.lib().lskum_flux_residual(self._ctx))

    def state_update(self, rk):
        s, m, p = C.c_double(), C.c_double(), C.c_int32()
        self._chk(L.lib().lskum_state_update(self._ctx, rk, C.byref(s), C.byref(m), C.byref(p)))
        self.sum_res_sqr, self.max_res, self.max_res_point = s.value, m.value, p.value
        return self.sum_res_sqr, self.max_res, self.max_res_point

    def compute_cl_cd_cm(self):
        cl, cd, cm = C.c_double(), C.c_double(), C.c_double()
        self._chk(L.lib().lskum_forces(self._ctx, C.byref(cl), C.byref(cd), C.byref(cm)))
        self.Cl, self.Cd, self.Cm = cl.value, cd.value, cm.value
        return self.Cl, self.Cd, self.Cm

    def q_lskum(self, n_iter):
        """n_iter fpi_solver iterations in one call; returns (res_new, residue)
        per iteration."""
        rn = np.zeros(max(n_iter, 1))
        rs = np.zeros(max(n_iter, 1))
        self._chk(L.lib().lskum_iterate(self._ctx, n_iter, L.dptr(rn), L.dptr(rs)))
        if n_iter:
            self.res_new, self.residue = rn[n_iter - 1], rs[n_iter - 1]
            self.res_old = self.iteration()[1]
        return rn[:n_iter], rs[:n_iter]

    # -------------------------------------------------- adjoint / tangent
    def flux_vjp(self, lam):
        """J^T lam for J = d flux_res / d prim at the current state (delta
        held fixed); lam [4, n] -> [4, n]."""
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(4, self.n)
        out = np.empty((4, self.n))
        self._chk(L.lib().lskum_flux_vjp(self._ctx, L.dptr(lam), L.dptr(out)))
        return out

    def flux_jvp(self, v):
        """J v for J = d flux_res / d prim at the current state."""
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(4, self.n)
        out = np.empty((4, self.n))
        self._chk(L.lib().lskum_flux_jvp(self._ctx, L.dptr(v), L.dptr(out)))
        return out

    def forces_grad(self):
        """(dCl, dCd, dCm) / d prim, each [4, n]."""
        g = [np.empty((4, self.n)) for _ in range(3)]
        self._chk(L.lib().lskum_forces_grad(self._ctx, *[L.dptr(a) for a in g]))
        return tuple(g)

    # -------------------------------------------------- diagnostics
    def compute_entropy(self):
        """(entropy [n], total) -- compute_entropy_mod.h:5-24"""
        e, t = np.empty(self.n), C.c_double()
        self._chk(L.lib().lskum_compute_entropy(self._ctx, L.dptr(e), C.byref(t)))
        return e, t.value

    def compute_enstrophy(self):
        """(vorticity [n], vorticity_sqr [n], total) -- compute_enstrophy_mod.h:5-76"""
        v, v2, t = np.empty(self.n), np.empty(self.n), C.c_double()
        self._chk(L.lib().lskum_compute_enstrophy(self._ctx, L.dptr(v), L.dptr(v2),
                                                  C.byref(t)))
        return v, v2, t.value

    def stagnation_pressure(self):
        """((pMin, pMax, pMin/p0_inf, pMax/p0_inf), (indexMin, indexMax)) --
        stagnation_values_mod.h:8-46, indices 0-based"""
        vals, idx = np.empty(4), np.empty(2, np.int32)
        self._chk(L.lib().lskum_stagnation_pressure(self._ctx, L.dptr(vals), L.iptr(idx)))
        return vals, idx

    def objective_J(self):
        """stagnation-pressure loss -- stagnation_values_mod.h:49-80"""
        J = C.c_double()
        self._chk(L.lib().lskum_objective_J(self._ctx, C.byref(J)))
        return J.value

    # -------------------------------------------------- shape sensitivities
    def set_coordinates(self, x, y):
        """Moves the points to (x, y), keeping the neighbour lists, their
        sub-stencil split and min_dist; normals and least-squares
        coefficients follow."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.n)
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(self.n)
        self._chk(L.lib().lskum_set_coordinates(self._ctx, L.dptr(x), L.dptr(y)))

    @staticmethod
    def _weights(w):
        """(Cl, Cd, Cm[, stagnation-pressure loss]) weights as 4 doubles"""
        out = np.zeros(4)
        w = np.asarray(w, dtype=np.float64).ravel()
        out[:len(w)] = w
        return out

    def shape_vjp(self, n_iter, w=(1.0, 0.0, 0.0)):
        """J = w . (Cl, Cd, Cm, J_p0) after n_iter iterations from the current
        state (J_p0 the stagnation-pressure loss of objective_J; w may omit
        trailing zeros), and its gradients with respect to the coordinates and
        the starting state by the discrete adjoint: (J, x_bar, y_bar, prim_bar)."""
        w = self._weights(w)
        J = C.c_double()
        xb, yb, pb = np.empty(self.n), np.empty(self.n), np.empty((4, self.n))
        self._chk(L.lib().lskum_shape_vjp(self._ctx, n_iter, L.dptr(w), C.byref(J),
                                          L.dptr(xb), L.dptr(yb), L.dptr(pb)))
        return J.value, xb, yb, pb

    def shape_jvp(self, n_iter, w=(1.0, 0.0, 0.0), dx=None, dy=None, dprim=None):
        """(J, dJ) with dJ the derivative of J along the coordinate direction
        (dx, dy) and the starting-state direction dprim (None: zero), by a
        forward-mode run of the same iterations."""
        w = self._weights(w)

        def arr(a, shape):
            return None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(shape)

        dx, dy, dprim = arr(dx, self.n), arr(dy, self.n), arr(dprim, (4, self.n))
        ptr = lambda a: None if a is None else L.dptr(a)
        J, dJ = C.c_double(), C.c_double()
        self._chk(L.lib().lskum_shape_jvp(self._ctx, n_iter, L.dptr(w), ptr(dx), ptr(dy),
                                          ptr(dprim), C.byref(J), C.byref(dJ)))
        return J.value, dJ.value

    def objective_after(self, n_iter, w=(1.0, 0.0, 0.0)):
        """J = w . (Cl, Cd, Cm, J_p0) after n_iter plain iterations from the
        current state, which is restored afterwards (the iteration counter is
        not)."""
        w = self._weights(w)
        p0 = self.prim
        self.q_lskum(n_iter)
        cl, cd, cm = self.compute_cl_cd_cm()
        jp0 = self.objective_J() if w[3] != 0.0 else 0.0
        self.set_state(p0)
        return float(np.dot(w, (cl, cd, cm, jp0)))

    def forces_after(self, n_iter):
        """(Cl, Cd, Cm) after n_iter plain iterations from the current state,
        which is restored afterwards (the iteration counter is not)."""
        p0 = self.prim
        self.q_lskum(n_iter)
        out = self.compute_cl_cd_cm()
        self.set_state(p0)
        return np.array(out)

    def residual_of(self, prim):
        """flux_res at the state prim with delta held at its last value: the
        map lskum_flux_vjp / lskum_flux_jvp differentiate."""
        self.set_state(prim)
        self.eval_q_variables()
        self.eval_q_derivatives()
        self.eval_q_inner_loop(self.inner_iterations)
        self.cal_flux_residual()
        return self.flux_res

    def fpi_solver(self, it=None):
        """One iteration (fpi_solver_mod.h:14-150); `it` must be the next
        iteration number when given."""
        if it is not None and it != self.iteration()[0] + 1:
            raise ValueError("fpi_solver(it) must be called with the next iteration number")
        rn, rs = self.q_lskum(1)
        return rn[0], rs[0]

    def iteration(self):
        it, ro = C.c_int32(), C.c_double()
        self._chk(L.lib().lskum_get_iteration(self._ctx, C.byref(it), C.byref(ro)))
        return it.value, ro.value

    def set_iteration(self, it, res_old, restart=True):
        self._chk(L.lib().lskum_set_iteration(self._ctx, it, res_old, 1 if restart else 0))

    def set_graph(self, iters_per_graph):
        """Run iterations as captured HIP graphs of this many iterations (0: off)."""
        self._chk(L.lib().lskum_set_graph(self._ctx, iters_per_graph))

    def write_restart(self, path):
        """restart.dat of the current state and iteration counter (the
        layout initial_conditions_mod.h:7-21 reads)."""
        it, res_old = self.iteration()
        restart_write(path, self.cloud, it, res_old, self.prim)

    def write_solution(self, path):
        """output.dat of the current state with its entropy
        (post_processing_mod.h:7-27, one point per line)."""
        it, res_old = self.iteration()
        view = self.cloud.view()
        entropy, _ = self.compute_entropy()
        L.check(L.lib().lskum_solution_write(str(path).encode(), C.byref(view), it, res_old,
                                             L.dptr(self.prim), L.dptr(entropy)))

    def write_cp(self, path):
        """Cp along the shape points (compute_force_coeffs_mod.h:20-21,54)."""
        view = self.cloud.view()
        L.check(L.lib().lskum_cp_write(str(path).encode(), C.byref(view), L.dptr(self.prim),
                                       self.params.mach))

    def read_restart(self, path):
        """Resume from a restart file: its state, iteration counter and
        res_old, with the res_old latch kept (the reference's restart = 1)."""
        it, res_old, prim = restart_read(path, self.n)
        self.set_state(prim)
        self.set_iteration(it, res_old, restart=True)

    def layout(self):
        """(mask of kernels using LDS tiles, tile_max, order) from lskum_create."""
        t, m, o = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(L.lib().lskum_layout(self._ctx, C.byref(t), C.byref(m), C.byref(o)))
        return t.value, m.value, o.value

    def flux_tile(self):
        """Tile entries the flux stages in LDS per block (lskum_flux_tile)."""
        cap = C.c_int32()
        self._chk(L.lib().lskum_flux_tile(self._ctx, C.byref(cap)))
        return cap.value

    # ------------------------------------------------------------ timing
    def set_timing(self, enable):
        self._chk(L.lib().lskum_set_timing(self._ctx, 1 if enable else 0))

    def kernel_time(self, kernel):
        ms, cnt = C.c_double(), C.c_int32()
        self._chk(L.lib().lskum_kernel_time(self._ctx, kernel, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


def kat_fluxes(rows6, mixed=False):
    """Device flux functions on rows (nx, ny, u1, u2, rho, pr) -> [n, 8, 4];
    mixed: the FP32 functions of the mixed form, on FP32-rounded inputs."""
    rows6 = np.ascontiguousarray(rows6, dtype=np.float64)
    n = rows6.shape[0]
    out = np.zeros((n, 8, 4))
    L.check(L.lib().lskum_kat_fluxes(n, L.dptr(rows6), L.dptr(out), 1 if mixed else 0))
    return out


def kat_qtilde(qt):
    """Device qtilde_to_primitive, in the form the flux kernels evaluate it,
    on rows q~[4] -> [n, 4] (u1, u2, rho, pr)."""
    qt = np.ascontiguousarray(qt, dtype=np.float64)
    n = qt.shape[0]
    out = np.zeros((n, 4))
    L.check(L.lib().lskum_kat_qtilde(n, L.dptr(qt), L.dptr(out)))
    return out


def restart_write(path, cloud, it, res_old, prim):
    """Host-only writer of a restart file in the layout restart_sol reads
    (initial_conditions_mod.h:7-21): prim [4][n] in the cloud's order."""
    view = cloud.view()
    prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(4, cloud.n)
    L.check(L.lib().lskum_restart_write(str(path).encode(), C.byref(view), int(it),
                                        float(res_old), L.dptr(prim)))


def restart_read(path, n):
    """(it, res_old, prim [4][n]) of a restart file."""
    it, ro = C.c_int32(), C.c_double()
    prim = np.zeros((4, n))
    L.check(L.lib().lskum_restart_read(str(path).encode(), n, C.byref(it), C.byref(ro),
                                       L.dptr(prim)))
    return it.value, ro.value, prim
