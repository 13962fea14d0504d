"""Shape sensitivities of the force coefficients in Hicks-Henne design
variables (no reference counterpart): the design fields, the gradient of
J = w . (Cl, Cd, Cm) after a number of iterations by the discrete adjoint
(Solver.shape_vjp), and the two checks it is held to: the tangent-linear run
(Solver.shape_jvp) and central finite differences of the plain iterations."""
import ctypes as C

import numpy as np

from . import _lib as L


def hicks_henne_fields(cloud, n_per_side=10, p_lo=0.05, p_hi=0.95, width_exponent=3,
                       decay_radius=0.5):
    """[2 n_per_side, n]: the y displacement of every point per unit bump
    amplitude, upper-surface bumps first (lskum_hicks_henne_fields)."""
    view = cloud.view()
    dy = np.empty((2 * n_per_side, cloud.n))
    L.check(L.lib().lskum_hicks_henne_fields(C.byref(view), n_per_side, p_lo, p_hi,
                                             width_exponent, decay_radius, L.dptr(dy)))
    return dy


def design_gradient(solver, fields, n_iter, w=(1.0, 0.0, 0.0)):
    """(J, dJ/d alpha) by one adjoint sweep: dJ/d alpha_j = y_bar . fields[j]
    (the fields move y only)."""
    J, _, yb, _ = solver.shape_vjp(n_iter, w)
    return J, fields @ yb


def design_tangent(solver, field, n_iter, w=(1.0, 0.0, 0.0)):
    """dJ along one design field by the tangent-linear run."""
    return solver.shape_jvp(n_iter, w, dy=field)[1]


def design_fd(solver, field, n_iter, w=(1.0, 0.0, 0.0), h=1e-6, one_sided=False):
    """Central difference (J(+h) - J(-h)) / 2h along one design field, the
    points moved by +-h field and put back; with one_sided also the forward
    and backward differences from J(0)."""
    x0 = np.ascontiguousarray(solver.cloud.x, dtype=np.float64)
    y0 = np.ascontiguousarray(solver.cloud.y, dtype=np.float64)
    J = {}
    for sgn in ((1, -1, 0) if one_sided else (1, -1)):
        solver.set_coordinates(x0, y0 + sgn * h * field)
        J[sgn] = solver.objective_after(n_iter, w)
    solver.set_coordinates(x0, y0)
    central = (J[1] - J[-1]) / (2 * h)
    if not one_sided:
        return central
    return central, (J[1] - J[0]) / h, (J[0] - J[-1]) / h


def design_fd_forces(solver, field, n_iter, h=1e-6):
    """Central, forward and backward differences of (Cl, Cd, Cm) along one
    design field from the same three runs (points moved by +h, -h and 0,
    then put back): three arrays of 3."""
    x0 = np.ascontiguousarray(solver.cloud.x, dtype=np.float64)
    y0 = np.ascontiguousarray(solver.cloud.y, dtype=np.float64)
    F = {}
    for sgn in (1, -1, 0):
        solver.set_coordinates(x0, y0 + sgn * h * field)
        F[sgn] = solver.forces_after(n_iter)
    solver.set_coordinates(x0, y0)
    return (F[1] - F[-1]) / (2 * h), (F[1] - F[0]) / h, (F[0] - F[-1]) / h
