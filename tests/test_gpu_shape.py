"""Shape sensitivities through whole iterations: lskum_set_coordinates,
lskum_shape_vjp (discrete adjoint), lskum_shape_jvp (forward-mode run in
dual numbers).

The map is (coordinates, starting state P0) -> J = w . (Cl, Cd, Cm) after
n_iter iterations, with the neighbour lists, their sub-stencil split and
min_dist held as built.  Checks:
  - moving the points and back reproduces the iterations bit for bit;
  - the dual-number run's values reproduce the FP64 iterations' J;
  - duality: <(x_bar, y_bar, prim_bar), (dx, dy, dprim)> of the reverse
    sweep equals the tangent run's dJ in random directions, to rounding --
    the two share no derivative code;
  - central differences (h = 1e-6) of the plain iterations along
    Hicks-Henne design fields, with the reference's limiter constant.
"""
import numpy as np
import pytest

import lskum
from lskum import shape
from lskum._lib import LskumError
from lskum.solver import Solver

pytestmark = pytest.mark.gpu


def _developed(name, inner, iters, mach=0.8, aoa=1.25):
    cloud = lskum.workload(name)
    s = Solver(cloud, mach=mach, aoa=aoa, cfl=0.2, vl_const=5.0, inner_iterations=inner)
    s.q_lskum(iters)
    return cloud, s


def test_set_coordinates_round_trip():
    cloud, s = _developed("naca0012_1480_k20", 0, 20)
    print("case: set_coordinates round trip", flush=True)
    p0 = s.prim
    s.q_lskum(5)
    ref = s.prim, s.compute_cl_cd_cm()
    s.set_state(p0)
    # a real move, then back: every derived quantity returns exactly
    fields = shape.hicks_henne_fields(cloud)
    s.set_coordinates(cloud.x, cloud.y + 1e-3 * fields[3])
    moved_normals = s.normals
    s.set_coordinates(cloud.x, cloud.y)
    s.q_lskum(5)
    assert np.array_equal(s.prim, ref[0])
    assert s.compute_cl_cd_cm() == ref[1]
    assert not np.array_equal(moved_normals, s.normals)


def test_set_coordinates_refuses_a_degenerate_move():
    cloud, s = _developed("naca0012_1480_k20", 0, 10)
    p0 = s.prim
    s.q_lskum(3)
    ref = s.prim
    s.set_state(p0)
    y = cloud.y.copy()
    # every neighbour of point 200 onto its horizontal line: collinear
    nb = cloud.conn[cloud.conn_off[200]:cloud.conn_off[201]]
    y[nb] = y[200]
    with pytest.raises(LskumError) as e:
        s.set_coordinates(cloud.x, y)
    assert e.value.code == -6  # LSKUM_ESTENCIL
    # the context kept its coordinates
    s.q_lskum(3)
    assert np.array_equal(s.prim, ref)


@pytest.mark.parametrize("inner", [0, 2])
def test_tangent_run_reproduces_the_iterations(inner):
    cloud, s = _developed("naca0012_1480_k20", inner, 30)
    J_plain = s.objective_after(6, (1.0, 0.5, 0.25))
    J_tan, _ = s.shape_jvp(6, (1.0, 0.5, 0.25))
    J_adj = s.shape_vjp(6, (1.0, 0.5, 0.25))[0]
    print(f"case: tangent values inner={inner}: {J_plain!r} {J_tan!r} {J_adj!r}", flush=True)
    assert J_adj == J_plain  # the adjoint's forward is the plain iteration
    assert abs(J_tan - J_plain) <= 1e-12 * abs(J_plain)


@pytest.mark.parametrize("name,inner,w", [("naca0012_1480_k20", 0, (1.0, 0.0, 0.0)),
                                          ("naca0012_1480_k20", 2, (0.0, 1.0, 0.0)),
                                          ("naca0012_3620_k24", 2, (0.3, -0.7, 1.1))])
def test_adjoint_tangent_duality(name, inner, w):
    cloud, s = _developed(name, inner, 40)
    rng = np.random.default_rng(3)
    n_iter = 4
    J, xb, yb, pb = s.shape_vjp(n_iter, w)
    p0 = s.prim
    for trial in range(3):
        dx = rng.standard_normal(cloud.n) * 1e-3
        dy = rng.standard_normal(cloud.n) * 1e-3
        dp = rng.standard_normal((4, cloud.n)) * np.abs(p0) * 1e-3
        if trial == 0:  # geometry only, as the design fields use it
            dp = np.zeros_like(dp)
        _, dJ = s.shape_jvp(n_iter, w, dx=dx, dy=dy, dprim=dp)
        rev = float(xb @ dx + yb @ dy + np.sum(pb * dp))
        scale = np.sqrt(np.sum(xb * xb * dx * dx) + np.sum(yb * yb * dy * dy) +
                        np.sum(pb * pb * dp * dp)) * np.sqrt(cloud.n)
        print(f"case: duality {name} inner={inner} trial {trial}: rev {rev:.15e} "
              f"tan {dJ:.15e}", flush=True)
        assert abs(rev - dJ) <= 1e-10 * max(abs(dJ), scale)


@pytest.mark.parametrize("inner", [0, 2])
def test_design_gradient_matches_central_differences(inner):
    # K = 150, the reference's VL_CONST: epsilon^2 = (K min_dist)^6 is large
    # next to the squared limiter increments, so phi stays within about
    # (del/epsilon)^2 of 1 and its jump at |del_neg| = 1e-5 is negligible --
    # the iterations are smooth enough in the coordinates for h = 1e-6
    # differences.  (With K = 5 that jump is of order one where the stencil
    # has an extremum, and differences across it measure the jump.)
    cloud = lskum.workload("naca0012_3620_k24")
    s = Solver(cloud, mach=0.8, aoa=1.25, cfl=0.2, vl_const=150.0, inner_iterations=inner)
    s.q_lskum(60)
    fields = shape.hicks_henne_fields(cloud, n_per_side=10)
    n_iter = 10
    for w in [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0)]:
        J, g = shape.design_gradient(s, fields, n_iter, w)
        for j in (2, 7, 12, 17):
            fd = shape.design_fd(s, fields[j], n_iter, w, h=1e-6)
            tan = shape.design_tangent(s, fields[j], n_iter, w)
            print(f"case: design FD inner={inner} w={w} var {j}: adjoint {g[j]:.12e} "
                  f"tangent {tan:.12e} central {fd:.12e} rel {abs(fd - g[j]) / abs(g[j]):.2e}",
                  flush=True)
            assert abs(tan - g[j]) <= 1e-10 * abs(g[j])
            assert abs(fd - g[j]) <= 1e-5 * abs(g[j])


def test_stagnation_loss_objective():
    # J_p0, the stagnation-pressure loss, as the adjoint's objective (alone
    # and mixed with Cl): duality with the tangent run and central
    # differences in the starting state and the coordinates
    cloud = lskum.workload("naca0012_1480_k20")
    s = Solver(cloud, mach=0.8, aoa=1.25, cfl=0.2, vl_const=150.0, inner_iterations=0)
    s.q_lskum(40)
    rng = np.random.default_rng(9)
    p0 = s.prim
    for w in [(0.0, 0.0, 0.0, 1.0), (0.5, 0.0, 0.0, 30.0)]:
        J, xb, yb, pb = s.shape_vjp(5, w)
        assert J == pytest.approx(s.objective_after(5, w), rel=1e-14)
        dy = rng.standard_normal(cloud.n) * 1e-3 * (np.hypot(cloud.x - 0.5, cloud.y) < 1)
        dp = rng.standard_normal((4, cloud.n)) * np.abs(p0) * 1e-3
        _, dJ = s.shape_jvp(5, w, dy=dy, dprim=dp)
        rev = float(yb @ dy + np.sum(pb * dp))
        print(f"case: stagnation-loss objective w={w}: rev {rev:.15e} tan {dJ:.15e}", flush=True)
        assert abs(rev - dJ) <= 1e-10 * abs(dJ)
        h = 1e-6
        s.set_state(p0 + h * dp)
        jp = s.objective_after(5, w)
        s.set_state(p0 - h * dp)
        jm = s.objective_after(5, w)
        s.set_state(p0)
        fd_p = (jp - jm) / (2 * h)
        assert fd_p == pytest.approx(float(np.sum(pb * dp)), rel=1e-5)
