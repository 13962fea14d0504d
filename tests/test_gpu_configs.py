"""The benchmark configurations at their own workloads, on the GPU.

  - 10M points (naca0012_10m_k24: wall 4,000, 1,000 x 10,000, k = 24), M 0.63,
    AoA 2 deg, no inner passes: the flux residual of the free stream is
    exactly zero in both forms, and the mixed form (FP32 gathers, fluxes,
    limiter and least-squares sums; FP64 coordinates, least-squares
    geometry, state and reductions) stays within 1e-6 of the FP64 form in Cl
    and Cd over 100 iterations from the same start;
  - the many-case workload (64 flow conditions over 8 NACA 4-digit clouds of
    50k-800k points, k = 20): the layout lskum_autotune picks, 50 iterations
    of all cases as captured graphs, and three cases rerun alone, which must
    match their batch results bit for bit;
  - the shape-gradient cloud (naca0012_1m_k24: 1,000,940 points, k = 24) at
    M 0.8, AoA 1.25 deg, K 150, after 2,000 iterations: the reverse sweep
    (lskum_shape_vjp) against the dual-number tangent run (lskum_shape_jvp)
    along a Hicks-Henne design field and a random state direction, and
    against central differences of the plain iterations along a design
    field on each surface.
"""
import numpy as np
import pytest

import lskum
from lskum import _lib as L
from lskum import shape
from lskum.batch import Batch, autotune, batch_workload
from lskum.solver import Solver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cloud_10m():
    return lskum.workload("naca0012_10m_k24")


@pytest.mark.timeout(600)
def test_10m_points_free_stream_is_exact_in_both_forms(cloud_10m):
    for precision in (L.FP64, L.MIXED):
        print("case: free stream naca0012_10m_k24, precision", precision, flush=True)
        s = Solver(cloud_10m, mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, inner_iterations=0,
                   precision=precision)
        p0 = s.prim
        s.func_delta()
        s.eval_q_variables()
        s.eval_q_derivatives()
        s.cal_flux_residual()
        assert not np.any(s.flux_res)
        np.testing.assert_array_equal(s.prim, p0)
        s.close()


@pytest.mark.timeout(600)
def test_10m_points_mixed_stays_within_1e6_of_fp64(cloud_10m):
    runs = {}
    for name, precision in (("fp64", L.FP64), ("mixed", L.MIXED)):
        s = Solver(cloud_10m, mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, inner_iterations=0,
                   precision=precision)
        s.set_graph(50)
        rn, _ = s.q_lskum(100)
        runs[name] = (rn, np.array(s.compute_cl_cd_cm()))
        s.close()
    (rf, ff), (rm, fm) = runs["fp64"], runs["mixed"]
    d = np.abs(fm - ff)
    e_res = float(np.abs(rm - rf).max() / np.abs(rf).max())
    print(f"case: 10M points, 100 iterations: |dCl| {d[0]:.2e} |dCd| {d[1]:.2e} "
          f"|dCm| {d[2]:.2e} residual history {e_res:.2e}", flush=True)
    assert np.isfinite(rf).all() and np.isfinite(rm).all()
    assert d[0] <= 1e-6 and d[1] <= 1e-6 and d[2] <= 1e-6
    assert e_res <= 1e-5


@pytest.mark.timeout(900)
def test_many_case_workload_autotuned_graphs_equal_single_runs():
    clouds, cases = batch_workload()
    assert len(clouds) == 8 and len(cases) == 64
    sizes = [c.n for c in clouds]
    print("case: many-case workload, clouds", sizes, flush=True)
    # the generator's ring counts approximate the recipe's sizes (50k-800k)
    assert min(sizes) >= 49_000 and max(sizes) <= 820_000
    num = dict(cfl=0.2, vl_const=5.0, power=0.0, inner_iterations=0)
    reps = [next(j for j, c in enumerate(cases) if c[0] == ci) for ci in range(len(clouds))]
    best, trials = autotune(clouds, [cases[j] for j in reps], iters=5, **num)
    print("case: autotune best", best, "of", len(trials), "trials", flush=True)
    # six block-shape/tile candidates, then 2 and 4 points per thread
    assert len(trials) == 8
    assert {t["points_per_thread"] for t in trials} == {1, 2, 4}
    assert best["tile_block"] in (64, 128, 256)
    assert best["ms_per_iter"] == min(t["ms_per_iter"] for t in trials)
    # all 50 iterations as one captured graph: one launch
    b = Batch(clouds, cases, tile_block=best["tile_block"], lds_tiles=best["lds_tiles"],
              points_per_thread=best["points_per_thread"], graph_iters=50, **num)
    rn, rs = b.q_lskum(50)
    forces = b.compute_cl_cd_cm()
    prims = b.prim()
    assert not any(b.failed_at())
    b.close()
    assert np.isfinite(rn).all() and np.isfinite(forces).all()
    # the single runs take another points-per-thread choice: results do
    # not depend on it
    other_ppt = 1 if best["points_per_thread"] != 1 else 4
    for j in (0, 27, 63):
        ci, mach, aoa = cases[j]
        s = Solver(clouds[ci], mach=mach, aoa=aoa, tile_block=best["tile_block"],
                   lds_tiles=best["lds_tiles"], points_per_thread=other_ppt, **num)
        srn, srs = s.q_lskum(50)
        print(f"case: case {j} (cloud {ci}, M {mach:.3f}, AoA {aoa:.2f}) alone", flush=True)
        np.testing.assert_array_equal(rn[j], srn)
        np.testing.assert_array_equal(rs[j], srs)
        np.testing.assert_array_equal(forces[j], np.array(s.compute_cl_cd_cm()))
        np.testing.assert_array_equal(prims[j], s.prim)
        s.close()


@pytest.mark.timeout(900)
def test_shape_gradient_cloud_duality_and_central_difference():
    cloud = lskum.workload("naca0012_1m_k24")
    assert cloud.n == 1_000_940
    s = Solver(cloud, mach=0.8, aoa=1.25, cfl=0.2, vl_const=150.0, inner_iterations=0)
    s.set_graph(50)
    s.q_lskum(2000)
    fields = shape.hicks_henne_fields(cloud, n_per_side=10, p_lo=0.05, p_hi=0.95,
                                      width_exponent=3, decay_radius=0.5)
    n_iter, w = 5, (1.0, 0.0, 0.0)
    J, xb, yb, pb = s.shape_vjp(n_iter, w)
    assert J == s.objective_after(n_iter, w)
    rng = np.random.default_rng(7)
    p0 = s.prim
    dp = rng.standard_normal(p0.shape) * np.abs(p0) * 1e-3
    for j, dprim in ((3, None), (14, dp)):
        Jt, dJ = s.shape_jvp(n_iter, w, dy=fields[j], dprim=dprim)
        rev = float(yb @ fields[j] + (np.sum(pb * dprim) if dprim is not None else 0.0))
        print(f"case: 1M points duality field {j}: adjoint {rev:.15e} tangent {dJ:.15e}",
              flush=True)
        assert abs(Jt - J) <= 1e-12 * abs(J)
        assert abs(rev - dJ) <= 1e-10 * abs(dJ)
    # differences along a field on each surface, h = 1e-8.  A limiter or
    # extremum switch of some point within +-h is a kink of J: the one-sided
    # difference taken on the side free of it is the slope of the piece the
    # state is on, the other one and the central difference are not (with
    # h = 1e-6 the central difference is off by 5e-5 to 8e-5 relative at
    # this state).  Rounding moves J (about 0.1, a sum over 1,500 wall
    # points) by some 1e-16 absolute, 1e-8 of the difference, which is
    # 5e-6 relative for the flatter field 14 (slope 7.7e-3); measured, the
    # better side agreed to 1.8e-6 and 9.2e-6.
    for j in (3, 14):
        g = float(yb @ fields[j])
        c8, f8, b8 = shape.design_fd(s, fields[j], n_iter, w, h=1e-8, one_sided=True)
        c6 = shape.design_fd(s, fields[j], n_iter, w, h=1e-6)
        rel = [abs(d - g) / abs(g) for d in (c8, f8, b8, c6)]
        print(f"case: 1M points differences field {j}: adjoint {g:.10e}, h=1e-8 central "
              f"{rel[0]:.1e} forward {rel[1]:.1e} backward {rel[2]:.1e}, h=1e-6 central "
              f"{rel[3]:.1e}", flush=True)
        assert min(rel[:3]) <= 3e-5
