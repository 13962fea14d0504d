"""Mixed-precision form vs the reference fixtures and the oracle, through the
C ABI.

The mixed form gathers neighbour data in FP32 and evaluates the limiter,
q~ -> primitive, the fluxes and each neighbour's local-step term in FP32,
keeping FP64 for coordinates, least-squares solves, sums, the state update
and the reductions.  Its bars
are therefore set by FP32 rounding (unit roundoff 6e-8) and by how the
stages amplify it, not by the FP64 parity bar:
  - q: one rounding of the FP64 value                          <= 1e-7
  - delta: FP32 offset, sqrtf and reciprocal per neighbour     <= 1e-6
  - dq: differences of neighbouring FP32 q, ~1e-3 apart in
    relative terms, so ~6e-8/1e-3 relative per term            <= 2e-5
  - flux residual: the same cancellation in G_k - G_i          <= 5e-5
  - prim after one stage (the residual is a 0.5*delta-sized
    increment of the state)                                    <= 1e-7
  - Cl, Cd after whole runs, against the FP64 reference:
    |dCl|, |dCd| <= 1e-6 (the project's stated bound)
Norms are normwise per component, max|a - b| / max|b|.
"""
import numpy as np
import pytest

import lskum
from lskum import _lib as L
from lskum.solver import Solver, kat_fluxes
from conftest import golden, normwise
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-6


def _solver(cloud, params, precision=L.MIXED, **kw):
    mach, aoa, cfl, vl, power, inner = params
    return Solver(cloud, mach=mach, aoa=aoa, cfl=cfl, vl_const=vl, power=power,
                  inner_iterations=int(inner), precision=precision, **kw)


def test_mixed_fluxes_on_reference_known_answers():
    g = golden("kat_data.npz")
    rows = g["rows"]
    print("case: mixed kat fluxes", rows.shape, flush=True)
    got = kat_fluxes(rows[:, 4:10], mixed=True)
    for f in range(8):
        assert normwise(got[:, f, :].T, g["fluxes"][:, f, :].T) <= 5e-6, f


def test_mixed_stages_against_oracle():
    g = golden("small_m063.npz")
    cloud = lskum.workload("naca0012_1480_k20")
    params = (0.63, 2.0, 0.2, 5.0, 0.0, 2)
    print("case: mixed stage errors naca0012_1480_k20", flush=True)
    s = _solver(cloud, params)
    o = O.Oracle(cloud, mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, inner_iterations=2,
                 threads=8)
    prim = g["rk1_prim_in"]
    s.set_state(prim)
    o.set("prim", prim)
    o.set("prim_old", prim)
    s.snapshot()
    s.func_delta()
    o.func_delta()
    # each neighbour's step cfl |x_k - x_i| / speed_k in FP32: a rounded
    # offset, sqrtf and a hardware reciprocal, a few FP32 ulps in all
    assert normwise(s.delta, o.get("delta")) <= 1e-6
    for rk in range(1, 5):
        s.eval_q_variables()
        o.eval_q_variables()
        assert normwise(s.q, o.get("q")) <= 1e-7, rk
        s.eval_q_derivatives()
        o.eval_q_derivatives()
        assert normwise(s.qm, o.get("qm")) <= 1e-7, rk
        assert normwise(s.dq, o.get("dq")) <= 2e-5, rk
        for _ in range(2):
            s.eval_q_inner_loop()
            o.eval_q_inner_loop()
            o.eval_update_innerloop()
        assert normwise(s.dq, o.get("dq")) <= 2e-5, rk
        s.cal_flux_residual()
        o.cal_flux_residual()
        assert normwise(s.flux_res, o.get("flux_res")) <= 5e-5, rk
        sr, _, _ = s.state_update(rk)
        o.state_update(rk)
        assert abs(sr - o.scalars()["sum_res_sqr"]) <= 1e-4 * o.scalars()["sum_res_sqr"], rk
        assert normwise(s.prim, o.get("prim")) <= 1e-7, rk
        s.set_state(o.get("prim"))


def test_mixed_free_stream_is_exact():
    # equal states at both ends of every pair give bit-equal FP32 fluxes
    cloud = lskum.workload("naca0012_3620_k24")
    s = _solver(cloud, (0.85, 1.25, 0.2, 5.0, 0.0, 2))
    s.func_delta()
    s.eval_q_variables()
    s.eval_q_derivatives()
    s.eval_q_inner_loop(2)
    assert not np.any(s.dq)
    s.cal_flux_residual()
    assert not np.any(s.flux_res)


def test_mixed_9760_points_1000_iterations_forces():
    # |dCl|, |dCd| against the reference's own FP64 run, every 100 iterations
    g = golden("naca0012_9760_m063_1000it.npz")
    cloud = lskum.workload("naca0012_9760_k20")
    print("case: mixed naca0012_9760_k20 M0.63, 1000 iterations", flush=True)
    s = _solver(cloud, g["params"])
    h = g["history"]
    rn_all = []
    worst = 0.0
    for chunk in range(10):
        rn, _ = s.q_lskum(100)
        rn_all.append(rn)
        cl, cd, _ = s.compute_cl_cd_cm()
        ref = h[100 * chunk + 99]
        worst = max(worst, abs(cl - ref[2]), abs(cd - ref[3]))
        assert abs(cl - ref[2]) <= FORCE_TOL and abs(cd - ref[3]) <= FORCE_TOL, chunk
    print(f"worst |dCl|, |dCd| = {worst:.2e}", flush=True)
    assert normwise(np.concatenate(rn_all), h[:, 0]) <= 1e-5
    assert normwise(s.prim, g["prim"]) <= 1e-5


def test_mixed_k24_inner_passes_forces():
    g = golden("k24_m085.npz")
    cloud = lskum.workload("naca0012_3620_k24")
    s = _solver(cloud, g["params"])
    rn, _ = s.q_lskum(20)
    h = g["history"]
    cl, cd, _ = s.compute_cl_cd_cm()
    assert abs(cl - h[-1, 2]) <= FORCE_TOL and abs(cd - h[-1, 3]) <= FORCE_TOL
    assert normwise(rn, h[:, 0]) <= 1e-5
    assert normwise(s.prim, g["prim"]) <= 1e-5


def test_mixed_layout_variants_are_bit_identical():
    cloud = lskum.workload("naca0012_3620_k24")
    params = (0.85, 1.25, 0.2, 5.0, 0.0, 2)
    runs = []
    for order, tiles, tb in [(L.REORDER_HILBERT, -1, 0), (L.REORDER_HILBERT, 0, 0),
                             (L.REORDER_HILBERT, 7, 128), (L.REORDER_RCM, -1, 64),
                             (L.REORDER_NONE, 15, 0)]:
        s = _solver(cloud, params, reorder=order, lds_tiles=tiles, tile_block=tb)
        s.q_lskum(8)
        runs.append((s.prim, s.flux_res, s.dq, s.qm))
        s.close()
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            np.testing.assert_array_equal(a, b)


def test_mixed_full_size_against_fp64():
    cloud = lskum.workload("naca0012_2p5m_k24")
    params = (0.85, 1.25, 0.2, 5.0, 0.0, 2)
    print("case: mixed vs FP64, 5 iterations naca0012_2p5m_k24", flush=True)
    a = _solver(cloud, params, precision=L.FP64)
    ra, _ = a.q_lskum(5)
    cla = a.compute_cl_cd_cm()
    pa = a.prim
    a.close()
    b = _solver(cloud, params)
    rb, _ = b.q_lskum(5)
    clb = b.compute_cl_cd_cm()
    assert abs(clb[0] - cla[0]) <= FORCE_TOL and abs(clb[1] - cla[1]) <= FORCE_TOL
    assert normwise(rb, ra) <= 1e-5
    assert normwise(b.prim, pa) <= 1e-6
