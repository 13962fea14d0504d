"""HIP path vs the oracle and the reference fixtures, through the C ABI.

Tolerances are normwise per component, max|gpu - ref| / max|ref| (pointwise
relative error is ill-posed where a component crosses zero).  The FP64
kernels keep the reference's formulas; differences come from FMA
contraction, the flux's reuse of shared terms and precomputed least-squares
coefficients, the ROCm device exp/log/erf (about an ulp from glibc) and the
order of the residual reduction:
  - one stage from identical inputs:          <= 1e-12
  - whole iterations vs the reference:        <= 1e-10 (the stated parity bar)
"""
import math

import numpy as np
import pytest

import lskum
from lskum import _lib as L
from lskum.solver import Solver, kat_fluxes, kat_qtilde
from conftest import cloud_sha, golden, normwise
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STAGE_TOL = 1e-12
ITER_TOL = 1e-10


def _solver(cloud, params, **kw):
    mach, aoa, cfl, vl, power, inner = params
    return Solver(cloud, mach=mach, aoa=aoa, cfl=cfl, vl_const=vl, power=power,
                  inner_iterations=int(inner), **kw)


def _oracle(cloud, params, threads=8):
    mach, aoa, cfl, vl, power, inner = params
    return O.Oracle(cloud, mach=mach, aoa=aoa, cfl=cfl, vl_const=vl, power=power,
                    inner_iterations=int(inner), threads=threads)


def test_device_fluxes_match_reference_known_answers():
    g = golden("kat_data.npz")
    rows = g["rows"]
    print("case: kat fluxes on data.dat rows", rows.shape, flush=True)
    got = kat_fluxes(rows[:, 4:10])
    ref = g["fluxes"]
    for f in range(8):
        assert normwise(got[:, f, :].T, ref[:, f, :].T) <= 1e-14, f
    qt = np.stack([rows[:, 0], rows[:, 1], rows[:, 2], -rows[:, 3]], axis=1)
    got = kat_qtilde(qt)
    ref = g["qtilde_prim"]
    # the reference's exp overflows in rho (and so p) on three rows: the
    # device must overflow on exactly those, to +inf, and agree elsewhere
    inf_rows = np.flatnonzero(~np.isfinite(ref).all(axis=1))
    assert inf_rows.tolist() == [3, 450, 788]
    np.testing.assert_array_equal(np.isposinf(got), np.isposinf(ref))
    assert np.isfinite(got[~np.isinf(ref)]).all()
    keep = np.setdiff1d(np.arange(ref.shape[0]), inf_rows)
    for comp in range(2):  # u1, u2 on every finite row
        assert normwise(got[keep, comp], ref[keep, comp]) <= 1e-14, comp
    # rho = exp(arg), arg = q~1 + beta |u|^2 - 2.5 ln beta (q_variables_mod.h:
    # 245-251), and p = rho / (2 beta).  The reference rounds arg (|arg| ulps
    # of exp's result, plus 2.5 |ln beta| from its logarithm); the device
    # forms exp(t) beta^-2.5 with t = q~1 + beta |u|^2 (|t| <= |arg| +
    # 2.5 |ln beta| ulps) and beta^-2.5 = y^5 from a refined 1/sqrt(beta)
    # (about a dozen ulps with the products), and p takes 0.5 y^2 (a few
    # more).  The data.dat rows reach arg = 186 (rho = 9e80), so rho and p
    # are held, row by row, to (16 + 2 |arg| + 5 |ln beta|) ulps of
    # relative error, the sum of those terms.
    beta = -qt[:, 3] * 0.5
    u1, u2 = qt[:, 1] * (0.5 / beta), qt[:, 2] * (0.5 / beta)
    arg = qt[:, 0] + beta * (u1 * u1 + u2 * u2) - np.log(beta) * 2.5
    bound = ((16.0 + 2.0 * np.abs(arg[keep]) + 5.0 * np.abs(np.log(beta[keep])))
             * np.finfo(np.float64).eps)
    for comp in (2, 3):
        rel = np.abs(got[keep, comp] / ref[keep, comp] - 1.0)
        worst = float((rel / bound).max())
        print(f"case: kat qtilde comp {comp}: worst relative error / bound {worst:.3f}",
              flush=True)
        assert worst <= 1.0, comp


def test_stages_match_oracle_from_a_developed_state():
    g = golden("small_m063.npz")
    cloud = lskum.workload("naca0012_1480_k20")
    params = (0.63, 2.0, 0.2, 5.0, 0.0, 2)
    print("case: stage parity naca0012_1480_k20", flush=True)
    s = _solver(cloud, params)
    o = _oracle(cloud, params)
    prim = g["rk1_prim_in"]
    s.set_state(prim)
    o.set("prim", prim)
    o.set("prim_old", prim)
    s.snapshot()
    np.testing.assert_array_equal(s.get(L.A_PRIM_OLD), prim)
    s.func_delta()
    o.func_delta()
    assert normwise(s.delta, o.get("delta")) <= STAGE_TOL
    for rk in range(1, 5):
        s.eval_q_variables()
        o.eval_q_variables()
        assert normwise(s.q, o.get("q")) <= STAGE_TOL, rk
        s.eval_q_derivatives()
        o.eval_q_derivatives()
        assert normwise(s.qm, o.get("qm")) <= STAGE_TOL, rk
        assert normwise(s.dq, o.get("dq")) <= STAGE_TOL, rk
        for _ in range(2):
            s.eval_q_inner_loop()
            o.eval_q_inner_loop()
            o.eval_update_innerloop()
        assert normwise(s.dq, o.get("dq")) <= STAGE_TOL, rk
        # feed both the oracle's derivative state so the flux sees identical inputs
        s.cal_flux_residual()
        o.cal_flux_residual()
        assert normwise(s.flux_res, o.get("flux_res")) <= STAGE_TOL, rk
        sr, mr, mp = s.state_update(rk)
        o.state_update(rk)
        sc = o.scalars()
        assert abs(sr - sc["sum_res_sqr"]) <= 1e-12 * sc["sum_res_sqr"], rk
        assert abs(mr - sc["max_res"]) <= 1e-12 * sc["max_res"], rk
        assert mp == sc["max_res_point"], rk
        assert normwise(s.prim, o.get("prim")) <= STAGE_TOL, rk
        # re-synchronise so that each stage is compared from equal inputs
        s.set_state(o.get("prim"))


@pytest.mark.parametrize("reorder", [L.REORDER_NONE, L.REORDER_RCM, L.REORDER_HILBERT])
def test_iterations_match_reference_small(reorder):
    g = golden("small_m063.npz")
    cloud = lskum.workload("naca0012_1480_k20")
    print("case: 40 iterations naca0012_1480_k20 reorder", reorder, flush=True)
    s = _solver(cloud, g["params"], reorder=reorder)
    rn, rs = s.q_lskum(40)
    h = g["history"]
    assert normwise(rn, h[:, 0]) <= ITER_TOL
    np.testing.assert_allclose(rs[2:], h[2:, 1], rtol=0, atol=1e-9)
    cl, cd, cm = s.compute_cl_cd_cm()
    np.testing.assert_allclose([cl, cd, cm], h[-1, 2:], rtol=ITER_TOL, atol=1e-13)
    e_prim, e_flux = normwise(s.prim, g["prim"]), normwise(s.flux_res, g["flux_res"])
    print(f"case: 40 iterations reorder {reorder}: prim {e_prim:.2e} flux_res {e_flux:.2e}",
          flush=True)
    assert e_prim <= ITER_TOL
    assert e_flux <= ITER_TOL


def test_point_orders_give_bit_identical_states():
    # per-point arithmetic does not depend on the device order or on whether
    # neighbours come from LDS tiles or from HBM; only the residual sum's
    # reduction order does
    g = golden("k24_m085.npz")
    cloud = lskum.workload("naca0012_3620_k24")
    runs = []
    for order, tiles, tb in [(L.REORDER_NONE, 0, 0), (L.REORDER_RCM, -1, 0),
                             (L.REORDER_HILBERT, 15, 0), (L.REORDER_HILBERT, 0, 0),
                             (L.REORDER_HILBERT, -1, 0), (L.REORDER_HILBERT, -1, 64),
                             (L.REORDER_HILBERT, 15, 128)]:
        s = _solver(cloud, g["params"], reorder=order, lds_tiles=tiles, tile_block=tb)
        s.q_lskum(10)
        runs.append((s.layout(), s.prim, s.flux_res, s.dq, s.qm))
    assert runs[2][0][0] == 15, "Hilbert order should let every kernel use LDS tiles"
    assert runs[3][0][0] == 0
    for r in runs[1:]:
        for a, b in zip(runs[0][1:], r[1:]):
            np.testing.assert_array_equal(a, b)


def test_iterations_match_reference_k24_inner_passes():
    g = golden("k24_m085.npz")
    cloud = lskum.workload("naca0012_3620_k24")
    print("case: 20 iterations naca0012_3620_k24 M0.85 inner 2", flush=True)
    s = _solver(cloud, g["params"])
    rn, rs = s.q_lskum(20)
    h = g["history"]
    assert normwise(rn, h[:, 0]) <= ITER_TOL
    cl, cd, cm = s.compute_cl_cd_cm()
    np.testing.assert_allclose([cl, cd, cm], h[-1, 2:], rtol=ITER_TOL, atol=1e-13)
    e_prim, e_flux = normwise(s.prim, g["prim"]), normwise(s.flux_res, g["flux_res"])
    print(f"case: k24 20 iterations: prim {e_prim:.2e} flux_res {e_flux:.2e}", flush=True)
    assert e_prim <= ITER_TOL
    assert e_flux <= ITER_TOL


def test_9760_points_1000_iterations_match_reference():
    g = golden("naca0012_9760_m063_1000it.npz")
    cloud = lskum.workload("naca0012_9760_k20")
    print("case: naca0012_9760_k20 M0.63, 1000 iterations", flush=True)
    s = _solver(cloud, g["params"])
    h = g["history"]
    rn_all = []
    for chunk in range(10):
        rn, rs = s.q_lskum(100)
        rn_all.append(rn)
        cl, cd, cm = s.compute_cl_cd_cm()
        ref = h[100 * chunk + 99]
        np.testing.assert_allclose([cl, cd], ref[2:4], rtol=ITER_TOL, atol=0)
        assert abs(cm - ref[4]) <= ITER_TOL * max(abs(ref[4]), abs(ref[2]))
    rn_all = np.concatenate(rn_all)
    assert normwise(rn_all, h[:, 0]) <= ITER_TOL
    e_prim, e_flux = normwise(s.prim, g["prim"]), normwise(s.flux_res, g["flux_res"])
    print(f"case: 1000 iterations: prim {e_prim:.2e} flux_res {e_flux:.2e}", flush=True)
    assert e_prim <= ITER_TOL
    assert e_flux <= ITER_TOL


def test_runs_are_deterministic():
    cloud = lskum.workload("naca0012_3620_k24")
    params = (0.85, 1.25, 0.2, 5.0, 0.0, 2)
    a = _solver(cloud, params)
    b = _solver(cloud, params)
    ra = a.q_lskum(15)
    rb = b.q_lskum(15)
    np.testing.assert_array_equal(ra[0], rb[0])
    np.testing.assert_array_equal(a.prim, b.prim)


def test_nonfinite_state_is_reported():
    cloud = lskum.workload("naca0012_1480_k20")
    s = _solver(cloud, (0.63, 2.0, 0.2, 5.0, 0.0, 0))
    p = s.prim
    p[3, 700] = -1.0  # negative pressure: log of a negative beta downstream
    s.set_state(p)
    with pytest.raises(L.LskumError) as e:
        s.q_lskum(3)
    assert e.value.code == -5


def test_free_stream_preserved_at_full_size():
    cloud = lskum.workload("naca0012_2p5m_k24")
    print("case: free stream, naca0012_2p5m_k24", cloud.n, flush=True)
    s = _solver(cloud, (0.85, 1.25, 0.2, 5.0, 0.0, 2))
    # every stencil kernel tiled: the flux's partial tile (605 entries of
    # 128 bytes) fits twice in a CU's LDS
    assert s.layout()[0] == 15
    s.func_delta()
    s.eval_q_variables()
    s.eval_q_derivatives()
    s.eval_q_inner_loop(2)
    s.cal_flux_residual()
    assert not np.any(s.flux_res)


def test_full_size_iterations_match_oracle():
    # three iterations at 2.5M points, k = 24, two inner passes, vs the oracle
    # on all host cores
    cloud = lskum.workload("naca0012_2p5m_k24")
    params = (0.85, 1.25, 0.2, 5.0, 0.0, 2)
    print("case: 3 iterations naca0012_2p5m_k24 vs oracle", flush=True)
    s = _solver(cloud, params)
    rn, _ = s.q_lskum(3)
    o = _oracle(cloud, params, threads=16)
    orn, _ = o.q_lskum(3)
    assert normwise(rn, orn) <= STAGE_TOL
    assert normwise(s.prim, o.get("prim")) <= STAGE_TOL
    assert normwise(s.flux_res, o.get("flux_res")) <= STAGE_TOL
    s.compute_cl_cd_cm()
    o.compute_cl_cd_cm()
    sc = o.scalars()
    np.testing.assert_allclose([s.Cl, s.Cd], [sc["cl"], sc["cd"]], rtol=1e-11)


def test_full_size_500_iterations_match_oracle():
    # the configuration's own length: 500 iterations at 2.5M points, k = 24,
    # two inner passes, M 0.85, against the oracle's run of the same cloud
    # (tests/golden/make_golden.py --only c2): the residual history, Cl, Cd
    # and Cm every 50 iterations, and the final prim and flux_res -- their
    # per-component max and L2 norms over all points and their values at
    # every wall point plus 4,000 others, normwise against the whole array
    g = golden("naca0012_2p5m_m085_500it.npz")
    cloud = lskum.workload("naca0012_2p5m_k24")
    assert cloud_sha(cloud) == str(g["cloud_sha"]), "generator changed"
    print("case: 500 iterations naca0012_2p5m_k24 vs the oracle's run", flush=True)
    s = _solver(cloud, g["params"])
    h = g["history"]
    rn_all, rs_all = [], []
    for chunk, (it, cl_o, cd_o, cm_o) in enumerate(g["forces"]):
        rn, rs = s.q_lskum(50)
        rn_all.append(rn)
        rs_all.append(rs)
        assert s.iteration()[0] == int(it)
        cl, cd, cm = s.compute_cl_cd_cm()
        print(f"case: it {int(it)} dCl {abs(cl - cl_o) / abs(cl_o):.2e} "
              f"dCd {abs(cd - cd_o) / abs(cd_o):.2e}", flush=True)
        np.testing.assert_allclose([cl, cd], [cl_o, cd_o], rtol=ITER_TOL, atol=0)
        assert abs(cm - cm_o) <= ITER_TOL * max(abs(cm_o), abs(cl_o))
    rn_all, rs_all = np.concatenate(rn_all), np.concatenate(rs_all)
    assert normwise(rn_all, h[:, 0]) <= ITER_TOL
    np.testing.assert_allclose(rs_all[2:], h[2:, 1], rtol=0, atol=1e-9)
    prim, fres = s.prim, s.flux_res
    smp = g["sample"]
    for name, a, ref, amax, l2 in (("prim", prim, g["prim_sample"], g["prim_absmax"],
                                    g["prim_l2"]),
                                   ("flux_res", fres, g["flux_res_sample"],
                                    g["flux_res_absmax"], g["flux_res_l2"])):
        assert np.isfinite(a).all(), name
        e_pts = max(float(np.abs(a[r, smp] - ref[r]).max() / amax[r]) for r in range(4))
        e_max = float(np.abs(np.abs(a).max(axis=1) - amax).max() / amax.min())
        e_l2 = float(np.abs(np.sqrt((a * a).sum(axis=1)) / l2 - 1.0).max())
        print(f"case: 500 iterations {name}: sample {e_pts:.2e} max {e_max:.2e} "
              f"l2 {e_l2:.2e}", flush=True)
        assert e_pts <= ITER_TOL and e_max <= ITER_TOL and e_l2 <= ITER_TOL, name


def test_restart_resumes_bit_for_bit(tmp_path):
    # q_lskum_mod.h:29-33 / fpi_solver_mod.h:139 with restart = 1: the
    # iteration count and res_old come from the file, the latch is kept
    cloud = lskum.workload("naca0012_3620_k24")
    params = (0.85, 1.25, 0.2, 5.0, 0.0, 2)
    a = _solver(cloud, params)
    a.q_lskum(15)
    a.write_restart(tmp_path / "restart.dat")
    ra = a.q_lskum(15)
    b = _solver(cloud, params)
    b.read_restart(tmp_path / "restart.dat")
    assert b.iteration()[0] == a.iteration()[0] - 15, b.iteration()
    rb = b.q_lskum(15)
    np.testing.assert_array_equal(ra[0], rb[0])
    np.testing.assert_array_equal(ra[1], rb[1])
    np.testing.assert_array_equal(a.prim, b.prim)


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_iteration_delta_equals_the_stand_alone_pass(precision):
    # lskum_iterate takes func_delta's minimum inside stage 1's
    # q_derivatives (where it can); the delta an iteration leaves is the
    # stand-alone pass's on the iteration's starting state, bit for bit
    g = golden("small_m063.npz")
    cloud = lskum.workload("naca0012_1480_k20")
    params = (0.63, 2.0, 0.2, 5.0, 0.0, 2)
    prim = g["rk1_prim_in"]
    print("case: iteration delta vs func_delta, precision", precision, flush=True)
    a = _solver(cloud, params, precision=precision)
    a.set_state(prim)
    a.q_lskum(1)
    da = a.delta
    a.close()
    b = _solver(cloud, params, precision=precision)
    b.set_state(prim)
    b.func_delta()
    db = b.delta
    b.close()
    np.testing.assert_array_equal(da, db)


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_flux_tile_cap_reads_the_rest_through_l2(precision):
    # with enough neighbours per point the largest tiles exceed the entries
    # the flux stages per block (64 neighbours: 700-entry tiles, beyond the
    # FP64 flux's cap; the mixed flux's smaller records need more); the
    # blocks beyond the cap read their neighbours through L2, and the states
    # equal those of the untiled layout bit for bit
    for k in (64, 96, 128):
        cloud = lskum.Cloud.generate(naca="0012", n_wall=160, n_radial=64, n_circ=150,
                                     stretch=1.08, k=k, radius_factor=6.0, seed=1)
        runs = []
        for tiles in (L.TILES_AUTO, 0):
            s = Solver(cloud, mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, inner_iterations=1,
                       precision=precision, lds_tiles=tiles)
            print("case: flux tile cap, precision", precision, "k", k, "layout", s.layout(),
                  "flux tile", s.flux_tile(), flush=True)
            runs.append((s.layout()[1], s.flux_tile()))
            if tiles == L.TILES_AUTO and not 0 < runs[0][1] < runs[0][0]:
                s.close()
                break  # every tile fits: more neighbours
            s.q_lskum(10)
            runs[-1] += (s.prim, s.flux_res)
            s.close()
        if len(runs) == 2:
            break
    assert len(runs) == 2, "no neighbour count made a tile exceed the flux's cap"
    (tmax, cap), (_, cap0) = runs[0][:2], runs[1][:2]
    assert 0 < cap < tmax and cap0 == 0
    for a, b in zip(runs[0][2:], runs[1][2:]):
        np.testing.assert_array_equal(a, b)
