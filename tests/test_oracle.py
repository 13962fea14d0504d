"""The oracle pinned against the compiled reference and against mathematics.

CPU only.  The golden fixtures were written by tests/golden/make_golden.py
from the reference's own functions; the oracle must reproduce them bit for
bit (same operation order, same glibc libm, no FMA).  The remaining tests pin
properties the mathematics fixes, independently of both.
"""
import hashlib
import math
import os

import numpy as np
import pytest

import lskum
from conftest import golden
from oracle import oracle as O

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _grid_sha(cloud, tmp_path):
    p = tmp_path / "g.txt"
    cloud.write_partgrid(p)
    return hashlib.sha256(p.read_bytes()).hexdigest()


def make_oracle(cloud, params, threads=4):
    mach, aoa, cfl, vl, power, inner = params
    return O.Oracle(cloud, mach=mach, aoa=aoa, cfl=cfl, vl_const=vl, power=power,
                    inner_iterations=int(inner), threads=threads)


# ----------------------------------------------------------------- goldens
def test_kat_fluxes_match_reference():
    g = golden("kat_data.npz")
    rows = g["rows"]
    for r, row in enumerate(rows):
        nx, ny, u1, u2, rho, pr = row[4:10]
        for f in range(8):
            got = O.flux(f, nx, ny, u1, u2, rho, pr)
            assert np.array_equal(got, g["fluxes"][r, f]), (r, f)
        qt = np.array([row[0], row[1], row[2], -row[3]])
        assert np.array_equal(O.qtilde_to_primitive(qt), g["qtilde_prim"][r]), r


def test_small_cloud_stage_by_stage_bit_exact(tmp_path):
    g = golden("small_m063.npz")
    cloud = lskum.workload("naca0012_1480_k20")
    assert _grid_sha(cloud, tmp_path) == str(g["grid_sha"])
    o = make_oracle(cloud, g["params"])
    # normals and sub-stencils
    assert np.array_equal(np.stack([o.get("nx"), o.get("ny")]), g["normals"])
    counts = np.stack([o.substencil(s)[0] for s in range(4)], axis=1)
    assert np.array_equal(counts, g["sub_counts"])
    per_point = []
    subs = [o.substencil(s) for s in range(4)]
    offs = [np.concatenate([[0], np.cumsum(c)]) for c, _ in subs]
    for i in range(cloud.n):
        for s in range(4):
            per_point.append(subs[s][1][offs[s][i]:offs[s][i + 1]])
    members = np.concatenate(per_point).astype(np.int32)
    assert _sha(members) == str(g["sub_members_sha"])
    # 40 iterations with iteration 30 done stage by stage
    hist = []
    for it in range(1, 41):
        if it == 30:
            o.snapshot()
            o.func_delta()
            assert np.array_equal(o.get("delta"), g["delta30"])
            for rk in range(1, 5):
                if rk == 1:
                    assert np.array_equal(o.get("prim"), g["rk1_prim_in"])
                o.eval_q_variables()
                if rk == 1:
                    assert np.array_equal(o.get("q"), g["rk1_q"])
                o.eval_q_derivatives()
                if rk == 1:
                    assert np.array_equal(o.get("qm"), g["rk1_qm"])
                    assert np.array_equal(o.get("dq"), g["rk1_dq"])
                o.cal_flux_residual()
                assert np.array_equal(o.get("flux_res"), g[f"rk{rk}_flux_res"]), rk
                o.state_update(rk)
                s = o.scalars()
                red = g[f"rk{rk}_red"]
                assert s["sum_res_sqr"] == red[0] and s["max_res"] == red[1]
                assert s["max_res_point"] == int(red[2])
                assert np.array_equal(o.get("prim"), g[f"rk{rk}_prim"]), rk
            s = o.scalars()
            # fpi_solver_mod.h:138,146 with the C libm (numpy's log10 may
            # round differently)
            res_new = math.sqrt(s["sum_res_sqr"]) / cloud.n
            row = [res_new, math.log10(res_new / s["res_old"])]
            o.set_iteration(30, s["res_old"], 0)
        else:
            o.fpi_solver(it)
            s = o.scalars()
            row = [s["res_new"], s["residue"]]
        o.compute_cl_cd_cm()
        s = o.scalars()
        hist.append(row + [s["cl"], s["cd"], s["cm"]])
    assert np.array_equal(np.array(hist), g["history"])
    for nm in ["prim", "q", "dq", "qm", "flux_res", "delta"]:
        assert np.array_equal(o.get(nm), g[nm]), nm


def test_k24_inner_passes_other_mach_bit_exact(tmp_path):
    g = golden("k24_m085.npz")
    cloud = lskum.workload("naca0012_3620_k24")
    assert _grid_sha(cloud, tmp_path) == str(g["grid_sha"])
    o = make_oracle(cloud, g["params"])
    rn, rs = o.q_lskum(20)
    h = g["history"]
    assert np.array_equal(rn, h[:, 0]) and np.array_equal(rs, h[:, 1])
    o.compute_cl_cd_cm()
    s = o.scalars()
    # the reference's force routine divides by its compile-time Mach; the
    # harness rescales, which costs at most an ulp or two
    np.testing.assert_allclose([s["cl"], s["cd"], s["cm"]], h[-1, 2:], rtol=1e-14, atol=0)
    assert np.array_equal(o.get("prim"), g["prim"])
    assert np.array_equal(o.get("flux_res"), g["flux_res"])


@pytest.mark.slow
def test_9760_points_1000_iterations_bit_exact(tmp_path):
    g = golden("naca0012_9760_m063_1000it.npz")
    cloud = lskum.workload("naca0012_9760_k20")
    assert _grid_sha(cloud, tmp_path) == str(g["grid_sha"])
    o = make_oracle(cloud, g["params"], threads=8)
    rn, rs = o.q_lskum(1000)
    assert np.array_equal(rn, g["history"][:, 0])
    assert np.array_equal(rs, g["history"][:, 1])
    o.compute_cl_cd_cm()
    s = o.scalars()
    assert [s["cl"], s["cd"], s["cm"]] == list(g["history"][-1, 2:])
    for nm in ["prim", "q", "dq", "qm", "flux_res", "delta"]:
        assert _sha(o.get(nm)) == str(g[nm + "_sha"]), nm


# ------------------------------------------------------- mathematical pins
def _rand_state(rng, n):
    rho = rng.uniform(0.5, 2.0, n)
    u1 = rng.uniform(-1.5, 1.5, n)
    u2 = rng.uniform(-1.5, 1.5, n)
    pr = rng.uniform(0.3, 2.0, n)
    return rho, u1, u2, pr


def _q_of(rho, u1, u2, pr):
    """q-variables from their definition: q1 = ln rho + ln(beta)/(gamma-1) -
    beta |u|^2, (q2, q3) = 2 beta u, q4 = -2 beta, beta = rho/(2p)."""
    beta = rho / (2 * pr)
    return np.stack([np.log(rho) + np.log(beta) / 0.4 - beta * (u1**2 + u2**2),
                     2 * beta * u1, 2 * beta * u2, -2 * beta])


def test_qtilde_to_primitive_inverts_q_variables():
    rng = np.random.default_rng(0)
    for rho, u1, u2, pr in zip(*_rand_state(rng, 200)):
        qt = _q_of(rho, u1, u2, pr)
        got = O.qtilde_to_primitive(qt)
        np.testing.assert_allclose(got, [u1, u2, rho, pr], rtol=1e-12, atol=1e-13)


def _euler_flux(nx, ny, u1, u2, rho, pr, normal):
    tx, ty = ny, -nx
    ut, un = u1 * tx + u2 * ty, u1 * nx + u2 * ny
    e = 2.5 * pr + 0.5 * rho * (ut**2 + un**2)
    if not normal:
        return np.array([rho * ut, pr + rho * ut * ut, rho * ut * un, ut * (e + pr)])
    return np.array([rho * un, rho * ut * un, pr + rho * un * un, un * (e + pr)])


def test_split_fluxes_sum_to_euler_flux():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(a), np.sin(a)
        rho, u1, u2, pr = [v[0] for v in _rand_state(rng, 1)]
        args = (nx, ny, u1, u2, rho, pr)
        np.testing.assert_allclose(O.flux(0, *args) + O.flux(1, *args),
                                   _euler_flux(*args, normal=False), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(O.flux(2, *args) + O.flux(3, *args),
                                   _euler_flux(*args, normal=True), rtol=1e-12, atol=1e-12)


def test_quadrant_fluxes_sum_to_half_fluxes():
    rng = np.random.default_rng(2)
    for _ in range(200):
        a = rng.uniform(0, 2 * np.pi)
        args = (np.cos(a), np.sin(a)) + tuple(v[0] for v in _rand_state(rng, 1))
        np.testing.assert_allclose(O.flux(5, *args) + O.flux(6, *args), O.flux(0, *args),
                                   rtol=1e-12, atol=1e-12)  # GxII + GxIII = Gx+
        np.testing.assert_allclose(O.flux(4, *args) + O.flux(7, *args), O.flux(1, *args),
                                   rtol=1e-12, atol=1e-12)  # GxI + GxIV = Gx-


def test_split_fluxes_limits():
    # supersonic to the right: Gx- -> 0 and Gx+ -> Euler flux
    args = (0.0, 1.0, 8.0, 0.1, 1.0, 0.7)
    scale = np.abs(_euler_flux(*args, normal=False)).max()
    assert np.abs(O.flux(1, *args)).max() < 1e-15 * scale
    np.testing.assert_allclose(O.flux(0, *args), _euler_flux(*args, normal=False), rtol=1e-13)


@pytest.fixture(scope="module")
def small():
    cloud = lskum.workload("naca0012_1480_k20")
    return cloud, make_oracle(cloud, (0.63, 2.0, 0.2, 5.0, 0.0, 0))


def test_q_derivatives_exact_for_linear_fields(small):
    cloud, o = small
    coef = np.array([[0.3, -1.2, 2.0], [1.5, 0.7, -0.4], [-0.9, 0.2, 0.5], [-2.0, 0.1, 0.3]])
    q = coef[:, :1] + coef[:, 1:2] * cloud.x + coef[:, 2:3] * cloud.y
    o.set("q", q)
    o.eval_q_derivatives()
    dq = o.get("dq")
    np.testing.assert_allclose(dq[0], np.repeat(coef[:, 1:2], cloud.n, 1), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(dq[1], np.repeat(coef[:, 2:3], cloud.n, 1), rtol=1e-9, atol=1e-9)
    # qmax/qmin by brute force over {i} and its neighbours
    qm = o.get("qm")
    for i in range(0, cloud.n, 97):
        nb = np.concatenate([[i], cloud.conn[cloud.conn_off[i]:cloud.conn_off[i + 1]]])
        assert np.array_equal(qm[0][:, i], q[:, nb].max(1))
        assert np.array_equal(qm[1][:, i], q[:, nb].min(1))


def test_inner_pass_keeps_exact_gradients_of_quadratic_fields(small):
    # for q = a + b.x + x^T C x the defect-corrected difference
    # (q_k - D.grad_k/2) - (q_i - D.grad_i/2) equals grad_i . D exactly
    cloud, o = small
    x, y = cloud.x, cloud.y
    q = np.stack([0.5 + x - y + 0.3 * x * x + 0.2 * x * y - 0.1 * y * y,
                  1.0 + 0.1 * x * x, -0.5 + 0.4 * x * y, -1 + 0.2 * y * y])
    grad = np.stack([np.stack([1 + 0.6 * x + 0.2 * y, 0.2 * x, 0.4 * y, 0 * x]),
                     np.stack([-1 + 0.2 * x - 0.2 * y, 0 * x, 0.4 * x, 0.4 * y])])
    o.set("q", q)
    o.set("dq", grad)
    o.eval_q_inner_loop()
    o.eval_update_innerloop()
    np.testing.assert_allclose(o.get("dq"), grad, rtol=0, atol=1e-9)


def test_limiter_properties(small):
    cloud, o = small
    rng = np.random.default_rng(3)
    o.set("q", rng.normal(size=(4, cloud.n)))
    o.eval_q_derivatives()
    q, qm = o.get("q"), o.get("qm")
    import ctypes as C
    lib = O.lib()
    lib.orc_venkat_limiter.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
    for k in range(0, cloud.n, 37):
        for scale in [1e-7, 1e-3, 0.1, 10.0]:
            qt = q[:, k] + scale * rng.normal(size=4)
            phi = np.zeros(4)
            lib.orc_venkat_limiter(o.c, qt.ctypes.data_as(C.POINTER(C.c_double)),
                                   phi.ctypes.data_as(C.POINTER(C.c_double)), k)
            dn = qt - q[:, k]
            assert np.all(phi <= 1.0) and np.all(phi > 0.0)
            assert np.all(phi[np.abs(dn) <= 1e-5] == 1.0)
            # extrapolation inside [qmin, qmax] by a small fraction is not limited much
            inside = (qt < qm[0][:, k]) & (qt > qm[1][:, k]) & (np.abs(dn) > 1e-5)
            dp = np.where(dn > 0, qm[0][:, k] - q[:, k], qm[1][:, k] - q[:, k])
            strong = inside & (np.abs(dn) < 0.1 * np.abs(dp))
            assert np.all(phi[strong] > 0.9)


def test_free_stream_is_preserved(small):
    cloud = small[0]
    for mach, aoa in [(0.63, 2.0), (0.85, 1.25)]:
        o = make_oracle(cloud, (mach, aoa, 0.2, 5.0, 0.0, 2))
        o.eval_q_variables()
        o.eval_q_derivatives()
        o.eval_q_inner_loop()
        o.eval_update_innerloop()
        o.func_delta()
        o.cal_flux_residual()
        assert np.array_equal(o.get("flux_res"), np.zeros((4, cloud.n)))


def test_state_update_wall_and_outer_boundaries(small):
    cloud, _ = small
    o = make_oracle(cloud, (0.63, 2.0, 0.2, 5.0, 0.0, 0))
    o.snapshot()
    o.set("flux_res", np.zeros((4, cloud.n)))
    o.state_update(1)
    p = o.get("prim")
    nx, ny = o.get("nx"), o.get("ny")
    wall = cloud.flag_1 == 0
    outer = cloud.flag_1 == 2
    # wall: normal velocity removed (state_update_mod.h:150)
    assert np.abs(p[1][wall] * nx[wall] + p[2][wall] * ny[wall]).max() < 1e-15
    # far field: Ubar of the free stream is the free stream (A+ + A- = 1, B terms cancel)
    qinf = np.array([1.0, 0.63 * np.cos(2 * np.pi / 180), 0.63 * np.sin(2 * np.pi / 180), 1 / 1.4])
    np.testing.assert_allclose(p[:, outer], np.repeat(qinf[:, None], outer.sum(), 1), rtol=1e-14, atol=1e-15)
    # interior: zero residual leaves the state (to rounding)
    inner = cloud.flag_1 == 1
    np.testing.assert_allclose(p[:, inner], np.repeat(qinf[:, None], inner.sum(), 1), rtol=1e-14, atol=1e-15)


def test_func_delta_brute_force(small):
    cloud, o = small
    rng = np.random.default_rng(4)
    rho, u1, u2, pr = _rand_state(rng, cloud.n)
    o.set("prim", np.stack([rho, u1, u2, pr]))
    o.func_delta()
    got = o.get("delta")
    for i in range(0, cloud.n, 53):
        nb = cloud.conn[cloud.conn_off[i]:cloud.conn_off[i + 1]]
        d = np.hypot(cloud.x[nb] - cloud.x[i], cloud.y[nb] - cloud.y[i])
        dt = 0.2 * d / (np.hypot(u1[nb], u2[nb]) + 3 * np.sqrt(pr[nb] / rho[nb]))
        assert abs(got[i] - min(1.0, dt.min())) <= 1e-15 * got[i]


def test_forces_vanish_at_free_stream_pressure(small):
    cloud, o = small
    o2 = make_oracle(cloud, (0.63, 2.0, 0.2, 5.0, 0.0, 0))
    o2.compute_cl_cd_cm()
    s = o2.scalars()
    assert s["cl"] == 0.0 and s["cd"] == 0.0 and s["cm"] == 0.0


def test_forces_of_uniform_cp_vanish_on_closed_contour(small):
    # a uniform pressure offset integrates to (nearly) zero force on the closed
    # wall contour: sum(n ds) is the discrete closed-curve identity
    cloud, _ = small
    o2 = make_oracle(cloud, (0.63, 2.0, 0.2, 5.0, 0.0, 0))
    p = o2.get("prim")
    p[3] += 0.05
    o2.set("prim", p)
    o2.compute_cl_cd_cm()
    s = o2.scalars()
    assert abs(s["cl"]) < 2e-3 and abs(s["cd"]) < 2e-3


def test_thread_count_does_not_change_results(small):
    cloud = small[0]
    a = make_oracle(cloud, (0.63, 2.0, 0.2, 5.0, 0.0, 1), threads=1)
    b = make_oracle(cloud, (0.63, 2.0, 0.2, 5.0, 0.0, 1), threads=4)
    ra = a.q_lskum(5)
    rb = b.q_lskum(5)
    assert np.array_equal(ra[0], rb[0]) and np.array_equal(a.get("prim"), b.get("prim"))


# ---------------------------------------------------------------- diagnostics
def test_diagnostics_match_reference(tmp_path):
    """compute_entropy, compute_enstrophy and stagnation_pressure of the
    reference run on its own final state (diag_small.npz)."""
    g = golden("diag_small.npz")
    cloud = lskum.workload("naca0012_1480_k20")
    assert _grid_sha(cloud, tmp_path) == str(g["grid_sha"])
    o = make_oracle(cloud, g["params"])
    o.set("prim", g["prim"])
    e, _ = o.compute_entropy()
    v, v2, tot = o.compute_enstrophy()
    assert np.array_equal(e, g["entropy"])
    assert np.array_equal(v, g["vorticity"]) and np.array_equal(v2, g["vorticity_sqr"])
    assert tot == float(g["total_enstrophy"])
    stag, idx = o.stagnation_pressure()
    # the reference prints these four with %f
    assert np.allclose(stag, g["stagnation_printed"], rtol=0, atol=5e-7)


def _uniform(cloud, rho, u1, u2, p):
    return np.tile(np.array([rho, u1, u2, p])[:, None], (1, cloud.n))


def test_stagnation_pressure_closed_forms():
    cloud = lskum.workload("naca0012_1480_k20")
    o = make_oracle(cloud, (0.63, 2.0, 0.2, 5.0, 0.0, 0))
    g = 1.4
    # isentropic stagnation ratio p0/p = (1 + (g-1)/2 M^2)^(g/(g-1)): at
    # M = 1, 1.2^3.5 = 1.89293; at M = 2, 1.8^3.5 = 7.82445
    for M, ratio in ((1.0, 1.8929291587378819), (2.0, 7.824449066867264)):
        p, rho = 0.7, 1.3
        a = math.sqrt(g * p / rho)
        o.set("prim", _uniform(cloud, rho, M * a * 0.6, M * a * 0.8, p))
        stag, idx = o.stagnation_pressure()
        assert stag[0] == pytest.approx(ratio * p, rel=1e-13)
        assert stag[0] == stag[1] and tuple(idx) == (0, 0)  # ties keep the first point
    # free stream: p0 = p0_inf everywhere, J = 0
    o.set("prim", _uniform(cloud, 1.0, 0.63 * math.cos(math.radians(2)),
                           0.63 * math.sin(math.radians(2)), 1 / 1.4))
    stag, _ = o.stagnation_pressure()
    assert stag[2] == pytest.approx(1.0, abs=1e-14) and abs(o.objective_J()) < 1e-28
    # one point brought to rest: its p0 is its p, and J is that one term
    prim = o.get("prim")
    prim[1:3, 5] = 0.0
    o.set("prim", prim)
    p0_inf = (1 / 1.4) * (1 + 0.2 * 0.63 ** 2) ** 3.5
    assert o.objective_J() == pytest.approx((p0_inf - 1 / 1.4) ** 2 / (p0_inf ** 2 * cloud.n),
                                            rel=1e-10)
    stag, idx = o.stagnation_pressure()
    assert idx[0] == 5 and stag[0] == pytest.approx(1 / 1.4, rel=1e-15)


def test_entropy_and_vorticity_closed_forms():
    cloud = lskum.workload("naca0012_1480_k20")
    o = make_oracle(cloud, (0.63, 2.0, 0.2, 5.0, 0.0, 0))
    x, y = cloud.x, cloud.y
    # along the isentrope p = p_inf rho^gamma the entropy measure is 0; off
    # it, |ln(p / rho^g) - ln p_inf|
    rho = 1.0 + 0.2 * np.sin(x)
    prim = np.stack([rho, 0 * x, 0 * x, (1 / 1.4) * rho ** 1.4])
    o.set("prim", prim)
    e, tot = o.compute_entropy()
    assert np.abs(e).max() < 1e-14
    prim[3] *= math.e ** 0.25
    o.set("prim", prim)
    e, tot = o.compute_entropy()
    assert np.allclose(e, 0.25, atol=1e-14) and tot == pytest.approx(0.25 * cloud.n)
    # solid-body rotation u = w (-y, x): vorticity 2w at every point (the
    # least squares are exact for linear fields); a strain u = (x, -y) has none
    w = 0.37
    o.set("prim", np.stack([1 + 0 * x, -w * y, w * x, 0.7 + 0 * x]))
    v, v2, tot = o.compute_enstrophy()
    assert np.allclose(v, 2 * w, rtol=0, atol=1e-11)
    assert tot == pytest.approx(cloud.n * (2 * w) ** 2, rel=1e-10)
    o.set("prim", np.stack([1 + 0 * x, x, -y, 0.7 + 0 * x]))
    v, v2, tot = o.compute_enstrophy()
    assert np.abs(v).max() < 1e-11


def test_reference_restart_from_a_library_written_file_is_exact():
    # the reference's restart_sol() (initial_conditions_mod.h:7-21) read the
    # file lskum_restart_write made from its own state after 10 iterations,
    # and its q_lskum (q_lskum_mod.h:29-37) ran iterations 11..20: bit for
    # bit its uninterrupted run, so the writer's layout is the one the
    # reference reads (17 significant digits round-trip exactly)
    r = golden("restart_k24_m085.npz")
    k = golden("k24_m085.npz")
    assert str(r["grid_sha"]) == str(k["grid_sha"])
    np.testing.assert_array_equal(r["history"], k["history"][10:])
    np.testing.assert_array_equal(r["prim"], k["prim"])
    np.testing.assert_array_equal(r["flux_res"], k["flux_res"])
    assert [ln.split()[1] for ln in r["printed"]] == [str(i) for i in range(11, 21)]
    # the oracle resumes from the same file, read by the library's reader
    import lskum
    from lskum.solver import restart_read, restart_write
    import tempfile
    cloud = lskum.workload("naca0012_3620_k24")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "restart.dat")
        restart_write(path, cloud, int(r["it"]), float(r["res_old"]), r["prim10"])
        it, res_old, prim = restart_read(path, cloud.n)
    assert it == 10 and res_old == float(r["res_old"])
    np.testing.assert_array_equal(prim, r["prim10"])
    o = O.Oracle(cloud, mach=0.85, aoa=1.25, cfl=0.2, vl_const=5.0, power=0.0,
                 inner_iterations=2, threads=4)
    o.set("prim", prim)
    o.set_iteration(it, res_old, 1)
    rn, rs = o.q_lskum(10)
    np.testing.assert_array_equal(rn, r["history"][:, 0])
    np.testing.assert_array_equal(rs, r["history"][:, 1])
    np.testing.assert_array_equal(o.get("prim"), r["prim"])
