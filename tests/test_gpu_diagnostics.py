"""Diagnostics on the GPU (lskum_compute_entropy, lskum_compute_enstrophy,
lskum_stagnation_pressure, lskum_objective_J) against the reference's own
values (diag_small.npz: the reference's diagnostics of its final state after
40 iterations) and against the oracle; a batch gives each case what a
context of its own gives."""
import numpy as np
import pytest

import lskum
from conftest import golden, normwise
from lskum.batch import Batch
from lskum.solver import Solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_diagnostics_match_reference_fixture():
    g = golden("diag_small.npz")
    cloud = lskum.workload("naca0012_1480_k20")
    mach, aoa, cfl, vl, power, inner = g["params"]
    s = Solver(cloud, mach=mach, aoa=aoa, cfl=cfl, vl_const=vl, power=power,
               inner_iterations=int(inner))
    s.set_state(g["prim"])
    print("case: diagnostics on the reference's final state", flush=True)
    e, te = s.compute_entropy()
    v, v2, tv = s.compute_enstrophy()
    assert normwise(e, g["entropy"]) <= 1e-14
    assert normwise(v, g["vorticity"]) <= 1e-12 and normwise(v2, g["vorticity_sqr"]) <= 1e-12
    assert tv == pytest.approx(float(g["total_enstrophy"]), rel=1e-12)
    assert te == pytest.approx(np.sum(g["entropy"]), rel=1e-12)
    stag, idx = s.stagnation_pressure()
    assert np.allclose(stag, g["stagnation_printed"], rtol=0, atol=5e-7)
    o = O.Oracle(cloud, mach=mach, aoa=aoa, cfl=cfl, vl_const=vl, power=power,
                 inner_iterations=int(inner))
    o.set("prim", g["prim"])
    ostag, oidx = o.stagnation_pressure()
    assert np.allclose(stag, ostag, rtol=1e-14, atol=0) and np.array_equal(idx, oidx)
    assert s.objective_J() == pytest.approx(o.objective_J(), rel=1e-12)


def test_diagnostics_developed_k24_against_oracle():
    cloud = lskum.workload("naca0012_3620_k24")
    s = Solver(cloud, mach=0.85, aoa=1.25, cfl=0.2, vl_const=5.0, inner_iterations=2)
    s.q_lskum(30)
    o = O.Oracle(cloud, mach=0.85, aoa=1.25, cfl=0.2, vl_const=5.0, inner_iterations=2)
    o.set("prim", s.prim)
    print("case: diagnostics k24 developed", flush=True)
    e, te = s.compute_entropy()
    oe, ote = o.compute_entropy()
    assert normwise(e, oe) <= 1e-13 and te == pytest.approx(ote, rel=1e-12)
    v, v2, tv = s.compute_enstrophy()
    ov, ov2, otv = o.compute_enstrophy()
    assert normwise(v, ov) <= 1e-11 and tv == pytest.approx(otv, rel=1e-11)
    stag, idx = s.stagnation_pressure()
    ostag, oidx = o.stagnation_pressure()
    assert np.allclose(stag, ostag, rtol=1e-14, atol=0) and np.array_equal(idx, oidx)
    assert s.objective_J() == pytest.approx(o.objective_J(), rel=1e-11)


def test_batch_diagnostics_equal_single_contexts():
    clouds = [lskum.workload("naca0012_1480_k20"), lskum.workload("naca0012_3620_k24")]
    cases = [(0, 0.63, 2.0), (1, 0.8, 1.25), (0, 0.5, -1.0)]
    num = dict(cfl=0.2, vl_const=5.0, power=0.0, inner_iterations=0)
    b = Batch(clouds, cases, **num)
    b.q_lskum(20)
    print("case: batch diagnostics", flush=True)
    ent, tent = b.diagnostics("entropy")
    vor, tvor = b.diagnostics("enstrophy")
    stag, idx = b.diagnostics("stagnation")
    J = b.diagnostics("J")
    for j, (ci, mach, aoa) in enumerate(cases):
        s = Solver(clouds[ci], mach=mach, aoa=aoa, **num)
        s.q_lskum(20)
        e, te = s.compute_entropy()
        v, v2, tv = s.compute_enstrophy()
        st, ix = s.stagnation_pressure()
        assert np.array_equal(ent[j], e) and tent[j] == te
        assert np.array_equal(vor[j], v) and tvor[j] == tv
        assert np.array_equal(stag[j], st) and np.array_equal(idx[j], ix)
        assert J[j] == s.objective_J()


def test_solution_and_cp_files_of_a_state(tmp_path):
    cloud = lskum.workload("naca0012_1480_k20")
    s = Solver(cloud, mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, inner_iterations=0)
    s.q_lskum(25)
    print("case: solution and cp files", flush=True)
    s.write_solution(tmp_path / "output.dat")
    s.write_cp(tmp_path / "cp_file")
    body = np.loadtxt(tmp_path / "output.dat", skiprows=1)
    np.testing.assert_array_equal(body[:, 3:7].T, s.prim)
    np.testing.assert_array_equal(body[:, 8], s.compute_entropy()[0])
    cp = np.loadtxt(tmp_path / "cp_file")
    nw = int(np.sum(cloud.flag_2 > 0))
    np.testing.assert_allclose(cp[:, 2], -(s.prim[3, :nw] - 1 / 1.4) / (0.5 * 0.63 ** 2),
                               rtol=1e-14)
    # the force integral of these Cp values gives the library's coefficients
    # (compute_force_coeffs_mod.h:25-66: panel length from the left and right
    # neighbours, the wall normal)
    m = np.arange(nw)
    l, r = cloud.left[m], cloud.right[m]
    ds = 0.5 * (np.hypot(cloud.x[m] - cloud.x[l], cloud.y[m] - cloud.y[l]) +
                np.hypot(cloud.x[r] - cloud.x[m], cloud.y[r] - cloud.y[m]))
    nx, ny = s.normals[0, m], s.normals[1, m]
    H, V = np.sum(cp[:, 2] * nx * ds), np.sum(cp[:, 2] * ny * ds)
    th = np.radians(2.0)
    cl, cd, _ = s.compute_cl_cd_cm()
    assert V * np.cos(th) - H * np.sin(th) == pytest.approx(cl, rel=1e-12)
    assert H * np.cos(th) + V * np.sin(th) == pytest.approx(cd, rel=1e-12)
