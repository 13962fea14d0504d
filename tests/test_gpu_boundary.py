"""The C++ boundary on the GPU: the reference's host loop (q_lskum /
fpi_solver, q_lskum_mod.h:13-38, fpi_solver_mod.h:14-150) run stage by stage
through the C ABI by compiled C++ programs, against the reference's own runs.

  - c-_serial_amd/lskum/lskum_fpi: the library's program over
    include/lskum_fpi.hpp (partGrid in, the reference's per-iteration line
    out);
  - oracle/_ref/fpi_dropin: the reference's own main() -- its partGrid
    reader, normals, sub-stencils and initial conditions -- with q_lskum()
    replaced by include/lskum_fpi_reference.hpp (built only where the
    reference sources are present; skipped elsewhere);
  - restart: the reference's restart_sol() read a file the library wrote and
    ran 10 more iterations; the library resumes from the same file.

Bars as in test_gpu_parity: 1e-10 normwise for states, residual histories and
forces; the printed residue (13 digits) to 1e-9 absolute.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import lskum
from lskum.solver import Solver, restart_write
from conftest import REPO, golden, normwise

pytestmark = pytest.mark.gpu

ITER_TOL = 1e-10
FPI = os.path.join(REPO, "c-_serial_amd", "lskum", "lskum_fpi")
DROPIN = os.path.join(REPO, "oracle", "_ref", "fpi_dropin")
LINE = re.compile(r"^iterations (\d+) residue : (-?\d\.\d{13}e[+-]\d\d|0\.0{13}e\+00)$")


def _parse(lines):
    its, res = [], []
    for ln in lines:
        m = LINE.match(ln)
        assert m, f"not the reference's line format: {ln!r}"
        its.append(int(m.group(1)))
        res.append(float(m.group(2)))
    return np.array(its), np.array(res)


def _check_lines(got_text, ref_lines):
    got = [ln for ln in got_text.splitlines() if ln.startswith("iterations ")]
    gi, gr = _parse(got)
    ri, rr = _parse(list(ref_lines))
    np.testing.assert_array_equal(gi, ri)
    np.testing.assert_allclose(gr, rr, rtol=0, atol=1e-9)
    return got


def _run(cmd, cwd):
    print("case:", " ".join(os.path.basename(c) if i == 0 else c for i, c in enumerate(cmd)),
          flush=True)
    proc = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stderr
    return proc.stdout


def test_per_stage_program_matches_reference_small(tmp_path):
    g = golden("small_m063.npz")
    cloud = lskum.workload("naca0012_1480_k20")
    grid = tmp_path / "grid.txt"
    cloud.write_partgrid(grid)
    out = _run([FPI, str(grid), "--iters", "40", "--mach", "0.63", "--aoa", "2.0",
                "--cfl", "0.2", "--vl", "5.0", "--inner", "0", "--out", str(tmp_path / "run")],
               tmp_path)
    got = _check_lines(out, g["printed"])
    # the iterations at which the reference latches res_old print 0 exactly
    assert got[0] == g["printed"][0] == "iterations 1 residue : 0.0000000000000e+00"
    prim = np.fromfile(tmp_path / "run.prim.bin").reshape(4, cloud.n)
    fres = np.fromfile(tmp_path / "run.flux_res.bin").reshape(4, cloud.n)
    forces = np.fromfile(tmp_path / "run.forces.bin")
    e = normwise(prim, g["prim"]), normwise(fres, g["flux_res"])
    print(f"case: lskum_fpi 40 iterations: prim {e[0]:.2e} flux_res {e[1]:.2e}", flush=True)
    assert e[0] <= ITER_TOL and e[1] <= ITER_TOL
    np.testing.assert_allclose(forces, g["history"][-1, 2:], rtol=ITER_TOL, atol=1e-13)


def test_per_stage_loop_equals_lskum_iterate(tmp_path):
    # the stage-by-stage loop and lskum_iterate run the same kernels; only
    # q_variables is a separate launch in one and fused into the state
    # update in the other, with the same arithmetic
    g = golden("k24_m085.npz")
    cloud = lskum.workload("naca0012_3620_k24")
    grid = tmp_path / "grid.txt"
    cloud.write_partgrid(grid)
    _run([FPI, str(grid), "--iters", "20", "--mach", "0.85", "--aoa", "1.25", "--cfl", "0.2",
          "--vl", "5.0", "--inner", "2", "--out", str(tmp_path / "run")], tmp_path)
    prim = np.fromfile(tmp_path / "run.prim.bin").reshape(4, cloud.n)
    s = Solver(cloud, mach=0.85, aoa=1.25, cfl=0.2, vl_const=5.0, inner_iterations=2)
    s.q_lskum(20)
    np.testing.assert_array_equal(prim, s.prim)
    assert normwise(prim, g["prim"]) <= ITER_TOL


def test_reference_main_with_dropin_matches_reference(tmp_path):
    if not os.path.isfile(DROPIN):
        pytest.skip("oracle/_ref/fpi_dropin is built only where the reference sources are")
    g = golden("small_m063.npz")
    cloud = lskum.workload("naca0012_1480_k20")
    grid = tmp_path / "grid.txt"
    cloud.write_partgrid(grid)
    out = _run([DROPIN, str(grid), str(tmp_path), "--iters", "40", "--cfl", "0.2",
                "--vl", "5.0", "--inner", "0"], tmp_path)
    _check_lines(out, g["printed"])
    prim = np.fromfile(tmp_path / "prim.bin").reshape(4, cloud.n)
    forces = np.fromfile(tmp_path / "forces.bin")
    e = normwise(prim, g["prim"])
    print(f"case: reference main + drop-in, 40 iterations: prim {e:.2e}", flush=True)
    assert e <= ITER_TOL
    np.testing.assert_allclose(forces, g["history"][-1, 2:], rtol=ITER_TOL, atol=1e-13)


def _restart_file(tmp_path, g, cloud):
    path = tmp_path / "restart.dat"
    restart_write(path, cloud, int(g["it"]), float(g["res_old"]), g["prim10"])
    return path


def test_restart_file_resumes_as_the_reference_does(tmp_path):
    g = golden("restart_k24_m085.npz")
    cloud = lskum.workload("naca0012_3620_k24")
    path = _restart_file(tmp_path, g, cloud)
    s = Solver(cloud, mach=0.85, aoa=1.25, cfl=0.2, vl_const=5.0, inner_iterations=2)
    s.read_restart(path)
    assert s.iteration() == (10, float(g["res_old"]))
    rn, rs = s.q_lskum(10)
    h = g["history"]
    assert normwise(rn, h[:, 0]) <= ITER_TOL
    np.testing.assert_allclose(rs, h[:, 1], rtol=0, atol=1e-9)
    assert s.iteration()[0] == 20
    e = normwise(s.prim, g["prim"]), normwise(s.flux_res, g["flux_res"])
    print(f"case: restart resume: prim {e[0]:.2e} flux_res {e[1]:.2e}", flush=True)
    assert e[0] <= ITER_TOL and e[1] <= ITER_TOL


def test_restart_through_the_per_stage_program(tmp_path):
    g = golden("restart_k24_m085.npz")
    cloud = lskum.workload("naca0012_3620_k24")
    path = _restart_file(tmp_path, g, cloud)
    grid = tmp_path / "grid.txt"
    cloud.write_partgrid(grid)
    out = _run([FPI, str(grid), "--iters", "10", "--mach", "0.85", "--aoa", "1.25",
                "--cfl", "0.2", "--vl", "5.0", "--inner", "2", "--restart", str(path),
                "--write-restart", str(tmp_path / "after.dat"), "--out",
                str(tmp_path / "run")], tmp_path)
    got = _check_lines(out, g["printed"])
    assert got[0].startswith("iterations 11 ")
    prim = np.fromfile(tmp_path / "run.prim.bin").reshape(4, cloud.n)
    assert normwise(prim, g["prim"]) <= ITER_TOL
    # the file written after the run carries iteration 20 and the latched res_old
    with open(tmp_path / "after.dat") as f:
        head = f.readline().split()
    assert int(head[0]) == cloud.n and int(head[1]) == 20
    assert float(head[2]) == float(g["res_old"])
