"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in the build container only (it needs the reference sources to build
oracle/_ref/ref_driver):  python tests/golden/make_golden.py [--long]

Every fixture is data: the clouds come from the project generator (their
partGrid SHA-256 is stored so a generator change is caught), the outputs
come from the reference's own functions run by oracle/ref_driver.cpp.
  kat_data.npz       first 1000 rows of the reference's data.dat and, per
                     row, the eight flux functions and qtilde_to_primitive
  small_m063.npz     naca0012_1480_k20, M 0.63, AoA 2, CFL 0.2, K 5, no inner
                     passes, 40 iterations: history, final arrays,
                     stage-by-stage flux_res/prim at iteration 30, and
                     the printed lines of an uninterrupted run
  k24_m085.npz       naca0012_3620_k24, M 0.85, AoA 1.25, 2 inner passes,
                     20 iterations: history, final prim and flux_res
  naca0012_9760_m063_1000it.npz
                     naca0012_9760_k20 with the same numerics, 1000
                     iterations: history, final prim and flux_res and
                     SHA-256 digests of every final array
  diag_small.npz     naca0012_1480_k20, the same numerics, 40 iterations:
                     final prim and the reference's diagnostics of it
                     (entropy, vorticity, its square and total enstrophy;
                     the first four values stagnation_pressure() prints)
  restart_k24_m085.npz
                     naca0012_3620_k24 as k24_m085: the state after 10
                     iterations written by the library's restart writer,
                     and the reference's own restart run from that file
                     (iterations 11..20: history, final prim and flux_res,
                     the printed lines)
  naca0012_2p5m_m085_500it.npz  (--only c2; hours on all host threads)
                     naca0012_2p5m_k24, M 0.85, AoA 1.25, 2 inner passes,
                     500 iterations by the oracle (the reference itself
                     would take about a day on one core): history, forces
                     every 50 iterations, final-array norms and a sample
"""
import argparse
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "c-_serial_amd"))
import lskum  # noqa: E402

REF = os.environ.get("LSKUM_REFERENCE", "/root/reference")
DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_driver")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def grid_sha(cloud, path):
    cloud.write_partgrid(path)
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def run_ref(cloud, iters, cfl, vl, inner, extra=(), stdout=None):
    d = tempfile.mkdtemp(prefix="lskum_ref_")
    g = os.path.join(d, "grid.txt")
    gsha = grid_sha(cloud, g)
    cmd = [DRIVER, "run", g, d, "--iters", str(iters), "--cfl", repr(cfl), "--vl", repr(vl),
           "--inner", str(inner), "--dump", str(iters), *extra]
    proc = subprocess.run(cmd, check=True, cwd=d, capture_output=True, text=True)
    if stdout is not None:
        stdout.append(proc.stdout)
    load = lambda name: np.fromfile(os.path.join(d, name + ".bin"))
    return d, gsha, load


def kat():
    rows = np.loadtxt(os.path.join(REF, "data.dat"), max_rows=1000)
    d = tempfile.mkdtemp(prefix="lskum_kat_")
    out = os.path.join(d, "kat.bin")
    subprocess.run([DRIVER, "kat", os.path.join(REF, "data.dat"), "1000", out], check=True, cwd=d)
    res = np.fromfile(out).reshape(1000, 36)
    np.savez_compressed(os.path.join(HERE, "kat_data.npz"), rows=rows,
                        fluxes=res[:, :32].reshape(1000, 8, 4), qtilde_prim=res[:, 32:])


def printed_lines(text):
    return np.array([ln for ln in text.splitlines() if ln.startswith("iterations ")])


def small_m063():
    cloud = lskum.workload("naca0012_1480_k20")
    n = cloud.n
    d, gsha, load = run_ref(cloud, 40, 0.2, 5.0, 0, ("--stages", "30"))
    stage = {}
    for rk in range(1, 5):
        stage[f"rk{rk}_flux_res"] = load(f"it30_rk{rk}_flux_res").reshape(4, n)
        stage[f"rk{rk}_prim"] = load(f"it30_rk{rk}_prim").reshape(4, n)
        stage[f"rk{rk}_red"] = load(f"it30_rk{rk}_red")
    stage["rk1_prim_in"] = load("it30_rk1_prim_in").reshape(4, n)
    stage["rk1_q"] = load("it30_rk1_q").reshape(4, n)
    stage["rk1_dq"] = load("it30_rk1_dq").reshape(2, 4, n)
    stage["rk1_qm"] = load("it30_rk1_qm").reshape(2, 4, n)
    stage["delta30"] = load("it30_delta")
    # the reference's own per-iteration lines (fpi_solver_mod.h:150) of an
    # uninterrupted run (the stage-by-stage iteration above prints none)
    out = []
    run_ref(cloud, 40, 0.2, 5.0, 0, ("--print", "1"), stdout=out)
    stage["printed"] = printed_lines(out[0])
    np.savez_compressed(
        os.path.join(HERE, "small_m063.npz"), grid_sha=gsha,
        params=np.array([0.63, 2.0, 0.2, 5.0, 0.0, 0]),
        history=load("history").reshape(-1, 5),
        prim=load("it40_prim").reshape(4, n), q=load("it40_q").reshape(4, n),
        dq=load("it40_dq").reshape(2, 4, n), qm=load("it40_qm").reshape(2, 4, n),
        flux_res=load("it40_flux_res").reshape(4, n), delta=load("it40_delta"),
        normals=load("normals").reshape(2, n),
        sub_counts=load("substencil_counts").astype(np.int32).reshape(n, 4),
        sub_members_sha=sha(load("substencil_members").astype(np.int32)),
        **stage)


def k24():
    cloud = lskum.workload("naca0012_3620_k24")
    n = cloud.n
    d, gsha, load = run_ref(cloud, 20, 0.2, 5.0, 2, ("--mach", "0.85", "--aoa", "1.25"))
    np.savez_compressed(
        os.path.join(HERE, "k24_m085.npz"), grid_sha=gsha,
        params=np.array([0.85, 1.25, 0.2, 5.0, 0.0, 2]),
        history=load("history").reshape(-1, 5),
        prim=load("it20_prim").reshape(4, n), flux_res=load("it20_flux_res").reshape(4, n),
        sub_counts_sha=sha(load("substencil_counts").astype(np.int32)),
        sub_members_sha=sha(load("substencil_members").astype(np.int32)))


def restart_k24():
    """naca0012_3620_k24 as k24_m085 (M 0.85, 2 inner passes): the reference
    runs 10 iterations; the library's writer puts that state in a restart
    file; the reference's own restart path (initial_conditions_mod.h:7-21,
    q_lskum_mod.h:29-37) reads it and runs iterations 11..20."""
    from lskum.solver import restart_write
    cloud = lskum.workload("naca0012_3620_k24")
    n = cloud.n
    flow = ("--mach", "0.85", "--aoa", "1.25")
    d, gsha, load = run_ref(cloud, 10, 0.2, 5.0, 2, flow)
    h10 = load("history").reshape(-1, 5)
    prim10 = load("it10_prim").reshape(4, n)
    res_old = h10[1, 0]  # latched at iteration 2 (fpi_solver_mod.h:139-143)
    path = os.path.join(d, "restart_written_by_library.dat")
    restart_write(path, cloud, 10, res_old, prim10)
    out = []
    d2, _, load2 = run_ref(cloud, 10, 0.2, 5.0, 2,
                           flow + ("--restart", path, "--dump", "20", "--print", "1"),
                           stdout=out)
    np.savez_compressed(
        os.path.join(HERE, "restart_k24_m085.npz"), grid_sha=gsha,
        params=np.array([0.85, 1.25, 0.2, 5.0, 0.0, 2]), it=10, res_old=res_old,
        prim10=prim10, history=load2("history").reshape(-1, 5),
        prim=load2("it20_prim").reshape(4, n), flux_res=load2("it20_flux_res").reshape(4, n),
        printed=printed_lines(out[0]))


def long_9760():
    cloud = lskum.workload("naca0012_9760_k20")
    n = cloud.n
    d, gsha, load = run_ref(cloud, 1000, 0.2, 5.0, 0)
    final = {nm: load("it1000_" + nm) for nm in ["prim", "q", "dq", "qm", "flux_res", "delta"]}
    np.savez_compressed(
        os.path.join(HERE, "naca0012_9760_m063_1000it.npz"), grid_sha=gsha,
        params=np.array([0.63, 2.0, 0.2, 5.0, 0.0, 0]),
        history=load("history").reshape(-1, 5), prim=final["prim"].reshape(4, n),
        flux_res=final["flux_res"].reshape(4, n),
        **{nm + "_sha": sha(v) for nm, v in final.items()},
        sub_counts_sha=sha(load("substencil_counts").astype(np.int32)),
        sub_members_sha=sha(load("substencil_members").astype(np.int32)))


C2_PARAMS = dict(mach=0.85, aoa=1.25, cfl=0.2, vl_const=5.0, power=0.0, inner_iterations=2)
C2_ITERS = 500
C2_FORCE_EVERY = 50


def c2_sample(cloud, n_random=4000, seed=2502):
    """Every wall point and n_random others drawn without replacement."""
    wall = np.flatnonzero(cloud.flag_1 == 0)
    rest = np.flatnonzero(cloud.flag_1 != 0)
    pick = np.random.default_rng(seed).choice(rest, size=n_random, replace=False)
    return np.sort(np.concatenate([wall, pick])).astype(np.int32)


def c2_500(threads=8, checkpoint_dir=None):
    """naca0012_2p5m_k24 (2,502,000 points), M 0.85, AoA 1.25, CFL 0.2, K 5,
    2 inner passes, 500 iterations, run by the oracle (the reference at this
    size takes about a day on one core; the oracle reproduces it bit for bit
    on the fixtures above).  Stored: the residual history, Cl/Cd/Cm every 50
    iterations, per component max|.| and the L2 norm of the final prim and
    flux_res, and both arrays at every wall point plus 4,000 others.
    checkpoint_dir keeps the state every 25 iterations, so an interrupted run
    resumes exactly (the state is all an iteration reads: q, the derivatives
    and delta are recomputed from prim)."""
    sys.path.insert(0, REPO)
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from conftest import cloud_sha
    from oracle.oracle import Oracle
    cloud = lskum.workload("naca0012_2p5m_k24")
    n = cloud.n
    o = Oracle(cloud, threads=threads, **C2_PARAMS)
    hist = np.zeros((C2_ITERS, 2))
    forces = []
    done = 0
    ck = os.path.join(checkpoint_dir, "c2_checkpoint.npz") if checkpoint_dir else None
    if ck and os.path.isfile(ck):
        z = np.load(ck)
        done = int(z["it"])
        o.set("prim", z["prim"])
        o.set_iteration(done, float(z["res_old"]), 1)
        hist[:done] = z["hist"][:done]
        forces = [tuple(r) for r in z["forces"]]
        print("resumed at iteration", done, flush=True)
    while done < C2_ITERS:
        step = min(25, C2_ITERS - done)
        rn, rs = o.q_lskum(step)
        hist[done:done + step, 0] = rn
        hist[done:done + step, 1] = rs
        done += step
        if done % C2_FORCE_EVERY == 0:
            o.compute_cl_cd_cm()
            sc = o.scalars()
            forces.append((done, sc["cl"], sc["cd"], sc["cm"]))
        print(f"iteration {done} res_new {hist[done - 1, 0]:.6e}", flush=True)
        if ck:
            tmp = ck + ".tmp.npz"
            np.savez(tmp, it=done, prim=o.get("prim"), res_old=o.scalars()["res_old"],
                     hist=hist, forces=np.array(forces).reshape(-1, 4))
            os.replace(tmp, ck)
    prim = o.get("prim")
    fres = o.get("flux_res")
    sample = c2_sample(cloud)
    np.savez_compressed(
        os.path.join(HERE, "naca0012_2p5m_m085_500it.npz"), cloud_sha=cloud_sha(cloud),
        params=np.array([0.85, 1.25, 0.2, 5.0, 0.0, 2]), history=hist,
        forces=np.array(forces), n_points=n, sample=sample,
        prim_sample=prim[:, sample], flux_res_sample=fres[:, sample],
        prim_absmax=np.abs(prim).max(axis=1), prim_l2=np.sqrt((prim * prim).sum(axis=1)),
        flux_res_absmax=np.abs(fres).max(axis=1),
        flux_res_l2=np.sqrt((fres * fres).sum(axis=1)))


def diag():
    cloud = lskum.workload("naca0012_1480_k20")
    n = cloud.n
    out = []
    d, gsha, load = run_ref(cloud, 40, 0.2, 5.0, 0, ("--diag", "1"), stdout=out)
    line = [ln for ln in out[0].splitlines() if ln.startswith("Stagnation values are")][0]
    stag = np.array([float(v) for v in line.split()[3:7]])
    np.savez_compressed(
        os.path.join(HERE, "diag_small.npz"), grid_sha=gsha,
        params=np.array([0.63, 2.0, 0.2, 5.0, 0.0, 0]),
        prim=load("it40_prim").reshape(4, n), entropy=load("diag_entropy"),
        vorticity=load("diag_vorticity"), vorticity_sqr=load("diag_vorticity_sqr"),
        total_enstrophy=load("diag_total_enstrophy")[0], stagnation_printed=stag)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--long", action="store_true",
                    help="also the 1000-iteration run on the 9,760-point cloud")
    ap.add_argument("--only", default="")
    ap.add_argument("--threads", type=int, default=8, help="oracle threads (c2 job)")
    ap.add_argument("--checkpoint-dir", default=None, help="resumable state (c2 job)")
    a = ap.parse_args()
    if not os.path.isfile(DRIVER) and a.only != "c2":
        sys.exit("build the reference harness first: make -C oracle ref")
    jobs = {"kat": kat, "small": small_m063, "k24": k24, "long": long_9760, "diag": diag,
            "restart": restart_k24,
            "c2": lambda: c2_500(a.threads, a.checkpoint_dir)}
    todo = (a.only.split(",") if a.only else
            ["kat", "small", "k24", "diag", "restart"] + (["long"] if a.long else []))
    for j in todo:
        jobs[j]()
        print("wrote", j, flush=True)
