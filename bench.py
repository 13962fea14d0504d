#!/usr/bin/env python3
"""Benchmark: q-LSKUM iterations on one MI355X, FP64 and mixed precision.

A step is one fpi_solver iteration (snapshot, local time step, four SSPRK
stages of q-variables, q-derivatives, 2 inner passes, flux residual and state
update) over the 2,502,000-point NACA0012 cloud (wall 2,000, 500 x 5,000
O-grid, k = 24), M 0.85, AoA 1.25 deg, CFL 0.2, K 5.  Inputs are
resident in HBM before the timed region.  `value` is the FP64 form; the
`mixed` object times the mixed-precision form on the same cloud from the
same start and states its Cl/Cd/residual differences from the FP64 run.
Prints one JSON line.

  python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-cpu-baseline]
                  [--no-mixed] [--workload batch|shape|mixed10m]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "c-_serial_amd"))
sys.path.insert(0, REPO)

import lskum  # noqa: E402
from lskum import _lib as L  # noqa: E402
from lskum.solver import Solver  # noqa: E402

METRIC = "point-iterations/sec + HBM GB/s, FP64 vs mixed at 2.5M pts; adjoint grads/sec"
HBM_PEAK_GBPS = 8000.0
CASE = dict(mach=0.85, aoa=1.25, cfl=0.2, vl_const=5.0, power=0.0, inner_iterations=2)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_point_iteration(k, n_inner):
    """Compulsory HBM bytes of one iteration per point (FP64 SoA, int32
    indices, each array once per kernel, neighbour gathers counted once):
    snapshot 64, func_delta 60 + 4k, and per stage q_variables 64,
    q_derivatives 180 + 4k, n_inner x (180 + 4k), flux 260 + 8k, state 148."""
    return 64 + (60 + 4 * k) + 4 * (64 + (180 + 4 * k) + n_inner * (180 + 4 * k) + (260 + 8 * k) + 148)


def flux_bytes_per_point(k, mixed=False):
    """Compulsory bytes of one flux-residual launch per point: q, dq, qm, x,
    y, nx, ny, min_dist-derived eps, delta, flag in; flux_res out; plus the
    2k sub-stencil index entries.  The mixed form stores q, dq, qm (80 of
    the 160 bytes) and eps (4 of 8) in FP32."""
    return (176 if mixed else 260) + 8 * k


def run_form(cloud, precision, args, k_mean):
    """Warmup, then args.steps timed iterations of one form; returns the
    measurements (the solver is closed)."""
    reorder = {"hilbert": L.REORDER_HILBERT, "rcm": L.REORDER_RCM,
               "none": L.REORDER_NONE}[args.reorder]
    t = time.perf_counter()
    solver = Solver(cloud, reorder=reorder, lds_tiles=args.lds_tiles, tile_block=args.tile_block,
                    points_per_thread=args.points_per_thread, precision=precision, **CASE)
    tiles, tile_max, _ = solver.layout()
    log(f"{'mixed' if precision else 'fp64'} context created in "
        f"{time.perf_counter() - t:.1f}s; LDS tile mask {tiles}, largest tile {tile_max} points")
    if args.warmup:
        solver.q_lskum(args.warmup)
    solver.set_timing(True)
    t0 = time.perf_counter()
    rn, rs = solver.q_lskum(args.steps)  # synchronises its stream at the end
    elapsed = time.perf_counter() - t0
    flux_ms, flux_n = solver.kernel_time(L.K_FLUX)
    kt = {name: solver.kernel_time(kid)[0] for name, kid in
          [("flux", L.K_FLUX), ("q_derivatives", L.K_QDERIV), ("q_inner", L.K_QINNER),
           ("state_update", L.K_STATE), ("func_delta", L.K_DELTA), ("q_variables", L.K_QVAR)]}
    solver.set_timing(False)
    cl, cd, cm = solver.compute_cl_cd_cm()
    solver.close()
    if not np.all(np.isfinite(rn)):
        raise SystemExit("non-finite residual during the benchmark")
    n = cloud.n
    flux_bytes = n * flux_bytes_per_point(k_mean, mixed=bool(precision))
    return {"elapsed": elapsed, "value": n * args.steps / elapsed,
            "ms_per_step": 1e3 * elapsed / args.steps, "flux_ms": flux_ms, "flux_n": flux_n,
            "flux_bytes": flux_bytes, "kernel_avg_ms": kt, "tiles": tiles,
            "tile_max": tile_max, "res_new": rn, "cl": cl, "cd": cd, "cm": cm}


def cpu_baseline(args):
    """The reference's own serial fpi_solver (oracle/_ref/ref_driver, built
    from the reference sources) on a bounded sample of the workload: a
    100,400-point cloud of the same generator family (wall 400, 100 x 1,000
    O-grid, k = 24) with the same numerics, 8 iterations (about 20 s), 1 core."""
    sample = dict(naca="0012", n_wall=400, n_radial=100, n_circ=1000, stretch=0.0, k=24, seed=2)
    cloud = lskum.Cloud.generate(**sample)
    iters = 8
    driver = os.path.join(REPO, "oracle", "_ref", "ref_driver")
    desc = (f"{cloud.n}-point NACA0012 cloud of the same generator family (wall 400, "
            f"100 x 1,000, k 24), same numerics, {iters} iterations, 1 thread")
    with tempfile.TemporaryDirectory() as d:
        if os.path.isfile(driver):
            grid = os.path.join(d, "grid")
            cloud.write_partgrid(grid)
            cmd = [driver, "run", grid, d, "--iters", str(iters), "--cfl", "0.2", "--vl", "5",
                   "--inner", "2", "--mach", "0.85", "--aoa", "1.25", "--dump", "0"]
            subprocess.run(cmd, check=True, cwd=d, stdout=subprocess.DEVNULL,
                           env=dict(os.environ, OMP_NUM_THREADS="1"))
            secs = float(np.fromfile(os.path.join(d, "loop_seconds.bin"))[0])
            kind = "reference"
        else:
            from oracle.oracle import Oracle
            o = Oracle(cloud, mach=0.85, aoa=1.25, cfl=0.2, vl_const=5.0, inner_iterations=2,
                       threads=1)
            t = time.perf_counter()
            o.q_lskum(iters)
            secs = time.perf_counter() - t
            kind = "port"
    return {"value": cloud.n * iters / secs, "unit": "point-iterations/s", "cores": 1,
            "kind": kind, "sample": desc, "seconds": round(secs, 3),
            "size_check": "the reference's rate per point-iteration on one core was the "
                          "same within about 5 % on this sample and on the full 2,502,000-point "
                          "cloud (one iteration), measured side by side on one host",
            "cpu": platform.processor() or platform.machine(), "cpu_model": cpu_model(),
            "host_cpus": os.cpu_count(),
            "build": ref_build_flags() if kind == "reference" else "oracle C restatement"}


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def ref_build_flags():
    """The compiler line the reference's serial solver was built with
    (RFLAGS of oracle/Makefile)."""
    try:
        with open(os.path.join(REPO, "oracle", "Makefile")) as f:
            for line in f:
                if line.startswith("RFLAGS"):
                    return "g++ " + line.split("=", 1)[1].strip()
    except OSError:
        pass
    return None


def load_traffic(name, key="traffic_bytes_per_launch"):
    """A figure per flux launch measured by rocprofv3 --pmc and committed
    under profiles/: HBM bytes (FETCH_SIZE and WRITE_SIZE runs, copy-kernel
    calibrated) or VALU wave-instructions (SQ_INSTS_VALU)."""
    p = os.path.join(REPO, "profiles", name)
    if os.path.isfile(p):
        with open(p) as f:
            return json.load(f).get(key)
    return None


# MI355X: 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock; a wave64 VALU
# instruction occupies a 16-lane SIMD for 4 cycles
SIMDS, CLOCK_HZ, VALU_CYCLES = 1024, 2.4e9, 4


def valu_bound(name, avg_ms):
    """The flux's VALU issue time at its counted wave-instructions against its
    measured average duration (the bound the FP64 flux actually sits at)."""
    insts = load_traffic(name, "valu_wave_instructions_per_launch")
    if not insts:
        return None
    issue_ms = insts * VALU_CYCLES / SIMDS / CLOCK_HZ * 1e3
    return {"wave_instructions": insts, "issue_ms": issue_ms, "frac": issue_ms / avg_ms}


def run_batch(args):
    """The many-case workload: 64 flow conditions over 8 NACA 4-digit clouds
    of 50k-800k points (k = 20, inner passes 0, CFL 0.2, K 5), all cases
    advanced together as one captured graph of all --steps iterations,
    replayed once (--graph-iters K: one replay per K iterations), with the
    block shape, flux tiles and state-update points per thread chosen by
    lskum_autotune on one case per cloud.  value = sum of case points x
    iterations / s.  Two cases are
    rerun alone and must match their batch results bit for bit; cases whose
    residual turned non-finite are listed (they run on as NaN without
    touching the others).  With args.graph_sweep, also the time per
    iteration with host launches (0) and with graphs of 10, 100 and 1,000
    iterations, after the run."""
    from lskum.batch import Batch, autotune, batch_workload
    num = dict(cfl=0.2, vl_const=5.0, power=0.0, inner_iterations=0)
    t = time.perf_counter()
    clouds, cases = batch_workload()
    log(f"batch: {len(clouds)} clouds {[c.n for c in clouds]}, {len(cases)} cases, "
        f"{time.perf_counter() - t:.1f}s")
    reps = [next(j for j, c in enumerate(cases) if c[0] == ci) for ci in range(len(clouds))]
    best, trials = autotune(clouds, [cases[j] for j in reps], iters=5, **num)
    log(f"autotune: best {best}")
    graph_iters = args.graph_iters or args.steps
    b = Batch(clouds, cases, tile_block=best["tile_block"], lds_tiles=best["lds_tiles"],
              points_per_thread=best["points_per_thread"], graph_iters=graph_iters, **num)
    if args.warmup:
        b.q_lskum(args.warmup)
    t0 = time.perf_counter()
    rn, _ = b.q_lskum(args.steps)
    elapsed = time.perf_counter() - t0
    forces = b.compute_cl_cd_cm()
    prims = b.prim()
    failed = {int(j): int(it) for j, it in enumerate(b.failed_at()) if it}
    pts = int(b.points.sum())
    sweep = {}
    if args.graph_sweep:
        # each length timed twice, in the order 1000, 100, 10, 0, 0, 10, 100,
        # 1000, so that a drift of the clock over the sweep weighs on every
        # length alike; the mean of the two is reported
        lengths = (1000, 100, 10, 0)
        for K in lengths + lengths[::-1]:
            iters = max(K, 200)
            b.set_graph(K)
            b.q_lskum(K if K else 10)  # capture and first replay, untimed
            t = time.perf_counter()
            b.q_lskum(iters)
            ms = 1e3 * (time.perf_counter() - t) / iters
            sweep.setdefault(str(K), {"iterations": iters, "ms_per_iter_runs": []})
            sweep[str(K)]["ms_per_iter_runs"].append(ms)
            log(f"batch: graphs of {K} iterations: {ms:.3f} ms per iteration")
        for v in sweep.values():
            v["ms_per_iter"] = float(np.mean(v["ms_per_iter_runs"]))
    b.close()
    ok = np.isfinite(forces).all(axis=1)
    same = {}
    for j in (0, len(cases) - 1):
        ci, mach, aoa = cases[j]
        s = Solver(clouds[ci], mach=mach, aoa=aoa, tile_block=best["tile_block"],
                   lds_tiles=best["lds_tiles"], points_per_thread=best["points_per_thread"],
                   **num)
        s.q_lskum(args.warmup)
        srn, _ = s.q_lskum(args.steps)
        same[j] = bool(np.array_equal(srn, rn[j]) and
                       np.array_equal(np.array(s.compute_cl_cd_cm()), forces[j]) and
                       np.array_equal(s.prim, prims[j]))
        s.close()
    return {
        "metric": METRIC, "value": pts * args.steps / elapsed, "unit": "point-iterations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batch_64cases_8clouds_50k-800k_k20", "cases": len(cases),
                   "clouds": [c.n for c in clouds], "case_points": pts,
                   "graph_iters": graph_iters,
                   "tile_block": best["tile_block"], "lds_tile_mask": best["lds_tiles"],
                   "points_per_thread": best["points_per_thread"]},
        "autotune": {"best": best, "trials": trials},
        "cases_bit_identical_to_single_runs": same,
        "graph_length_sweep": sweep,
        "cases_failed": failed,
        "cl_cd": {"min_cl": float(forces[ok, 0].min()), "max_cl": float(forces[ok, 0].max()),
                  "min_cd": float(forces[ok, 1].min()), "max_cd": float(forces[ok, 1].max())},
    }


def run_mixed10m(args):
    """The 10,004,000-point NACA0012 cloud (wall 4,000, 1,000 x 10,000
    O-grid, k = 24) at M 0.63, AoA 2 deg, CFL 0.2, K 5, no inner passes,
    `--steps` iterations from free stream in FP64 and then in the mixed form
    from the same start.  value = the mixed form's point-iterations/s; the
    line states its speed-up and its Cl, Cd, Cm and residual-history
    differences from the FP64 run, which is what the mixed form is checked
    against at this size."""
    t = time.perf_counter()
    cloud = lskum.workload("naca0012_10m_k24")
    log(f"mixed10m: cloud {cloud.n} points, {time.perf_counter() - t:.1f}s")
    runs = {}
    for name, precision in (("fp64", L.FP64), ("mixed", L.MIXED)):
        s = Solver(cloud, mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, power=0.0,
                   inner_iterations=0, precision=precision)
        s.set_graph(50)
        rn_all, elapsed, done = [], 0.0, 0
        while done < args.steps:
            step = min(250, args.steps - done)
            t0 = time.perf_counter()
            rn, _ = s.q_lskum(step)
            elapsed += time.perf_counter() - t0
            done += step
            rn_all.append(rn)
            log(f"mixed10m {name}: {done}/{args.steps} res_new {rn[-1]:.4e} "
                f"{1e3 * elapsed / done:.2f} ms/iteration")
        runs[name] = {"res_new": np.concatenate(rn_all), "elapsed": elapsed,
                      "forces": s.compute_cl_cd_cm()}
        s.close()
    f, m = runs["fp64"], runs["mixed"]
    n = cloud.n
    return {
        "metric": METRIC, "value": n * args.steps / m["elapsed"], "unit": "point-iterations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": 1e3 * m["elapsed"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 gathers/fluxes/limiter/least-squares sums, f64 coordinates/least-squares geometry/state/reductions", "data": "synthetic",
        "config": {"workload": "naca0012_10m_k24", "points": n, "k": 24, "inner_iterations": 0,
                   "mach": 0.63, "aoa_deg": 2.0, "cfl": 0.2, "vl_const": 5.0,
                   "graph_iters": 50},
        "fp64": {"value": n * args.steps / f["elapsed"],
                 "ms_per_step": 1e3 * f["elapsed"] / args.steps,
                 "cl_cd_cm": list(f["forces"]), "final_res_new": float(f["res_new"][-1])},
        "speedup_vs_fp64": f["elapsed"] / m["elapsed"],
        "error_vs_fp64": {"abs_dCl": abs(m["forces"][0] - f["forces"][0]),
                          "abs_dCd": abs(m["forces"][1] - f["forces"][1]),
                          "abs_dCm": abs(m["forces"][2] - f["forces"][2]),
                          "res_new_normwise": float(np.abs(m["res_new"] - f["res_new"]).max()
                                                    / np.abs(f["res_new"]).max())},
    }


# CFL 1: the iteration reaches its plateau (residue oscillating between
# about -1.8 and -2.5, Cl 0.31-0.34) within 30,000 iterations, where CFL
# 0.2 is still drifting after 120,000
SHAPE_FLOW = dict(mach=0.8, aoa=1.25, cfl=1.0, vl_const=150.0, power=0.0, inner_iterations=0)
FD_VARS = (1, 3, 5, 7, 9, 10, 12, 14, 16, 18)  # five bumps on each surface


def shape_measurement(args, n_fd=len(FD_VARS)):
    """Shape gradients (the adjoint workload): the 1,000,940-point NACA0012
    cloud naca0012_1m_k24 (wall 1,500 with blended spacing, 806 x 1,240
    O-grid, k = 24) at M 0.8, AoA 1.25 deg, the reference's limiter constant
    K = 150, no inner passes, CFL 1, advanced args.converge iterations from
    free stream as captured graphs (the state to linearise about); 20
    Hicks-Henne bumps (10 per surface, peaks over x/c 0.05..0.95, width
    exponent 3, inverse-distance propagation within 0.5 chord of the wall).
    Timed: one adjoint sweep over args.adjoint_iters iterations per
    objective (Cl, Cd), each giving all 20 derivatives; value = gradients
    (objectives x design variables) per second of sweep time.  Checked:
    central and one-sided differences (h = 1e-8) of the plain iterations
    along n_fd design fields, five on each surface, for both objectives from
    the same runs (and central differences with h = 1e-6, for comparison).
    A limiter or extremum switch within h of the state is a
    kink of J: the one-sided difference on the side free of it is the slope
    the adjoint takes, the central one is not; the one-sided spread says
    where that happened.  Also the tangent-linear run along one field per
    surface."""
    from lskum import shape
    t = time.perf_counter()
    cloud = lskum.workload("naca0012_1m_k24")
    log(f"shape: cloud {cloud.n} points, {time.perf_counter() - t:.1f}s")
    s = Solver(cloud, **SHAPE_FLOW)
    if args.shape_graph:
        s.set_graph(args.shape_graph)
    t = time.perf_counter()
    done, forces = 0, []
    while done < args.converge:
        step = min(1000, args.converge - done)
        rn, rs = s.q_lskum(step)
        done += step
        forces.append(s.compute_cl_cd_cm())
        log(f"shape: converge {done}/{args.converge}: res_new {rn[-1]:.3e} residue "
            f"{rs[-1]:.2f} cl {forces[-1][0]:.8f} cd {forces[-1][1]:.8f}")
    conv_s = time.perf_counter() - t
    fields = shape.hicks_henne_fields(cloud, n_per_side=10, p_lo=0.05, p_hi=0.95,
                                      width_exponent=3, decay_radius=0.5)
    K = args.adjoint_iters
    objectives = {"cl": (1.0, 0.0, 0.0), "cd": (0.0, 1.0, 0.0)}
    grads, Js, sweep_s = {}, {}, {}
    for name, w in objectives.items():
        t = time.perf_counter()
        Js[name], grads[name] = shape.design_gradient(s, fields, K, w)
        sweep_s[name] = time.perf_counter() - t
        log(f"shape: adjoint sweep {name} over {K} iterations in {sweep_s[name]:.1f}s")
    t = time.perf_counter()
    s.forces_after(K)
    fwd_s = time.perf_counter() - t
    fd, h_fd = {}, 1e-8
    for j in FD_VARS[:n_fd]:
        c, f_, b_ = shape.design_fd_forces(s, fields[j], K, h=h_fd)
        c6 = shape.design_fd_forces(s, fields[j], K, h=1e-6)[0]
        for k, name in enumerate(("cl", "cd")):
            g = float(grads[name][j])
            rel = [abs(float(d[k]) - g) / abs(g) for d in (c, f_, b_)]
            fd[f"{name}_var{j}"] = {"adjoint": g, "fd_central": float(c[k]),
                                    "rel_err": rel[0], "rel_err_best": min(rel),
                                    "rel_err_central_h1e-6": abs(float(c6[k]) - g) / abs(g),
                                    "one_sided_spread_rel": abs(float(f_[k] - b_[k])) /
                                    abs(float(c[k]))}
        log(f"shape: fd var {j}: " + ", ".join(
            f"{n} {fd[f'{n}_var{j}']['rel_err']:.2e} (best side "
            f"{fd[f'{n}_var{j}']['rel_err_best']:.2e})" for n in ("cl", "cd")))
    tan = {}
    for j in (3, 14):
        t = time.perf_counter()
        d = shape.design_tangent(s, fields[j], K, objectives["cl"])
        g = float(grads["cl"][j])
        tan[f"cl_var{j}"] = {"adjoint": g, "tangent": d, "rel_err": abs(d - g) / abs(g),
                             "seconds": time.perf_counter() - t}
    s.close()
    total = sum(sweep_s.values())
    n_grads = len(objectives) * fields.shape[0]
    last = np.array(forces[-2:]) if len(forces) >= 2 else np.zeros((2, 3))
    return {
        "value": n_grads / total, "unit": "gradients/s",
        "workload": "shape_naca0012_1m_k24", "points": cloud.n,
        "objectives": list(objectives), "design_variables": int(fields.shape[0]),
        "adjoint_iterations": K, "adjoint_sweep_s": sweep_s,
        "adjoint_point_iterations_per_s": cloud.n * K * len(objectives) / total,
        "forward_s_same_iterations": fwd_s,
        "adjoint_cost_vs_forward": total / len(objectives) / fwd_s,
        "state": {"iterations": done, "res_new": float(s.res_new), "residue": float(s.residue),
                  "abs_dCl_last_1000": float(abs(last[1][0] - last[0][0])),
                  "abs_dCd_last_1000": float(abs(last[1][1] - last[0][1])),
                  "converge_s": conv_s},
        "J": Js,
        "fd_checks": {"h": h_fd, "count": len(fd),
                      "worst_rel_err": max(v["rel_err"] for v in fd.values()),
                      "worst_rel_err_best": max(v["rel_err_best"] for v in fd.values()),
                      "worst_rel_err_central_h1e-6": max(v["rel_err_central_h1e-6"]
                                                         for v in fd.values()),
                      "median_rel_err": float(np.median([v["rel_err"] for v in fd.values()])),
                      "per_variable": fd},
        "tangent_checks": {"worst_rel_err": max(v["rel_err"] for v in tan.values()),
                           "per_variable": tan},
    }


def run_shape(args):
    """The shape-gradient workload alone, as its own JSON line."""
    r = shape_measurement(args)
    return {
        "metric": METRIC, "value": r["value"], "unit": "gradients/s", "n_gpus": 1,
        "steps": args.adjoint_iters, "warmup": 0,
        "ms_per_step": 1e3 * sum(r["adjoint_sweep_s"].values()) / (2 * args.adjoint_iters),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "shape_naca0012_1m_k24", "points": r["points"], "k": 24,
                   **{k: v for k, v in SHAPE_FLOW.items()},
                   "converge_iterations": args.converge, "adjoint_iterations": args.adjoint_iters,
                   "converge_graph_iters": args.shape_graph,
                   "design_variables": r["design_variables"],
                   "bumps": "10 upper + 10 lower, x/c 0.05..0.95, exponent 3",
                   "propagation": "inverse distance |d|^-3, faded to 0 at 0.5 chord"},
        "adjoint": r,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="naca0012_2p5m_k24")
    ap.add_argument("--reorder", default="hilbert", choices=["hilbert", "rcm", "none"])
    ap.add_argument("--lds-tiles", type=int, default=-1,
                    help="-1 library default, else mask: 1 q_derivatives, 2 q_inner, "
                         "4 func_delta, 8 flux")
    ap.add_argument("--tile-block", type=int, default=0, choices=[0, 64, 128, 256])
    ap.add_argument("--points-per-thread", type=int, default=0, choices=[0, 1, 2, 4],
                    help="state update points per thread (0: library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="leave the shape-gradient (adjoint) measurement out of the line")
    ap.add_argument("--graph-iters", type=int, default=0,
                    help="many-case workload: iterations per captured graph (0: all --steps "
                         "iterations in one graph, one launch)")
    ap.add_argument("--graph-sweep", action="store_true",
                    help="many-case workload: also time graphs of 0, 10, 100, 1000 iterations")
    ap.add_argument("--converge", type=int, default=30000,
                    help="shape workload: iterations before the adjoint")
    ap.add_argument("--adjoint-iters", type=int, default=300,
                    help="shape workload: iterations the adjoint differentiates")
    ap.add_argument("--shape-graph", type=int, default=50,
                    help="shape workload: iterations per captured graph while converging "
                         "(0: host launches, as a kernel trace under rocprofv3 needs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the library runs on one GPU only; the q-LSKUM iteration would shard "
              "by point partition with a one-ring halo, which is not built. Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.workload == "shape":
        print(json.dumps(run_shape(args)), flush=True)
        return
    if args.workload == "mixed10m":
        print(json.dumps(run_mixed10m(args)), flush=True)
        return
    t = time.perf_counter()
    cloud = lskum.workload(args.workload)
    log(f"cloud {args.workload}: {cloud.n} points, {cloud.conn.size} neighbour entries, "
        f"{time.perf_counter() - t:.1f}s")
    k_mean = cloud.conn.size / cloud.n
    f = run_form(cloud, L.FP64, args, k_mean)
    n = cloud.n
    achieved = f["flux_bytes"] / (f["flux_ms"] * 1e-3) / 1e9
    it_bytes = n * bytes_per_point_iteration(k_mean, CASE["inner_iterations"])
    out = {
        "metric": METRIC,
        "value": f["value"],
        "unit": "point-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "points": n, "k": 24, "k_mean": round(k_mean, 3),
                   "inner_iterations": 2, "mach": 0.85, "aoa_deg": 1.25, "cfl": 0.2,
                   "vl_const": 5.0, "precision": "fp64", "order": args.reorder,
                   "lds_tile_mask": f["tiles"], "tile_max": f["tile_max"],
                   "tile_block": args.tile_block or 256,
                   "points_per_thread": args.points_per_thread or 1},
        "roofline": {"bound": "hbm", "kernel": "k_flux", "achieved": achieved,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": load_traffic("flux_naca0012_2p5m_k24_traffic.json"),
                     "avg_ms": f["flux_ms"], "launches": f["flux_n"],
                     "algorithmic_bytes_per_launch": f["flux_bytes"],
                     "valu": valu_bound("flux_naca0012_2p5m_k24_traffic.json", f["flux_ms"])},
        "iteration_hbm_GBps": it_bytes * args.steps / f["elapsed"] / 1e9,
        "kernel_avg_ms": f["kernel_avg_ms"],
        "final": {"res_new": float(f["res_new"][-1]), "cl": f["cl"], "cd": f["cd"],
                  "cm": f["cm"]},
    }
    if not args.no_mixed:
        m = run_form(cloud, L.MIXED, args, k_mean)
        ach_m = m["flux_bytes"] / (m["flux_ms"] * 1e-3) / 1e9
        rn64 = f["res_new"]
        out["mixed"] = {
            "value": m["value"], "ms_per_step": m["ms_per_step"],
            "speedup_vs_fp64": m["value"] / f["value"],
            "dtype": "f32 gathers/fluxes/limiter/least-squares sums, f64 coordinates/least-squares geometry/state/reductions",
            "lds_tile_mask": m["tiles"],
            "flux_roofline": {"achieved": ach_m, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                              "frac": ach_m / HBM_PEAK_GBPS, "avg_ms": m["flux_ms"],
                              "algorithmic_bytes_per_launch": m["flux_bytes"]},
            "kernel_avg_ms": m["kernel_avg_ms"],
            "error_vs_fp64": {"iterations": args.warmup + args.steps,
                              "abs_dCl": abs(m["cl"] - f["cl"]),
                              "abs_dCd": abs(m["cd"] - f["cd"]),
                              "abs_dCm": abs(m["cm"] - f["cm"]),
                              "res_new_normwise": float(np.abs(m["res_new"] - rn64).max()
                                                        / np.abs(rn64).max())},
        }
    if not args.no_adjoint:
        out["adjoint"] = shape_measurement(args)
    if not args.no_cpu_baseline:
        log("cpu baseline: reference serial fpi_solver on a bounded sample ...")
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
