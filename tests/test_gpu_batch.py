"""Many cases in one context (lskum_create_batch) and captured HIP graphs.

Each case of a batch must reproduce a context of its own bit for bit: the
stencil kernels see only the case's points, every case starts on a
256-point block boundary so its residual partials cover the same points in
the same order, and its force sum runs over its own panels in the same
order.  Graph-replayed iterations must equal host-launched ones bit for bit,
on every replay, with host work between the replays.
"""
import numpy as np
import pytest

import lskum
from lskum import _lib as L
from lskum.batch import Batch, batch_workload
from lskum.solver import Solver

pytestmark = pytest.mark.gpu

NUMERICS = dict(cfl=0.2, vl_const=5.0, power=0.0, inner_iterations=1)


def _small_batch():
    # the many-case workload's recipe at small sizes: 8 airfoils of
    # 2k-16k points, 4 conditions each
    return batch_workload(conditions_per_cloud=4, min_points=2000, max_points=16000)


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_batch_cases_equal_individual_runs(precision):
    clouds, cases = _small_batch()
    print("case: batch of", len(cases), "cases, precision", precision, flush=True)
    b = Batch(clouds, cases, precision=precision, graph_iters=10, **NUMERICS)
    rn, rs = b.q_lskum(25)
    forces = b.compute_cl_cd_cm()
    prims = b.prim()
    b.close()
    for j in (0, 5, 13, 22, 31):
        ci, mach, aoa = cases[j]
        s = Solver(clouds[ci], mach=mach, aoa=aoa, precision=precision, **NUMERICS)
        srn, srs = s.q_lskum(25)
        np.testing.assert_array_equal(rn[j], srn)
        np.testing.assert_array_equal(rs[j], srs)
        np.testing.assert_array_equal(forces[j], np.array(s.compute_cl_cd_cm()))
        np.testing.assert_array_equal(prims[j], s.prim)
        s.close()


def test_graph_replays_equal_host_launches():
    clouds, cases = _small_batch()
    eager = Batch(clouds, cases, **NUMERICS)
    graph = Batch(clouds, cases, graph_iters=4, **NUMERICS)
    for rep in range(3):
        # 9 = two graph launches + one host-launched iteration
        a = eager.q_lskum(9)
        g = graph.q_lskum(9)
        np.testing.assert_array_equal(a[0], g[0])
        np.testing.assert_array_equal(eager.compute_cl_cd_cm(), graph.compute_cl_cd_cm())
        for pa, pg in zip(eager.prim(), graph.prim()):
            np.testing.assert_array_equal(pa, pg)


def test_single_case_graph_equals_host_launches():
    cloud = lskum.workload("naca0012_3620_k24")
    kw = dict(mach=0.85, aoa=1.25, cfl=0.2, vl_const=5.0, inner_iterations=2)
    a = Solver(cloud, **kw)
    g = Solver(cloud, **kw)
    g.set_graph(5)
    for _ in range(2):
        ra = a.q_lskum(12)
        rg = g.q_lskum(12)
        np.testing.assert_array_equal(ra[0], rg[0])
        np.testing.assert_array_equal(a.prim, g.prim)


def test_split_calls_equal_one_call():
    # q and the time-step speed carry over between calls
    cloud = lskum.workload("naca0012_1480_k20")
    kw = dict(mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, inner_iterations=2)
    a = Solver(cloud, **kw)
    b = Solver(cloud, **kw)
    ra = a.q_lskum(20)[0]
    rb = np.concatenate([b.q_lskum(7)[0], b.q_lskum(13)[0]])
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(a.prim, b.prim)


def test_a_failing_case_leaves_the_others_untouched():
    clouds, cases = _small_batch()
    cases = cases[:6]
    b = Batch(clouds, cases, **NUMERICS)
    ref = Batch(clouds, cases, **NUMERICS)
    prims = b.prim()
    p = prims[2].copy()
    p[3, 700] = -1.0  # negative pressure in case 2: its residual turns NaN
    full = np.concatenate([(q if j != 2 else p).ravel() for j, q in enumerate(prims)])
    L.check(L.lib().lskum_set_state(b._ctx, L.dptr(full)), b._ctx)
    rn, _ = b.q_lskum(6)
    rr, _ = ref.q_lskum(6)
    fa = b.failed_at()
    assert fa[2] == 1 and not np.delete(fa, 2).any()
    assert not np.isfinite(rn[2]).all()
    np.testing.assert_array_equal(np.delete(rn, 2, axis=0), np.delete(rr, 2, axis=0))
    f, fr = b.compute_cl_cd_cm(), ref.compute_cl_cd_cm()
    np.testing.assert_array_equal(np.delete(f, 2, axis=0), np.delete(fr, 2, axis=0))


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_points_per_thread_leave_results_unchanged(precision):
    # the state update at 1, 2 and 4 points per thread (256-point blocks in
    # all three): the same per-point arithmetic and the same reduction tree
    # over each block's slots, so the same bits -- on a batch whose cases
    # pad their last blocks, and on a single case whose point count is not
    # a multiple of 256
    clouds, cases = _small_batch()
    cloud = lskum.workload("naca0012_1480_k20")
    print("case: points per thread 1/2/4, precision", precision, flush=True)
    runs = []
    for ppt in (1, 2, 4):
        b = Batch(clouds, cases, precision=precision, points_per_thread=ppt, **NUMERICS)
        rn, rs = b.q_lskum(12)
        s = Solver(cloud, precision=precision, points_per_thread=ppt, **NUMERICS)
        srn, _ = s.q_lskum(12)
        runs.append((rn, rs, b.compute_cl_cd_cm(), b.prim(), srn, s.prim))
        b.close()
        s.close()
    for other in runs[1:]:
        np.testing.assert_array_equal(runs[0][0], other[0])
        np.testing.assert_array_equal(runs[0][1], other[1])
        np.testing.assert_array_equal(runs[0][2], other[2])
        for pa, pb in zip(runs[0][3], other[3]):
            np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(runs[0][4], other[4])
        np.testing.assert_array_equal(runs[0][5], other[5])
