"""lskum_iterate_noalloc (the entry point that allocates nothing) through the
C ABI, against lskum_iterate on a context of its own: the same kernels in the
same order, so residual histories and states must be equal bit for bit."""
import numpy as np
import pytest

import lskum
from lskum import _lib as L
from lskum.solver import Solver

pytestmark = pytest.mark.gpu

HIST_MAX = 131072  # LSKUM_NOALLOC_HIST_MAX


def _noalloc(s, n_iter):
    rn = np.zeros(n_iter)
    rs = np.zeros(n_iter)
    rc = L.lib().lskum_iterate_noalloc(s._ctx, n_iter, L.dptr(rn), L.dptr(rs))
    L.check(rc, s._ctx)
    return rn, rs


def _pair(name, params, n_iter):
    print(f"case: noalloc vs iterate, {name}, {n_iter} iterations", flush=True)
    cloud = lskum.workload(name)
    a = Solver(cloud, **params)
    b = Solver(cloud, **params)
    ra = a.q_lskum(n_iter)
    rb = _noalloc(b, n_iter)
    np.testing.assert_array_equal(ra[0], rb[0])
    np.testing.assert_array_equal(ra[1], rb[1])
    np.testing.assert_array_equal(a.prim, b.prim)
    assert a.iteration() == b.iteration()
    return a, b


def test_noalloc_equals_iterate_on_two_calls_in_a_row():
    # first a 3,620-point cloud with inner passes, then a smaller cloud with
    # fewer iterations: the second call reuses the device's scratch array,
    # which still holds the first call's sums
    a, b = _pair("naca0012_3620_k24",
                 dict(mach=0.85, aoa=1.25, cfl=0.2, vl_const=5.0, inner_iterations=2), 12)
    _pair("naca0012_1480_k20",
          dict(mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, inner_iterations=0), 5)
    # the first contexts go on: mixing the two entry points keeps a context's
    # history equal to one that only used lskum_iterate
    ra = a.q_lskum(4)
    rb = _noalloc(b, 4)
    np.testing.assert_array_equal(ra[0], rb[0])
    np.testing.assert_array_equal(ra[1], rb[1])
    rb2 = b.q_lskum(3)
    ra2 = a.q_lskum(3)
    np.testing.assert_array_equal(ra2[0], rb2[0])
    np.testing.assert_array_equal(a.prim, b.prim)


def test_noalloc_rejects_what_it_cannot_hold():
    cloud = lskum.workload("naca0012_1480_k20")
    s = Solver(cloud, mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, inner_iterations=0)
    before = s.prim
    it0 = s.iteration()
    big = HIST_MAX + 1
    rn = np.zeros(big)
    rc = L.lib().lskum_iterate_noalloc(s._ctx, big, L.dptr(rn), None)
    assert rc == L.LSKUM_EINVAL
    assert b"exceeds the scratch array" in L.lib().lskum_last_error(s._ctx)
    rc = L.lib().lskum_iterate_noalloc(s._ctx, 3, None, None)
    assert rc == L.LSKUM_EINVAL
    s.set_timing(True)
    rc = L.lib().lskum_iterate_noalloc(s._ctx, 3, L.dptr(np.zeros(3)), None)
    assert rc == L.LSKUM_EINVAL
    s.set_timing(False)
    # nothing ran
    assert s.iteration() == it0
    np.testing.assert_array_equal(s.prim, before)
    # after the refusals the context still runs
    rn, rs = _noalloc(s, 2)
    assert np.isfinite(rn).all()
