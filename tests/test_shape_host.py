"""Hicks-Henne design fields and their inverse-distance propagation
(lskum_hicks_henne_fields, host code), pinned by what the definitions fix:
  - a bump is 1 at its own peak and 0 at the chord's ends and beyond;
  - upper-surface variables move the upper wall points by +b, lower ones the
    lower points by -b; the trailing- and leading-edge points do not move;
  - an interior point next to a wall point takes that point's displacement;
    points farther than the decay radius from the wall stay put;
  - on a small synthetic cloud the propagation equals a brute-force sum of
    |p - w|^-3 weights with the (1 - d/R)^2 fade.
"""
import ctypes as C

import numpy as np
import pytest

import lskum
from lskum import _lib as L
from lskum import shape
from lskum.cloud import Cloud


def _synthetic(wall_x, wall_y, pts):
    """A cloud with the given wall points first (flag_1 0) and interior
    points after; neighbour lists are irrelevant to the design fields."""
    nw = len(wall_x)
    x = np.concatenate([wall_x, pts[:, 0]])
    y = np.concatenate([wall_y, pts[:, 1]])
    n = len(x)
    flag = np.concatenate([np.zeros(nw, np.int32), np.ones(n - nw, np.int32)])
    left = np.full(n, -1, np.int32)
    right = np.full(n, -1, np.int32)
    left[:nw] = (np.arange(nw) - 1) % nw
    right[:nw] = (np.arange(nw) + 1) % nw
    return Cloud(x, y, left, right, flag, (flag == 0).astype(np.int32), np.ones(n),
                 np.zeros(n + 1, np.int32), np.zeros(0, np.int32)).normalized()


def _wall(nw):
    """Clockwise from the trailing edge, lower surface first, as the
    generator orders it: x = (1 + cos t)/2, y = -/+ 0.06 sin t."""
    t = 2 * np.pi * np.arange(nw) / nw
    return 0.5 * (1 + np.cos(t)), np.where(np.sin(t) >= 0, -0.06, 0.06) * np.abs(np.sin(t))


def test_bump_peaks_and_ends():
    nps = 5
    peaks = np.linspace(0.05, 0.95, nps)
    # wall points placed exactly at the peaks (and at the ends) on both sides
    wx = np.concatenate([[1.0], peaks[::-1], [0.0], peaks])
    nw = len(wx)
    wy = np.zeros(nw)
    c = _synthetic(wx, wy, np.array([[0.5, 3.0]]))
    f = shape.hicks_henne_fields(c, n_per_side=nps, p_lo=0.05, p_hi=0.95, width_exponent=3,
                                 decay_radius=0.5)
    assert f.shape == (2 * nps, c.n)
    lower = np.arange(1, 1 + nps)[::-1]  # lower wall point at peaks[j] is lower[j]
    upper = np.arange(nw - nps, nw)
    for j in range(nps):
        assert f[j, upper[j]] == pytest.approx(1.0, abs=1e-14)
        assert f[nps + j, lower[j]] == pytest.approx(-1.0, abs=1e-14)
        # an upper variable does not move the lower surface, and vice versa
        assert np.all(f[j, 1:1 + nps] == 0.0) and np.all(f[nps + j, nw - nps:nw] == 0.0)
    # trailing edge (x = 1) and leading edge (x = 0) points never move
    assert np.all(f[:, 0] == 0.0) and np.all(f[:, nps + 1] == 0.0)
    # a bump is below 1 away from its peak
    assert np.all(np.abs(f[:nps, upper]) <= 1.0 + 1e-15)
    # the interior point is 2.94 chords from the wall: beyond the decay radius
    assert np.all(f[:, nw] == 0.0)


def test_propagation_against_brute_force():
    wx, wy = _wall(40)
    rng = np.random.default_rng(2)
    pts = np.column_stack([rng.uniform(-0.4, 1.4, 60), rng.uniform(-0.6, 0.6, 60)])
    pts[0] = [wx[7] + 1e-9, wy[7]]  # next to a lower wall point
    c = _synthetic(wx, wy, pts)
    nps, R = 4, 0.5
    f = shape.hicks_henne_fields(c, n_per_side=nps, p_lo=0.1, p_hi=0.9, width_exponent=2,
                                 decay_radius=R)
    nw = len(wx)
    wall = f[:, :nw]
    # the point next to wall point 7 carries its displacement
    assert np.allclose(f[:, nw], wall[:, 7], rtol=0, atol=1e-8)
    for i in range(1, len(pts)):
        d = np.hypot(pts[i, 0] - wx, pts[i, 1] - wy)
        fade = max(0.0, 1 - d.min() / R)
        expect = (wall @ d ** -3) / np.sum(d ** -3) * fade ** 2
        assert np.allclose(f[:, nw + i], expect, rtol=1e-12, atol=1e-15)


def test_generated_cloud_fields():
    c = lskum.workload("naca0012_1480_k20")
    f = shape.hicks_henne_fields(c)
    nw = int(np.sum(c.flag_1 == 0))
    assert f.shape == (20, c.n)
    # upper variables move y up on the upper surface only; far field fixed
    upper = np.arange(nw // 2 + 1, nw)
    assert np.all(f[:10][:, upper] >= 0) and np.all(f[10:][:, upper] == 0)
    assert np.all(f[:, c.flag_1 == 2] == 0.0)
    r = np.hypot(c.x - 0.5, c.y)
    assert np.all(f[:, r > 2.0] == 0.0)


def test_fields_reject_bad_arguments():
    c = lskum.workload("naca0012_1480_k20")
    view = c.view()
    out = np.empty((2, c.n))
    for args in [(0, 0.05, 0.95, 3, 0.5), (1, 0.0, 0.95, 3, 0.5), (1, 0.05, 1.0, 3, 0.5),
                 (1, 0.6, 0.4, 3, 0.5), (1, 0.05, 0.95, 0, 0.5), (1, 0.05, 0.95, 3, 0.0)]:
        assert L.lib().lskum_hicks_henne_fields(C.byref(view), *args, L.dptr(out)) == -1
