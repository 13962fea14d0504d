"""Host side of the product on the CPU: ABI exports, clouds, partGrid I/O,
set-up (normals, sub-stencils, RCM) and the create-time validation paths.
No GPU call is made here."""
import ctypes as C
import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

import lskum
from lskum import _lib as L
from conftest import REPO, golden
from oracle import oracle as O


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(REPO, "include", "lskum.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char|void \*)\s*\*?\s*(lskum_\w+)\(",
                              header, re.M))
    assert len(declared) >= 30
    handle = C.CDLL(L.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(handle, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (lskum_\w+)$", out, re.M))
    assert declared <= exported
    # and the Python binding covers the whole ABI
    assert declared == set(L.exported_symbols())


def test_generator_is_deterministic_and_seeded():
    a = lskum.workload("naca0012_1480_k20")
    b = lskum.workload("naca0012_1480_k20")
    for f in a.__dataclass_fields__:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    c = lskum.Cloud.generate(**dict(lskum.WORKLOADS["naca0012_1480_k20"], seed=99))
    assert not np.array_equal(a.x, c.x)


def test_generated_cloud_layout():
    c = lskum.workload("naca0012_9760_k20")
    assert c.n == 160 + 64 * 150
    assert (c.flag_1[:160] == 0).all() and (c.flag_2[:160] == 1).all()
    assert (c.flag_1[160:160 + 63 * 150] == 1).all() and (c.flag_1[-150:] == 2).all()
    nb = c.nbhs()
    assert nb.max() == 20 and nb.min() >= 15
    # neighbours sorted by distance, no self, no duplicates, min_dist = nearest
    for i in range(0, c.n, 101):
        k = c.conn[c.conn_off[i]:c.conn_off[i + 1]]
        dx, dy = c.x[k] - c.x[i], c.y[k] - c.y[i]
        d = np.sqrt(dx * dx + dy * dy)
        assert i not in k and len(set(k)) == len(k)
        assert np.all(np.diff(d) >= 0)
        assert c.min_dist[i] == d[0]
    # outer ring on the far-field circle, clockwise
    r = np.hypot(c.x[-150:] - 0.5, c.y[-150:])
    np.testing.assert_allclose(r, 20.0, rtol=1e-12)
    lo, hi, dets = c.stats()
    assert min(dets) > 0.01


def test_no_neighbour_sees_through_the_body():
    c = lskum.workload("naca0012_9760_k20")
    nw = 160
    side = np.where(np.arange(nw) == 0, 0, np.where(np.arange(nw) < nw // 2, -1, 1))
    side[nw // 2] = 0
    for i in range(nw):
        k = c.conn[c.conn_off[i]:c.conn_off[i + 1]]
        kw = k[k < nw]
        assert np.all(side[i] * side[kw] >= 0), i


def test_partgrid_round_trip(tmp_path):
    c = lskum.workload("naca0012_1480_k20")
    p = tmp_path / "grid"
    c.write_partgrid(p)
    d = lskum.Cloud.read_partgrid(p)
    for f in c.__dataclass_fields__:
        assert np.array_equal(getattr(c, f), getattr(d, f)), f
    # 1-based in the file, as main.cpp:30-58 reads it
    first = p.read_text().splitlines()[1].split()
    assert int(first[2]) == c.left[0] + 1 and int(first[7]) == c.nbhs()[0]


def test_partgrid_reader_errors(tmp_path):
    with pytest.raises(L.LskumError) as e:
        lskum.Cloud.read_partgrid(tmp_path / "missing")
    assert e.value.code == -4
    c = lskum.workload("naca0012_1480_k20")
    p = tmp_path / "grid"
    c.write_partgrid(p)
    text = p.read_text()
    (tmp_path / "trunc").write_text(text[: len(text) // 2])
    with pytest.raises(L.LskumError) as e:
        lskum.Cloud.read_partgrid(tmp_path / "trunc")
    assert e.value.code == -4
    lines = text.splitlines()
    f = lines[5].split()
    f[4] = "7"  # flag_1 outside 0..2
    lines[5] = " ".join(f)
    (tmp_path / "badflag").write_text("\n".join(lines) + "\n")
    with pytest.raises(L.LskumError):
        lskum.Cloud.read_partgrid(tmp_path / "badflag")


def _sub_lists(mask, cloud, which):
    bit = 1 << which
    counts = np.zeros(cloud.n, np.int32)
    mem = []
    for i in range(cloud.n):
        s, e = cloud.conn_off[i], cloud.conn_off[i + 1]
        sel = cloud.conn[s:e][(mask[s:e] & bit) != 0]
        counts[i] = len(sel)
        mem.append(sel)
    return counts, np.concatenate(mem)


@pytest.mark.parametrize("name,fixture", [("naca0012_1480_k20", "small_m063.npz"),
                                          ("naca0012_3620_k24", "k24_m085.npz")])
def test_host_substencils_equal_oracle_and_reference(name, fixture):
    cloud = lskum.workload(name)
    nx, ny, mask = cloud.connectivity()
    o = O.Oracle(cloud)
    assert np.array_equal(nx, o.get("nx")) and np.array_equal(ny, o.get("ny"))
    per_point = [[] for _ in range(cloud.n)]
    all_counts = []
    for s in range(4):
        counts, mem = _sub_lists(mask, cloud, s)
        oc, om = o.substencil(s)
        assert np.array_equal(counts, oc) and np.array_equal(mem, om), s
        all_counts.append(counts)
        off = np.concatenate([[0], np.cumsum(counts)])
        for i in range(cloud.n):
            per_point[i].append(mem[off[i]:off[i + 1]])
    g = golden(fixture)
    members = np.concatenate([m for p in per_point for m in p]).astype(np.int32)
    assert hashlib.sha256(members.tobytes()).hexdigest() == str(g["sub_members_sha"])


def test_rcm_is_a_permutation_and_narrows_a_shuffled_cloud():
    c = lskum.workload("naca0012_1480_k20")
    rng = np.random.default_rng(0)
    shuf = rng.permutation(c.n).astype(np.int32)  # new -> old
    inv = np.empty_like(shuf)
    inv[shuf] = np.arange(c.n, dtype=np.int32)
    # shuffled copy of the cloud (left/right and conn relabelled)
    conn_lists = [inv[c.conn[c.conn_off[o]:c.conn_off[o + 1]]] for o in shuf]
    off = np.concatenate([[0], np.cumsum([len(x) for x in conn_lists])]).astype(np.int32)
    relabel = lambda a: np.where(a >= 0, inv[np.maximum(a, 0)], -1).astype(np.int32)
    s = lskum.Cloud(c.x[shuf], c.y[shuf], relabel(c.left[shuf]), relabel(c.right[shuf]),
                    c.flag_1[shuf], c.flag_2[shuf], c.min_dist[shuf], off,
                    np.concatenate(conn_lists).astype(np.int32))
    perm = s.rcm()
    assert np.array_equal(np.sort(perm), np.arange(c.n))

    def bandwidth(cloud, order):
        pos = np.empty(cloud.n, np.int64)
        pos[order] = np.arange(cloud.n)
        rows = np.repeat(np.arange(cloud.n), np.diff(cloud.conn_off))
        return np.abs(pos[rows] - pos[cloud.conn]).max()

    assert bandwidth(s, perm) < bandwidth(s, np.arange(c.n)) / 5


def _params(**kw):
    p = dict(mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, power=0.0, inner_iterations=0,
             precision=0, reorder=0, lds_tiles=-1, tile_block=0, points_per_thread=0)
    p.update(kw)
    return L.Params(p["mach"], p["aoa"], p["cfl"], p["vl_const"], p["power"],
                    p["inner_iterations"], p["precision"], p["reorder"], p["lds_tiles"],
                    p["tile_block"], p["points_per_thread"])


def _create(cloud, params):
    v = cloud.view()
    ctx = C.c_void_p()
    rc = L.lib().lskum_create(C.byref(v), C.byref(params), C.byref(ctx))
    msg = L.lib().lskum_last_error(ctx).decode() if ctx else ""
    if ctx:
        L.lib().lskum_destroy(ctx)
    return rc, msg


def test_create_rejects_bad_parameters_before_touching_the_device():
    c = lskum.workload("naca0012_1480_k20")
    for bad in [dict(mach=-1.0), dict(cfl=0.0), dict(inner_iterations=-1), dict(precision=7),
                dict(reorder=5), dict(lds_tiles=16), dict(lds_tiles=-2), dict(tile_block=100),
                dict(tile_block=512), dict(points_per_thread=3), dict(points_per_thread=8),
                dict(points_per_thread=-1)]:
        rc, msg = _create(c, _params(**bad))
        assert rc == -1, bad
        assert msg


def test_create_rejects_degenerate_substencils():
    c = lskum.workload("naca0012_1480_k20")
    # make interior point 500 see only neighbours to its left (x+ only, x- empty)
    i = 500
    s, e = c.conn_off[i], c.conn_off[i + 1]
    k = c.conn[s:e]
    left = k[c.x[k] < c.x[i]]
    conn = np.concatenate([c.conn[:s], np.resize(left, e - s), c.conn[e:]]).astype(np.int32)
    bad = lskum.Cloud(c.x, c.y, c.left, c.right, c.flag_1, c.flag_2, c.min_dist, c.conn_off, conn)
    rc, msg = _create(bad, _params())
    assert rc == -6 and "point 500" in msg


def test_create_rejects_out_of_range_neighbours():
    c = lskum.workload("naca0012_1480_k20")
    conn = c.conn.copy()
    conn[10] = c.n + 5
    bad = lskum.Cloud(c.x, c.y, c.left, c.right, c.flag_1, c.flag_2, c.min_dist, c.conn_off, conn)
    rc, _ = _create(bad, _params())
    assert rc == -1


def _inside_wall_polygon(c, pts):
    nw = int((c.flag_1 == 0).sum())
    ax, ay = c.x[:nw], c.y[:nw]
    bx, by = np.roll(ax, -1), np.roll(ay, -1)
    inside = []
    for i in pts:
        px, py = c.x[i], c.y[i]
        span = (ay > py) != (by > py)
        xc = ax + (py - ay) * (bx - ax) / np.where(span, by - ay, 1.0)
        inside.append(int(np.count_nonzero(span & (px < xc))) % 2 == 1)
    return np.array(inside)


def _used_rel_dets(c):
    """Per point, the smallest det/(sxx*syy) over the sub-stencils the flux
    uses (interior: x+, x-, y+, y-; wall and outer: x+, x-)."""
    nx, ny, mask = c.connectivity()
    n = c.n
    pt = np.repeat(np.arange(n), np.diff(c.conn_off))
    k = c.conn
    dx, dy = c.x[k] - c.x[pt], c.y[k] - c.y[pt]
    ds = dx * ny[pt] - dy * nx[pt]
    dn = dx * nx[pt] + dy * ny[pt]
    out = np.full(n, np.inf)
    for sb in range(4):
        m = (mask & (1 << sb)) != 0
        sxx = np.bincount(pt[m], ds[m] * ds[m], n)
        syy = np.bincount(pt[m], dn[m] * dn[m], n)
        sxy = np.bincount(pt[m], ds[m] * dn[m], n)
        with np.errstate(divide="ignore", invalid="ignore"):
            r = (sxx * syy - sxy * sxy) / (sxx * syy)
        used = np.ones(n, bool) if sb < 2 else c.flag_1 == 1
        out = np.where(used, np.minimum(out, r), out)
    return out


def test_cambered_cloud_has_no_points_inside_the_body():
    # a cambered section's lower surface descends toward the trailing edge,
    # where blending toward the far field would put first-ring points inside
    c = lskum.Cloud.generate(naca="4612", n_wall=180, n_radial=45, n_circ=450, stretch=0.0,
                             k=20, seed=4007)
    nw = 180
    box = ((c.x >= c.x[:nw].min()) & (c.x <= c.x[:nw].max()) & (c.y >= c.y[:nw].min()) &
           (c.y <= c.y[:nw].max()) & (c.flag_1 == 1))
    pts = np.nonzero(box)[0]
    assert len(pts) > 0
    assert not _inside_wall_polygon(c, pts).any()


def test_cambered_cloud_substencils_meet_the_balance_target():
    c = lskum.Cloud.generate(naca="4612", n_wall=180, n_radial=45, n_circ=450, stretch=0.0,
                             k=20, seed=4007)
    assert _used_rel_dets(c).min() >= 0.05 * (1 - 1e-9)


def test_create_batch_validates_cases_and_names_the_cloud():
    a = lskum.workload("naca0012_1480_k20")
    b = lskum.workload("naca0012_3620_k24")
    conn = b.conn.copy()
    conn[10] = b.n + 5
    bad = lskum.Cloud(b.x, b.y, b.left, b.right, b.flag_1, b.flag_2, b.min_dist, b.conn_off, conn)

    def create(clouds, cases):
        views = (L.CloudView * len(clouds))(*[c.view() for c in clouds])
        cs = (L.Case * len(cases))(*[L.Case(*x) for x in cases])
        ctx = C.c_void_p()
        rc = L.lib().lskum_create_batch(views, len(clouds), cs, len(cases),
                                        C.byref(_params()), C.byref(ctx))
        msg = L.lib().lskum_last_error(ctx).decode() if ctx else ""
        if ctx:
            L.lib().lskum_destroy(ctx)
        return rc, msg

    rc, msg = create([a], [(0, 0.8, 1.0), (1, 0.8, 1.0)])
    assert rc == -1 and "case 1" in msg
    rc, msg = create([a], [(0, -0.5, 1.0)])
    assert rc == -1 and "case 0" in msg
    rc, msg = create([a, bad], [(0, 0.8, 1.0), (1, 0.7, 2.0)])
    assert rc == -1 and "cloud 1" in msg


def _restart_write(cloud, path, it, res_old, prim):
    v = cloud.view()
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    return L.lib().lskum_restart_write(str(path).encode(), C.byref(v), it, res_old, L.dptr(prim))


def _restart_read(path, n):
    it, ro = C.c_int32(), C.c_double()
    prim = np.zeros((4, n))
    rc = L.lib().lskum_restart_read(str(path).encode(), n, C.byref(it), C.byref(ro),
                                    L.dptr(prim))
    return rc, it.value, ro.value, prim


def test_restart_file_round_trip_and_layout(tmp_path):
    c = lskum.workload("naca0012_1480_k20")
    rng = np.random.default_rng(3)
    prim = np.stack([1 + 0.1 * rng.random(c.n), rng.random(c.n), rng.random(c.n) - 0.5,
                     0.7 + 0.1 * rng.random(c.n)]) * (1 + 1e-13 * rng.standard_normal(c.n))
    p = tmp_path / "restart.dat"
    assert _restart_write(c, p, 137, 3.25e-7, prim) == 0
    rc, it, ro, back = _restart_read(p, c.n)
    assert rc == 0 and it == 137 and ro == 3.25e-7
    np.testing.assert_array_equal(back, prim)
    # the layout restart_sol reads (initial_conditions_mod.h:12-17): a header
    # "n itr res_old", then nine numbers per point, prim in the last four
    tok = p.read_text().split()
    assert int(float(tok[0])) == c.n and int(tok[1]) == 137 and float(tok[2]) == 3.25e-7
    body = np.array(tok[3:], dtype=float).reshape(c.n, 9)
    np.testing.assert_array_equal(body[:, 5:].T, prim)
    np.testing.assert_array_equal(body[:, 0], c.x)


def test_restart_reader_errors(tmp_path):
    c = lskum.workload("naca0012_1480_k20")
    prim = np.ones((4, c.n))
    assert _restart_read(tmp_path / "missing.dat", c.n)[0] == -4
    p = tmp_path / "restart.dat"
    assert _restart_write(c, p, 5, 1e-6, prim) == 0
    assert _restart_read(p, c.n + 1)[0] == -1       # a different cloud
    lines = p.read_text().splitlines()
    (tmp_path / "short.dat").write_text("\n".join(lines[:100]) + "\n")
    assert _restart_read(tmp_path / "short.dat", c.n)[0] == -4


def test_solution_file_layout_and_mach(tmp_path):
    """output.dat (post_processing_mod.h:7-27, defects fixed): one point per
    line, full precision, Mach = |u|/c."""
    c = lskum.workload("naca0012_1480_k20")
    rng = np.random.default_rng(5)
    rho = 1 + 0.2 * rng.random(c.n)
    p = 0.7 + 0.1 * rng.random(c.n)
    a = np.sqrt(1.4 * p / rho)
    mach = 0.1 + 2 * rng.random(c.n)
    ang = 2 * np.pi * rng.random(c.n)
    prim = np.ascontiguousarray(np.stack([rho, mach * a * np.cos(ang), mach * a * np.sin(ang), p]))
    ent = rng.random(c.n)
    v = c.view()
    path = tmp_path / "output.dat"
    assert L.lib().lskum_solution_write(str(path).encode(), C.byref(v), 42, 1.5e-7,
                                        L.dptr(prim), L.dptr(ent)) == 0
    lines = path.read_text().splitlines()
    assert lines[0].split() == [str(c.n), "42", "1.4999999999999999e-07"]
    body = np.array([ln.split() for ln in lines[1:]], dtype=float)
    assert body.shape == (c.n, 9)
    np.testing.assert_array_equal(body[:, 0], c.x)
    np.testing.assert_array_equal(body[:, 2], c.flag_1)
    np.testing.assert_array_equal(body[:, 3:7].T, prim)
    np.testing.assert_allclose(body[:, 7], mach, rtol=1e-14)   # |u| / c
    np.testing.assert_array_equal(body[:, 8], ent)
    # no entropy given: zeros, as before compute_entropy has run
    assert L.lib().lskum_solution_write(str(path).encode(), C.byref(v), 0, 0.0, L.dptr(prim),
                                        None) == 0
    assert np.all(np.loadtxt(path, skiprows=1)[:, 8] == 0.0)
    assert L.lib().lskum_solution_write(str(tmp_path / "no" / "dir.dat").encode(), C.byref(v),
                                        0, 0.0, L.dptr(prim), None) == -4


def test_cp_file(tmp_path):
    """Cp file (compute_force_coeffs_mod.h:20-21,54, defect fixed): flag_2, x,
    cp = -(p - p_inf)/(M^2/2) for the shape points in the force loop's order."""
    c = lskum.workload("naca0012_1480_k20")
    nw = int(np.sum(c.flag_2 > 0))
    p = np.full(c.n, 1 / 1.4)
    p[:nw] = np.linspace(0.3, 1.2, nw)
    prim = np.ascontiguousarray(np.stack([np.ones(c.n), np.zeros(c.n), np.zeros(c.n), p]))
    v = c.view()
    path = tmp_path / "cp_file"
    M = 0.63
    assert L.lib().lskum_cp_write(str(path).encode(), C.byref(v), L.dptr(prim), M) == 0
    t = np.loadtxt(path)
    assert t.shape == (nw, 3)
    np.testing.assert_array_equal(t[:, 0], c.flag_2[:nw])
    np.testing.assert_array_equal(t[:, 1], c.x[:nw])
    np.testing.assert_allclose(t[:, 2], -(p[:nw] - 1 / 1.4) / (0.5 * M * M), rtol=1e-15, atol=0)
    assert L.lib().lskum_cp_write(str(path).encode(), C.byref(v), L.dptr(prim), 0.0) == -1


def test_erf_from_exponential_constants():
    # physics.hpp's double erf_x, 1 - exp(-s^2) P(t)/Q(t) (t = min(|s|, 6)),
    # evaluated here in double with the constants read from the source and
    # compared with the C library's erf: within 6e-16 absolute on [-8, 8]
    import math
    import re
    src = open(os.path.join(REPO, "c-_serial_amd", "csrc", "physics.hpp")).read()
    body = src[src.index("__device__ __forceinline__ double erf_x("):]
    body = body[:body.index("\n}\n")]
    p = [float(x) for x in re.findall(r"p = fma\(p, t, ([0-9.e+-]+)\)", body)]
    q = [float(x) for x in re.findall(r"q = fma\(q, t, ([0-9.e+-]+)\)", body)]
    p0 = float(re.search(r"double p = ([0-9.e+-]+);", body).group(1))
    q0 = float(re.search(r"double q = ([0-9.e+-]+);", body).group(1))
    assert len(p) == 8 and len(q) == 8
    worst = 0.0
    for s in np.linspace(-8.0, 8.0, 40001):
        t = min(abs(s), 6.0)
        pp, qq = p0, q0
        for c in p:
            pp = pp * t + c
        for c in q:
            qq = qq * t + c
        e = 1.0 - math.exp(-s * s) * (pp / qq)
        e = -e if s < 0 else e
        worst = max(worst, abs(e - math.erf(s)))
    assert worst <= 6e-16, worst


def test_exp_reduction_constants():
    # physics.hpp's double exponential, a = n ln2 + r and 2^n (1 + r + c2 r^2
    # + ... + c10 r^10), evaluated here in double (math.fma needs Python
    # 3.13, so the FMAs are emulated exactly with fractions) with the
    # constants read from the source and compared with the C library's exp:
    # within 4 ulps relative on [-700, 700] (the fit's own error is 3.3e-16),
    # the clamps of its two entry points at -708, the cap of n at 1100, and
    # n from the 1.5 * 2^52 shifter in both variants
    import math
    import re
    import struct
    from fractions import Fraction
    src = open(os.path.join(REPO, "c-_serial_amd", "csrc", "physics.hpp")).read()

    def fn(sig):
        b = src[src.index(sig):]
        return b[:b.index("\n}\n")]

    poly = fn("__device__ __forceinline__ double exp_poly(double r)")
    shifted = fn("__device__ __forceinline__ double exp_shifted(double a, double nb)")
    body = fn("__device__ __forceinline__ double exp_reduced(double a)")
    nonpos = fn("__device__ __forceinline__ double exp_reduced_nonpos(double a)")
    c = [float(x) for x in re.findall(r"p = fma\(p, r, ([0-9.e+-]+)\)", poly)]
    c0 = float(re.search(r"double p = ([0-9.e+-]+);", poly).group(1))
    hi, lo = (float(x) for x in re.findall(r"r = fma\(-n, ([0-9.e+-]+), [ar]\)", shifted))
    assert len(c) == 10 and c[-2:] == [1.0, 1.0]
    assert abs(hi + lo - math.log(2)) < 1e-30 + 2 ** -60
    shift = float(re.search(r"constexpr double kExpShift = ([0-9.]+);", src).group(1))
    assert shift == 1.5 * 2.0 ** 52
    assert re.search(r"const double n = nb - kExpShift;", shifted)
    assert re.search(r"ldexp\(exp_poly\(r\), \(int\)\(uint32_t\)__double_as_longlong\(nb\)\)",
                     shifted)
    # both variants shift a / ln2 by the same constant; exp_reduced caps nb
    pat = r"fma\(a, ([0-9.e+-]+), kExpShift\)"
    inv_ln2 = float(re.search(pat, body).group(1))
    assert float(re.search(pat, nonpos).group(1)) == inv_ln2
    assert abs(inv_ln2 * math.log(2) - 1) < 2e-16
    assert re.search(r"exp_shifted\(a, fmin\(fma\(a, [0-9.e+-]+, kExpShift\), "
                     r"kExpShift \+ 1100\.0\)\)", body)
    assert re.search(r"exp_shifted\(a, fma\(a, [0-9.e+-]+, kExpShift\)\)", nonpos)

    def fma(a, b, d):
        if not all(math.isfinite(x) for x in (a, b, d)):
            return a * b + d  # IEEE rules for inf and NaN operands, no rounding involved
        v = Fraction(a) * Fraction(b) + Fraction(d)
        try:
            return float(v)
        except OverflowError:  # rounds past the largest double: +-inf
            return math.inf if v > 0 else -math.inf

    def tail(n, a, n_int):
        r = fma(-n, lo, fma(-n, hi, a))
        p = c0
        for k in c:
            p = fma(p, r, k)
        try:
            return math.ldexp(p, n_int)
        except OverflowError:  # the device's ldexp overflows to inf
            return math.inf

    def shifted(a, nb):
        n = nb - shift
        low = struct.unpack("<q", struct.pack("<d", nb))[0] & 0xFFFFFFFF
        n_int = low - (1 << 32) if low >= 1 << 31 else low
        assert n_int == n  # the shifter's low word is n itself
        return tail(n, a, n_int)

    def fmin(x, y):  # IEEE minNum, as the device's: a NaN operand gives the other
        return y if math.isnan(x) else min(x, y)

    def exp_reduced(a):
        return shifted(a, fmin(fma(a, inv_ln2, shift), shift + 1100.0))

    def exp_reduced_nonpos(a):
        return shifted(a, fma(a, inv_ln2, shift))

    def exp_k(a):
        return exp_reduced(-708.0 if a < -708.0 else a)

    def exp_neg_sq(a):
        return exp_reduced_nonpos(max(a, -708.0))

    worst = 0.0
    for a in np.concatenate([np.linspace(-700.0, 700.0, 4001), np.linspace(-2.0, 2.0, 4001)]):
        a = float(a)
        worst = max(worst, abs(exp_k(a) / math.exp(a) - 1.0))
        if a <= 0.0:
            worst = max(worst, abs(exp_neg_sq(a) / math.exp(a) - 1.0))
    assert worst <= 4 * 2.0 ** -53, worst
    # ties of a / ln2 at half-integers, where the shifter's n may differ
    # from rint's by one
    for k in range(-1020, 1020, 7):
        a = (k + 0.5) * math.log(2)
        assert abs(exp_k(a) / math.exp(a) - 1.0) <= 4 * 2.0 ** -53
        if a <= 0.0:
            assert abs(exp_neg_sq(a) / math.exp(a) - 1.0) <= 4 * 2.0 ** -53
    assert exp_k(-800.0) == exp_k(-708.0) > 0.0
    assert exp_neg_sq(-800.0) == exp_neg_sq(-708.0) > 0.0
    assert exp_neg_sq(0.0) == 1.0
    # the entry points: exp_k clamps by a comparison (a NaN passes through
    # it), the kinetic terms' exp_neg_sq by one fmax
    ek = fn("__device__ __forceinline__ double exp_k(double a)")
    assert re.search(r"return exp_reduced\(a < -708\.0 \? -708\.0 : a\);", ek)
    en = fn("__device__ __forceinline__ double exp_neg_sq(double a)")
    assert re.search(r"return exp_reduced_nonpos\(fmax\(a, -708\.0\)\);", en)
    # the upper end: nb is capped at n = 1100 (keeping n's low word n
    # itself) and the reduction uses the capped n, so past a = 762 r grows
    # with a; every coefficient being positive, p(r) then grows without
    # bound and 2^n p overflows: finite up to the largest double's
    # logarithm, inf past it, for huge finite a and for +inf (where
    # r = inf - 1100 ln2 = inf, not the inf - inf of an uncapped n); a NaN
    # a makes r, p and the result NaN
    assert c0 > 0 and all(k > 0 for k in c)
    assert abs(exp_k(709.78) / math.exp(709.78) - 1.0) <= 4 * 2.0 ** -53
    for a in (709.79, 710.0, 762.0, 763.0, 1e4, 1e15, 1e16, 1e300, math.inf):
        assert exp_k(a) == math.inf, a
    assert math.isnan(exp_k(math.nan))


def test_reciprocal_and_rsqrt_refinements():
    # physics.hpp's rcp_refined and rsqrt_nr refine the hardware estimate
    # (relative error well under 2^-22) in one higher-order step.  Evaluated
    # here with exact FMAs from estimates perturbed by up to 2^-22 either
    # way, the results must be within 1.5 ulps of 1/b and 1/sqrt(b); the
    # steps' constants are read from the source
    import math
    import re
    from fractions import Fraction
    src = open(os.path.join(REPO, "c-_serial_amd", "csrc", "physics.hpp")).read()
    rc = src[src.index("__device__ __forceinline__ double rcp_refined(double b)"):]
    rc = rc[:rc.index("\n}\n")]
    rs = src[src.index("__device__ __forceinline__ double rsqrt_nr(double b)"):]
    rs = rs[:rs.index("\n}\n")]
    assert "const double e = fma(-b, r, 1.0);" in rc and "return fma(r, fma(e, e, e), r);" in rc
    m = re.search(r"return fma\(y, e \* fma\(([0-9.]+), e, ([0-9.]+)\), y\);", rs)
    assert m and "const double e = fma(-b * y, y, 1.0);" in rs
    c1, c0 = float(m.group(1)), float(m.group(2))

    def fma(a, b, d):
        return float(Fraction(a) * Fraction(b) + Fraction(d))

    def ulps(x, exact):
        return abs(Fraction(x) - exact) / Fraction(math.ulp(float(exact)))

    worst_r = worst_s = 0
    rng = np.random.default_rng(3)
    bs = np.concatenate([rng.uniform(1e-3, 1e3, 300), 10.0 ** rng.uniform(-30, 30, 300)])
    for b in bs.tolist():
        for d in (-2.0 ** -22, -2.0 ** -24, 0.0, 2.0 ** -24, 2.0 ** -22):
            r0 = (1.0 / b) * (1.0 + d)
            e = fma(-b, r0, 1.0)
            r = fma(r0, fma(e, e, e), r0)
            worst_r = max(worst_r, ulps(r, 1 / Fraction(b)))
            y0 = (1.0 / math.sqrt(b)) * (1.0 + d)
            e = fma(-(b * y0), y0, 1.0)
            y = fma(y0, e * fma(c1, e, c0), y0)
            exact = Fraction(1.0 / math.sqrt(b))  # within half an ulp of 1/sqrt(b)
            worst_s = max(worst_s, ulps(y, exact) - Fraction(1, 2))
    assert worst_r <= 1.5 and worst_s <= 1.5, (float(worst_r), float(worst_s))


def test_normwise_is_strict_about_non_finite_values():
    # the parity measure fails on any NaN and on an infinity the other side
    # lacks, in every component, and masks equal infinities
    from conftest import normwise
    ref = np.arange(1.0, 13.0).reshape(3, 4)
    assert normwise(ref, ref) == 0.0
    for r in range(3):
        bad = ref.copy()
        bad[r, 2] = np.nan
        assert normwise(bad, ref) == np.inf, r
        assert normwise(ref, bad) == np.inf, r
        one_sided = ref.copy()
        one_sided[r, 1] = np.inf
        assert normwise(one_sided, ref) == np.inf, r
        assert normwise(ref, one_sided) == np.inf, r
        both = one_sided.copy()
        assert normwise(both, one_sided) == 0.0, r
        off = ref.copy()
        off[r, 3] += 1e-3 * np.abs(ref[r]).max()
        assert abs(normwise(off, ref) - 1e-3) < 1e-12, r
    # a later component's error counts as much as the first's
    shifted = ref.copy()
    shifted[2] *= 2.0
    assert normwise(shifted, ref) == 1.0
    # signs of infinity matter
    a = np.array([np.inf, 1.0])
    assert normwise(a, np.array([-np.inf, 1.0])) == np.inf
    with pytest.raises(ValueError):
        normwise(np.zeros(3), np.zeros(4))


FPI = os.path.join(REPO, "c-_serial_amd", "lskum", "lskum_fpi")


def test_fpi_program_usage_and_input_errors(tmp_path):
    # the per-stage program fails before touching a device on bad usage or
    # a missing grid
    assert subprocess.run([FPI], capture_output=True).returncode == 2
    assert subprocess.run([FPI, "grid", "--bogus", "1"], capture_output=True).returncode == 2
    proc = subprocess.run([FPI, str(tmp_path / "missing.txt")], capture_output=True, text=True)
    assert proc.returncode == 3 and "cannot read" in proc.stderr


def test_fpi_header_compiles_on_its_own(tmp_path):
    src = tmp_path / "tu.cpp"
    src.write_text('#include "lskum_fpi.hpp"\n'
                   'int run(lskum_ctx *c) { lskum_fpi::Globals g;\n'
                   '  return lskum_fpi::q_lskum(c, 3, 2, 100, g, stdout); }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                    "-I" + os.path.join(REPO, "include"), str(src)], check=True)


def test_reference_binding_builds_against_reference_headers():
    # include/lskum_fpi_reference.hpp on the reference's own globals: the
    # reference's main() with the drop-in, compiled from the reference's
    # headers where they lie and linked against the library
    ref = os.environ.get("LSKUM_REFERENCE", "/root/reference")
    if not os.path.isfile(os.path.join(ref, "data_structure_mod.h")):
        pytest.skip("reference sources absent")
    subprocess.run(["make", "-B", "-C", os.path.join(REPO, "oracle"), "dropin", "REF=" + ref],
                   check=True, capture_output=True)
    exe = os.path.join(REPO, "oracle", "_ref", "fpi_dropin")
    proc = subprocess.run([exe], capture_output=True, text=True)
    assert proc.returncode == 2 and "usage" in proc.stderr
