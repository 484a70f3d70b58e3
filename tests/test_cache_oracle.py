"""The cooling-table cache on the CPU: the oracle's restatement of getvalues()
and Akima's spline (oracle/cache_oracle.c) and the library's file format and
axis generator, pinned by

  - the reference's own cache4d.c and akima.c, compiled from where they lie
    (oracle/_ref/libcacheref.so): the packed header size, the axes of every
    axis type, spline/derivative/linear values on random windows, and whole
    lookups (both rules, z planes and z interpolation, component masks,
    derivatives) on caches written by harm_cache_write and loaded by the
    reference's load(), which also pins the file layout;
  - closed forms: Akima's spline returns the data at the nodes (to rounding)
    and reproduces straight lines, as does the linear rule.

Bar: bitwise.  The reference is skipped where its library was not built.
Two behaviours are the oracle's alone: the points the reference would compute
first (status 3, its worker processes) or re-centre a rolling axis for
(status 2); and the linear rule at x == x_max, where the reference reads one
point past the axis.
"""
import ctypes

import numpy as np
import pytest

import harm_cool_amd as hc
from oracle_lib import CacheRef, OracleCache, cache_lib, cacheref_available

needs_ref = pytest.mark.skipif(not cacheref_available(), reason="oracle/_ref/libcacheref.so not built")


def random_cache(rng, X=7, Y=6, Z=4, D=5, types=(hc.AXIS_LOG, hc.AXIS_LOG, hc.AXIS_LINEAR)):
    desc = hc.cache_desc(X, 1e6, 1e13, Y, 1e9, 1e12, Z, 0.5, 2.0, d=D, xtype=types[0], ytype=types[1],
                         ztype=types[2], xrat=0.2 if types[0] == hc.AXIS_ROLLING else 1.1, yrat=1.2, zrat=1.05)
    vx, vy, vz = hc.cache_axes(desc)
    # smooth in log rho, log T plus noise, of the sizes the generator writes
    lx, ly = np.log10(vx), np.log10(vy)
    base = (lx[None, :, None] * 0.7 + ly[None, None, :] * 1.3 + vz[:, None, None] * 0.2)
    values = np.empty((Z, X, Y, D))
    for k in range(D):
        values[..., k] = 10.0 ** (base + 0.1 * k) * (1 + 0.05 * rng.standard_normal((Z, X, Y)))
    counted = np.ones((Z, X, Y), dtype=np.uint8)
    return desc, vx, vy, vz, values, counted


def query_points(rng, vx, vy, vz, n):
    """Random interior points plus nodes and both ends of every axis."""
    pts = []
    for _ in range(n):
        pts.append((10 ** rng.uniform(np.log10(vx[0]), np.log10(vx[-1])),
                    10 ** rng.uniform(np.log10(vy[0]), np.log10(vy[-1])), rng.uniform(vz[0], vz[-1])))
    for a in (0, 1, 2, len(vx) - 2, len(vx) - 1):
        pts.append((vx[a], vy[min(a, len(vy) - 1)], vz[min(a, len(vz) - 1)]))
    z1 = vz[min(1, len(vz) - 1)]
    pts.append((vx[0] * 0.999, vy[2], z1))      # outside in x
    pts.append((vx[3], vy[-1] * 1.001, z1))     # outside in y
    pts.append((vx[3], vy[2], vz[-1] + 0.1))    # outside in z (matters only when interpolating in z)
    return pts


@needs_ref
def test_header_is_the_references_packed_record():
    assert CacheRef().desc_size() == ctypes.sizeof(hc.CacheDesc) == 229


@needs_ref
@pytest.mark.parametrize("kind", [hc.AXIS_LOG, hc.AXIS_LINEAR, hc.AXIS_NONUNIFORM, hc.AXIS_ROLLING, hc.AXIS_CUSTOM])
def test_axes_match_reference(tmp_path, kind):
    cust = np.sort(np.random.default_rng(1).uniform(1, 50, 9))
    d = hc.cache_desc(9, 1e3, 1e11, 9, 2.0, 40.0, 9, 0.3, 3.0, xtype=kind, ytype=kind, ztype=kind, xrat=1.07,
                      yrat=0.05, zrat=1.3)
    if kind == hc.AXIS_CUSTOM:
        d.xcust = d.ycust = d.zcust = cust.ctypes.data
    got = [hc.cache_axis(kind, 9, lo, hi, r, cust if kind == hc.AXIS_CUSTOM else None)
           for lo, hi, r in ((1e3, 1e11, 1.07), (2.0, 40.0, 0.05), (0.3, 3.0, 1.3))]
    want = CacheRef().axes(d, str(tmp_path / "absent.dat"))
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@needs_ref
def test_akima_windows_match_reference():
    rng = np.random.default_rng(7)
    L, ref = cache_lib(), CacheRef()
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for trial in range(3000):
        n = int(rng.integers(3, 9))
        x = np.cumsum(rng.uniform(0.05, 2.0, n)) + rng.uniform(-5, 5)
        y = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 3)
        if trial % 5 == 0:
            y[1:3] = y[0]  # flat stretches take the 0.5 (m1 + m2) branch
        t = rng.uniform(x[0] - 1.0, x[-1] + 1.0) if trial % 4 else x[int(rng.integers(n))]
        out = np.empty(1)
        assert L.orc_akima(n, p(x), p(y), 1, p(np.array([t])), p(out)) == 0
        assert out[0] == ref.spline_akima(x, y, t)[0], (trial, n)
        v, dv = np.empty(1), np.empty(1)
        assert L.orc_akima_der(n, p(x), p(y), t, p(v), p(dv)) == 0
        rv, rd = ref.spline_akima_der(x, y, t)
        assert (v[0], dv[0]) == (rv[0], rd[0]), trial  # t < 0 keeps the search of t - h
        assert L.orc_lin(n, p(x), p(y), t) == ref.linear(x, y, t)[0]


def test_akima_closed_forms():
    rng = np.random.default_rng(3)
    L = cache_lib()
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for _ in range(500):
        n = int(rng.integers(3, 9))
        x = np.cumsum(rng.uniform(0.1, 1.0, n))
        y = rng.standard_normal(n)
        out = np.empty(n)
        L.orc_akima(n, p(x), p(y), n, p(x.copy()), p(out))
        # the nodes themselves: node 0 exactly, node k > 0 as the right end of
        # interval k-1 (the search stops at the first x[i+1] >= t), to rounding
        assert out[0] == y[0]
        np.testing.assert_allclose(out, y, rtol=0, atol=1e-13 * np.abs(y).max())
        a, b = rng.standard_normal(2)
        line = a + b * x
        t = np.sort(rng.uniform(x[0], x[-1], 5))
        out5 = np.empty(5)
        L.orc_akima(n, p(x), p(line), 5, p(t), p(out5))
        np.testing.assert_allclose(out5, a + b * t, rtol=0, atol=1e-13 * (abs(a) + abs(b) * x[-1] + 1))
        assert abs(L.orc_lin(n, p(x), p(line), t[2]) - (a + b * t[2])) <= 1e-13 * (abs(a) + abs(b) * x[-1] + 1)


@needs_ref
@pytest.mark.parametrize("spline", [True, False])
def test_lookups_match_reference(tmp_path, spline):
    rng = np.random.default_rng(11 + spline)
    desc, vx, vy, vz, values, counted = random_cache(rng)
    path = str(tmp_path / "cache.dat")
    hc.cache_write(path, desc, vx, vy, vz, values, counted)
    ref = CacheRef()
    h = ref.open(path)
    try:
        got = ref.arrays(h, hc.CacheDesc)
        want = hc.CacheDesc.from_buffer_copy(bytes(desc))
        want.wholecachecalculated = 1  # load() marks a cache whose every flag is 1
        assert bytes(got["desc"]) == bytes(want)
        for k in ("vx", "vy", "vz", "values", "counted"):
            assert np.array_equal(got[k], {"vx": vx, "vy": vy, "vz": vz, "values": values, "counted": counted}[k]), k
        o = OracleCache(desc, vx, vy, vz, values, counted)
        D = desc.d
        n = 0
        for (x, y, z) in query_points(rng, vx, vy, vz, 150):
            if not spline and (x == vx[-1] or y == vy[-1] or z == vz[-1]):
                continue  # the reference reads past the axis there
            for zi in (-1, 0, desc.z - 1):
                for mask in (None, [1, 0, 1, 0, 0]):
                    for der in ((False, True) if spline else (False,)):
                        rs, ro, rd = ref.get(h, D, x, y, z, zi, spline, mask, der)
                        os_, oo, od = o.get(x, y, z, zi, spline, mask, der)
                        assert rs == os_, (x, y, z, zi)
                        if rs:
                            sel = np.ones(D, bool) if (mask is None or not spline) else np.array(mask, bool)
                            assert np.array_equal(ro[sel], oo[sel]), (x, y, z, zi, mask)
                            if der:
                                assert np.array_equal(rd[sel], od[sel]), (x, y, z, zi, mask)
                            n += 1
        assert n > 500
    finally:
        ref.close(h)


@needs_ref
def test_version_1_flags_read_as_the_reference_reads_them(tmp_path):
    rng = np.random.default_rng(5)
    desc, vx, vy, vz, values, counted = random_cache(rng, X=4, Y=5, Z=2, D=2)
    counted[1, 2, 3] = 0
    desc.ver = b"AGACACHE_1.0"
    path = str(tmp_path / "v1.dat")
    with open(path, "wb") as f:
        f.write(bytes(desc))
        for a in (vx, vy, vz, values):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        f.write(counted.astype(np.int32).tobytes())
    mine = hc.cache_read(path)
    ref = CacheRef()
    h = ref.open(path)
    try:
        theirs = ref.arrays(h, hc.CacheDesc)
    finally:
        ref.close(h)
    assert np.array_equal(mine["counted"], theirs["counted"]) and mine["counted"][1, 2, 3] == 0
    assert np.array_equal(mine["values"], theirs["values"])


def test_file_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(9)
    desc, vx, vy, vz, values, counted = random_cache(rng, X=5, Y=4, Z=3, D=3)
    counted[0, 1, 1] = 0
    path = str(tmp_path / "c.dat")
    hc.cache_write(path, desc, vx, vy, vz, values, counted)
    raw = open(path, "rb").read()
    assert raw[:229] == bytes(desc)
    assert len(raw) == 229 + 8 * (5 + 4 + 3) + 8 * 5 * 4 * 3 * 3 + 5 * 4 * 3
    assert raw[229:229 + 40] == vx.tobytes() and raw[-60:] == counted.tobytes()
    back = hc.cache_read(path)
    assert bytes(back["desc"]) == bytes(desc)
    for k, a in (("vx", vx), ("vy", vy), ("vz", vz), ("values", values), ("counted", counted)):
        assert np.array_equal(back[k], a), k
    with open(path, "r+b") as f:
        f.truncate(len(raw) - 1)
    with pytest.raises(hc.HarmError):
        hc.cache_read(path)


def test_missing_rolling_and_broken_points():
    rng = np.random.default_rng(13)
    desc, vx, vy, vz, values, counted = random_cache(rng)
    counted[1, 3, 2] = 0
    o = OracleCache(desc, vx, vy, vz, values, counted)
    assert not o.whole
    # a window that holds the missing point needs it computed; one that does not is served
    assert o.get(vx[3], vy[2], zindex=1)[0] == hc.CACHE_MISSING
    assert o.get(vx[3], vy[2], zindex=0)[0] == hc.CACHE_OK
    assert o.get(vx[0], vy[-1], zindex=1)[0] == hc.CACHE_OK
    counted[1, 3, 2] = 7
    assert OracleCache(desc, vx, vy, vz, values, counted).get(vx[3], vy[2], zindex=1)[0] == hc.CACHE_OUT_OF_RANGE
    desc.wholecachecalculated = 1  # a header that says whole is taken at its word
    assert OracleCache(desc, vx, vy, vz, values, counted).get(vx[3], vy[2], zindex=1)[0] == hc.CACHE_OK
    d2, vx2, vy2, vz2, v2, c2 = random_cache(rng, types=(hc.AXIS_ROLLING, hc.AXIS_LOG, hc.AXIS_LINEAR))
    o2 = OracleCache(d2, vx2, vy2, vz2, v2, c2)
    assert o2.get(0.5 * (vx2[0] + vx2[1]), vy2[2], zindex=0)[0] == hc.CACHE_NEEDS_RECENTRE
    assert o2.get(vx2[2], vy2[2], zindex=0)[0] == hc.CACHE_OK
    assert o.get(vx[2], vy[2], zindex=desc.z)[0] == hc.CACHE_BAD_ZINDEX


def _solve_cases(rng, desc, vx, vy, values, n):
    """(rho, e, z plane, T start) with e taken from the table near a random T,
    starts near and far, some e outside what the table reaches."""
    lo, hi = vy[0] * (1 + 1e-6), vy[-1] * (1 - 1e-6)
    out = []
    for k in range(n):
        ix = int(rng.integers(0, desc.x))
        zi = int(rng.integers(0, desc.z))
        rho = vx[ix] * (1 + 0.3 * rng.random()) if ix < desc.x - 1 else vx[ix]
        e = values[zi, ix, int(rng.integers(0, desc.y)), 0] * (1 + 0.02 * rng.standard_normal())
        if k % 10 == 0:
            e *= 1e3  # beyond the table: no root inside the bounds
        T0 = 10 ** rng.uniform(np.log10(lo), np.log10(hi))
        out.append((rho, e, zi, T0))
    return out, lo, hi


@needs_ref
def test_temperature_solve_matches_reference(tmp_path):
    from oracle_lib import oracle_solve_T
    rng = np.random.default_rng(17)
    desc, vx, vy, vz, values, counted = random_cache(rng, X=12, Y=20, Z=3, D=5)
    # e increasing with T, as an energy density is, plus wiggles
    values[..., 0] = np.sort(values[..., 0], axis=2)
    path = str(tmp_path / "c.dat")
    hc.cache_write(path, desc, vx, vy, vz, values, counted)
    o = OracleCache(desc, vx, vy, vz, values, counted)
    ref = CacheRef()
    h = ref.open(path)
    cases, lo, hi = _solve_cases(rng, desc, vx, vy, values, 400)
    codes = []
    try:
        for rho, e, zi, T0 in cases:
            r = ref.solve_T(h, rho, e, zi, T0, lo, hi)
            mine = oracle_solve_T(o, rho, e, zi, T0, lo, hi)
            assert mine == r, (rho, e, zi, T0)
            codes.append(r[0])
            if r[0] == 1:  # converged: the table's e at T matches to the tolerance
                s, v, _ = o.get(rho, r[1], zindex=zi)
                assert s == 1 and abs(e - v[0]) / e <= 1e-6
    finally:
        ref.close(h)
    assert codes.count(1) > 200 and codes.count(1) < len(codes), np.bincount(codes)
