"""Cooling-table cache lookups on the GPU (harm_cache_lookup, one getvalues()
per thread) against the oracle's restatement (oracle/cache_oracle.c, itself
pinned bitwise to the reference's cache4d.c and akima.c).

Bar: statuses equal and every written component bitwise equal, for the spline
(with and without the y derivative, with component masks) and the linear rule,
z planes and z interpolation, windows at both ends of every axis, points
outside the grid and windows holding points not yet computed; on random caches
and on one filled by the device's cooling-table generator.
"""
import numpy as np
import pytest

import harm_cool_amd as hc
from oracle_lib import OracleCache
from test_cache_oracle import query_points, random_cache

pytestmark = pytest.mark.gpu


def _check(cache, o, pts, zidx, spline, mask, der):
    import torch

    x, y, z = (torch.tensor([p[k] for p in pts], dtype=torch.float64, device="cuda:0") for k in range(3))
    zi = torch.tensor(zidx, dtype=torch.int32, device="cuda:0")
    out, dv, st = cache.lookup(x, y, z, zi, hc.INTERP_SPLINE if spline else hc.INTERP_LIN, mask, der)
    out, st = out.cpu().numpy(), st.cpu().numpy()
    dv = dv.cpu().numpy() if der else None
    D = cache.desc.d
    sel = np.ones(D, bool) if (mask is None or not spline) else np.array(mask, bool)
    for i, (p, q) in enumerate(zip(pts, zidx)):
        s, oo, od = o.get(p[0], p[1], p[2], q, spline, mask, der)
        assert st[i] == s, (i, p, q)
        if s == hc.CACHE_OK:
            assert np.array_equal(out[sel, i], oo[sel]), (i, p, q)
            assert np.all(np.isnan(out[~sel, i]))  # masked components are not written
            if der:
                assert np.array_equal(dv[sel, i], od[sel]), (i, p, q)
        else:
            assert np.all(np.isnan(out[:, i]))
    return st


@pytest.mark.parametrize("spline,der,mask", [(True, False, None), (True, True, None), (True, True, [1, 0, 0, 1, 0, 1]),
                                             (False, False, None)])
def test_random_cache_matches_oracle(spline, der, mask):
    rng = np.random.default_rng(21)
    desc, vx, vy, vz, values, counted = random_cache(rng, X=40, Y=30, Z=5, D=6)
    counted[2, 17, 11] = 0  # one point not yet computed
    c = hc.Cache.create(desc, vx, vy, vz, values, counted)
    o = OracleCache(desc, vx, vy, vz, values, counted)
    pts = query_points(rng, vx, vy, vz, 3000)
    pts += [(vx[17], vy[11], vz[2])] * 4 + [(vx[16], vy[12], 0.5 * (vz[2] + vz[3]))] * 4
    if not spline:  # the reference reads past the axis at the top end of the linear rule
        pts = [p for p in pts if p[0] != vx[-1] and p[1] != vy[-1] and p[2] != vz[-1]]
    zidx = [int(v) for v in rng.integers(-1, desc.z, len(pts))]
    zidx[-8:] = [2, -1, 2, -1, -1, -1, 2, -1]
    st = _check(c, o, pts, zidx, spline, mask, der)
    counts = np.bincount(st, minlength=5)
    assert counts[hc.CACHE_OK] > 2000 and counts[hc.CACHE_OUT_OF_RANGE] > 0 and counts[hc.CACHE_MISSING] > 0, counts
    c.close()


def test_file_cache_and_generator_values(tmp_path):
    import torch

    # a cache whose values are the device generator's cooling-table points
    desc = hc.cache_desc(9, 1e8, 1e13, 8, 5e9, 2e11, 1, 0.5, 0.5, d=6, ztype=hc.AXIS_CUSTOM)
    vx, vy, vz = hc.cache_axes(desc)
    rho, T = np.meshgrid(vx, vy, indexing="ij")
    n = rho.size
    d = [torch.as_tensor(a.ravel(), device="cuda:0") for a in (rho, T, np.full(n, 0.5))]
    out = torch.empty((6, n), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    hc._check(hc.lib().harm_eos_point_batch(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), out.data_ptr(), n,
                                            None), "harm_eos_point_batch")
    values = out.cpu().numpy().T.reshape(1, 9, 8, 6)
    assert np.isfinite(values).all()
    counted = np.ones((1, 9, 8), np.uint8)
    path = str(tmp_path / "cache_gen.dat")
    hc.cache_write(path, desc, vx, vy, vz, values, counted)
    c = hc.Cache.open(path)
    assert c.whole
    o = OracleCache(desc, vx, vy, vz, values, counted)
    rng = np.random.default_rng(4)
    pts = query_points(rng, vx, vy, vz, 1500)
    _check(c, o, pts, [0] * len(pts), True, [1, 0, 0, 0, 0, 0], True)  # e and de/dT, as the T solve asks
    _check(c, o, pts, [0] * len(pts), True, None, False)
    c.close()


def test_temperature_solve_matches_oracle():
    """harm_cache_solve_T against the oracle's run of every Newton trial (which
    is bitwise the reference's own mnewtc<1>): code, T and |F| bitwise."""
    import torch
    from oracle_lib import oracle_solve_T
    from test_cache_oracle import _solve_cases

    rng = np.random.default_rng(23)
    desc, vx, vy, vz, values, counted = random_cache(rng, X=30, Y=40, Z=3, D=5)
    values[..., 0] = np.sort(values[..., 0], axis=2)
    c = hc.Cache.create(desc, vx, vy, vz, values, counted)
    o = OracleCache(desc, vx, vy, vz, values, counted)
    cases, lo, hi = _solve_cases(rng, desc, vx, vy, values, 3000)
    cases.append((vx[0] * 0.5, 1e20, 0, vy[3]))  # rho outside the cache: the lookup fails
    dev = lambda k, dt=torch.float64: torch.tensor([q[k] for q in cases], dtype=dt, device="cuda:0")
    T = dev(3)
    berr, ret = c.solve_T(dev(0), dev(1), T, dev(2, torch.int32), lo, hi)
    T, berr, ret = T.cpu().numpy(), berr.cpu().numpy(), ret.cpu().numpy()
    for i, (rho, e, zi, T0) in enumerate(cases):
        r, t, b = oracle_solve_T(o, rho, e, zi, T0, lo, hi)
        assert (ret[i], T[i], berr[i]) == (r, t, b), (i, rho, e, zi, T0)
    counts = np.bincount(ret + 1, minlength=4)
    assert counts[2] > 1500 and counts[0] == 1, counts  # converged, and the one failed lookup
    c.close()
