"""The gated learned solve of the oracle (orc_utoprim_2d_gated), pinned by what
its composition fixes (CPU, on the reference's own Utoprim_2d vectors):

  gate closed   a threshold no first step reaches leaves every zone on the plain
                solve: Utoprim_2d's codes, iteration counts and primitives, bit
                for bit (the gate only reads the first iteration);
  gate open     a threshold every first step passes sends every zone with a start
                to the seed: where the seeded solve converges the result is
                orc_utoprim_2d_seeded's, with one more iteration (the plain first
                step), and where it fails the plain result with both counts;
  no start      a guess with u^2 above 1e20 gives Utoprim_2d code 2 without iterating;
                the gated solve takes the seed there whatever the threshold;
  threshold     the set of zones the gate sends to the seed grows as tau falls.
"""
import ctypes
import os

import numpy as np

from oracle_lib import _ptr, lib

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fns():
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.orc_utoprim_2d_gated.argtypes = [dp, dp, dp, ctypes.c_double, dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_double, ip, ip, ip]
    L.orc_utoprim_2d_seeded.argtypes = [dp, dp, dp, ctypes.c_double, dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, ip, ip]
    return L


def _vectors():
    d = np.load(os.path.join(GOLD, "utoprim.npz"))
    return d, float(d["gam"])


def _plain(L, d, z, gam, guess=None):
    p = np.ascontiguousarray(d["guess"][z] if guess is None else guess, dtype=np.float64).copy()
    n, k = ctypes.c_int(), ctypes.c_int()
    r = L.orc_utoprim_2d(_ptr(np.ascontiguousarray(d["U"][z])), _ptr(np.ascontiguousarray(d["gcov"][z])),
                         _ptr(np.ascontiguousarray(d["gcon"][z])), float(d["gdet"][z]), _ptr(p), gam,
                         ctypes.byref(n), ctypes.byref(k))
    return r, n.value, k.value, p


def _seeded(L, d, z, gam, W0, guess=None):
    p = np.ascontiguousarray(d["guess"][z] if guess is None else guess, dtype=np.float64).copy()
    n, k = ctypes.c_int(), ctypes.c_int()
    r = L.orc_utoprim_2d_seeded(_ptr(np.ascontiguousarray(d["U"][z])), _ptr(np.ascontiguousarray(d["gcov"][z])),
                                _ptr(np.ascontiguousarray(d["gcon"][z])), float(d["gdet"][z]), _ptr(p), gam, W0, 0.0,
                                1, ctypes.byref(n), ctypes.byref(k))
    return r, n.value, k.value, p


def _gated(L, d, z, gam, W0, tau, guess=None):
    p = np.ascontiguousarray(d["guess"][z] if guess is None else guess, dtype=np.float64).copy()
    n, k, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    r = L.orc_utoprim_2d_gated(_ptr(np.ascontiguousarray(d["U"][z])), _ptr(np.ascontiguousarray(d["gcov"][z])),
                               _ptr(np.ascontiguousarray(d["gcon"][z])), float(d["gdet"][z]), _ptr(p), gam, W0, 0.0,
                               1, tau, ctypes.byref(n), ctypes.byref(k), ctypes.byref(s))
    return r, n.value, k.value, p, s.value


def _seed_W(d, z, gam, rel):
    """A seed near the reference's converged W (W = (rho + gam u) gamma^2)."""
    pr = d["prim"][z]
    gc = d["gcov"][z].reshape(4, 4)
    u = pr[2:5]
    utsq = float(u @ gc[1:, 1:] @ u)
    return (pr[0] + gam * pr[1]) * (1 + utsq) * (1 + rel)


def test_gate_closed_is_the_plain_solve():
    L = _fns()
    d, gam = _vectors()
    for z in range(0, 1200, 3):
        r0, n0, k0, p0 = _plain(L, d, z, gam)
        r, n, k, p, s = _gated(L, d, z, gam, 1.0, 1e300)
        assert (r, n, k, s) == (r0, n0, k0, 0) and np.array_equal(p, p0), z


def test_gate_open_takes_the_seed_with_the_first_step_counted():
    L = _fns()
    d, gam = _vectors()
    nseed = nfall = 0
    for z in range(0, 1200, 3):
        if d["U"][z][0] <= 0:
            continue
        W0 = _seed_W(d, z, gam, 1e-3)
        r0, n0, k0, p0 = _plain(L, d, z, gam)
        r1, n1, k1, p1 = _seeded(L, d, z, gam, W0)
        r, n, k, p, s = _gated(L, d, z, gam, W0, 1e-300)
        first = 1 if n0 > 0 else 0
        if r1 == 0:
            assert (r, n, k, s) == (0, first + n1, first + k1, 1) and np.array_equal(p, p1), z
            nseed += 1
        else:
            assert s == 0 and r == r0 and n == n1 + n0 and k == n1 + k0 and np.array_equal(p, p0), z
            nfall += 1
    assert nseed > 300


def test_no_start_takes_the_seed():
    L = _fns()
    d, gam = _vectors()
    ok = 0
    for z in range(0, 1200, 7):
        if d["ret"][z] != 0:
            continue
        bad = d["guess"][z].copy()
        bad[2:5] = 1e12  # u^2 beyond UTSQ_TOO_BIG: no start (code 2)
        r0, n0, _, _ = _plain(L, d, z, gam, bad)
        assert (r0, n0) == (2, 0)
        W0 = _seed_W(d, z, gam, 1e-6)
        r1, n1, k1, p1 = _seeded(L, d, z, gam, W0, bad)
        r, n, k, p, s = _gated(L, d, z, gam, W0, 1e300, bad)
        if r1 == 0:
            assert (r, n, k, s) == (0, n1, k1, 1) and np.array_equal(p, p1), z
            ok += 1
    assert ok > 50


def test_threshold_splits_zones_by_their_first_step():
    """As tau falls, the set of seeded zones only grows (the gate compares one
    number per zone, its first step, with tau), from none at 1e300 (above) to
    some at 1e-8."""
    L = _fns()
    d, gam = _vectors()
    taus = [1e-1, 1e-2, 1e-4, 1e-6, 1e-8]
    prev = set()
    for tau in taus:
        cur = set()
        for z in range(0, 1200, 5):
            if d["U"][z][0] <= 0:
                continue
            W0 = _seed_W(d, z, gam, 1e-9)
            if _gated(L, d, z, gam, W0, tau)[4]:
                cur.add(z)
        assert prev <= cur
        prev = cur
    assert 0 < len(prev)
