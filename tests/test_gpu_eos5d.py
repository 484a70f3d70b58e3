"""Tabulated EOS and the batched 5D inversion on the GPU against the oracle,
bit for bit (the interpolant and the solver use no transcendental function
whose last bit could differ between libm and the device)."""
import ctypes
import os

import numpy as np
import pytest

import harm_cool_amd as hc
from oracle_lib import _ptr, lib as olib, orc_eos

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GAM = 13. / 9.


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.fixture(scope="module")
def table():
    return hc.eos_table_build(hc.eos_spec())


def test_eos_eval_matches_oracle_bitwise(table):
    tab = hc.DeviceEos(table, GAM)
    rng = np.random.default_rng(0)
    n = 20000
    rho = 10 ** rng.uniform(np.log10(table["rho"][0]) - 0.5, np.log10(table["rho"][-1]) + 0.5, n)
    u = 10 ** rng.uniform(-9, 2.2, n)
    p, dr, du = (x.cpu().numpy() for x in hc.eos_eval(tab, _dev(rho), _dev(u)))
    eos = orc_eos(table, GAM)
    L = olib()
    n_tab = 0
    for z in range(n):
        pp, a, b, T = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        r = L.orc_eos_table_p(ctypes.byref(eos), rho[z], u[z], ctypes.byref(pp), ctypes.byref(a), ctypes.byref(b),
                              ctypes.byref(T))
        assert (pp.value, a.value, b.value) == (p[z], dr[z], du[z]), z
        n_tab += r == 0
    assert n_tab > n // 4


def _batch_vs_oracle(U, gcov, gcon, gdet, guess, eos_host, tab):
    """Run the GPU batch, then the oracle zone by zone; return (ret, iters, prim)."""
    prim = _dev(guess.T)
    ret, it = hc.utoprim5d_batch(_dev(U.T), _dev(gcov.T), _dev(gcon.T), _dev(gdet), prim, GAM, tab)
    ret, it, prim = ret.cpu().numpy(), it.cpu().numpy(), prim.cpu().numpy().T
    L = olib()
    for z in range(len(gdet)):
        p = np.ascontiguousarray(guess[z].copy())
        ni = ctypes.c_int()
        r = L.orc_utoprim_5d(_ptr(np.ascontiguousarray(U[z])), _ptr(np.ascontiguousarray(gcov[z])),
                             _ptr(np.ascontiguousarray(gcon[z])), float(gdet[z]), _ptr(p), GAM,
                             ctypes.byref(eos_host) if eos_host is not None else None, ctypes.byref(ni))
        assert np.int8(r) == ret[z] and ni.value == it[z], (z, r, ret[z], ni.value, it[z])
        assert np.array_equal(p, prim[z]), z
    return ret, it, prim


def test_5d_gamma_law_batch_matches_oracle_on_reference_vectors():
    d = np.load(os.path.join(GOLD, "utoprim.npz"))
    ret, it, prim = _batch_vs_oracle(d["U"], d["gcov"], d["gcon"], d["gdet"], d["guess"], None, None)
    ok = (ret == 0) & (d["ret"] == 0)
    assert ok.sum() > 600
    assert np.abs(prim[ok, 0] / d["prim"][ok, 0] - 1).max() < 1e-8


def test_5d_tabulated_batch_matches_oracle(table):
    tab = hc.DeviceEos(table, GAM)
    eos = orc_eos(table, GAM)
    L = olib()
    rng = np.random.default_rng(1)
    spec = hc.stress_spec(0, rho_lo=table["rho"][0], rho_hi=table["rho"][-1], urho_lo=1e-2, urho_hi=3.,
                          lorentz_hi=5., sigma_hi=1.)
    n = 3000
    b = {k: v.cpu().numpy() for k, v in hc.stress_batch(n, spec).items()}
    # conserved variables of the same primitives with the tabulated EOS
    U = np.zeros((n, 8))
    P = b["prim_true"].T.copy()
    gcov, gcon, gdet = b["gcov"].T.copy(), b["gcon"].T.copy(), b["gdet"].copy()
    for z in range(n):
        L.orc_primtoU_eos(_ptr(np.ascontiguousarray(P[z])), _ptr(np.ascontiguousarray(gcov[z])),
                          _ptr(np.ascontiguousarray(gcon[z])), float(gdet[z]), GAM, ctypes.byref(eos),
                          _ptr(U[z]))
    guess = P.copy()
    guess[:, :5] *= 1 + 0.01 * (2 * rng.random((n, 5)) - 1)
    ret, it, prim = _batch_vs_oracle(U, gcov, gcon, gdet, guess, eos, tab)
    ok = ret == 0
    assert ok.mean() > 0.95
    rel = np.abs(prim[ok, 0] / P[ok, 0] - 1)
    assert np.median(rel) < 1e-12


def _exact_vs_oracle(U, gcov, gcon, gdet, guess, eos_host):
    """The exact-signature Utoprim_5d zone by zone against the oracle: the full
    return code (not the char the step stores), primitives bitwise."""
    L = olib()
    for z in range(len(gdet)):
        p = np.ascontiguousarray(guess[z].copy())
        r = hc.Utoprim_5d(U[z], gcov[z].reshape(4, 4), gcon[z].reshape(4, 4), float(gdet[z]), p)
        q = np.ascontiguousarray(guess[z].copy())
        ni = ctypes.c_int()
        ro = L.orc_utoprim_5d(_ptr(np.ascontiguousarray(U[z])), _ptr(np.ascontiguousarray(gcov[z])),
                              _ptr(np.ascontiguousarray(gcon[z])), float(gdet[z]), _ptr(q), GAM,
                              ctypes.byref(eos_host) if eos_host is not None else None, ctypes.byref(ni))
        assert r == ro, (z, r, ro)
        assert np.array_equal(p, q), z


def test_exact_signature_utoprim_5d_gamma_law_and_table(table):
    """Utoprim_5d (utoprim_5d.c:179, decs.h:610-611) with the thread's Gamma and
    tabulated EOS (harm_set_gamma, harm_set_u2p_eos)."""
    _torch()
    d = np.load(os.path.join(GOLD, "utoprim.npz"))
    hc.set_gamma(GAM)
    hc.set_u2p_eos(None)
    _exact_vs_oracle(d["U"][:200], d["gcov"][:200], d["gcon"][:200], d["gdet"][:200], d["guess"][:200], None)
    tab = hc.DeviceEos(table, GAM)
    eos = orc_eos(table, GAM)
    L = olib()
    spec = hc.stress_spec(0, rho_lo=table["rho"][0], rho_hi=table["rho"][-1], urho_lo=1e-2, urho_hi=3.,
                          lorentz_hi=5., sigma_hi=1.)
    n = 200
    b = {k: v.cpu().numpy() for k, v in hc.stress_batch(n, spec).items()}
    P = b["prim_true"].T.copy()
    gcov, gcon, gdet = b["gcov"].T.copy(), b["gcon"].T.copy(), b["gdet"].copy()
    U = np.zeros((n, 8))
    for z in range(n):
        L.orc_primtoU_eos(_ptr(np.ascontiguousarray(P[z])), _ptr(np.ascontiguousarray(gcov[z])),
                          _ptr(np.ascontiguousarray(gcon[z])), float(gdet[z]), GAM, ctypes.byref(eos), _ptr(U[z]))
    guess = P * (1 + 0.01 * np.random.default_rng(2).uniform(-1, 1, P.shape))
    hc.set_u2p_eos(tab)
    try:
        _exact_vs_oracle(U, gcov, gcon, gdet, guess, eos)
    finally:
        hc.set_u2p_eos(None)
