"""Pins of the CPU restatement of the learned Newton seed (oracle/mlp_oracle.py)
by mathematics: finite differences, Adam's closed form, BF16 rounding cases and
an explicit scalar forward pass.  CPU only."""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

import oracle_lib as ol  # noqa: E402


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, mo.NIN))
    T = rng.standard_normal((n, mo.NOUT))
    return X, T


def test_param_count_matches_header():
    hdr = open(os.path.join(ol.ROOT, "include", "harm_cool.h")).read()
    assert f"#define HARM_MLP_NPARAMS {mo.NPARAMS}" in hdr


def test_backward_matches_finite_differences():
    p = mo.init_params(3).astype(np.float64)
    p[mo.NPARAMS - 2:] = [0.3, -0.2]
    p += 0.05 * np.random.default_rng(4).standard_normal(p.size)  # nonzero biases too
    X, T = _batch(16, 5)
    mask = np.ones(16)
    mask[3] = 0.
    _, g = mo.loss_grad(p, X, T, mask)
    rng = np.random.default_rng(6)
    idx = rng.choice(mo.NPARAMS, 300, replace=False)
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_grad(pp, X, T, mask)[0] - mo.loss_grad(pm, X, T, mask)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(g[k]), (k, fd, g[k])


def test_output_weights():
    # a zero weight removes the output from loss and gradient; the weighted loss
    # equals the per-output losses combined; its gradient matches finite differences
    p = mo.init_params(5).astype(np.float64)
    p += 0.05 * np.random.default_rng(7).standard_normal(p.size)
    X, T = _batch(12, 8)
    l0, g0 = mo.loss_grad(p, X, T, None, (1.0, 0.0))
    T2 = T.copy()
    T2[:, 1] += 50.0
    l0b, g0b = mo.loss_grad(p, X, T2, None, (1.0, 0.0))
    assert l0 == l0b and np.array_equal(g0, g0b)
    la, _ = mo.loss_grad(p, X, T, None, (1.0, 0.0))
    lb, _ = mo.loss_grad(p, X, T, None, (0.0, 1.0))
    lw, gw = mo.loss_grad(p, X, T, None, (0.3, 2.0))
    assert abs(lw - (0.3 * la + 2.0 * lb)) <= 1e-12 * lw
    h = 1e-6
    for k in np.random.default_rng(9).choice(mo.NPARAMS, 100, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_grad(pp, X, T, None, (0.3, 2.0))[0] - mo.loss_grad(pm, X, T, None, (0.3, 2.0))[0]) / (2 * h)
        assert abs(fd - gw[k]) <= 1e-7 + 1e-5 * abs(gw[k]), (k, fd, gw[k])


def test_masked_rows_carry_no_gradient():
    p = mo.init_params(1).astype(np.float64)
    X, T = _batch(8, 2)
    mask = np.array([1, 1, 1, 1, 0, 0, 0, 0], float)
    l1, g1 = mo.loss_grad(p, X, T, mask)
    T2 = T.copy()
    T2[4:] += 100.0
    l2, g2 = mo.loss_grad(p, X, T2, mask)
    assert l1 == l2 and np.array_equal(g1, g2)


def test_adam_closed_form():
    rng = np.random.default_rng(0)
    p0 = rng.standard_normal(50)
    g1 = rng.standard_normal(50)
    g2 = rng.standard_normal(50)
    lr = 1e-3
    m = np.zeros(50)
    v = np.zeros(50)
    p1, m, v = mo.adam(p0, m, v, 1, g1, lr)
    # first step: m_hat = g, v_hat = g^2
    np.testing.assert_allclose(p1, p0 - lr * g1 / (np.abs(g1) + 1e-8), rtol=0, atol=1e-15)
    p2, m, v = mo.adam(p1, m, v, 2, g2, lr)
    mh = (0.9 * 0.1 * g1 + 0.1 * g2) / (1 - 0.81)
    vh = (0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2) / (1 - 0.999 ** 2)
    np.testing.assert_allclose(p2, p1 - lr * mh / (np.sqrt(vh) + 1e-8), rtol=1e-12, atol=1e-15)


def test_bf16_round_to_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7  # bf16 spacing at 1.0 (8 significand bits incl. the hidden one)
    cases = [
        (one, one),
        (one + ulp / 2, one),                      # tie -> even (1.0)
        (one + 3 * ulp / 2, one + 2 * ulp),        # tie -> even (1 + 2 ulp)
        (one + ulp / 2 + 2 ** -20, one + ulp),     # above the tie
        (-(one + 0.75 * ulp), -(one + ulp)),
        (3.0e38, 2.99e38),                         # large values keep their scale
    ]
    for x, want in cases[:-1]:
        assert float(mo.bf16_round(np.float32(x))) == want, x
    assert abs(float(mo.bf16_round(np.float32(3.0e38))) / 3.0e38 - 1) < 2 ** -8


def test_forward_matches_scalar_loop():
    p = mo.init_params(7)
    p[mo.NPARAMS - 2:] = [0.1, -0.4]
    X, _ = _batch(1, 9)
    W1, b1, W2, b2, W3, b3, W4, b4 = [np.asarray(a, np.float64) for a in mo.split(p)]
    h = list(X[0])
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        h = [math.tanh(sum(W[o][k] * h[k] for k in range(len(h))) + b[o]) for o in range(W.shape[0])]
    out = [sum(W4[o][k] * h[k] for k in range(64)) + b4[o] for o in range(2)]
    O, _ = mo.forward(p, X)
    np.testing.assert_allclose(O[0], out, rtol=1e-13, atol=1e-14)


def test_bf16_forward_close_to_fp64():
    p = mo.init_params(11)
    X, _ = _batch(512, 12)
    O64, _ = mo.forward(p, X)
    Ob = mo.forward_bf16(p, X)
    err = np.abs(Ob - O64)
    assert err.max() < 0.1 and np.median(err) < 0.02


def test_features_closed_form_flat_space():
    """Flat space, B = 0, fluid at rest: D = rho, E = rho + u, Q^2 = 0."""
    L = ol.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.orc_u2p_features.argtypes = [dp, dp, dp, ctypes.c_double, dp]
    g = np.diag([-1.0, 1.0, 1.0, 1.0])
    rho, u, gam = 0.7, 0.3, 13. / 9.
    prim = np.array([rho, u, 0, 0, 0, 0, 0, 0], float)
    U = np.zeros(8)
    L.orc_primtoU(ol._ptr(prim), ol._ptr(g), ol._ptr(g), 1.0, gam, ol._ptr(U))
    f = np.zeros(8)
    assert L.orc_u2p_features(ol._ptr(U), ol._ptr(g), ol._ptr(g), 1.0, ol._ptr(f)) == 0
    want = [math.log(rho), math.asinh((rho + u) / rho), 0., 0., 0., 1.0, math.asinh(u / rho), 0.]
    np.testing.assert_allclose(f, want, rtol=1e-14, atol=1e-15)


def _oracle_seeded():
    L = ol.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.orc_utoprim_2d_seeded.argtypes = [dp, dp, dp, ctypes.c_double, dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, ip, ip]
    L.orc_utoprim_2d_learned.argtypes = [dp, dp, dp, ctypes.c_double, dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, ip, ip, ip]
    return L


def test_seeded_solve_from_exact_root_and_plain_equivalence():
    """Seeded at the plain guess's own start the seeded solve is Utoprim_2d
    itself; seeded at the converged root it stops after the minimum count."""
    L = _oracle_seeded()
    d = np.load(os.path.join(ol.ROOT, "tests", "golden", "utoprim.npz"))
    gam = float(d["gam"])
    n_ok = 0
    for z in range(200):
        U = np.ascontiguousarray(d["U"][z])
        gc = np.ascontiguousarray(d["gcov"][z])
        gn = np.ascontiguousarray(d["gcon"][z])
        g = float(d["gdet"][z])
        p_ref = np.ascontiguousarray(d["guess"][z].copy())
        ni, kc = ctypes.c_int(), ctypes.c_int()
        r = L.orc_utoprim_2d(ol._ptr(U), ol._ptr(gc), ol._ptr(gn), g, ol._ptr(p_ref), gam, ctypes.byref(ni),
                             ctypes.byref(kc))
        if r != 0:
            continue
        # W and v^2 of the converged primitives
        ut = p_ref[2:5]
        utsq = ut @ gc.reshape(4, 4)[1:, 1:] @ ut
        W = (p_ref[0] + gam * p_ref[1]) * (1 + utsq)
        p2 = np.ascontiguousarray(d["guess"][z].copy())
        ni2, kc2 = ctypes.c_int(), ctypes.c_int()
        r2 = L.orc_utoprim_2d_seeded(ol._ptr(U), ol._ptr(gc), ol._ptr(gn), g, ol._ptr(p2), gam, W, 0., 1,
                                     ctypes.byref(ni2), ctypes.byref(kc2))
        assert r2 == 0
        assert kc2.value == 1 and ni2.value == 3, (z, ni2.value, kc2.value)
        scale = np.abs(p_ref[:5]).max()
        np.testing.assert_allclose(p2[:5], p_ref[:5], rtol=1e-11, atol=1e-12 * scale)
        n_ok += 1
    assert n_ok > 100
