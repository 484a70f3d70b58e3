"""The learned Newton seed on the GPU: FP32 training kernels against the CPU
restatement (oracle/mlp_oracle.py), the BF16 MFMA inference against its BF16
emulation, the seeded inversion against the oracle bit for bit, and the
learned step against the plain step."""
import ctypes
import os
import sys

import numpy as np
import pytest

import harm_cool_amd as hc
from harm_cool_amd.params import torus_params
from oracle_lib import _ptr, lib as olib

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _dev(a, dtype=None):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda:0")


def _params(seed):
    p = mo.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    p = p + (0.05 * rng.standard_normal(p.size)).astype(np.float32)  # nonzero biases
    return p.astype(np.float32)


def _per_tensor_close(g_gpu, g_ref, rel):
    o = 0
    for s in mo.SHAPES:
        n = int(np.prod(s))
        a, b = g_gpu[o:o + n].astype(np.float64), g_ref[o:o + n]
        scale = np.abs(b).max() + 1e-30
        assert np.abs(a - b).max() <= rel * scale, (s, np.abs(a - b).max(), scale)
        o += n


@pytest.mark.parametrize("n", [3000, 65536])
def test_fp32_forward_and_gradient_match_oracle(n):
    torch = _torch()
    m = hc.Mlp()
    p = _params(1)
    m.set_params(p)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((n, 8)).astype(np.float32)
    T = rng.standard_normal((n, 2)).astype(np.float32)
    mask = (rng.random(n) > 0.1).astype(np.uint8)
    Xd, Td, Md = _dev(X), _dev(T), _dev(mask)
    Y = m.forward(Xd, bf16=False).cpu().numpy()
    O, _ = mo.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.abs(Y - O).max() <= 2e-5 * (np.abs(O).max() + 1)
    loss, g = m.grad(Xd, Td, Md)
    l_ref, g_ref = mo.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64), mask)
    assert abs(loss - l_ref) <= 1e-5 * l_ref
    _per_tensor_close(g, g_ref, 1e-4)
    # weighted outputs (the W-only training of the seed uses (1, 0))
    for ow in ((1.0, 0.0), (0.25, 2.0)):
        m.set_loss_weights(*ow)
        loss, g = m.grad(Xd, Td, Md)
        l_ref, g_ref = mo.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64), mask, ow)
        assert abs(loss - l_ref) <= 1e-5 * l_ref
        _per_tensor_close(g, g_ref, 1e-4)
    m.close()
    torch.cuda.synchronize()


def test_adam_steps_match_oracle():
    m = hc.Mlp()
    p0 = _params(3)
    m.set_params(p0)
    rng = np.random.default_rng(4)
    n = 4096
    X = _dev(rng.standard_normal((n, 8)).astype(np.float32))
    T = _dev(rng.standard_normal((n, 2)).astype(np.float32))
    pm = np.zeros(mo.NPARAMS)
    pv = np.zeros(mo.NPARAMS)
    p = p0.astype(np.float64)
    for t, lr in ((1, 1e-3), (2, 5e-4), (3, 5e-4)):
        _, g = m.grad(X, T)  # the gradient the step below uses (deterministic)
        m.train_batch(X, T, lr=lr)
        p, pm, pv = mo.adam(p, pm, pv, t, g.astype(np.float64), lr)
        got = m.params().astype(np.float64)
        assert np.abs(got - p).max() <= 1e-6 * np.abs(p).max() + 1e-3 * lr, t
        p = got  # continue from the device's float32 parameters
    m.close()


def test_bf16_inference_matches_emulation():
    """The BF16 MFMA forward (register-chained transposed tiles) against the
    CPU emulation with the same BF16 roundings; a lane-layout mistake would
    show as O(1) errors on most rows."""
    m = hc.Mlp()
    p = _params(5)
    m.set_params(p)
    rng = np.random.default_rng(6)
    n = 4096 + 37
    X = rng.standard_normal((n, 8)).astype(np.float32)
    Y = m.forward(_dev(X), bf16=True).cpu().numpy()
    E = mo.forward_bf16(p, X)
    d = np.abs(Y - E)
    assert np.quantile(d, 0.999) <= 1e-3, np.quantile(d, 0.999)
    assert d.max() <= 3e-2, d.max()
    O, _ = mo.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.median(np.abs(Y - O)) <= 2e-2
    m.close()


def _golden():
    g = np.load(os.path.join(GOLD, "utoprim.npz"))
    return g


def test_features_match_oracle():
    torch = _torch()
    g = _golden()
    L = olib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.orc_u2p_features.argtypes = [dp, dp, dp, ctypes.c_double, dp]
    m = hc.Mlp()
    feat, _, _ = m.seed(_dev(g["U"].T), _dev(g["gcov"].T), _dev(g["gcon"].T), _dev(g["gdet"]), want_seed=False)
    feat = feat.cpu().numpy().T
    for z in range(len(g["gdet"])):
        f = np.zeros(8)
        r = L.orc_u2p_features(_ptr(np.ascontiguousarray(g["U"][z])), _ptr(np.ascontiguousarray(g["gcov"][z])),
                               _ptr(np.ascontiguousarray(g["gcon"][z])), float(g["gdet"][z]), _ptr(f))
        np.testing.assert_allclose(feat[z], f, rtol=1e-13, atol=1e-300 if r == 0 else 0)
    m.close()
    torch.cuda.synchronize()


def _trained_on_golden(vsq_from_w):
    """A network fitted briefly to the golden inversions (enough to give
    seeds that are neither exact nor wild)."""
    torch = _torch()
    g = _golden()
    L = olib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.orc_u2p_features.argtypes = [dp, dp, dp, ctypes.c_double, dp]
    gam = float(g["gam"])
    ok = g["ret"] == 0
    U = g["U"][ok]
    P = g["prim"][ok]
    gc = g["gcov"][ok].reshape(-1, 4, 4)
    feat = np.zeros((len(U), 8))
    for z in range(len(U)):
        L.orc_u2p_features(_ptr(np.ascontiguousarray(U[z])), _ptr(np.ascontiguousarray(g["gcov"][ok][z])),
                           _ptr(np.ascontiguousarray(g["gcon"][ok][z])), float(g["gdet"][ok][z]), _ptr(feat[z]))
    ut = P[:, 2:5]
    utsq = np.einsum("ni,nij,nj->n", ut, gc[:, 1:, 1:], ut)
    wref = np.exp(feat[:, 0]) * (np.sinh(feat[:, 1]) - 0.5 * np.sinh(feat[:, 3]))  # E - B^2/2
    assert (wref > 0).all()
    targ = np.stack([np.log((P[:, 0] + gam * P[:, 1]) * (1 + utsq)) - np.log(wref), utsq / (1 + utsq)])
    m = hc.Mlp(seed=9)
    fd, td = _dev(feat.T), _dev(targ)
    vd = _dev(np.ones(len(U), np.uint8))
    m.fit_norm(fd, td, vd)
    m.train(fd, td, vd, batch=1024, steps=300, lr0=3e-3, lr1=3e-4)
    m.set_vsq_from_w(vsq_from_w)
    torch.cuda.synchronize()
    return m


@pytest.mark.parametrize("vsq_from_w", [1, 0])
def test_learned_batch_inversion_matches_oracle_bitwise(vsq_from_w):
    torch = _torch()
    g = _golden()
    m = _trained_on_golden(vsq_from_w)
    gam = float(g["gam"])
    Ud, gcd, gnd, gdd = _dev(g["U"].T), _dev(g["gcov"].T), _dev(g["gcon"].T), _dev(g["gdet"])
    _, W0, v0 = m.seed(Ud, gcd, gnd, gdd)
    W0, v0 = W0.cpu().numpy(), v0.cpu().numpy()
    prim = _dev(g["guess"].T)
    ret, it, kc = hc.utoprim2d_batch(Ud, gcd, gnd, gdd, prim, gam, mlp=m)
    torch.cuda.synchronize()
    ret, it, kc, prim = ret.cpu().numpy(), it.cpu().numpy(), kc.cpu().numpy(), prim.cpu().numpy().T
    L = olib()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.orc_utoprim_2d_learned.argtypes = [dp, dp, dp, ctypes.c_double, dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, ip, ip, ip]
    n_fb = 0
    for z in range(len(g["gdet"])):
        p = np.ascontiguousarray(g["guess"][z].copy())
        ni, kcz, fb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        r = L.orc_utoprim_2d_learned(_ptr(np.ascontiguousarray(g["U"][z])), _ptr(np.ascontiguousarray(g["gcov"][z])),
                                     _ptr(np.ascontiguousarray(g["gcon"][z])), float(g["gdet"][z]), _ptr(p), gam,
                                     float(W0[z]), float(v0[z]), vsq_from_w, ctypes.byref(ni), ctypes.byref(kcz),
                                     ctypes.byref(fb))
        assert np.int8(r) == ret[z], z
        assert ni.value == it[z] and kcz.value == kc[z], z
        assert np.array_equal(p, prim[z]), z
        n_fb += fb.value
    # the seeded solves agree with the reference's converged results
    ok = (ret == 0) & (g["ret"] == 0)
    rel = np.abs(prim[ok, :5] - g["prim"][ok, :5]) / (np.abs(g["prim"][ok, :5]).max(axis=1, keepdims=True))
    assert rel.max() <= 1e-10
    m.close()


def _small_torus():
    return torus_params(32, 32, 64)


def test_capture_is_passive_and_consistent():
    torch = _torch()
    P = _small_torus()
    a = hc.Harm(P)
    b = hc.Harm(P)
    a.init()
    b.init()
    cap = 200000
    feat = torch.zeros((8, cap), dtype=torch.float64, device="cuda:0")
    targ = torch.zeros((2, cap), dtype=torch.float64, device="cuda:0")
    valid = torch.zeros(cap, dtype=torch.uint8, device="cuda:0")
    stride = 7
    b.set_capture(feat, targ, valid, stride)
    a.step_ch(3)
    b.step_ch(3)
    assert np.array_equal(a.prims(), b.prims())
    na = 32 * 32 * 64
    want, phase = 0, 0
    for _ in range(6):
        want += (na - phase + stride - 1) // stride
        phase = (phase + 7919) % stride
    assert b.capture_count() == want
    f = feat[:, :want].cpu().numpy()
    t = targ[:, :want].cpu().numpy()
    v = valid[:want].cpu().numpy().astype(bool)
    assert v.mean() > 0.99
    assert np.isfinite(f[:, v]).all()
    # log(W/Wref) lies within a few units of 0; v^2 in [0, 1)
    assert (np.abs(t[0, v]) < 10).all() and (t[1, v] >= 0).all() and (t[1, v] < 1).all()
    a.close()
    b.close()


def test_learned_step_matches_plain_step():
    """Ten steps with the learned seed against ten plain steps: the Newton
    solves converge to the same roots (field-normalized 1e-10)."""
    torch = _torch()
    P = _small_torus()
    a = hc.Harm(P)
    a.init()
    cap = 400000
    feat = torch.zeros((8, cap), dtype=torch.float64, device="cuda:0")
    targ = torch.zeros((2, cap), dtype=torch.float64, device="cuda:0")
    valid = torch.zeros(cap, dtype=torch.uint8, device="cuda:0")
    a.set_capture(feat, targ, valid, 1)
    p0 = a.prims()
    a.step_ch(2)
    n = a.capture_count()
    a.set_capture(None)
    m = hc.Mlp(seed=3)
    fd, td, vd = feat[:, :n].contiguous(), targ[:, :n].contiguous(), valid[:n].contiguous()
    with pytest.raises(ValueError):
        m.fit_norm(feat[:, :n], targ[:, :n], valid[:n])  # strided views are refused
    m.fit_norm(fd, td, vd)
    lt = m.train(fd, td, vd, batch=8192, steps=400, lr0=3e-3, lr1=2e-4, trace=True)
    # FP64 numpy on the same data and schedule reaches ~1.4e-3 from ~1.0
    assert lt[-20:].mean() < 0.01 * lt[:5].mean(), (lt[:5], lt[-20:])
    sl0 = (slice(None), slice(2, 34), slice(2, 34), slice(None))
    plain = hc.Harm(P)
    learned = hc.Harm(P)
    closed = hc.Harm(P)
    gated = hc.Harm(P)
    for s in (plain, learned, closed, gated):
        s.init()
        s.set_prims(p0)
    learned.set_guess(1, m)
    # the gated seed with a threshold no first step reaches: the plain step, bit for bit
    closed.set_guess(2, m, gate=1e300)
    # a low threshold: zones whose previous-step start is off by more than 1e-9 go
    # through the seed list (k_seedfix)
    gated.set_guess(2, m, gate=1e-9)
    plain.step_ch(10)
    learned.step_ch(10)
    closed.step_ch(10)
    gated.step_ch(10)
    assert np.array_equal(closed.prims(), plain.prims())
    assert closed.counters() == plain.counters()
    pg = gated.prims()
    for v in range(8):
        x, y = plain.prims()[sl0][v], pg[sl0][v]
        assert np.abs(x - y).max() <= 1e-10 * np.abs(x).max(), v
    cg = gated.counters()
    # some zones were seeded (their Newton counts differ from the plain solve's)
    assert cg["newton_iters"] != plain.counters()["newton_iters"]
    assert cg["c2p_calls"] == plain.counters()["c2p_calls"] and cg["fatal"] == 0
    closed.close()
    gated.close()
    pa, pb = plain.prims(), learned.prims()
    sl = (slice(None), slice(2, 34), slice(2, 34), slice(None))
    for v in range(8):
        x, y = pa[sl][v], pb[sl][v]
        assert np.abs(x - y).max() <= 1e-10 * np.abs(x).max(), v
    ca, cb = plain.counters(), learned.counters()
    assert ca["c2p_calls"] == cb["c2p_calls"]
    ta, tb = plain.time()[0], learned.time()[0]
    assert abs(ta - tb) <= 1e-12 * ta
    for s in (plain, learned, a):
        s.close()
    m.close()
