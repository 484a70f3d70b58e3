"""The inversion stress batch (KS/MKS metric, relativistic and magnetized
states): the generator's U inverts back to its primitives on the CPU oracle,
and the GPU batch inversion, plain and with the learned seed, matches the
oracle bit for bit on the same inputs."""
import ctypes

import numpy as np
import pytest

import harm_cool_amd as hc
from oracle_lib import _ptr, lib as olib

pytestmark = pytest.mark.gpu
N = 1 << 16


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _host(b):
    return {k: v.cpu().numpy() for k, v in b.items()}


def _oracle_fns():
    L = olib()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.orc_utoprim_2d_learned.argtypes = [dp, dp, dp, ctypes.c_double, dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, ip, ip, ip]
    return L


def _zone(h, z):
    return (np.ascontiguousarray(h["U"][:, z]), np.ascontiguousarray(h["gcov"][:, z]),
            np.ascontiguousarray(h["gcon"][:, z]), float(h["gdet"][z]))


def test_generator_metric_and_round_trip():
    """gcov gcon = 1, gdet^2 = -det gcov, u.u = -1, and Utoprim_2d (oracle)
    started from the perturbed guess recovers the generated primitives."""
    spec = hc.stress_spec(0)
    h = _host(hc.stress_batch(N, spec))
    gc = h["gcov"].T.reshape(-1, 4, 4)
    gn = h["gcon"].T.reshape(-1, 4, 4)
    eye = np.einsum("nij,njk->nik", gc, gn)
    assert np.abs(eye - np.eye(4)).max() < 1e-9
    det = np.linalg.det(gc)
    assert np.abs(np.sqrt(-det) / h["gdet"] - 1).max() < 1e-10
    L = olib()
    gam = spec.gam
    rng = np.random.default_rng(0)
    bad, errs = 0, []
    for z in rng.choice(N, 2000, replace=False):
        U, gcz, gnz, g = _zone(h, z)
        p = np.ascontiguousarray(h["guess"][:, z].copy())
        ni, kc = ctypes.c_int(), ctypes.c_int()
        r = L.orc_utoprim_2d(_ptr(U), _ptr(gcz), _ptr(gnz), g, _ptr(p), gam, ctypes.byref(ni), ctypes.byref(kc))
        if r != 0:
            bad += 1
            continue
        t = h["prim_true"][:, z]
        errs.append(abs(p[0] / t[0] - 1))
    errs = np.array(errs)
    print(f"round trip: {bad} of 2000 failed; rho rel err median {np.median(errs):.2e}, "
          f"99% {np.quantile(errs, 0.99):.2e}")
    # extreme sigma / Lorentz-factor corners fail or lose digits in the reference solver too
    assert bad < 100
    assert np.median(errs) < 1e-12 and np.quantile(errs, 0.95) < 1e-8


@pytest.mark.parametrize("kind", [0, 1])
def test_plain_batch_matches_oracle_bitwise(kind):
    torch = _torch()
    spec = hc.stress_spec(kind)
    b = hc.stress_batch(N, spec)
    prim = b["guess"].clone()
    ret, it, kc = hc.utoprim2d_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], prim, spec.gam)
    torch.cuda.synchronize()
    h = _host(b)
    ret, it, kc, prim = ret.cpu().numpy(), it.cpu().numpy(), kc.cpu().numpy(), prim.cpu().numpy()
    L = olib()
    rng = np.random.default_rng(1)
    for z in rng.choice(N, 1500, replace=False):
        U, gcz, gnz, g = _zone(h, z)
        p = np.ascontiguousarray(h["guess"][:, z].copy())
        ni, kcz = ctypes.c_int(), ctypes.c_int()
        r = L.orc_utoprim_2d(_ptr(U), _ptr(gcz), _ptr(gnz), g, _ptr(p), spec.gam, ctypes.byref(ni), ctypes.byref(kcz))
        assert np.int8(r) == ret[z] and ni.value == it[z] and kcz.value == kc[z], z
        assert np.array_equal(p, prim[:, z]), z


def test_learned_batch_on_stress_matches_oracle_bitwise():
    """Train a seed network on solver-generated pairs of a disjoint batch, then
    the learned batch inversion equals the oracle's seeded solve from the same
    seeds, and converged zones agree with the plain solve to 1e-10."""
    torch = _torch()
    spec = hc.stress_spec(0)
    tr = hc.stress_batch(N, hc.stress_spec(0, seed=(777, 888, 999, 1111)))
    ptr = tr["guess"].clone()
    rtr, _, _ = hc.utoprim2d_batch(tr["U"], tr["gcov"], tr["gcon"], tr["gdet"], ptr, spec.gam)
    feat, targ, valid = hc.mlp_pairs(tr["U"], tr["gcov"], tr["gcon"], tr["gdet"], ptr, rtr, spec.gam)
    m = hc.Mlp(seed=5)
    m.fit_norm(feat, targ, valid)
    lt = m.train(feat, targ, valid, batch=8192, steps=600, lr0=3e-3, lr1=2e-4, trace=True)
    assert lt[-20:].mean() < 0.15 * lt[:5].mean()  # a wide distribution: training works, not converged
    b = hc.stress_batch(N, spec)
    _, W0, v0 = m.seed(b["U"], b["gcov"], b["gcon"], b["gdet"])
    prim = b["guess"].clone()
    ret, it, kc = hc.utoprim2d_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], prim, spec.gam, mlp=m)
    p_plain = b["guess"].clone()
    ret0, it0, _ = hc.utoprim2d_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], p_plain, spec.gam)
    torch.cuda.synchronize()
    h = _host(b)
    W0, v0 = W0.cpu().numpy(), v0.cpu().numpy()
    ret, it, kc, prim = ret.cpu().numpy(), it.cpu().numpy(), kc.cpu().numpy(), prim.cpu().numpy()
    ret0, p_plain = ret0.cpu().numpy(), p_plain.cpu().numpy()
    L = _oracle_fns()
    rng = np.random.default_rng(2)
    for z in rng.choice(N, 1500, replace=False):
        U, gcz, gnz, g = _zone(h, z)
        p = np.ascontiguousarray(h["guess"][:, z].copy())
        ni, kcz, fb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        r = L.orc_utoprim_2d_learned(_ptr(U), _ptr(gcz), _ptr(gnz), g, _ptr(p), spec.gam, float(W0[z]), float(v0[z]),
                                     1, ctypes.byref(ni), ctypes.byref(kcz), ctypes.byref(fb))
        assert np.int8(r) == ret[z] and ni.value == it[z] and kcz.value == kc[z], z
        assert np.array_equal(p, prim[:, z]), z
    both = (ret == 0) & (ret0 == 0)
    rel = (np.abs(prim[:2, both] - p_plain[:2, both]) / np.abs(p_plain[:2, both])).max(axis=0)
    print(f"learned vs plain on {both.sum()} zones: rel diff median {np.median(rel):.2e}, "
          f"99% {np.quantile(rel, 0.99):.2e}, max {rel.max():.2e}; "
          f"iters plain {it0.float().mean().item():.3f}, learned {it.mean():.3f}")
    assert np.median(rel) <= 1e-14 and np.quantile(rel, 0.99) <= 1e-10
    m.close()


@pytest.mark.parametrize("tau", [1e-2, 1e-12])
def test_gated_batch_on_stress_matches_oracle_bitwise(tau):
    """The gated learned inversion (guess_mode 2): each zone starts from the
    1%-perturbed guess; only zones whose first Newton step exceeds tau take the
    network's seed (a wave runs the network only when one of its zones asks).
    Codes, iteration counts and primitives equal the oracle's gated solve fed the
    same seeds; zones the gate leaves alone equal the plain solve bit for bit."""
    torch = _torch()
    spec = hc.stress_spec(0)
    tr = hc.stress_batch(N, hc.stress_spec(0, seed=(777, 888, 999, 1111)))
    ptr = tr["guess"].clone()
    rtr, _, _ = hc.utoprim2d_batch(tr["U"], tr["gcov"], tr["gcon"], tr["gdet"], ptr, spec.gam)
    feat, targ, valid = hc.mlp_pairs(tr["U"], tr["gcov"], tr["gcon"], tr["gdet"], ptr, rtr, spec.gam)
    m = hc.Mlp(seed=5)
    m.fit_norm(feat, targ, valid)
    m.set_loss_weights(1.0, 0.0)
    m.train(feat, targ, valid, batch=8192, steps=300, lr0=3e-3, lr1=2e-4)
    b = hc.stress_batch(N, spec)
    _, W0, v0 = m.seed(b["U"], b["gcov"], b["gcon"], b["gdet"])
    prim = b["guess"].clone()
    ret, it, kc = hc.utoprim2d_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], prim, spec.gam, mlp=m, gate=tau)
    p_plain = b["guess"].clone()
    ret0, it0, kc0 = hc.utoprim2d_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], p_plain, spec.gam)
    torch.cuda.synchronize()
    h = _host(b)
    W0, v0 = W0.cpu().numpy(), v0.cpu().numpy()
    ret, it, kc, prim = ret.cpu().numpy(), it.cpu().numpy(), kc.cpu().numpy(), prim.cpu().numpy()
    ret0, it0, kc0, p_plain = ret0.cpu().numpy(), it0.cpu().numpy(), kc0.cpu().numpy(), p_plain.cpu().numpy()
    L = olib()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.orc_utoprim_2d_gated.argtypes = [dp, dp, dp, ctypes.c_double, dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_double, ip, ip, ip]
    rng = np.random.default_rng(4)
    nseed = 0
    for z in rng.choice(N, 1500, replace=False):
        U, gcz, gnz, g = _zone(h, z)
        p = np.ascontiguousarray(h["guess"][:, z].copy())
        ni, kcz, sd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        r = L.orc_utoprim_2d_gated(_ptr(U), _ptr(gcz), _ptr(gnz), g, _ptr(p), spec.gam, float(W0[z]), float(v0[z]), 1,
                                   tau, ctypes.byref(ni), ctypes.byref(kcz), ctypes.byref(sd))
        assert np.int8(r) == ret[z] and ni.value == it[z] and kcz.value == kc[z], z
        assert np.array_equal(p, prim[:, z]), z
        nseed += sd.value
        if not sd.value and ni.value == it0[z]:
            assert ret[z] == ret0[z] and np.array_equal(prim[:, z], p_plain[:, z]), z
    print(f"tau {tau:g}: {nseed} of 1500 sampled zones seeded; iters plain {it0.mean():.3f}, gated {it.mean():.3f}")
    if tau < 1e-10:
        assert nseed > 1400  # a 1% start never takes a first step below 1e-12
    m.close()


def test_learned_seed_gain_on_the_stress_distribution():
    """What the trained network gains over the method without it, on the stress
    distribution: trained on solver-generated pairs of a disjoint batch, its
    seed takes >= 15% fewer Newton iterations per zone than a cold start and
    fewer than a start 1% off the root (measured at 2^26 zones with 3000
    training steps: 7.51 against 10.39 and 8.21); the converged primitives agree
    with the plain solve's to 1e-14 (median) and 1e-10 (99%), no zone fails."""
    torch = _torch()
    n = 1 << 18
    tr = hc.stress_batch(n, hc.stress_spec(0, seed=(777, 888, 999, 1111)))
    ptr = tr["guess"].clone()
    gam = hc.stress_spec(0).gam
    rtr, _, _ = hc.utoprim2d_batch(tr["U"], tr["gcov"], tr["gcon"], tr["gdet"], ptr, gam)
    feat, targ, valid = hc.mlp_pairs(tr["U"], tr["gcov"], tr["gcon"], tr["gdet"], ptr, rtr, gam)
    del tr, ptr
    m = hc.Mlp(seed=13)
    m.fit_norm(feat, targ, valid)
    m.set_loss_weights(1.0, 0.0)
    m.train(feat, targ, valid, batch=65536, steps=1500, lr0=2e-3, lr1=5e-5)
    its = {}
    for kind, kname in ((0, "pert_1pct"), (1, "cold")):
        b = hc.stress_batch(n, hc.stress_spec(kind))
        p0 = b["guess"].clone()
        r0, it0, _ = hc.utoprim2d_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], p0, gam)
        p1 = b["guess"].clone()
        r1, it1, _ = hc.utoprim2d_batch(b["U"], b["gcov"], b["gcon"], b["gdet"], p1, gam, mlp=m)
        torch.cuda.synchronize()
        assert int((r0 != 0).sum()) == 0 and int((r1 != 0).sum()) == 0, kname
        its[kname] = (it0.double().mean().item(), it1.double().mean().item())
        rel = ((p1[:2] - p0[:2]).abs() / p0[:2].abs()).max(dim=0).values
        q = torch.quantile(rel, torch.tensor([0.5, 0.99], device=rel.device, dtype=rel.dtype)).tolist()
        assert q[0] <= 1e-14 and q[1] <= 1e-10, (kname, q)
        del b, p0, p1
    print("iterations/zone (plain, learned):", its)
    assert its["cold"][1] <= 0.85 * its["cold"][0]
    assert its["pert_1pct"][1] < its["pert_1pct"][0]
    m.close()
