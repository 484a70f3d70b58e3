"""The cooling-table generator's gas physics on the GPU (harm_fermi_batch,
harm_eos_species_batch) against the reference's own Fortran (oracle/_ref/
libeosref.so: dfermi, eos_npe, eos_rad, eos_neu, eos_he).

Bar: 1e-13 relative (the device's exp, log, pow and sqrt against the host's; the
same code compiled for the host agrees with the reference bit for bit)."""
import numpy as np
import pytest

import harm_cool_amd as hc
from oracle_lib import EosRef, eosref_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not eosref_available(), reason="oracle/_ref/libeosref.so not built")]
TOL = 1e-13


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda:0")


def test_dfermi_grid_matches_reference():
    torch = _torch()
    k, eta, th = np.meshgrid([0.5, 1.5, 2.5], np.linspace(-80, 400, 49), np.geomspace(1e-9, 80, 23), indexing="ij")
    k, eta, th = k.ravel(), eta.ravel(), th.ravel()
    n = k.size
    out = [torch.empty(n, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    ins = [_dev(torch, a) for a in (k, eta, th)]  # alive until the call returns
    rc = hc.lib().harm_fermi_batch(*(ctypes_ptr(t) for t in ins), *(ctypes_ptr(o) for o in out), n, None)
    assert rc == 0
    got = np.stack([o.cpu().numpy() for o in out])
    R = EosRef()
    ref = np.array([R.dfermi(a, b, c) for a, b, c in zip(k, eta, th)]).T
    rel = np.abs(got - ref) / np.abs(ref)
    print(f"dfermi: {n} points, {int(np.all(got == ref, axis=0).sum())} bitwise, max rel {rel.max():.2e}")
    assert rel.max() <= TOL


def ctypes_ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("kind,nin", [(0, 8), (1, 1), (2, 3), (3, 2)])
def test_species_eos_matches_reference(kind, nin):
    torch = _torch()
    rng = np.random.default_rng(kind)
    n = 2000
    kt = np.geomspace(0.01, 80.0, n)
    cols = [kt]
    if kind == 0:  # eta_n, eta_p, eta_e, n_n, n_p, n_e-, n_e+
        cols += [rng.uniform(-40, 60, n), rng.uniform(-40, 60, n), rng.uniform(-30, 200, n)]
        cols += [10.0 ** rng.uniform(-8, -1, n) for _ in range(3)]
        cols += [np.where(rng.random(n) < 0.2, 0.0, 10.0 ** rng.uniform(-9, -2, n))]
        cols[6] = np.where(rng.random(n) < 0.1, 0.0, cols[6])  # n_e- = 0: no positrons
    elif kind == 2:
        cols += [rng.uniform(0, 1, n), rng.uniform(0, 1, n)]
    elif kind == 3:
        cols += [np.where(rng.random(n) < 0.1, 0.0, 10.0 ** rng.uniform(-12, -1, n))]
    inp = np.stack(cols)
    out = torch.empty(3 * n, dtype=torch.float64, device="cuda:0")
    din = _dev(torch, inp)
    assert hc.lib().harm_eos_species_batch(kind, ctypes_ptr(din), ctypes_ptr(out), n, None) == 0
    got = out.cpu().numpy().reshape(3, n)
    R = EosRef()
    fn = [R.eos_npe, R.eos_rad, R.eos_neu, R.eos_he][kind]
    ref = np.array([fn(*inp[:, i]) for i in range(n)]).T
    scale = np.maximum(np.abs(ref), 1e-300)
    # the entropy (e + p)/kT - n (eta + 1/beta) cancels: measure it against (e + p)/kT
    scale[2] = np.maximum(scale[2], (np.abs(ref[0]) + np.abs(ref[1])) / kt)
    ok = (np.abs(got - ref) <= TOL * scale) | (got == ref)
    print(f"kind {kind}: {int(np.all(got == ref, axis=0).sum())}/{n} bitwise")
    assert ok.all(), np.argwhere(~ok)[:5]


def test_table_points_match_reference():
    """Whole table points (abundance.f90): the chemical-equilibrium Newton solve,
    ten trapping passes and the rates, one thread per point.  The reference's
    Newton solves stop at residuals of 1e-8 of the density, so an ulp of the
    device's exp or pow can move a solve's stopping iterate: the bar is 1e-6
    relative, and the test reports how many points agree bit for bit."""
    import time
    torch = _torch()
    rho, T = np.meshgrid(np.geomspace(1e8, 1e13, 9), np.geomspace(5e9, 2e11, 8), indexing="ij")
    rho, T = rho.ravel(), T.ravel()
    hd = np.full(rho.size, 0.5)
    n = rho.size
    ins = [_dev(torch, a) for a in (rho, T, hd)]
    out = torch.empty(6 * n, dtype=torch.float64, device="cuda:0")
    t0 = time.perf_counter()
    assert hc.lib().harm_eos_point_batch(*(ctypes_ptr(t) for t in ins), ctypes_ptr(out), n, None) == 0
    el = time.perf_counter() - t0
    got = out.cpu().numpy().reshape(6, n)
    R = EosRef()
    t0 = time.perf_counter()
    ref = np.array([R.abundance(a, b, 0.5) for a, b in zip(rho, T)]).T
    el_ref = time.perf_counter() - t0
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    print(f"table points: {n}, {int(np.all(got == ref, axis=0).sum())} bitwise, max rel {rel.max():.2e}; "
          f"GPU {el * 1e3:.1f} ms, reference on one core {el_ref * 1e3:.0f} ms")
    assert np.all(np.isfinite(got))
    assert rel.max() <= 1e-6


def test_wave_per_point_equals_thread_per_point():
    """The wave-per-point generator (each dfermi's nodes across the wavefront,
    the sums in the serial order) gives the one-thread-per-point kernel's bits
    at every point of a grid across the table's range, including points whose
    Newton solves run to their iteration limit."""
    import time
    torch = _torch()
    rho, T = np.meshgrid(np.geomspace(1e6, 1e13, 16), np.geomspace(1e9, 1e12, 16), indexing="ij")
    rho, T = rho.ravel(), T.ravel()
    hd = np.full(rho.size, 6.28e-4)
    n = rho.size
    ins = [_dev(torch, a) for a in (rho, T, hd)]
    outs = []
    for fn in (hc.lib().harm_eos_point_batch, hc.lib().harm_eos_point_batch_thread):
        out = torch.empty(6 * n, dtype=torch.float64, device="cuda:0")
        t0 = time.perf_counter()
        assert fn(*(ctypes_ptr(t) for t in ins), ctypes_ptr(out), n, None) == 0
        print(f"{fn.__name__}: {(time.perf_counter() - t0) * 1e3:.0f} ms")
        outs.append(out.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
