"""The Fermi-Dirac machinery of the cooling-table generator, pinned on the CPU.

  - The 20-point rules the library computes (harm_fermi_rules, a host call) are
    Gauss rules: the Legendre rule integrates x^(2k) on (-1, 1) exactly for
    2k <= 38, and the Laguerre rule (weights times exp(node)) integrates
    x^k exp(-x) on (0, inf) to k! for k <= 30.
  - The reference's own dfermi (gauss_integral.f90:241-349, built from its
    Fortran by oracle/Makefile into oracle/_ref/libeosref.so) against the
    mathematics: the non-degenerate limit Gamma(k+1) e^eta, the complete
    integrals at eta = 0 ((1 - 2^-k) zeta(k+1) Gamma(k+1)), the Sommerfeld
    expansion at eta = 200, and its eta and theta derivatives against finite
    differences.  The GPU batch is compared with this reference in
    tests/test_gpu_fermi.py.
"""
import ctypes
import math

import numpy as np
import pytest

import harm_cool_amd as hc
from oracle_lib import EosRef, eosref_available


def _rules():
    out = np.zeros(100)
    assert hc.lib().harm_fermi_rules(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return out[:10], out[10:20], out[20:40], out[40:60]


def test_rules_are_gauss_rules():
    lx, lw, gx, gw = _rules()
    assert np.all(np.diff(lx) > 0) and 0 < lx[0] and lx[-1] < 1
    assert np.all(np.diff(gx) > 0) and gx[0] > 0
    for k in range(20):
        assert abs(2 * np.sum(lw * lx ** (2 * k)) - 2.0 / (2 * k + 1)) <= 1e-14, k
    w = gw * np.exp(-gx)
    for k in range(31):
        exact = math.factorial(k)
        assert abs(math.fsum(w * gx ** k) - exact) <= 1e-11 * exact, k


need_ref = pytest.mark.skipif(not eosref_available(), reason="oracle/_ref/libeosref.so not built (no reference tree)")


@need_ref
def test_reference_dfermi_against_limits():
    from scipy.special import zeta
    R = EosRef()
    for k in (0.5, 1.5, 2.5):
        g = math.gamma(k + 1)
        fd, _, _ = R.dfermi(k, -40.0, 1e-12)
        assert abs(fd - g * math.exp(-40.0)) <= 1e-12 * g * math.exp(-40.0)
        fd, _, _ = R.dfermi(k, 0.0, 1e-14)
        exact = (1 - 2.0 ** (-k)) * zeta(k + 1) * g
        assert abs(fd - exact) <= 1e-12 * exact, (k, fd, exact)
        eta = 200.0
        fd, _, _ = R.dfermi(k, eta, 1e-14)
        somm = (eta ** (k + 1) / (k + 1) + math.pi ** 2 / 6 * k * eta ** (k - 1)
                + 7 * math.pi ** 4 / 360 * k * (k - 1) * (k - 2) * eta ** (k - 3))
        assert abs(fd - somm) <= 1e-9 * somm, (k, fd, somm)


@need_ref
def test_reference_dfermi_derivatives():
    R = EosRef()
    for k, eta, th in [(1.5, -3.0, 0.02), (2.5, 5.0, 0.3), (1.5, 30.0, 2.0), (0.5, 0.7, 1e-3)]:
        fd, fde, fdt = R.dfermi(k, eta, th)
        h = 1e-5 * max(1.0, abs(eta))
        num_e = (R.dfermi(k, eta + h, th)[0] - R.dfermi(k, eta - h, th)[0]) / (2 * h)
        ht = 1e-5 * th
        num_t = (R.dfermi(k, eta, th + ht)[0] - R.dfermi(k, eta, th - ht)[0]) / (2 * ht)
        assert abs(fde - num_e) <= 1e-7 * abs(fde)
        assert abs(fdt - num_t) <= 1e-6 * abs(fdt)
