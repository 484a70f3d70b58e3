"""The step with the tabulated EOS (the reference's COOL_EOS build: table
pressures in fluxes, sources and U, the tabulated branch of vchar, Utoprim_5d
without Font fix) on the GPU against the oracle running the same table.

Two tables: the closed-form one of harm_eos_table_build, and the cooling-table
generator's own (harm_eos_table_fermi: abundance() per point, the reference's
Fermi-gas / NSE chain) over the 192x192x128 workload's range, log10 rho in
[6, 13] g/cm^3, log10 T in [9, 12] K, density unit 1e11 g/cm^3, at the disk
height factor of the torus's pressure maximum."""
import ctypes
import functools

import numpy as np
import pytest

import harm_cool_amd as hc
from harm_cool_amd.params import torus_params
from oracle_lib import EosRef, Oracle, eosref_available, lib as olib, orc_eos

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


CASES = {
    # density unit chosen so that the torus interior lies inside the table
    "3d_32x32x64": dict(n=(32, 32, 64), rho_unit=1e8, refd=0),
    "2d_64x64_refderivs": dict(n=(64, 64, 1), rho_unit=1e8, refd=1),
    "3d_32x32x64_unit1e11": dict(n=(32, 32, 64), rho_unit=1e11, refd=0),
    # the generator's table, the 192x192x128 workload's units and range
    "fermi_3d_32x32x64": dict(n=(32, 32, 64), rho_unit=1e11, refd=0, fermi=True),
    "fermi_2d_64x64_refderivs": dict(n=(64, 64, 1), rho_unit=1e11, refd=1, fermi=True),
}
FERMI_NR, FERMI_NT = 40, 40


@functools.lru_cache(maxsize=1)
def _fermi_table():
    return hc.fermi_eos_table(hc.eos_spec(nr=FERMI_NR, nt=FERMI_NT))


def test_fermi_table_units_and_reference_points():
    """The generator's table in code units: rho / 1e11, e and p over
    rho_unit c^2; e grows with T at every density; sampled points agree with
    the reference's own Fortran abundance() to the generator's device
    tolerance."""
    _torch()
    t = _fermi_table()
    spec = hc.eos_spec(nr=FERMI_NR, nt=FERMI_NT)
    assert np.allclose(t["rho"] * 1e11, np.logspace(6, 13, FERMI_NR), rtol=1e-14)
    assert np.allclose(t["T"], np.logspace(9, 12, FERMI_NT), rtol=1e-14)
    assert np.all(np.diff(t["e"], axis=1) > 0) and np.all(t["p"] > 0)
    if not eosref_available():
        pytest.skip("reference Fortran build (oracle/_ref/libeosref.so) absent")
    R = EosRef()
    u_unit = spec.rho_unit * 2.99792458e10 ** 2
    for i, j in [(0, 0), (5, 30), (17, 11), (25, 39), (33, 2), (39, 20)]:
        e, p = R.abundance(t["rho"][i] * 1e11, t["T"][j], hc.c3_hdfac())[:2]
        assert abs(t["e"][i, j] * u_unit - e) <= 1e-6 * e, (i, j)
        assert abs(t["p"][i, j] * u_unit - p) <= 1e-6 * p, (i, j)


@pytest.mark.parametrize("case", list(CASES))
def test_tabulated_step_matches_oracle(case):
    _torch()
    c = CASES[case]
    P = torus_params(*c["n"])
    if c.get("fermi"):
        host = _fermi_table()
    else:
        host = hc.eos_table_build(hc.eos_spec(rho_unit=c["rho_unit"], nr=120, nt=100))
    tab = hc.DeviceEos(host, P.gam)
    g = hc.Harm(P)
    g.set_eos(tab, bool(c["refd"]))
    o = Oracle(P)
    olib().orc_set_threads(8)
    e = orc_eos(host, P.gam)
    olib().orc_set_eos(o.h, ctypes.byref(e), c["refd"])
    g.init()
    o.init_torus()
    assert np.array_equal(g.prims(), o.prims())
    nsteps = 4
    for s in range(nsteps):
        g.step_ch(1)
        assert o.step() == 0
        tg, dtg = g.time()
        to, dto = o.time()
        assert tg == to and dtg == dto, (s, tg, to, dtg, dto)
    pg, po = g.prims(), o.prims()
    n1, n2 = c["n"][0], c["n"][1]
    sl = (slice(None), slice(2, n1 + 2), slice(2, n2 + 2), slice(None))
    for v in range(8):
        x, y = pg[sl][v], po[sl][v]
        assert np.array_equal(x, y) or np.abs(x - y).max() <= 1e-10 * np.abs(y).max(), v
    cg, co = g.counters(), o.counters()
    for k in ("c2p_calls", "newton_iters", "fail_2d", "fixup_repairs", "floor_zones"):
        assert cg[k] == co[k], (k, cg[k], co[k])
    assert np.array_equal(g.pflag(), o.pflag())
    # how much of the torus the table covers (the rest uses the Gamma-law fallback)
    rho = po[0][sl[1:]]
    u = po[1][sl[1:]]
    inside = (rho >= host["rho"][0]) & (rho <= host["rho"][-1])
    on_table = np.zeros_like(inside)
    for z in zip(*np.nonzero(inside)):
        pp, a, b, tt = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        r = olib().orc_eos_table_p(ctypes.byref(e), rho[z], u[z], ctypes.byref(pp), ctypes.byref(a), ctypes.byref(b),
                                   ctypes.byref(tt))
        on_table[z] = r == 0
    print(f"{case}: {inside.mean():.3f} of zones inside the table's density range, {on_table.mean():.3f} on the "
          f"table (rho and u); Newton iterations/zone {cg['newton_iters'] / max(cg['c2p_calls'], 1):.2f}")
    if c.get("fermi"):
        assert on_table.mean() > 0.05  # the torus body runs on the generator's table
    g.close()
