"""The step with the tabulated EOS (the reference's COOL_EOS build: table
pressures in fluxes, sources and U, the tabulated branch of vchar, Utoprim_5d
without Font fix) on the GPU against the oracle running the same table."""
import ctypes

import numpy as np
import pytest

import harm_cool_amd as hc
from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, lib as olib, orc_eos

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
}


@pytest.mark.parametrize("case", list(CASES))
def test_tabulated_step_matches_oracle(case):
    _torch()
    c = CASES[case]
    P = torus_params(*c["n"])
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
    print(f"{case}: {inside.mean():.3f} of zones inside the table's density range; "
          f"Newton iterations/zone {cg['newton_iters'] / max(cg['c2p_calls'], 1):.2f}")
    g.close()
