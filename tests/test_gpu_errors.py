"""harm_step stops where the reference exits (GPU against the oracle).

step_ch.c:126-129 exits with code 11 when the step's dt is below 1e-9, after
diag_mass and before t += dt; fail() (diag.c:381-397) exits inside the step.
A batch of harm_step returns 11 or 14 with the state of the step where that
happened, t not advanced, and launches nothing that changes the context for the
steps after it; the context then refuses further steps until the clock is reset.
The oracle's orc_step stops the same way (tests/test_oracle_errors.py).
"""
import numpy as np
import pytest

import harm_cool_amd as hc
from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, lib as olib
from test_gpu_parity import assert_state_close

pytestmark = pytest.mark.gpu

CASES = {
    "2d_64": (64, 64, 1),
    "3d_uniform_32x32x64": (32, 32, 64),
}


@pytest.mark.parametrize("case", list(CASES))
def test_dt_below_1e9_in_step_2_of_5(case):
    # the second step's dt is tf - t = 5e-10 (step_ch.c:136)
    P = torus_params(*CASES[case], tf=1.0e-5 + 5.0e-10)
    olib().orc_set_threads(16)
    g, o = hc.Harm(P), Oracle(P)
    g.init()
    assert o.init_torus() == 0
    assert o.step() == 0
    assert o.step() == 11
    with pytest.raises(hc.HarmError) as e:
        g.step_ch(5)
    assert e.value.code == 11
    assert g.time() == o.time() == (1.0e-5, o.time()[1])
    assert g.counters() == o.counters()
    assert_state_close(g, o, case)
    mg, mo = g.diag()["mass_in"], o.diag_log()["mass_in"]
    assert abs(mg - mo) <= 1e-12 * abs(mo)
    # the stopped context refuses more steps and leaves everything as it is
    p = g.prims()
    with pytest.raises(hc.HarmError) as e:
        g.step_ch(1)
    assert e.value.code == 11
    assert np.array_equal(g.prims(), p)
    # a new clock restarts it
    g.set_time(0.0, 1.0e-5)
    g.step_ch(1)
    assert g.time()[0] == 1.0e-5


@pytest.mark.parametrize("case", list(CASES))
def test_fail_event_stops_batch(case):
    P = torus_params(*CASES[case])
    olib().orc_set_threads(16)
    g, o = hc.Harm(P), Oracle(P)
    g.init()
    assert o.init_torus() == 0
    p = o.prims()
    # negative internal energy in one torus zone: vchar's fail(FAIL_COEFF_NEG)
    p[1, 22, 18, :] = -p[0, 22, 18, :]
    for side in (g, o):
        side.set_prims(p)
    assert o.step() == 14
    with pytest.raises(hc.HarmError) as e:
        g.step_ch(3)
    assert e.value.code == 14
    assert g.time() == o.time() == (0.0, P.dt0)
    cg, co = g.counters(), o.counters()
    assert cg == co and cg["fatal"] > 0
    a, b = g.prims(), o.prims()
    assert np.array_equal(a, b, equal_nan=True)
    assert np.array_equal(g.pflag(), o.pflag())
    with pytest.raises(hc.HarmError) as e:
        g.step_ch(1)
    assert e.value.code == 14


def test_fixup_utoprim_beyond_former_list_size():
    """More than 65,536 clustered bad zones in one call: every plane's walk is
    repaired in the reference's (i, j) order (fixup.c:493-534)."""
    P = torus_params(64, 64, 128)
    olib().orc_set_threads(16)
    g, o = hc.Harm(P), Oracle(P)
    assert o.init_torus() == 0
    p = o.prims()
    rng = np.random.default_rng(3)
    flags = np.zeros(p.shape[1:], dtype=np.int8)
    flags[2:-2, 2:-2, :] = np.where(rng.random(flags[2:-2, 2:-2, :].shape) < 0.3, 1, 0)
    for side in (g, o):
        side.set_prims(p, flags)
        side.debug_bound()
    c0g, c0o = g.counters(), o.counters()
    g.debug_fixup_utoprim()
    o.debug_fixup_utoprim()
    assert np.array_equal(g.pflag(), o.pflag())
    assert np.array_equal(g.prims(), o.prims())
    cg, co = g.counters(), o.counters()
    rep = cg["fixup_repairs"] - c0g["fixup_repairs"]
    assert rep == co["fixup_repairs"] - c0o["fixup_repairs"] > 100000
    assert cg["fixup_nostencil"] - c0g["fixup_nostencil"] == co["fixup_nostencil"] - c0o["fixup_nostencil"]
