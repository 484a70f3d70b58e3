"""Where the reference's step exits, the oracle's step stops (CPU).

step_ch.c:126-129 exits with code 11 when the step's dt is below 1e-9, after
diag_mass and before t += dt; fail() (diag.c:381-397) exits inside the step.
The oracle returns 11 with the step's state and t unchanged, or 14 with the
state of the completed step, t unchanged and no mass-budget update.
"""
import numpy as np

from harm_cool_amd.params import torus_params
from oracle_lib import Oracle


def test_dt_below_1e9_stops_before_the_clock():
    # tf one half-nanosecond past the first step: the second step's dt is
    # tf - t = 5e-10 (step_ch.c:136), below the 1e-9 limit
    P = torus_params(32, 32, 1, tf=1.0e-5 + 5.0e-10)
    o = Oracle(P)
    assert o.init_torus() == 0
    assert o.step() == 0
    t1, dt1 = o.time()
    assert t1 == 1.0e-5 and 0.0 < dt1 < 1.0e-9
    p1 = o.prims()
    assert o.step() == 11
    t2, dt2 = o.time()
    assert (t2, dt2) == (t1, dt1)
    # the step itself ran: the state moved by a step of dt1
    assert not np.array_equal(o.prims(), p1)


def test_fail_event_stops_the_step():
    P = torus_params(32, 32, 1)
    o = Oracle(P)
    assert o.init_torus() == 0
    p = o.prims()
    # negative internal energy: the sound speed squared of the faces next to this
    # zone is negative, which vchar treats as fail(FAIL_COEFF_NEG) (phys.c:409)
    i, j = 2 + 20, 2 + 16
    p[1, i, j, 0] = -p[0, i, j, 0]
    o.set_prims(p)
    t0, dt0 = o.time()
    c0 = o.counters()["fatal"]
    assert o.step() == 14
    assert o.counters()["fatal"] > c0
    assert o.time() == (t0, dt0)


def test_fail_event_leaves_mass_budget():
    P = torus_params(32, 32, 1)
    o = Oracle(P)
    assert o.init_torus() == 0
    o.set_mass_budget(1.0, 0.25, 0.5)
    p = o.prims()
    p[1, 22, 18, 0] = -p[0, 22, 18, 0]
    o.set_prims(p)
    assert o.step() == 14
    r = o.diag_log()
    assert (r["mass_in"], r["mass_out"]) == (0.25, 0.5)
