"""The oracle's per-stage entries (orc_debug_advance, orc_debug_fixup) composed
into step_ch's order (step_ch.c:101-118: advance, fixup, bound_prim,
fixup_utoprim, bound_prim, twice) give orc_step's state bit for bit (CPU).

The GPU tests hold the device stages to these entries one at a time; this pins
that the stages, each taken alone, are the pieces of the step and not some
other split of it (a predictor that also bounded ph, a corrector that read the
wrong midpoint, a fixup that ran twice would all break the equality).
"""
import numpy as np
import pytest

from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, lib


@pytest.mark.parametrize("shape,kw", [((32, 32, 1), {}), ((16, 12, 8), {}), ((24, 24, 1), {"lim": 3})])
def test_stages_compose_to_the_step(shape, kw):
    lib().orc_set_threads(4)
    P = torus_params(*shape, **kw)
    ref, o = Oracle(P), Oracle(P)
    ref.init_torus()
    for _ in range(2):
        ref.step()
    p, f = ref.prims(), ref.pflag()
    t, dt = ref.time()
    o.set_prims(p, f)
    o.set_time(t, dt)
    ref.step()

    ph = o.debug_advance(0)
    assert np.array_equal(o.prims(), p), "the predictor changed the state"
    o.set_prims(ph, o.pflag())
    o.debug_bound()
    o.debug_fixup_utoprim()  # fixup_utoprim, then bound_prim
    ph_mid, f_mid = o.prims(), o.pflag()
    o.set_prims(p, f_mid)
    o.debug_advance(1, ph_mid)
    o.debug_bound()
    o.debug_fixup_utoprim()
    assert np.array_equal(o.prims(), ref.prims())
    assert np.array_equal(o.pflag(), ref.pflag())


def test_fixup_stage_floors_active_zones_only():
    """orc_debug_fixup on a state with rho = u = 0 in a tenth of the active
    zones: each of them is floored (counted, rho and u raised above 0); zones of
    the torus body, far above every floor, keep rho and u bit for bit; the ghost
    zones are not fixup's to touch (fixup.c:51-60 walks the active zones)."""
    lib().orc_set_threads(4)
    P = torus_params(24, 24, 1)
    o = Oracle(P)
    o.init_torus()
    o.step()
    p = o.prims()
    p2 = p.copy()
    act = (slice(2, P.n1 + 2), slice(2, P.n2 + 2), slice(None))
    rng = np.random.default_rng(3)
    m = rng.random(p2[0][act].shape) < 0.1
    p2[0][act][m] = 0.0
    p2[1][act][m] = 0.0
    o.set_prims(p2, o.pflag())
    c0 = o.counters()
    o.debug_fixup()
    q = o.prims()
    assert o.counters()["floor_zones"] - c0["floor_zones"] >= int(m.sum())
    assert np.all(q[0][act][m] > 0) and np.all(q[1][act][m] > 0)
    # zones left alone by the floors keep rho and u bit for bit
    keep = ~m & (p[0][act] > 1e-3) & (p[1][act] > 1e-3)
    assert np.array_equal(q[0][act][keep], p2[0][act][keep])
    assert np.array_equal(q[1][act][keep], p2[1][act][keep])
    # ghosts are not fixup's
    assert np.array_equal(q[:, :2], p2[:, :2]) and np.array_equal(q[:, -2:], p2[:, -2:])
