"""The step with the cache EOS (the reference's COOL_EOS pressure: a
temperature solve per call on the cooling-table cache of the zone's radius,
from the zone's last temperature) on the GPU against the oracle, bit for bit:
primitives, flags, counters, t and dt, and every zone's last temperature after
each step -- from unset temperatures (the first sweep of each direction runs in
the reference's j order), with one cache for every radius and one per radial
index, on 2D and 3D grids, with the Gamma-law branch below the cache's density
range; a restart carries the temperatures (restart_write_CoolEOS).
"""
import os

import numpy as np
import pytest

import ceos_cases as cc
import harm_cool_amd as hc
from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, lib as olib, orc_ceos

pytestmark = pytest.mark.gpu

U = hc.reference_units()


def _dev_caches(vx, vy, vals):
    X, Y = vx.size, vy.size
    out = []
    for k in range(vals.shape[0]):
        desc = hc.cache_desc(X, vx[0], vx[-1], Y, vy[0], vy[-1], z=1, d=vals.shape[3], whole=1)
        out.append(hc.Cache.create(desc, vx, vy, np.array([1.0]), vals[k][None], np.ones((1, X, Y), np.uint8)))
    return out


def _pair(P, vals_of, ranges=None, units=None):
    vx, vy, vals = vals_of(P)
    ranges = ranges or {}
    un = dict(U)
    un.update(units or {})
    olib().orc_set_threads(16)
    o = Oracle(P)
    o.set_eos_cache(orc_ceos(vx, vy, vals, P.gam, un["rho_unit"], un["u_unit"], un["p_unit"], **ranges))
    g = hc.Harm(P)
    caches = _dev_caches(vx, vy, vals)
    g.set_eos_cache(caches, dict(ranges, **un))
    return g, o, caches


def _same(g, o, label):
    assert np.array_equal(g.prims(), o.prims()), label
    assert np.array_equal(g.pflag(), o.pflag()), label
    assert np.array_equal(g.last_T(), o.last_T()), label
    assert g.counters() == o.counters(), label
    assert g.time() == o.time(), label


# density unit 1e3 g/cm^3: the finite-difference steps rho0 1e-6 / RHO_UNIT and u 1e-6 /
# U_UNIT are resolved, so dp/drho and dp/du take real solves at rho0 +- h, u +- h (a
# nonzero sound speed in vchar and the 5D Jacobian)
SMALL_UNITS = {"rho_unit": 1e3, "u_unit": 1e3 * cc.CL * cc.CL, "p_unit": 1e3 * cc.CL * cc.CL}

CASES = {
    "2d_shared": (lambda: torus_params(48, 48, 1), lambda P: cc.gas_rad(), None),
    "2d_resolved_derivative_steps": (lambda: torus_params(32, 32, 1), lambda P: cc.gas_rad(rho=(1e-5, 1e5)),
                                     None, SMALL_UNITS),
    "3d_per_radius": (lambda: torus_params(16, 12, 8), lambda P: cc.gas_rad(n=P.n1 + 4), None),
    "2d_vanleer_llf_rho_min": (lambda: torus_params(32, 32, 1, lim=1, hllf=0.0, laxf=1.0),
                               lambda P: cc.gas_rad(), {"rho_min": 1e6}),
    "3d_uniform_64": (lambda: torus_params(12, 8, 64), lambda P: cc.gas_rad(), None),
    # brackets walked from the axis tables' starts over axes no formula indexes
    "2d_irregular_axes": (lambda: torus_params(32, 32, 1), lambda P: cc.gas_rad(irregular=7), None),
}


@pytest.mark.parametrize("case", list(CASES))
def test_steps_match_oracle(case):
    mk, vals_of, ranges, *units = CASES[case]
    P = mk()
    g, o, caches = _pair(P, vals_of, ranges, units[0] if units else None)
    o.init_torus()
    g.init()
    assert np.array_equal(g.prims(), o.prims())
    for s in range(3):
        assert o.step() == 0
        g.step_ch(1)
        _same(g, o, f"{case} step {s}")
    assert g.eos_solves() > 0
    T = g.last_T()
    assert np.all(T[2:P.n1 + 2, 2:P.n2 + 2] > 0)
    g.close()


def test_set_last_T_reruns_the_ordered_sweep():
    """Temperatures set by hand, some unset: the next sweep of every direction
    takes unset ones from the j-1 neighbour in the reference's order."""
    P = torus_params(24, 24, 1)
    g, o, caches = _pair(P, lambda P: cc.gas_rad())
    o.init_torus()
    g.init()
    o.step()
    g.step_ch(1)
    T = o.last_T()
    rng = np.random.default_rng(4)
    T[rng.random(T.shape) < 0.3] = 0.0
    for side in (g, o):
        side.set_last_T(T)
    for s in range(2):
        assert o.step() == 0
        g.step_ch(1)
        _same(g, o, f"step {s}")
    g.close()


def test_restart_carries_the_temperatures(tmp_path):
    P = torus_params(24, 24, 1)
    g, o, caches = _pair(P, lambda P: cc.gas_rad())
    g.init()
    g.step_ch(2)
    path = os.path.join(tmp_path, "rdump")
    g.restart_write(path)
    g2 = hc.Harm(P)
    g2.set_eos_cache(caches)
    g2.restart_read(path)
    assert np.array_equal(g2.last_T(), g.last_T())
    g.step_ch(1)
    g2.step_ch(1)
    assert np.array_equal(g2.prims(), g.prims()) and np.array_equal(g2.last_T(), g.last_T())
    g.close()
    g2.close()


def test_outputs_that_would_move_the_temperatures_are_refused():
    P = torus_params(16, 16, 1)
    g, o, caches = _pair(P, lambda P: cc.gas_rad())
    g.init()
    with pytest.raises(hc.HarmError):
        g.dump_columns(0, 4)
    d = g.diag()
    assert np.isnan(d["energy"]) and d["divb_max"] < 1e-12
    g.close()


def test_an_axis_that_does_not_increase_is_refused():
    """The bracket walks take increasing axes (the bisection's index is the
    walk's only there): a cache with a repeated T point is never made."""
    vx, vy, vals = cc.gas_rad()
    vy = vy.copy()
    vy[10] = vy[9]
    with pytest.raises(hc.HarmError, match="increasing"):
        _dev_caches(vx, vy, vals)
