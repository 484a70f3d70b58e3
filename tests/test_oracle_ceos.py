"""The oracle's cache EOS (cooleos.c's COOL_EOS pressure on the cooling-table
cache, with the zones' last temperatures), pinned on the CPU by what the
construction fixes:

  linear cache      e = A rho T, p = (gam - 1) A rho T: the pressure is the
                    Gamma-law (gam - 1) u to the solve's 1e-6, and the last T
                    is e / (A rho), in code units through RHO_UNIT and U_UNIT
                    (a dropped unit, a swapped component or a wrong cache
                    index fails it)
  range branches    below rho_min / above rho_max the pressure is (gam - 1) u
                    bit for bit, e < 0 gives 0, and an unset last T becomes the
                    first guess (gam - 1) u / rho0 / k m_av c^2 (cooleos.c:341-348)
  repeats           the same call twice returns the same bits and leaves the
                    last T as the first left it
  InitCache         an unset last T takes the j-1 neighbour's, except at
                    j = jstart (cooleos.c:784-799)
  no root           a cache whose e never reaches the zone's keeps the last T
                    and returns (gam - 1) u
  first sweep       with no root anywhere, after the fluxes of one half-step
                    from unset temperatures every (h, i) column of the X1 sweep
                    holds the first guess of its jstart-1 and jstart faces'
                    left states, and every later j the jstart value (the chain
                    the reference's j-ordered loop makes)
  schedule          the steps do not depend on how OpenMP splits the zones
  step              the step on the linear cache agrees with the step on the
                    same EOS as a bilinear table (with the reference's zero
                    finite-difference derivatives) to 1e-6; after a step on the
                    gas + radiation cache each zone's last T solves its own
                    final (rho, u)

test_pins_catch_planted_mistakes builds the oracle with each of its cache-EOS
planted mistakes (ORC_MUTANT 16-19) and shows that the pin named for it fails.
"""
import ctypes

import numpy as np
import pytest

import ceos_cases as cc
from harm_cool_amd import reference_units
from harm_cool_amd.params import torus_params
from concurrent.futures import ThreadPoolExecutor

from oracle_lib import Oracle, build_mutant, cache_lib, lib, lib_at, orc_ceos, orc_eos

U = reference_units()


def _oracle(P, ceos, L=None):
    L = L or lib()
    L.orc_set_threads(4)
    o = Oracle(P, L)
    o.set_eos_cache(ceos)
    return o


def _lin_ceos(P, A=1e10, n=1, **kw):
    vx, vy, vals = cc.linear(A, P.gam, n=n)
    return orc_ceos(vx, vy, vals, P.gam, U["rho_unit"], U["u_unit"], U["p_unit"], **kw)


def pin_linear_cache(L) -> bool:
    P = torus_params(16, 16, 1)
    o = _oracle(P, _lin_ceos(P), L)
    rng = np.random.default_rng(1)
    A = 1e10
    for _ in range(200):
        rho0 = 10 ** rng.uniform(-8, 3)
        T = 10 ** rng.uniform(5, 15)  # the first guess, ~38 T here, stays inside the cache
        u = A * rho0 * U["rho_unit"] * T / U["u_unit"]
        o.set_last_T(np.zeros(o.shape[1:]))
        p = o.ceos_pressure(0, 5, 6, rho0, u)
        if abs(p - (P.gam - 1.) * u) > 2e-6 * (P.gam - 1.) * u:
            return False
        if abs(o.last_T()[5, 6, 0] - T) > 2e-6 * T:
            return False
    return True


def test_linear_cache_gives_the_gamma_law_pressure():
    assert pin_linear_cache(lib())


def test_range_branches_and_first_guess():
    P = torus_params(16, 16, 1)
    ce = _lin_ceos(P, rho_min=1e6, rho_max=1e13)
    o = _oracle(P, ce)
    for rho0, u in ((1e-7, 3e-9), (1e3, 2e-1), (5e-6, 1e-12)):
        if not (rho0 * U["rho_unit"] < 1e6 or rho0 * U["rho_unit"] > 1e13):
            continue
        o.set_last_T(np.zeros(o.shape[1:]))
        assert o.ceos_pressure(0, 4, 2, rho0, u) == (P.gam - 1.) * u
        assert o.last_T()[4, 2, 0] == cc.first_guess(P.gam, rho0, u, ce.T_min)
    o.set_last_T(np.zeros(o.shape[1:]))
    assert o.ceos_pressure(0, 4, 2, 1e-2, -1e-5) == 0.0
    # a start outside the cache's T axis: the reference reads values its lookup
    # did not write; here the solve stops there (Gamma-law, the start kept)
    o.set_last_T(np.zeros(o.shape[1:]))
    rho0 = 1e-2
    u = 1e10 * rho0 * U["rho_unit"] * 1e17 / U["u_unit"]
    assert o.ceos_pressure(0, 4, 2, rho0, u) == (P.gam - 1.) * u
    assert o.last_T()[4, 2, 0] == cc.first_guess(P.gam, rho0, u, ce.T_min) > 1e18
    # a set last T is not replaced by the guess
    T0 = np.zeros(o.shape[1:])
    T0[4, 2, 0] = 1234.5
    o.set_last_T(T0)
    o.ceos_pressure(0, 4, 2, 1e-9, 1e-9)
    assert o.last_T()[4, 2, 0] == 1234.5


def test_repeated_call_changes_nothing():
    P = torus_params(16, 16, 1)
    vx, vy, vals = cc.gas_rad()
    o = _oracle(P, orc_ceos(vx, vy, vals, P.gam, U["rho_unit"], U["u_unit"], U["p_unit"]))
    rng = np.random.default_rng(2)
    n = 0
    for _ in range(100):
        rho0, u = 10 ** rng.uniform(-6, 1), 10 ** rng.uniform(-8, -1)
        T0 = np.zeros(o.shape[1:])
        T0[7, 9, 0] = 10 ** rng.uniform(6, 12)
        o.set_last_T(T0)
        p1 = o.ceos_pressure(0, 7, 9, rho0, u)
        T1 = o.last_T()[7, 9, 0]
        p2 = o.ceos_pressure(0, 7, 9, rho0, u)
        assert np.float64(p1).tobytes() == np.float64(p2).tobytes() and o.last_T()[7, 9, 0] == T1
        n += T1 != T0[7, 9, 0]
    assert n > 50  # most calls solved (moved T)


def test_initcache_takes_the_j_minus_1_neighbour():
    P = torus_params(16, 16, 1)
    o = _oracle(P, _lin_ceos(P, rho_min=1e6, rho_max=1e13))
    js = 2
    rho0, u = 1e-9, 1e-9  # below rho_min: no solve, only the start is set
    T0 = np.zeros(o.shape[1:])
    T0[5, 7, 0] = 777.0
    T0[5, js - 1, 0] = 555.0
    o.set_last_T(T0)
    o.ceos_pressure(0, 5, 8, rho0, u)
    o.ceos_pressure(0, 5, js, rho0, u)
    T = o.last_T()
    assert T[5, 8, 0] == 777.0
    assert T[5, js, 0] == cc.first_guess(P.gam, rho0, u, 1e4)


def test_no_root_keeps_the_last_T():
    P = torus_params(16, 16, 1)
    o = _oracle(P, _lin_ceos(P, A=1e-30))
    T0 = np.zeros(o.shape[1:])
    T0[6, 6, 0] = 3.0e8
    o.set_last_T(T0)
    u = 1e-3
    assert o.ceos_pressure(0, 6, 6, 1e-2, u) == (P.gam - 1.) * u
    assert o.last_T()[6, 6, 0] == 3.0e8


def pin_first_sweep(L) -> bool:
    P = torus_params(12, 10, 4)
    o = _oracle(P, _lin_ceos(P, A=1e-30), L)
    o.init_torus()
    o.debug_fluxes(0)
    T = o.last_T()
    pr = o.prims()
    is_, ie, js, je = 2, P.n1 + 1, 2, P.n2 + 1
    for h in range(P.n3):
        for i in range(is_, ie + 2):
            for j in (js - 1, js):
                # the first pressure call of zone (h, i, j): its X1 face's left state
                rho_l = pr[0, i - 1, j, h] + 0.5 * L.orc_slope_lim(P.lim, pr[0, i - 2, j, h], pr[0, i - 1, j, h],
                                                                     pr[0, i, j, h])
                u_l = pr[1, i - 1, j, h] + 0.5 * L.orc_slope_lim(P.lim, pr[1, i - 2, j, h], pr[1, i - 1, j, h],
                                                                   pr[1, i, j, h])
                if T[i, j, h] != cc.first_guess(P.gam, rho_l, u_l, 1e4):
                    return False
            if not np.all(T[i, js + 1:je + 2, h] == T[i, js, h]):
                return False
    return True


def test_first_sweep_chains_along_j():
    assert pin_first_sweep(lib())


def test_step_on_the_linear_cache_is_the_table_step():
    """The same EOS as a cache (T solve per call) and as a bilinear (rho, T)
    table in code units, both with zero dp/drho, dp/du in vchar: two steps from
    the torus agree to the solve's tolerance."""
    P = torus_params(24, 24, 1)
    A = 1e10
    oc = _oracle(P, _lin_ceos(P, A=A))
    vx, vy = cc.axes()
    rho_c = vx / U["rho_unit"]
    Ac = A * U["rho_unit"] / U["u_unit"]
    R, TT = np.meshgrid(rho_c, vy, indexing="ij")
    tab = orc_eos({"rho": rho_c, "T": vy, "e": Ac * R * TT, "p": (P.gam - 1.) * Ac * R * TT}, P.gam)
    ot = Oracle(P)
    lib().orc_set_eos(ot.h, ctypes.byref(tab), 1)
    for o in (oc, ot):
        o.init_torus()
        for _ in range(2):
            assert o.step() == 0
    a, b = oc.prims(), ot.prims()
    for k in range(8):
        err = np.abs(a[k] - b[k]).max() / np.abs(b[k]).max()
        assert err < 1e-6, (k, err)
    assert oc.time() == pytest.approx(ot.time(), rel=1e-8)


def pin_schedule(L) -> bool:
    """Each zone's last T is its own: the same steps on 1 OpenMP thread and
    on 2, 3, 4 and 6 (a pressure call outside a zone's selection, or two zones
    sharing a temperature, would make the result depend on which zones a
    thread ran).  The zone loop hands out columns dynamically, so a stray call
    lands on the zone a thread happened to run last: on one thread count it can
    be the zone of the one-thread run with a chance of about one in that count,
    on all four together about one in 150."""
    P = torus_params(24, 24, 4)
    vx, vy, vals = cc.gas_rad(n=P.n1 + 4)
    out = []
    for nt in (1, 2, 3, 4, 6):
        o = _oracle(P, orc_ceos(vx, vy, vals, P.gam, U["rho_unit"], U["u_unit"], U["p_unit"]), L)
        L.orc_set_threads(nt)
        o.init_torus()
        for _ in range(2):
            if o.step() != 0:
                return False
        out.append((o.prims(), o.last_T()))
    L.orc_set_threads(4)
    return all(np.array_equal(out[0][0], r[0]) and np.array_equal(out[0][1], r[1]) for r in out[1:])


def test_steps_do_not_depend_on_the_thread_schedule():
    assert pin_schedule(lib())


def test_last_T_solves_each_zones_final_state():
    P = torus_params(24, 24, 1)
    vx, vy, vals = cc.gas_rad()
    o = _oracle(P, orc_ceos(vx, vy, vals, P.gam, U["rho_unit"], U["u_unit"], U["p_unit"]))
    o.init_torus()
    assert o.step() == 0
    pr, T, f = o.prims(), o.last_T(), o.pflag()
    L = cache_lib()
    types = (ctypes.c_int * 3)(0, 0, 2)
    mask = (ctypes.c_int * 64)(1)
    zv = np.zeros(1)
    ok = bad = 0
    for i in range(2, P.n1 + 2):
        for j in range(2, P.n2 + 2):
            rho, u = pr[0, i, j, 0] * U["rho_unit"], pr[1, i, j, 0] * U["u_unit"]
            # inverted zones (flag 1 after fixup_utoprim's flip) of the torus body,
            # which no floor reaches (a floored zone's state moved after its solve)
            if f[i, j, 0] != 1 or pr[0, i, j, 0] < 1e-2 or not (vx[0] <= rho <= vx[-1]):
                continue
            out = np.zeros(64)
            r = L.orc_cache_get(72, 72, 1, 5, types, 1, vx.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                vy.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                zv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, rho, T[i, j, 0], 0.0, 0, 1,
                                mask, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, 1)
            assert r == 1
            if abs(out[0] - u) <= 1e-6 * u:
                ok += 1
            else:
                bad += 1
    assert bad == 0 and ok > 50, (ok, bad)


MUTANTS = {
    16: ("InitCache without the j-1 neighbour", pin_first_sweep),
    17: ("the X1 face takes its left zone's last T", pin_first_sweep),
    18: ("the pressure read from component 0 (e)", pin_linear_cache),
    19: ("a pressure call after the advance (the mass budget's) moves a zone's last T", pin_schedule),
}


def test_pins_catch_planted_mistakes(tmp_path):
    with ThreadPoolExecutor(max_workers=3) as ex:
        sos = dict(zip(MUTANTS, ex.map(lambda n: build_mutant(n, str(tmp_path)), MUTANTS)))
    missed = [f"{n}: {what}" for n, (what, pin) in MUTANTS.items() if pin(lib_at(sos[n]))]
    assert not missed, "pins that did not catch their planted mistake: " + "; ".join(missed)
