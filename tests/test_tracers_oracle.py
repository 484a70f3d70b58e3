"""The oracle's Lagrangian tracers (tracers.c:109-560) pinned by what the
mathematics fixes:

  - seeding: one tracer per active zone with rho > RHOMIN, at the zone centre in
    X1, X2 and the phi face in X3 (tracers.c:150-170);
  - advection: in a state whose coordinate velocity u^i/u^t is an affine
    function of (X1, X2, X3), the piecewise-linear interpolation is exact, so
    one update moves every tracer to X + v(X) dt;
  - the dump record of a state with rho affine in X and u = k rho: rho(X)
    rho_unit, the Gamma-law T of a constant u/rho, Ye = -1, and (r, theta) of
    bl_coord;
  - removals: r < Rin_tr (IN), r >= Rout_tr inside [0.02 pi, 0.98 pi] (OUT) or
    outside it (JET), each leaving one final record; phi wraps into [0, 2 pi].
"""
import math

import numpy as np
import pytest

from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, lib as olib

CENT = 3
RHO, UU, U1, U2, U3 = range(5)
NONE, LIVE, IN, OUT, JET, LOST = range(6)


def _oracle(n1=16, n2=12, n3=8, **kw):
    olib().orc_set_threads(8)
    o = Oracle(torus_params(n1, n2, n3, **kw))
    assert o.init_torus() == 0
    return o


def _centres(o):
    sx, dx = o.grid()
    x1 = sx[1] + (np.arange(o.n1 + 4) - 2 + 0.5) * dx[1]
    x2 = sx[2] + (np.arange(o.n2 + 4) - 2 + 0.5) * dx[2]
    x3 = sx[3] + (np.arange(o.n3) + 0.5) * dx[3] if o.n3 > 2 else sx[3] + np.arange(o.n3) * dx[3]
    return x1, x2, x3


def test_seeding():
    o = _oracle()
    n = o.tracers_init()
    X, st = o.tracers_get()
    p = o.prims()
    rho = np.transpose(p[RHO, 2:o.n1 + 2, 2:o.n2 + 2, :], (2, 0, 1))
    assert np.array_equal(st == LIVE, rho > 1e-6) and n == int((rho > 1e-6).sum())
    assert 0 < n < rho.size
    sx, dx = o.grid()
    i = np.arange(o.n1)[None, :, None]
    j = np.arange(o.n2)[None, None, :]
    h = np.arange(o.n3)[:, None, None]
    assert np.array_equal(X[0], np.broadcast_to(sx[1] + (i + 0.5) * dx[1], X[0].shape))
    assert np.array_equal(X[1], np.broadcast_to(sx[2] + (j + 0.5) * dx[2], X[1].shape))
    assert np.array_equal(X[2], np.broadcast_to(sx[3] + h * dx[3], X[2].shape))


def _velocity_state(o, coef, r_min=3.0):
    """Primitives whose u^i/u^t is coef[k] . (1, X1, X2, X3) at zone centres
    with r > r_min; normal observers elsewhere; u = 0.01 rho; B = 0."""
    gcov, gcon, _, _, coords = o.geometry()
    x1, x2, x3 = _centres(o)
    p = o.prims().copy()
    p[5:8] = 0.0
    p[UU] = 0.01 * p[RHO]
    p[U1:U3 + 1] = 0.0
    for i in range(o.n1 + 4):
        for j in range(o.n2 + 4):
            if coords[i, j, 1] <= r_min:
                continue
            gc, gn = gcov[i, j, CENT], gcon[i, j, CENT]
            alpha = 1.0 / math.sqrt(-gn[0, 0])
            beta = gn[0, 1:] * alpha * alpha
            for h in range(o.n3):
                v = coef @ np.array([1.0, x1[i], x2[j], x3[h]])
                vv = np.concatenate(([1.0], v))
                norm = vv @ gc @ vv
                assert norm < 0
                ut = 1.0 / math.sqrt(-norm)
                p[U1:U3 + 1, i, j, h] = v * ut + ut * beta
    o.set_prims(p)
    return p


def test_affine_velocity_field_is_advected_exactly():
    o = _oracle()
    o.tracers_init()
    coef = np.array([[2e-3, 1e-3, 4e-4, 1e-4],
                     [-1e-3, 2e-4, 5e-4, -2e-4],
                     [3e-3, -5e-4, 1e-3, 3e-4]])
    _velocity_state(o, coef)
    sx, dx = o.grid()
    rng = np.random.default_rng(3)
    X, st = o.tracers_get()
    st[:] = NONE
    # tracers with r in (4.5, 12), X2 in (-0.6, 0.6) and phi below the
    # last face (a tracer in the last plane interpolates towards the periodic
    # copy of plane 0, where the affine field does not continue)
    idx = rng.choice(st.size, 200, replace=False)
    X = X.reshape(3, -1)
    X[0, idx] = rng.uniform(math.log(4.5), math.log(12.0), idx.size)
    X[1, idx] = rng.uniform(-0.6, 0.6, idx.size)
    X[2, idx] = rng.uniform(0.0, sx[3] + (o.n3 - 1) * dx[3], idx.size)
    st.reshape(-1)[idx] = LIVE
    o.tracers_set(X, st)
    dt = 0.7
    o.tracers_update(dt)
    Xn, stn = o.tracers_get()
    Xn = Xn.reshape(3, -1)
    assert np.all(stn.reshape(-1)[idx] == LIVE)
    v = coef @ np.vstack((np.ones(idx.size), X[:, idx]))
    expect = X[:, idx] + v * dt
    assert np.abs(Xn[:, idx] - expect).max() <= 1e-13
    # the velocity must move the tracers by more than the tolerance
    assert np.abs(v * dt).min() > 1e-6


def test_phi_seam_wraps():
    o = _oracle()
    o.tracers_init()
    coef = np.array([[0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0], [0.005, 0.0, 0.0, 0.0]])
    _velocity_state(o, coef)
    X, st = o.tracers_get()
    st[:] = NONE
    X = X.reshape(3, -1)
    k = np.arange(20)
    X[0, k] = math.log(6.0)
    X[1, k] = 0.5
    X[2, k] = 2 * math.pi - np.linspace(0.01, 0.2, k.size)
    st.reshape(-1)[k] = LIVE
    o.tracers_set(X, st)
    o.tracers_update(20.0)
    Xn, _ = o.tracers_get()
    x = X[2, k] + 0.1
    expect = np.where(x > 2 * math.pi, x - 2 * math.pi, x)
    assert np.abs(Xn.reshape(3, -1)[2, k] - expect).max() <= 1e-13
    assert (Xn.reshape(3, -1)[2, k] < 0.1).sum() == (x > 2 * math.pi).sum() > 0


def test_dump_record_fields():
    o = _oracle()
    o.tracers_init(rho_unit=7.0, t_unit=3.0)
    x1, x2, x3 = _centres(o)
    p = o.prims().copy()
    c = np.array([1.0, 0.3, 0.2, 0.05])
    rho = (c[0] + c[1] * x1[:, None, None] + c[2] * x2[None, :, None] + c[3] * x3[None, None, :])
    p[RHO] = rho
    p[UU] = 0.01 * rho
    o.set_prims(p)
    o.set_time(1.5, 0.1)
    recs = o.tracers_dump()
    X, st = o.tracers_get()
    sx, dx = o.grid()
    live = np.flatnonzero(st.reshape(-1) == LIVE)
    assert recs.size == live.size
    X = X.reshape(3, -1)[:, live]
    gam = o.params.gam
    CL, KBOL = 2.99792458e10, 1.38e-16
    mav = (1.673e-24 + 1.67492729e-24 + 2 * 9.109565e-28 + 6.6465e-24) / 5.0
    P = o.params
    r = np.exp(X[0]) + P.R0
    th = math.pi / 2 * (1 + X[1]) + (1 - P.hslope) / 2 * np.sin(math.pi * (1 + X[1]))
    np.testing.assert_allclose(recs["t"], 4.5, rtol=0, atol=0)
    # affine in X, except in the last phi plane: its upper neighbour is the
    # periodic copy of plane 0 (x3 = centre 0) placed one plane further on
    fA = c[0] + c[1] * X[0] + c[2] * X[1] + c[3] * X[2]
    last = recs["h"] == o.n3 - 1
    cA = sx[3] + (o.n3 - 1 + 0.5) * dx[3]
    cB = sx[3] + (o.n3 + 0.5) * dx[3]
    rA = c[0] + c[1] * X[0] + c[2] * X[1] + c[3] * cA
    rB = c[0] + c[1] * X[0] + c[2] * X[1] + c[3] * x3[0]
    want = np.where(last, rA + (X[2] - cA) * (rB - rA) / (cB - cA), fA)
    assert last.sum() > 0
    np.testing.assert_allclose(recs["rho"], 7.0 * want, rtol=1e-14)
    np.testing.assert_allclose(recs["T"], (gam - 1) * 0.01 / KBOL * mav * CL * CL, rtol=1e-14)
    assert np.all(recs["ye"] == -1.0)
    np.testing.assert_allclose(recs["r"], r, rtol=1e-15)
    np.testing.assert_allclose(recs["th"], th, rtol=1e-15)
    np.testing.assert_array_equal(recs["ph"], X[2])
    k = live
    n12 = o.n1 * o.n2
    assert np.array_equal(recs["h"], k // n12) and np.array_equal(recs["i"], (k % n12) // o.n2)
    assert np.array_equal(recs["j"], k % o.n2) and np.all(recs["kind"] == LIVE)


def test_removals_and_lost_tracers():
    o = _oracle()
    o.tracers_init()
    P = o.params
    sx, dx = o.grid()
    rin = math.exp(sx[1] - 0.5 * dx[1]) + P.R0
    rout = 0.8 * (math.exp(sx[1] + o.n1 * dx[1]) + P.R0)
    # inward in the inner region, outward in the outer one
    x1, x2, x3 = _centres(o)
    X, st = o.tracers_get()
    st[:] = NONE
    X = X.reshape(3, -1)
    st = st.reshape(-1)
    cases = [  # (X1, X2, X3); X2 = -1, 0, 1 are theta = 0, pi/2, pi
        (math.log(rin * 1.02), 0.0, 1.0),    # -> IN through r
        (math.log(rout * 0.98), 0.0, 1.0),   # -> OUT (equator)
        (math.log(rout * 0.98), -0.99, 1.0),  # -> JET (theta < 0.02 pi)
        (math.log(6.0), 0.0, 1.0),           # stays
        (sx[1] - 5.0, 0.0, 1.0),             # outside the grid -> LOST
    ]
    for k, (a, b, cph) in enumerate(cases):
        X[:, k] = (a, b, cph)
        st[k] = LIVE
    coef = np.zeros((3, 4))
    coef[0, 0] = 0.01
    # normal observers inside r = 3 fall in; outward dX1/dt = 0.01 outside
    _velocity_state(o, coef, r_min=3.0)
    o.tracers_set(X, st)
    o.set_time(2.0, 0.1)
    o.tracers_update(4.0)
    ev = o.tracers_events()
    _, stn = o.tracers_get()
    stn = stn.reshape(-1)
    assert list(stn[:5]) == [IN, OUT, JET, LIVE, LOST]
    assert list(ev["kind"]) == [IN, OUT, JET] and list(ev["j"]) == [0, 1, 2]
    assert ev["r"][0] < rin and ev["r"][1] >= rout and ev["r"][2] >= rout
    assert 0.02 * math.pi <= ev["th"][1] <= 0.98 * math.pi and ev["th"][2] < 0.02 * math.pi
    assert np.all(ev["t"] == 2.0)
    assert o.tracers_events().size == 0
