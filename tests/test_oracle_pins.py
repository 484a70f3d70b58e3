"""Pins for the oracle's 3D-only and fix-up branches (CPU).

The whole reference step cannot be built here (it calls MPI unguarded and takes
its configuration from compile-time edits of its sources), so these branches are
pinned by what the mathematics fixes instead of by a reference run:

  phi translation  the step commutes with a shift of the state by s phi planes
                   (the metric does not depend on phi); pins the periodic seam of
                   the X3 reconstruction, the X3 divergence and the 3D flux-CT.
  phi reflection   at a = 0 the step commutes with phi -> -phi, U3 and B3 odd;
                   pins the phi+pi plane of the pole ghosts (an off-by-one plane
                   breaks it) and the phi-plane pairing of the 3D EMFs.
  pole ghosts      the theta ghost zones of a field that is smooth in Cartesian
                   coordinates hold its continuation through the pole,
                   (theta, phi) -> (-theta, phi): pins the phi+pi source plane and
                   the U2/B2 sign flip (bounds.c:161-275).
  polefix          a U2 that falls linearly to 0 at the pole is kept by the
                   POLEFIX rows, which the reference fills by linear interpolation
                   towards the pole from row jstart+2 (bounds.c:381-436).
  3D flux-CT       corner div B stays at round-off over steps of a rotating
                   magnetized 3D torus (step_ch.c:726-798, diag.c:156-202).
  fixup_utoprim    hand-built clusters of bad zones over a field linear in (i, j)
                   (4-point averages of a linear field are exact) with sentinel
                   values in the bad zones: the repaired values, which branch each
                   zone took and the order of the walk follow by hand from
                   fixup.c:493-534, including the repaired zone staying "bad".

test_pins_catch_planted_mistakes builds the oracle with each planted mistake
(ORC_MUTANT, harm_oracle.c) and shows that the pin named for it fails.
"""
import ctypes
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, _ptr, build_mutant, lib, lib_at

RHO, UU, U1, U2, U3, B1, B2, B3 = range(8)


def _evolved(L, P, nsteps):
    o = Oracle(P, L)
    assert o.init_torus() == 0
    for _ in range(nsteps):
        assert o.step() == 0
    return o.prims(), o.pflag(), o.time()


def _run(L, P, prim, flag, t, nsteps):
    o = Oracle(P, L)
    o.set_prims(prim, flag)
    o.set_time(*t)
    for _ in range(nsteps):
        assert o.step() == 0
    return o.prims(), o.pflag(), o.time(), o.counters()


def check_phi_translation(L) -> bool:
    """Shift by 3 planes, step twice, shift back: bitwise the unshifted run."""
    P = torus_params(16, 16, 8)
    p, f, t = _evolved(L, P, 1)
    pa, fa, ta, ca = _run(L, P, p, f, t, 2)
    pb, fb, tb, cb = _run(L, P, np.roll(p, 3, axis=3), np.roll(f, 3, axis=2), t, 2)
    return (np.array_equal(np.roll(pa, 3, axis=3), pb) and np.array_equal(np.roll(fa, 3, axis=2), fb)
            and ta == tb and ca == cb)


def _reflect(p):
    q = p[..., ::-1].copy()
    q[U3] *= -1.0
    q[B3] *= -1.0
    return q


def check_phi_reflection(L) -> bool:
    """a = 0: reflecting phi (U3, B3 odd) commutes with two steps, equal as
    values (signed zeros may differ)."""
    P = torus_params(16, 16, 8, a=0.0)
    p, f, t = _evolved(L, P, 1)
    pa, fa, ta, ca = _run(L, P, p, f, t, 2)
    pb, fb, tb, cb = _run(L, P, _reflect(p), f[..., ::-1].copy(), t, 2)
    return bool(np.all(_reflect(pa) == pb)) and np.array_equal(fa[..., ::-1], fb) and ta == tb and ca == cb


def _centres(L, P):
    """Code coordinates X1, X2, X3, r, theta and the Jacobian d(r,theta,phi)/dX at
    every zone centre, ghosts included."""
    o = Oracle(P, L)
    sx, dx = o.grid()
    n1, n2, n3 = P.n1 + 4, P.n2 + 4, P.n3
    X = np.zeros((4,))
    V = np.zeros((4,))
    J = np.zeros((16,))
    r = np.zeros((n1, n2))
    th = np.zeros((n1, n2))
    d11 = np.zeros((n1, n2))
    d22 = np.zeros((n1, n2))
    for i in range(n1):
        for j in range(n2):
            X[:] = (0.0, sx[1] + (i - 2 + 0.5) * dx[1], sx[2] + (j - 2 + 0.5) * dx[2], 0.0)
            L.orc_bl_coord_vec(ctypes.byref(P), _ptr(X), _ptr(V))
            L.orc_dxdxp_func(ctypes.byref(P), _ptr(X), _ptr(J))
            r[i, j], th[i, j] = V[1], V[2]
            d11[i, j], d22[i, j] = J[5], J[10]
            # r depends on X1 only, theta on X2 only, phi = J[15] X3
            assert J[6] == 0.0 and J[9] == 0.0 and J[7] == J[11] == J[13] == J[14] == 0.0
    phi = J[15] * (sx[3] + (np.arange(n3) + 0.5) * dx[3])
    return r, th, phi, d11, d22, J[15]


def _xhat_field(r, th, phi, d11, d22, d33):
    """A uniform Cartesian field along x in the code basis: B^r = sin th cos ph,
    B^th = cos th cos ph / r, B^ph = -sin ph / (r sin th)."""
    R, TH, PH = r[..., None], th[..., None], phi[None, None, :]
    b1 = np.sin(TH) * np.cos(PH) / d11[..., None]
    b2 = np.cos(TH) * np.cos(PH) / R / d22[..., None]
    b3 = -np.sin(PH) / (R * np.sin(TH)) / d33
    return b1, b2, b3


def check_pole_continuation(L) -> bool:
    P = torus_params(8, 16, 8)
    r, th, phi, d11, d22, d33 = _centres(L, P)
    shape = (8, P.n1 + 4, P.n2 + 4, P.n3)
    p = np.zeros(shape)
    p[RHO] = 1.0
    p[UU] = 0.1
    b = _xhat_field(r, th, phi, d11, d22, d33)
    js, je = 2, P.n2 + 1
    act = np.zeros(shape[1:], dtype=bool)
    act[:, js:je + 1, :] = True
    for k, v in zip((B1, B2, B3), b):
        p[k] = np.where(act, v, 12345.0)  # theta ghosts hold garbage before bound
    o = Oracle(P, L)
    o.set_prims(p, np.zeros(shape[1:], dtype=np.int8))
    L.orc_debug_bound(o.h)
    q = o.prims()
    ok = True
    for k, v in zip((B1, B2, B3), b):
        scale = np.abs(v[2:-2, js:je + 1]).max()
        for j in (js - 2, js - 1, je + 1, je + 2):
            err = np.abs(q[k, 2:-2, j] - v[2:-2, j]).max() / scale
            ok = ok and err < 1e-9
    return bool(ok)


def check_polefix_linear(L) -> bool:
    """U2 = c (distance in X2 to the nearer pole) is kept in the POLEFIX rows;
    rho, u, U1, U3 there take row jstart+2's values (the reference's copy)."""
    P = torus_params(8, 16, 8)
    o = Oracle(P, L)
    sx, dx = o.grid()
    shape = (8, P.n1 + 4, P.n2 + 4, P.n3)
    js, je = 2, P.n2 + 1
    j = np.arange(P.n2 + 4)
    x2 = sx[2] + (j - 2 + 0.5) * dx[2]
    dist = np.minimum(x2 - sx[2], sx[2] + P.n2 * dx[2] - x2)
    p = np.zeros(shape)
    p[RHO] = (1.0 + 0.01 * j)[None, :, None]
    p[UU] = (0.1 + 0.001 * j)[None, :, None]
    p[U1] = 0.02
    p[U2] = (0.05 * dist)[None, :, None]
    p[U3] = (0.03 + 0.0001 * j)[None, :, None]
    p[B1] = 1e-3
    before = p.copy()
    o.set_prims(p, np.zeros(shape[1:], dtype=np.int8))
    L.orc_debug_bound(o.h)
    q = o.prims()
    ok = True
    for jj in (js, js + 1, je - 1, je):
        ok = ok and np.allclose(q[U2, 2:-2, jj], before[U2, 2:-2, jj], rtol=1e-13, atol=0.0)
        src = js + 2 if jj < P.n2 // 2 else je - 2
        for k in (RHO, UU, U1, U3):
            ok = ok and np.array_equal(q[k, 2:-2, jj], before[k, 2:-2, src])
    return bool(ok)


def check_ct3_divb(L) -> bool:
    P = torus_params(24, 16, 8)
    o = Oracle(P, L)
    assert o.init_torus() == 0
    d0 = o.divb_max()
    for _ in range(4):
        o.step()
    return d0 < 1e-12 and o.divb_max() < 1e-12


# fixup_utoprim clusters: (i, j) zones of a 2D 16x16 grid (active 2..17)
ISOLATED = [(5, 5)]
QUIRK = [(9, 5), (10, 5), (11, 6)]    # P, X, Q
ORDER = [(5, 12), (5, 13), (6, 11)]   # X, Y, D
SENTINEL = np.array([1024.0, 512.0, 0.5, -0.25, 0.125])


def _linear(i, j):
    """rho, u, U1, U2, U3 linear in (i, j) with exactly representable values."""
    return np.array([2.0 + 0.25 * i + 0.125 * j, 1.0 + 0.0625 * i + 0.5 * j, (i + 2 * j) / 1024.0,
                     (3 * i - j) / 2048.0, (i + j) / 4096.0])


def _expected_fixup():
    """Repaired rho..U3 of each bad zone, worked out from fixup.c:493-534.

    Walk order (i, j).  A repaired zone keeps the flag "bad" (fixup.c:530), so
    the branch a zone takes depends only on which zones were bad; the values an
    average reads depend on whether a bad neighbour was repaired before it.
      (5,5)  isolated: "+" average = f.
      P=(9,5): "+" holds X=(10,5) bad; "x" neighbours good: "x" average = f.
      X=(10,5): "+" holds P (bad, repaired); "x" holds Q=(11,6), bad and not yet
             repaired: v = 0 branch, rho,u = (AVG4_1 + AVG4_2)/2 with P's repaired
             value (= f) in AVG4_1 and Q's sentinel in AVG4_2; U = 0.
      Q=(11,6): "+" neighbours good: f.
      X=(5,12): "+" holds Y=(5,13), "x" holds D=(6,11), both not yet repaired:
             v = 0 branch with Y's sentinel in AVG4_1 and D's in AVG4_2.
      Y=(5,13): "+" holds X (bad); "x" good: f.
      D=(6,11): "+" good: f.
    Two zones take the v = 0 branch, seven are repaired."""
    f = _linear
    S = SENTINEL
    out = {}
    for z in ISOLATED + [QUIRK[0], QUIRK[2], ORDER[1], ORDER[2]]:
        out[z] = f(*z)
    i, j = QUIRK[1]
    a1 = 0.25 * (f(i, j + 1) + f(i, j - 1) + f(i - 1, j) + f(i + 1, j))
    a2 = 0.25 * (S + f(i + 1, j - 1) + f(i - 1, j + 1) + f(i - 1, j - 1))
    v = 0.5 * (a1 + a2)
    v[2:] = 0.0
    out[QUIRK[1]] = v
    i, j = ORDER[0]
    a1 = 0.25 * (S + f(i, j - 1) + f(i - 1, j) + f(i + 1, j))
    a2 = 0.25 * (f(i + 1, j + 1) + S + f(i - 1, j + 1) + f(i - 1, j - 1))
    v = 0.5 * (a1 + a2)
    v[2:] = 0.0
    out[ORDER[0]] = v
    return out


def check_fixup_clusters(L) -> bool:
    P = torus_params(16, 16, 1)
    o = Oracle(P, L)
    shape = (8, 20, 20, 1)
    p = np.zeros(shape)
    for i in range(20):
        for j in range(20):
            p[:5, i, j, 0] = _linear(i, j)
    flags = np.zeros(shape[1:], dtype=np.int8)
    bad = ISOLATED + QUIRK + ORDER
    for (i, j) in bad:
        p[:5, i, j, 0] = SENTINEL
        flags[i, j, 0] = 1
    o.set_prims(p, flags)
    c0 = o.counters()
    L.orc_debug_fixup_utoprim(o.h)
    q = o.prims()
    c1 = o.counters()
    want = _expected_fixup()
    ok = all(np.array_equal(q[:5, i, j, 0], want[(i, j)]) for (i, j) in bad)
    good = np.ones(shape[1:], dtype=bool)
    for (i, j) in bad:
        good[i, j, 0] = False
    good[:2] = good[-2:] = False
    good[:, :2] = good[:, -2:] = False
    ok = ok and np.array_equal(q[:, good], p[:, good]) and not q[B1:].any()
    ok = ok and c1["fixup_repairs"] - c0["fixup_repairs"] == 7
    ok = ok and c1["fixup_nostencil"] - c0["fixup_nostencil"] == 2
    return bool(ok)


PINS = {
    "phi_translation": check_phi_translation,
    "phi_reflection": check_phi_reflection,
    "pole_continuation": check_pole_continuation,
    "polefix_linear": check_polefix_linear,
    "ct3_divb": check_ct3_divb,
    "fixup_clusters": check_fixup_clusters,
}

# planted mistake (harm_oracle.c MUT(n)) -> the pin that must catch it
MUTANTS = {
    1: ("pole ghosts from plane h + N3/2 + 1", "phi_reflection"),
    2: ("pole ghosts from plane h (no phi+pi shift)", "pole_continuation"),
    3: ("POLEFIX U2 weight (j+0.5)/(jref+1)", "polefix_linear"),
    4: ("POLEFIX rows copied from jstart+1", "polefix_linear"),
    5: ("3D EMF e1 pairs plane h with h+1", "phi_reflection"),
    6: ("3D flux-CT F3[B1] with the wrong sign", "ct3_divb"),
    7: ("fixup_utoprim marks a repaired zone good", "fixup_clusters"),
    8: ("fixup_utoprim walks (h, j, i)", "fixup_clusters"),
    9: ("fixup_utoprim AVG4_2 with (i+1, j) for (i+1, j+1)", "fixup_clusters"),
    10: ("3D EMFs clamp h-1 at the seam", "phi_translation"),
    11: ("X3 slope clamps h+1 at the seam", "phi_translation"),
    12: ("X3 flux divergence clamps h+1 at the seam", "phi_translation"),
}


@pytest.mark.parametrize("pin", list(PINS))
def test_pin_holds(pin):
    assert PINS[pin](lib())


def test_pins_catch_planted_mistakes(tmp_path):
    with ThreadPoolExecutor(max_workers=6) as ex:
        sos = dict(zip(MUTANTS, ex.map(lambda n: build_mutant(n, str(tmp_path)), MUTANTS)))
    caught = {}
    for n, (what, pin) in MUTANTS.items():
        caught[n] = not PINS[pin](lib_at(sos[n]))
    missed = [f"{n}: {MUTANTS[n][0]} (pin {MUTANTS[n][1]})" for n, c in caught.items() if not c]
    assert not missed, "pins that did not catch their planted mistake: " + "; ".join(missed)
