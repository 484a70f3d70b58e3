"""PPM reconstruction (lim 3) of the oracle, pinned by closed forms (CPU).

The reference has no PPM (interp.c:49-93 offers MC, van Leer and minmod); this
option follows Colella & Woodward (1984, J. Comput. Phys. 54, 174), eqs. 1.6 and
1.10, on the reference's MC slopes.  What the mathematics fixes:

  cubic averages  for cell averages of a cubic, where MC picks the centred slope
                  and the profile is monotone, the interface value
                  (y(m) + y(m+1))/2 - (dq(m+1) - dq(m))/6 is the cubic at the
                  interface (the fourth-order formula is exact for cubics);
  monotone        on step data every face value lies between the zone's value
                  and its neighbour's across that face (no new extrema);
  constraints     a zone at a local extremum is flattened (l = r = y0); where the
                  parabola would overshoot beyond one face, the value on the
                  other face becomes 3 y0 - 2 (that face's value), which puts the
                  parabola's extremum on the overshooting face (zero slope there);
  wiring          the oracle's fluxes at faces whose two zones both have the
                  five-zone stencil equal face_flux of the exact cubic values; at
                  the first and last radial faces the zone whose stencil leaves
                  the two ghost layers keeps the MC state.

test_ppm_pins_catch_planted_mistakes builds the oracle with each planted mistake
(ORC_MUTANT in harm_oracle.c) and shows that the pin named for it fails.
"""
import ctypes
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, _ptr, build_mutant, lib, lib_at

PPM, MC = 3, 0
FACE1, FACE2, CORN, CENT, FACE3 = range(5)


def _para(L, ym2, ym, y0, yp, yp2):
    """Face values of the middle zone of five, with the oracle's MC slopes."""
    dqm = L.orc_slope_lim(MC, ym2, ym, y0)
    dq0 = L.orc_slope_lim(MC, ym, y0, yp)
    dqp = L.orc_slope_lim(MC, y0, yp, yp2)
    l, r = ctypes.c_double(), ctypes.c_double()
    L.orc_para.restype = ctypes.c_double
    L.orc_para.argtypes = [ctypes.c_double] * 6 + [ctypes.POINTER(ctypes.c_double)] * 2
    L.orc_para(ym, y0, yp, dqm, dq0, dqp, ctypes.byref(l), ctypes.byref(r))
    return l.value, r.value


def _cubic(c):
    """f(x) = c0 + c1 x + c2 x^2 + c3 x^3, its cell average over [m-1/2, m+1/2]."""
    f = lambda x: c[0] + c[1] * x + c[2] * x**2 + c[3] * x**3
    F = lambda x: c[0] * x + c[1] * x**2 / 2 + c[2] * x**3 / 3 + c[3] * x**4 / 4
    avg = lambda m: F(m + 0.5) - F(m - 0.5)
    return f, avg


# gentle monotone cubics (x in zone units around the middle of the range):
# MC picks the centred slope and no monotonicity constraint is active
CUBICS = [(1.0, 0.02, 0.003, 0.0002), (0.5, -0.01, 0.001, -0.00005), (-0.1, 0.004, -0.0002, 0.00001)]


def check_cubic_averages(L) -> bool:
    for c in CUBICS:
        f, avg = _cubic(c)
        for m in range(-6, 7):
            y = [avg(m + d) for d in (-2, -1, 0, 1, 2)]
            l, r = _para(L, *y)
            if abs(l - f(m - 0.5)) > 1e-13 * abs(f(m - 0.5)) + 1e-15 or abs(r - f(m + 0.5)) > 1e-13 * abs(
                    f(m + 0.5)) + 1e-15:
                return False
    return True


def check_monotone_steps(L) -> bool:
    rng = np.random.default_rng(7)
    for _ in range(300):
        n = 12
        y = np.full(n, rng.uniform(-1, 1))
        s = rng.integers(3, n - 3)
        y[s:] += rng.uniform(-2, 2)
        y += rng.uniform(-1e-3, 1e-3) * (rng.random() < 0.5) * np.arange(n)  # a ramp under some steps
        for m in range(2, n - 2):
            l, r = _para(L, *y[m - 2:m + 3])
            lo_l, hi_l = sorted((y[m - 1], y[m]))
            lo_r, hi_r = sorted((y[m], y[m + 1]))
            if not (lo_l - 1e-15 <= l <= hi_l + 1e-15 and lo_r - 1e-15 <= r <= hi_r + 1e-15):
                return False
    return True


def check_constraints(L) -> bool:
    # extremum: flattened
    l, r = _para(L, 0.0, 0.5, 1.0, 0.5, 0.0)
    if not (l == 1.0 and r == 1.0):
        return False
    # the parabola through (l, r) with mean 1 overshoots beyond the right face:
    # the left value becomes 3 y0 - 2 r and the slope at the right face is 0
    # (MC slopes 0, 0.4, 0.2 by hand: l0 = 0.5 - 0.4/6, r = 1.1 + 0.2/6)
    l, r = _para(L, 0.0, 0.0, 1.0, 1.2, 1.4)
    ok = abs(r - (1.1 + 0.2 / 6)) < 1e-15 and l == 3.0 * 1.0 - 2.0 * r
    da, a6 = r - l, 6 * (1.0 - 0.5 * (l + r))
    ok = ok and abs(da - a6) < 1e-14
    # the mirror image: the right value becomes 3 y0 - 2 l, zero slope on the left face
    l2, r2 = _para(L, 1.4, 1.2, 1.0, 0.0, 0.0)
    ok = ok and l2 == r and r2 == 3.0 * 1.0 - 2.0 * l2
    da, a6 = r2 - l2, 6 * (1.0 - 0.5 * (l2 + r2))
    return bool(ok and abs(da + a6) < 1e-14)


def _face_flux(L, pl, pr, gcov, gcon, gdet, d, gam):
    F = np.empty(8)
    ct = np.empty(1)
    L.orc_face_flux(_ptr(np.ascontiguousarray(pl)), _ptr(np.ascontiguousarray(pr)), _ptr(gcov), _ptr(gcon), gdet, d,
                    gam, 1.0, 0.0, _ptr(F), _ptr(ct))
    return F


def _wiring_case(L, d):
    """Primitives that are cell averages of cubics along direction d (constant
    across it); returns the oracle's fluxes and the exact face values."""
    if d == 3:
        P = torus_params(8, 8, 16, lim=PPM)
    else:
        P = torus_params(16, 12, 1, lim=PPM)
    o = Oracle(P, L)
    shape = o.shape  # (8, N1+4, N2+4, N3)
    n = shape[1 + (d - 1) if d < 3 else 3]
    # f = a + b g(x/20), g(t) = t + 0.3 t^2 + 0.1 t^3: g' >= 0.7 on |t| <= 0.6, so
    # each profile is monotone, MC takes the centred slope and no constraint acts
    ab = [(1.0, 0.1), (0.4, -0.04), (-0.05, 0.01), (0.01, -0.005), (0.02, 0.004), (0.1, 0.02), (0.05, -0.01),
          (0.03, 0.006)]
    mid = (n - 1) / 2.0
    p = np.empty(shape)
    exact = np.empty((8, n + 1))  # exact[k][f]: value at the face between zones f-1 and f
    for k in range(8):
        a, b = ab[k]
        f, avg = _cubic((a, b / 20, 0.3 * b / 400, 0.1 * b / 8000))
        prof = np.array([avg(m - mid) for m in range(n)])
        exact[k] = [f(m - 0.5 - mid) for m in range(n + 1)]
        if d == 1:
            p[k] = prof[:, None, None]
        elif d == 2:
            p[k] = prof[None, :, None]
        else:
            p[k] = prof[None, None, :]
    o.set_prims(p)
    F1, F2, F3 = o.debug_fluxes(0)[:3]
    gcov, gcon, gdet = o.geometry()[:3]
    return o, P, p, (F1, F2, F3)[d - 1], exact, gcov, gcon, gdet


def check_wiring(L) -> bool:
    ok = True
    for d in (1, 2, 3):
        o, P, p, F, exact, gcov, gcon, gdet = _wiring_case(L, d)
        loc = (FACE1, FACE2, FACE3)[d - 1]
        is_, ie = 2, 2 + P.n1 - 1
        js, je = 2, 2 + P.n2 - 1
        if d == 1:
            faces = [(i, js + 3, 0) for i in range(is_ + 1, ie + 1)]
        elif d == 2:
            faces = [(is_ + 3, j, 0) for j in range(js + 1, je + 1)]
        else:
            faces = [(is_ + 3, js + 3, h) for h in range(3, P.n3 - 2)]
        for (i, j, h) in faces:
            f = (i, j, h)[d - 1]
            want = _face_flux(L, exact[:, f], exact[:, f], gcov[i, j, loc], gcon[i, j, loc], gdet[i, j, loc], d, P.gam)
            got = F[:5, i, j, h]
            ok = ok and np.allclose(got, want[:5], rtol=1e-10, atol=1e-14)
        if d == 1:
            # first radial face: zone is-1 keeps its MC state, zone is has PPM's;
            # last radial face: zone ie+1 keeps its MC state
            j = js + 3
            for i, lin_left in ((is_, True), (ie + 1, False)):
                zl, zr = i - 1, i
                if lin_left:
                    dq = np.array([L.orc_slope_lim(MC, *p[k, zl - 1:zl + 2, j, 0]) for k in range(8)])
                    pl, pr = p[:, zl, j, 0] + 0.5 * dq, exact[:, i]
                else:
                    dq = np.array([L.orc_slope_lim(MC, *p[k, zr - 1:zr + 2, j, 0]) for k in range(8)])
                    pl, pr = exact[:, i], p[:, zr, j, 0] - 0.5 * dq
                want = _face_flux(L, pl, pr, gcov[i, j, FACE1], gcon[i, j, FACE1], gdet[i, j, FACE1], 1, P.gam)
                # fix_flux may zero the mass flux at these two faces (coord.c:310-320)
                ok = ok and np.allclose(F[1:5, i, j, 0], want[1:5], rtol=1e-10, atol=1e-14)
                ok = ok and (F[0, i, j, 0] == 0 or np.isclose(F[0, i, j, 0], want[0], rtol=1e-10))
        del o
    return bool(ok)


PINS = {
    "cubic_averages": check_cubic_averages,
    "monotone_steps": check_monotone_steps,
    "constraints": check_constraints,
    "wiring": check_wiring,
}

MUTANTS = {
    13: ("interface value with (dq(m+1) - dq(m))/8", "cubic_averages"),
    14: ("overshoot branch sets l = 3 y0 - 2 l", "constraints"),
    15: ("the first active radial zone keeps its MC state", "wiring"),
}


@pytest.mark.parametrize("pin", list(PINS))
def test_ppm_pin_holds(pin):
    assert PINS[pin](lib())


def test_ppm_reduces_to_mc_state_on_linear_data():
    """On a line MC picks the centred slope and PPM's interfaces are the line's
    values: both give the same face states."""
    L = lib()
    y = [0.3 + 0.1 * m for m in range(5)]
    l, r = _para(L, *y)
    assert abs(l - 0.45) < 1e-15 and abs(r - 0.55) < 1e-15


def test_ppm_pins_catch_planted_mistakes(tmp_path):
    with ThreadPoolExecutor(max_workers=3) as ex:
        sos = dict(zip(MUTANTS, ex.map(lambda n: build_mutant(n, str(tmp_path)), MUTANTS)))
    missed = [f"{n}: {what} (pin {pin})" for n, (what, pin) in MUTANTS.items() if PINS[pin](lib_at(sos[n]))]
    assert not missed, "pins that did not catch their planted mistake: " + "; ".join(missed)
