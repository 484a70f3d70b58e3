"""The oracle's diag() log record (diag.c:59-326, diag_flux diag.c:437-465,
diag_mass diag.c:477-580) pinned by what the mathematics fixes:

  - |div B| of a field that is zero except one planted zone: the flux-CT corner
    stencil gives 0.25 (resp. 0.5 in 2D) delta gdet / dx on the eight (four)
    corners around it, all equal, so the reported zone is the first of them in
    ZLOOP order -- across the periodic phi seam when the plant sits in the last
    phi plane;
  - the conserved totals of a B = 0 state rotating in phi, against T^t_mu built
    in numpy from the metric tables (U[RHO] = gdet rho u^t, U[UU] = gdet (T^t_t +
    rho u^t), U[U3] = gdet T^t_phi);
  - mdot/edot/ldot of a dt = 0 step against the full-step fluxes of the same
    state (the face, the component and the signs of diag.c:452-457);
  - the ener.out line of the product library against C-format strings of
    diag.c:317-323 (Python's %-formatting follows C's printf).
"""
import ctypes
import math

import numpy as np
import pytest

from harm_cool_amd.params import torus_params
from oracle_lib import Oracle, lib as olib

CENT = 3  # decs.h:205-210: FACE1, FACE2, CORN, CENT, FACE3
RHO, UU, U1, U2, U3, B1, B2, B3 = range(8)


def _oracle(n1, n2, n3):
    olib().orc_set_threads(8)
    o = Oracle(torus_params(n1, n2, n3))
    assert o.init_torus() == 0
    return o


@pytest.mark.parametrize("n3,comp", [(4, B1), (4, B2), (4, B3), (1, B1), (1, B2)])
def test_planted_divb_value_and_first_zone(n3, comp):
    o = _oracle(12, 10, n3)
    p = o.prims().copy()
    p[B1:B3 + 1] = 0.0
    o.set_prims(p)
    d = o.diag_log()
    assert d["divbmax"] == 0.0 and d["imax"] == (-1, -1, -1)
    i0, j0, h0 = 6, 5, n3 - 1
    delta = 1.0e-3
    p[comp, i0, j0, h0] = delta
    o.set_prims(p)
    d = o.diag_log()
    _, _, gdet, _, _ = o.geometry()
    _, dx = o.grid()
    c = 0.25 if n3 > 2 else 0.5
    expect = abs(c * (delta * gdet[i0, j0, CENT]) / dx[1 + comp - B1])
    assert d["divbmax"] == expect
    assert d["divbmax"] == o.divb_max()
    # first corner in (h, i, j) order: in 3D the phi-seam corner h = 0, whose
    # h-1 is the plant's plane N3-1; in 2D the plant's own corner (h = 0)
    assert d["imax"] == (i0 - 2, j0 - 2, 0)


def _rotating_state(o, u3):
    p = o.prims().copy()
    p[U1] = p[U2] = 0.0
    p[U3] = u3
    p[B1:B3 + 1] = 0.0
    o.set_prims(p)
    return p


def test_conserved_totals_closed_form():
    o = _oracle(12, 10, 4)
    p = _rotating_state(o, 0.05)
    gcov, gcon, gdet, _, _ = o.geometry()
    _, dx = o.grid()
    gam = o.params.gam
    dV = dx[1] * dx[2] * dx[3]
    rmed = pp = e = 0.0
    terms = []
    for i in range(2, o.n1 + 2):
        for j in range(2, o.n2 + 2):
            gc, gn, g = gcov[i, j, CENT], gcon[i, j, CENT], gdet[i, j, CENT]
            alpha = 1.0 / math.sqrt(-gn[0, 0])
            beta = gn[0, 1:] * alpha * alpha
            for h in range(o.n3):
                ut = p[U1:U3 + 1, i, j, h]
                gamma = math.sqrt(1.0 + ut @ gc[1:, 1:] @ ut)
                ucon = np.empty(4)
                ucon[0] = gamma / alpha
                ucon[1:] = ut - gamma * beta / alpha
                ucov = gc @ ucon
                rho, u = p[RHO, i, j, h], p[UU, i, j, h]
                w = rho + u + (gam - 1.0) * u
                T_t = w * ucon[0] * ucov  # T^t_mu without the pressure term
                T_t[0] += (gam - 1.0) * u
                rmed += g * rho * ucon[0] * dV
                e += g * (T_t[0] + rho * ucon[0]) * dV
                pp += g * T_t[3] * dV
                terms.append(abs(g * w * ucon[0] * dV) * (1.0 + math.sqrt(gc[3, 3])))
    d = o.diag_log()
    scale = math.fsum(terms)
    assert abs(d["rmed"] - rmed) <= 1e-12 * abs(rmed)
    assert abs(d["e"] - e) <= 1e-13 * scale
    assert abs(d["pp"] - pp) <= 1e-13 * scale
    assert abs(pp) > 1e-6 * scale  # the rotation makes T^t_phi visible
    # diag_mass's total over the same zones
    m = sum(p[RHO, i, j, h] * gdet[i, j, CENT] * dx[1] * dx[2] * dx[3]
            for h in range(o.n3) for i in range(2, o.n1 + 2) for j in range(2, o.n2 + 2))
    assert abs(d["mass_curr"] - m) <= 1e-13 * m


def test_inner_face_fluxes_of_a_zero_dt_step():
    o = _oracle(16, 12, 4)
    d0 = o.diag_log()
    assert not d0["flux_valid"] and d0["mdot"] == d0["edot"] == d0["ldot"] == 0.0
    assert d0["mass_start"] == d0["mass_curr"]  # main.c:173-176 on the initial state
    _rotating_state(o, 0.05)  # angular momentum through the inner face
    F1, _, _, _ = o.debug_fluxes(1)
    _, dx = o.grid()
    # dt = 0: the half step leaves the state as it is (to the inversion's
    # tolerance), so the full-step fluxes are those of the initial state
    o2 = _oracle(16, 12, 4)
    _rotating_state(o2, 0.05)
    o2.set_time(0.0, 0.0)
    o2.step()  # reports dt < 1e-9 (step_ch.c:126) after completing the step
    d = o2.diag_log()
    assert d["flux_valid"]
    a = dx[2] * dx[3]
    mdot = F1[RHO, 2, 2:14, :].sum() * a
    edot = -(F1[UU, 2, 2:14, :] - F1[RHO, 2, 2:14, :]).sum() * a
    ldot = F1[U3, 2, 2:14, :].sum() * a
    for k, v in (("mdot", mdot), ("edot", edot), ("ldot", ldot)):
        assert v != 0.0
        assert abs(d[k] - v) <= 1e-8 * abs(v), k
    # the next face in would differ by far more than the tolerance
    assert abs(F1[RHO, 3, 2:14, :].sum() * a - mdot) > 1e-4 * abs(mdot)
    o2.set_prims(o2.prims())
    assert not o2.diag_log()["flux_valid"]


def test_mass_budget_and_ghost_fluxes_follow_the_state():
    o = _oracle(12, 10, 4)
    o.set_mass_budget(2.0, 0.25, -0.5)
    d = o.diag_log()
    assert (d["mass_start"], d["mass_in"], d["mass_out"]) == (2.0, 0.25, -0.5)
    dd = o.diag()
    assert d["mdot_in"] == dd["mdot_in"] and d["mdot_out"] == dd["mdot_out"]
    assert d["lnu"] == d["lnu_rho"] == d["lnu_rho_2"] == d["lnu2"] == 0.0


ENER_FMT = ("%10.5g %10.5g %10.5g %10.5g %15.8g %15.8g "
            "%15.8g %15.8g %15.8g %15.8g %15.8g %15.8g %15.8g"
            " -0 %15.8g %15.8g %15.8g %15.8g %15.8g %15.8g %15.8g\n")


def test_ener_line_format():
    import harm_cool_amd as hc
    L = hc.lib()
    rng = np.random.default_rng(5)
    for trial in range(20):
        r = hc.DiagRec()
        vals = rng.standard_normal(18) * 10.0 ** rng.integers(-20, 20, 18)
        for (n, _), v in zip(r._fields_[:18], vals):
            setattr(r, n, float(v))
        buf = ctypes.create_string_buffer(512)
        n = L.harm_ener_line(ctypes.byref(r), buf, len(buf))
        expect = ENER_FMT % (r.t, r.rmed, r.pp, r.e, 0.0, 0.0, r.mdot, r.edot, r.ldot, r.lnu, r.lnu_rho,
                             r.lnu_rho_2, r.lnu2, r.mdot_in, r.mdot_out, r.mass_start, r.mass_curr, r.mass_in,
                             r.mass_out, (r.mass_curr - r.mass_in - r.mass_out) / r.mass_start * 100.0)
        assert buf.value.decode() == expect and n == len(expect)
    small = ctypes.create_string_buffer(16)
    assert L.harm_ener_line(ctypes.byref(r), small, len(small)) == -1
