"""Cooling-table caches for the cache-EOS tests: closed-form e(rho, T) and
p(rho, T) (cgs) on the reference's constant grid (cooleos.c:884-930: 72 x 72
log-spaced points, rho in [1e2, 1e15] g/cm^3, T in [1e4, 1e18] K), one
(x, y) plane of D = 5 components (e, p, then three unused here).

  gas_rad   ideal gas of mean particle mass m_av with Gamma_g = 5/3 plus
            radiation: e = rho k T / (m_av (Gamma_g - 1)) + a T^4,
            p = rho k T / m_av + a T^4 / 3; per radial index i the pressure
            may be scaled by 1 + 0.02 i (distinct caches per radius)
  linear    e = A rho T, p = (gam - 1) A rho T: the T solve and the splines
            are exact up to round-off and the solve's 1e-6, so the pressure is
            the Gamma-law (gam - 1) u
"""
import numpy as np

KBOL = 1.3e-16  # cooleos.c:86
CL = 2.99792458e10
MAV = (1.67262163783e-24 + 1.67492729e-24 + 2. * 9.10956e-28 + 6.6465e-24) / 5.
A_RAD = 7.5657e-15
HH = 1.0e-6

RHO_RANGE = (1e2, 1e15)
T_RANGE = (1e4, 1e18)


def axes(nx=72, ny=72, rho=RHO_RANGE, T=T_RANGE, irregular=0):
    """Log-spaced axes; irregular = seed: the interior points moved at random
    within their log intervals and a run of closely spaced points in each (an
    increasing axis that no formula indexes)."""
    vx, vy = np.geomspace(rho[0], rho[1], nx), np.geomspace(T[0], T[1], ny)
    if irregular:
        rng = np.random.default_rng(irregular)
        out = []
        for v in (vx, vy):
            lv = np.log(v)
            d = lv[1] - lv[0]
            lv[1:-1] += rng.uniform(-0.45, 0.45, lv.size - 2) * d
            k = lv.size // 3
            lv[k:k + 6] = lv[k] + np.arange(6) * 0.02 * d
            out.append(np.exp(np.sort(lv)))
        vx, vy = out
    return vx, vy


def gas_rad(nx=72, ny=72, n=1, rho=RHO_RANGE, T=T_RANGE, gam_g=5.0 / 3.0, irregular=0):
    vx, vy = axes(nx, ny, rho, T, irregular)
    R, TT = np.meshgrid(vx, vy, indexing="ij")
    e = R * KBOL * TT / (MAV * (gam_g - 1.0)) + A_RAD * TT ** 4
    p = R * KBOL * TT / MAV + A_RAD * TT ** 4 / 3.0
    vals = np.zeros((n, nx, ny, 5))
    for i in range(n):
        vals[i, :, :, 0] = e
        vals[i, :, :, 1] = p * (1.0 + 0.02 * i)
    return vx, vy, vals


def linear(A, gam, nx=72, ny=72, n=1, rho=RHO_RANGE, T=T_RANGE):
    vx, vy = axes(nx, ny, rho, T)
    R, TT = np.meshgrid(vx, vy, indexing="ij")
    vals = np.zeros((n, nx, ny, 5))
    vals[:, :, :, 0] = A * R * TT
    vals[:, :, :, 1] = (gam - 1.0) * A * R * TT
    return vx, vy, vals


def first_guess(gam, rho0, u, T_min):
    """pressure_rho0_u_CoolEOS's start for an unset last T (cooleos.c:341-348),
    in the reference's order of operations."""
    T = (gam - 1.) * u / rho0 / KBOL * MAV * CL * CL * 1.0
    lo = T_min * (1.0 + HH)
    return lo if T < lo else T
