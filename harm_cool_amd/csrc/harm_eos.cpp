// harm_eos.cpp -- a closed-form EOS table for the tabulated step (host).
//
// The reference fills its per-radius tables from a Fortran Fermi-gas /
// nuclear-statistical-equilibrium code (abundance.f90, eos.f90,
// gauss_integral.f90), which this library restates and runs on the GPU
// (harm_eos_table_fermi, harm_fermi.hip).  This function fills the same kind of
// table -- e(rho, T) and p(rho, T) on a log-spaced grid, Ye fixed -- from a
// closed-form EOS instead, for tests of the interpolant and the tabulated step
// that need no generator run: ultra-relativistic degenerate electrons
// (p = K (Ye rho)^(4/3), e = 3p), an ideal gas of nucleons and electrons
// ((1 + Ye) rho/m_u k T, e = 3/2 p), and radiation (a T^4 / 3, e = a T^4).
// e grows strictly with T at fixed rho, which the inversion p(rho, u) needs.
// Values are returned in code units: rho / rho_unit, e and p / (rho_unit c^2).
#include <cmath>
#include <string>

#include "harm_host.hpp"

extern "C" int harm_eos_table_build(const harm_eos_spec *s, double *rho, double *T, double *e, double *p)
{
    if (!s || !rho || !T || !e || !p) return harm_host::set_error("null argument", 1);
    if (s->nr < 2 || s->nt < 2 || !(s->rho_unit > 0.) || !(s->lrho_hi > s->lrho_lo) || !(s->lT_hi > s->lT_lo) ||
        !(s->ye > 0. && s->ye <= 1.))
        return harm_host::set_error("bad EOS table spec (nr, nt >= 2, increasing ranges, 0 < ye <= 1)", 4);
    const double m_u = 1.66053906660e-24, k_B = 1.380649e-16, a_rad = 7.565723e-15, c = 2.99792458e10;
    const double h = 6.62607015e-27;
    const double K_rel = std::cbrt(3. / M_PI) * h * c / 8. / std::pow(m_u, 4. / 3.);
    const double u_unit = s->rho_unit * c * c;
    for (int i = 0; i < s->nr; i++)
        rho[i] = std::pow(10., s->lrho_lo + (s->lrho_hi - s->lrho_lo) * i / (s->nr - 1)) / s->rho_unit;
    for (int j = 0; j < s->nt; j++) T[j] = std::pow(10., s->lT_lo + (s->lT_hi - s->lT_lo) * j / (s->nt - 1));
    for (int i = 0; i < s->nr; i++) {
        const double r = rho[i] * s->rho_unit;
        const double p_deg = K_rel * std::pow(s->ye * r, 4. / 3.);
        for (int j = 0; j < s->nt; j++) {
            const double p_gas = (1. + s->ye) * r / m_u * k_B * T[j];
            const double p_rad = a_rad * std::pow(T[j], 4.) / 3.;
            p[(size_t)i * s->nt + j] = (p_deg + p_gas + p_rad) / u_unit;
            e[(size_t)i * s->nt + j] = (3. * p_deg + 1.5 * p_gas + 3. * p_rad) / u_unit;
        }
    }
    return 0;
}
