// harm_ceos.hpp -- the step's COOL_EOS pressure on the cooling-table cache
// (cooleos.c:314-444, 667-741): p(rho0, u) is found by a temperature Newton
// solve on the cache of the zone's radial index, started from the zone's last
// temperature, which every converged solve replaces (the reference's LastT,
// selected with the cache by InitCache_CoolEOS, cooleos.c:770-807).
//
// Every pressure call of a face or a zone threads one temperature through the
// calls in the reference's order.  A call repeated with the same (rho0, u)
// returns the same bits and leaves the temperature as it was: after a
// converged solve the start already meets the tolerance (the solve returns it
// without a step), after a failed one the temperature was not replaced, and
// the range branches do not depend on it.  So the last call's result is kept
// and a repeat costs nothing; this is what makes the finite-difference
// derivatives cheap, whose step rho0 hh / RHO_UNIT (u hh / U_UNIT) is far below
// one ulp of rho0 (u) at the reference's units, so both their pressure calls
// repeat the one before and the derivatives are exactly 0 (or NaN at u = 0,
// as in the reference).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "harm_cache.hpp"

namespace harm {

// cooleos.c:72, 86-95
constexpr double CEOS_HH = 1.0E-6;
constexpr double CEOS_KBOL = 1.3e-16;
constexpr double CEOS_CL = 2.99792458e10;
constexpr double CEOS_MAV = (1.67262163783e-24 + 1.67492729e-24 + 2. * 9.10956e-28 + 6.6465e-24) / 5.;
constexpr int CEOS_NTRIAL = 300;     // cooleos.c:395
constexpr double CEOS_TOL = 1.0E-6;  // cooleos.c:395, 406, 409

// Where a bracket walk along an axis of positive values starts: the axis's
// bracket (cache_bracket) of the smallest double of each key, key = the bits
// of t >> shift (the exponent and the top mantissa bits, so a key spans a
// fixed fraction of an octave).  The walk finds the same bracket from any
// start; a start from the table is one or two reads away from it where the
// bisection takes log2(n) dependent reads.  nk = 0: no table (bisection).
struct CeosHint {
    const int *idx;
    unsigned long long k0;  // key of idx[0]
    int nk, shift;
};

// One run's cache EOS: the grid every radial index's cache shares, the values
// of each (n caches: 1 shared by every radius, or one per radial index of the
// grid, ghosts included), the units and the ranges of cooleos.c:826-930.
struct CeosParams {
    CacheView view;                // axes, sizes, types; view.values is unused
    const double *const *values;   // device array of n value pointers ([1][X][Y][D] each)
    const double *const *yco;      // n coefficient tables of e and p along T ([2][X][Y][12] each)
    int n;
    double rho_unit, u_unit, p_unit, gam;
    double rho_min, rho_max;       // rhoMin, rhoMax (g/cm^3)
    double T_lo, T_hi;             // TMin (1 + hh) (also the floor of a first guess), TMax (1 - hh)
    CeosHint hx, hy;               // bracket starts along rho and T
};

// The state a face or a zone carries through its pressure calls.
struct CeosZone {
    double T;  // the zone's last temperature (0: never set)
    unsigned long long kr = 0, ku = 0;
    double p = 0.;
    bool has = false;
    int nsolve = 0;  // solves run (repeats excluded), for the counters
};

__device__ __forceinline__ int ceos_radius(const CeosParams &P, int i) { return P.n == 1 ? 0 : i; }

struct CeosOut {
    double p, T;
    int solved;
};

// The bracket of t (cache_bracket's) found by walking from b in [1, n-1]:
// down while the bracket below also holds t, else up to the first one that
// does (after a step down t < v[b] holds, so the up walk would not move).  The
// axis increases, so this is the bisection's index; both neighbours of the
// start are read at once, and a start near t needs no more.
// Reads through a global-memory pointer: the out-of-line solve sees generic
// pointers, whose loads each wait for every load before them.
template <class T>
__device__ __forceinline__ T ceos_ld(const T *p, int i)
{
    return ((const __attribute__((address_space(1))) T *)p)[i];
}

__device__ __forceinline__ int ceos_walk_at(int n, const double *v, double t, int b, double below, double at)
{
    if (b > 1 && t < below) {
        b--;
        while (b > 1 && t < ceos_ld(v, b - 1)) b--;
    } else if (b < n - 1 && !(t < at)) {
        b++;
        while (b < n - 1 && !(t < ceos_ld(v, b))) b++;
    }
    return b;
}

__device__ __forceinline__ int ceos_walk(int n, const double *v, double t, int b)
{
    return ceos_walk_at(n, v, t, b, ceos_ld(v, b - 1), ceos_ld(v, b));
}

// The hint table's start for t (a table with nk > 0): any t, NaN and values
// off the axis included, reads an entry inside the table.
__device__ __forceinline__ int ceos_hint_start(const CeosHint &h, double t)
{
    const unsigned long long k = (unsigned long long)__double_as_longlong(t) >> h.shift;
    const unsigned long long d = k < h.k0 ? 0ull : k - h.k0;
    return ceos_ld(h.idx, d >= (unsigned long long)h.nk ? h.nk - 1 : (int)d);
}

template <bool DER>
__device__ __forceinline__ bool ceos_win(const CacheView &c, const double *yco, int x0, int nx, int yb, double rho,
                                         double t, double &v, double &dv)
{
    const int y0 = yb - 2 < 0 ? 0 : yb - 2;
    const int y1 = yb + 1 >= c.Y ? c.Y - 1 : yb + 1;
    const int ny = y1 - y0 + 1;
    bool ok = true;
    if (nx == 4 && ny == 4)
        cache_spline_win_co<4, 4, DER>(c, yco, x0, y0, yb, rho, t, v, dv, ok);
    else if (nx == 4)
        cache_spline_win_co<4, 3, DER>(c, yco, x0, y0, yb, rho, t, v, dv, ok);
    else if (ny == 4)
        cache_spline_win_co<3, 4, DER>(c, yco, x0, y0, yb, rho, t, v, dv, ok);
    else
        cache_spline_win_co<3, 3, DER>(c, yco, x0, y0, yb, rho, t, v, dv, ok);
    return ok;
}

// cache_solve_T on component 0 of plane 0 of a whole, non-rolling cache with
// its coefficient tables (yco: e, then p), and at convergence p = P(rho, T):
// every trial's lookup as cache_lookup1 makes it, with the density's window
// formed once and the temperature's bracket walked from the trial before.
__device__ inline int ceos_solve_Tp(const CacheView &c, const CeosHint &hx, const CeosHint &hy, const double *yco,
                                    double rho, double e, double T0, double lo, double hi, double &T, double &berr,
                                    double &pv)
{
    double x = T0, best = T0, besterr = 1.0E100;
    int ret = 0;
    const double xlo = ceos_ld(c.vx, 0), xhi = ceos_ld(c.vx, c.X - 1);
    const double ylo = ceos_ld(c.vy, 0), yhi = ceos_ld(c.vy, c.Y - 1);
    // the brackets of rho and of the start T (cache_bracket's); with hint
    // tables both walks' first reads go out at once.  A rho off the axis fails
    // the first trial, with what the reference's lookup returns there (as for a
    // T off the axis: T0, no error reached, -1).
    int left, yb;
    if (hx.nk && hy.nk) {
        int sx = ceos_hint_start(hx, rho), sy = ceos_hint_start(hy, x);
        // (each empty asm takes its reads' values, so the reads of a group are
        // in flight together before any is used)
        asm volatile("" : "+v"(sx), "+v"(sy));
        double bx = ceos_ld(c.vx, sx - 1), ax = ceos_ld(c.vx, sx), by = ceos_ld(c.vy, sy - 1), ay = ceos_ld(c.vy, sy);
        asm volatile("" : "+v"(bx), "+v"(ax), "+v"(by), "+v"(ay));
        left = ceos_walk_at(c.X, c.vx, rho, sx, bx, ax);
        yb = ceos_walk_at(c.Y, c.vy, x, sy, by, ay);
    } else {
        left = cache_bracket(c.X, c.vx, rho);
        yb = cache_bracket(c.Y, c.vy, x);
    }
    const bool rho_in = !((rho < xlo) | (rho > xhi));
    // cache_window's spline window around rho's bracket
    const int x0 = left - 2 < 0 ? 0 : left - 2;
    const int x1 = left + 1 >= c.X ? c.X - 1 : left + 1;
    const int nx = x1 - x0 + 1;
    for (int k = 1; k <= CEOS_NTRIAL; k++) {
        double val = 0.0, der = 0.0;
        bool ok = rho_in & !((x < ylo) | (x > yhi));
        if (ok) {
            if (k > 1) yb = ceos_walk(c.Y, c.vy, x, yb);  // the first trial's T is T0, bracketed above
            ok = ceos_win<true>(c, yco, x0, nx, yb, rho, x, val, der);
        }
        if (!ok) {
            T = x;
            berr = besterr;
            return -1;
        }
        const double F = (e - val) / e;
        const double D = -der / e;
        double bet = -F;
        const double errf = 0.0 + fabs(bet);
        if (errf <= CEOS_TOL) {
            T = x;
            berr = errf;
            double unused;
            ceos_win<false>(c, yco + (size_t)c.X * c.Y * CACHE_YCO, x0, nx, yb, rho, x, pv, unused);
            return 1;
        }
        if (errf < besterr) {
            if (k > 1) ret = 2;
            besterr = errf;
            best = x;
        }
        if (fabs(D) > 0.0) bet = bet / D;
        const double nv = x + bet;
        if (nv >= lo && nv <= hi && nv != x)
            x = nv;
        else
            break;
    }
    T = best;
    berr = besterr;
    return ret;
}

// pressure_rho0_u_CoolEOS with the zone's last temperature T (cooleos.c:314-444;
// MAXLOS 1: one solve, no random restart): the pressure and the new last T.
// One out-of-line copy serves every call site of a kernel.
__device__ __noinline__ CeosOut ceos_solve(const CeosParams *P, int r, double T, double rho0, double u)
{
    // every parameter read at once, ahead of the first branch
    const double rho_unit = P->rho_unit, u_unit = P->u_unit, p_unit = P->p_unit, gam = P->gam;
    const double rho_min = P->rho_min, rho_max = P->rho_max, T_lo = P->T_lo, T_hi = P->T_hi;
    const CacheView c = P->view;
    const CeosHint hx = P->hx, hy = P->hy;
    const double *yco = P->yco[r];
    CeosOut o{0., T, 0};
    const double rho = rho0 * rho_unit;
    const double e = u * u_unit;
    if (o.T == 0.0) {
        o.T = (gam - 1.) * u / rho0 / CEOS_KBOL * CEOS_MAV * CEOS_CL * CEOS_CL;
        if (o.T < T_lo) o.T = T_lo;
    }
    if (rho < rho_min || rho > rho_max) {
        o.p = (gam - 1.) * u;
        return o;
    }
    if (e < 0.0) return o;
    o.solved = 1;
    double Tn, berr, pv = 0.;
    ceos_solve_Tp(c, hx, hy, yco, rho, e, o.T, T_lo, T_hi, Tn, berr, pv);
    if (berr > CEOS_TOL) {
        o.p = (gam - 1.) * u;
        return o;
    }
    // the P lookup shares the converged trial's windows (cooleos.c:432)
    o.T = Tn;
    o.p = pv / p_unit;
    return o;
}

// The tabulated EOS policy of the step's physics templates (face_flux, vchar,
// source, primtoflux) and, through Eos5Cache, of the 5D inversion.
// V: the finite-difference steps rho0 hh / RHO_UNIT and u hh / U_UNIT are below
// half an ulp of every rho0 and u (both units at least hh 2^55), so rho0 +- h
// is rho0 and u +- h is u bit for bit: each derivative's two pressure calls
// repeat the call before it, whose result the derivative then reuses.
template <bool V>
struct EosCacheT {
    static constexpr bool TAB = true;
    const CeosParams *P;
    int r;  // the zone's radial index's cache
    CeosZone *z;

    __device__ double pres(double rho0, double u) const
    {
        const unsigned long long kr = (unsigned long long)__double_as_longlong(rho0);
        const unsigned long long ku = (unsigned long long)__double_as_longlong(u);
        if (z->has && kr == z->kr && ku == z->ku) return z->p;
        const CeosOut o = ceos_solve(P, r, z->T, rho0, u);
        z->T = o.T;
        z->nsolve += o.solved;
        z->has = true;
        z->kr = kr;
        z->ku = ku;
        z->p = o.p;
        return o.p;
    }
    // dpdrho_calc_CoolEOS (cooleos.c:704-741)
    __device__ double dpdrho(double rho0, double u) const
    {
        const double h = rho0 * CEOS_HH / P->rho_unit;
        const double rhoP = rho0 + h;
        double rhoM = rho0 - h;
        if (V) {
            const double p = pres(rho0, u);
            return (rhoM > 0) ? (p - p) / (2.0 * h) : (p - p) / (h);
        }
        if (rhoM > 0) {
            const double pP = pres(rhoP, u);
            const double pM = pres(rhoM, u);
            return (pP - pM) / (2.0 * h);
        }
        rhoM = rho0;
        const double pP = pres(rhoP, u);
        const double pM = pres(rhoM, u);
        return (pP - pM) / (h);
    }
    // dpdu_calc_CoolEOS (cooleos.c:667-701)
    __device__ double dpdu(double rho0, double u) const
    {
        const double h = u * CEOS_HH / P->u_unit;
        const double uP = u + h;
        double uM = u - h;
        if (V) {
            const double p = pres(rho0, u);
            return (uM > 0) ? (p - p) / (2.0 * h) : (p - p) / (h);
        }
        if (uM > 0) {
            const double pP = pres(rho0, uP);
            const double pM = pres(rho0, uM);
            return (pP - pM) / (2.0 * h);
        }
        uM = u;
        const double pP = pres(rho0, uP);
        const double pM = pres(rho0, uM);
        return (pP - pM) / (h);
    }
    // vchar's COOL_EOS sound speed (phys.c:365-377), after its pressure call
    __device__ double cs2(double rho, double u, double p) const
    {
        const double e = u / rho;
        const double enthalpy = 1.0 + e + p / rho;
        const double dpdr = dpdrho(rho, u);
        const double dpde = dpdu(rho, u) * rho;
        return (dpdr + p / rho / rho * dpde) / enthalpy;
    }
};

// The 5D inversion's view of the cache EOS (utoprim_5d.c:388-389, 454-481):
// the pressure, then dp/drho and dp/du, each by its calls above.
template <bool V>
struct Eos5CacheT {
    EosCacheT<V> e;
    double gam;
    __device__ static constexpr bool tabulated() { return true; }
    __device__ double p(double rho, double u) const { return e.pres(rho, u); }
    __device__ void p_derivs(double rho, double u, double &p, double &dpdrho, double &dpdu) const
    {
        p = e.pres(rho, u);
        dpdrho = e.dpdrho(rho, u);
        dpdu = e.dpdu(rho, u);
    }
};

constexpr bool ceos_vanishing_steps(double rho_unit, double u_unit)
{
    return CEOS_HH / rho_unit <= 0x1p-55 && CEOS_HH / u_unit <= 0x1p-55;
}

}  // namespace harm
