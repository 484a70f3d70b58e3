// harm_phys.hpp -- per-zone GRMHD physics for the gfx950 kernels (FP64).
//
// Every function keeps the reference's operation order (cited file:line) because
// parity with the CPU path is bitwise: the library is compiled with
// -ffp-contract=off, IEEE division and correctly rounded sqrt.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace harm {

enum { RHO = 0, UU, U1, U2, U3, B1, B2, B3, NPR };
enum { FACE1 = 0, FACE2 = 1, CORN = 2, CENT = 3, FACE3 = 4, NPG = 5 };

// decs.h:147-171, u2p_defs.h:79-96
constexpr double RHOMIN = 1.e-6, UUMIN = 1.e-8, RHOMINLIMIT = 1.e-20, UUMINLIMIT = 1.e-20;
constexpr double BSQORHOMAX = 50., BSQOUMAX = 2500., UORHOMAX = 50.;
constexpr double SMALL = 1.e-20, GAMMAMAX = 50.;
constexpr int MAX_NEWT_ITER = 30, EXTRA_NEWT_ITER = 2;
constexpr double NEWT_TOL = 1.0e-10, MIN_NEWT_TOL = 1.0e-10, NEWT_TOL2 = 1.0e-15;
constexpr double W_TOO_BIG = 1.e20, UTSQ_TOO_BIG = 1.e20, FAIL_VAL = 1.e30;

#define HD __device__ __forceinline__

struct Geom {
    double gcov[4][4];
    double gcon[4][4];
    double g;
    // Factors of the metric alone that the physics divides by or takes roots of,
    // formed once per location by geom_derived() with the expressions of
    // ucon_calc (phys.c:232-252) and utoprim_2d.c:145,176.
    double alpha;    // 1 / sqrt(-g^00)
    double ralpha;   // 1 / alpha, correctly rounded
    double lapse;    // sqrt(-1 / g^00)
    double rg;       // 1 / gdet, correctly rounded
    double beta[3];  // g^0j alpha alpha, j = 1..3
};

// Geometry tables hold GEO_BASE rows per location (gcov[16], gcon[16], gdet)
// followed by the GEO_DERIVED rows of geom_derived().
constexpr int GEO_BASE = 33, GEO_DERIVED = 7, GEO_ROWS = GEO_BASE + GEO_DERIVED;

// Host and device evaluate these with the same IEEE operations (correctly
// rounded sqrt and division), so a table built on the host equals what a kernel
// would compute for the same metric.
__host__ __device__ inline void geom_derived(double gcon00, double gcon01, double gcon02, double gcon03, double gdet,
                                             double *d)
{
    const double alpha = 1. / sqrt(-gcon00);
    d[0] = alpha;
    d[1] = 1. / alpha;
    d[2] = sqrt(-1. / gcon00);
    d[3] = 1. / gdet;
    d[4] = gcon01 * alpha * alpha;
    d[5] = gcon02 * alpha * alpha;
    d[6] = gcon03 * alpha * alpha;
}

HD void geom_set_derived(Geom &g, const double *d)
{
    g.alpha = d[0];
    g.ralpha = d[1];
    g.lapse = d[2];
    g.rg = d[3];
    g.beta[0] = d[4];
    g.beta[1] = d[5];
    g.beta[2] = d[6];
}

// For metrics formed inside a kernel (batch entry points, generated states).
HD void geom_derive(Geom &g)
{
    double d[GEO_DERIVED];
    geom_derived(g.gcon[0][0], g.gcon[0][1], g.gcon[0][2], g.gcon[0][3], g.g, d);
    geom_set_derived(g, d);
}

// a / b, correctly rounded, given rb = 1 / b correctly rounded.  Markstein's
// theorem: if rb is within half an ulp of 1/b and q within one ulp of a/b, the
// FMA residual r = a - b q is exact and RN(q + r rb) = RN(a/b).  q0 = RN(a rb)
// can be 1.5 ulp off, so one correction makes it faithful and a second rounds
// it.  Holds for finite a, b with residuals clear of underflow (|a| > 2^-960, or
// a = 0); the sign of a zero quotient is that of a rb.  Six operations in place
// of a full division where one denominator serves several numerators or its
// reciprocal is tabulated.
__host__ __device__ inline double div_r(double a, double b, double rb)
{
    const double q0 = a * rb;
    const double q1 = fma(fma(-b, q0, a), rb, q0);
    const double q2 = fma(fma(-b, q1, a), rb, q1);
    return copysign(q2, q0);
}

struct State {
    double ucon[4], ucov[4], bcon[4], bcov[4];
};

HD double dot4(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }

// ---------------------------------------------------------------------------
// Correctly rounded pow for x > 0 (used where the reference calls glibc's pow in
// the Font-fix solvers, utoprim_1dvsq2fix1.c:592, utoprim_1dfix1.c:832,849).
// glibc's pow is accurate to 0.52 ulp, so a correctly rounded result equals it
// except when the exact value lies within ~2^-68 of a rounding midpoint.  Method:
// double-double arithmetic with Dekker/Knuth error-free transformations;
// log(m) = 2 atanh((m-1)/(m+1)) by its series, exp by Taylor series on r/2^5 and
// five squarings; total relative error ~2^-100 before the final rounding.
// ---------------------------------------------------------------------------
struct dd {
    double hi, lo;
};
HD dd dd_two_sum(double a, double b)
{
    double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
HD dd dd_fast(double a, double b)
{
    double s = a + b;
    return {s, b - (s - a)};
}
HD dd dd_add(dd a, dd b)
{
    dd s = dd_two_sum(a.hi, b.hi);
    dd t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_fast(s.hi, s.lo);
    s.lo += t.lo;
    return dd_fast(s.hi, s.lo);
}
HD dd dd_mul(dd a, dd b)
{
    double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return dd_fast(p, e);
}
HD dd dd_mul_d(dd a, double b)
{
    double p = a.hi * b;
    double e = fma(a.hi, b, -p);
    e += a.lo * b;
    return dd_fast(p, e);
}
HD dd dd_div(dd a, dd b)
{
    double q1 = a.hi / b.hi;
    dd r = dd_add(a, dd_mul_d(b, -q1));
    double q2 = r.hi / b.hi;
    r = dd_add(r, dd_mul_d(b, -q2));
    double q3 = r.hi / b.hi;
    dd q = dd_fast(q1, q2);
    return dd_add(q, dd{q3, 0.});
}
HD dd dd_recip_int(int n)  // 1/n to ~2^-106
{
    double h = 1.0 / n;
    double r = fma(-h, (double)n, 1.0);
    return dd_fast(h, r / n);
}
HD dd dd_log(double x)  // x > 0, finite
{
    int e;
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < 0.70710678118654752) {
        m *= 2.;
        e -= 1;
    }
    dd num = {m - 1., 0.};  // exact (Sterbenz)
    dd den = dd_two_sum(m, 1.);
    dd s = dd_div(num, den);
    dd s2 = dd_mul(s, s);
    dd t = dd_recip_int(2 * 23 + 1);
    for (int k = 22; k >= 0; k--) t = dd_add(dd_recip_int(2 * k + 1), dd_mul(s2, t));
    dd lm = dd_mul_d(dd_mul(s, t), 2.);
    const dd ln2 = {6.93147180559945286227e-01, 2.31904681384629955842e-17};
    return dd_add(dd_mul_d(ln2, (double)e), lm);
}
HD double dd_exp_round(dd x)  // exp(x) rounded to double; |x.hi| < 700
{
    const dd ln2 = {6.93147180559945286227e-01, 2.31904681384629955842e-17};
    double k = rint(x.hi / ln2.hi);
    dd r = dd_add(x, dd_mul_d(ln2, -k));
    r.hi = ldexp(r.hi, -5);
    r.lo = ldexp(r.lo, -5);
    dd t = {1., 0.};
    for (int n = 13; n >= 1; n--) t = dd_add(dd{1., 0.}, dd_mul(dd_mul(r, t), dd_recip_int(n)));
    for (int q = 0; q < 5; q++) t = dd_mul(t, t);
    return ldexp(t.hi + t.lo, (int)k);
}
HD double cr_pow(double x, double y)
{
    if (!(x > 0.) || !isfinite(x) || !isfinite(y) || y == 0.) return pow(x, y);
    dd l = dd_mul_d(dd_log(x), y);
    if (!(fabs(l.hi) < 700.)) return pow(x, y);
    return dd_exp_round(l);
}
HD double dmax(double a, double b) { return (a > b) ? a : b; }
HD double dmin(double a, double b) { return (a < b) ? a : b; }

// Read-only table access through the constant address space: when the column
// index is wave-uniform the loads become scalar (s_load) and the values live in
// SGPRs, so one metric fetch serves the 64 phi-zones of a wave.
typedef const __attribute__((address_space(4))) double cdouble;
HD double ldc(const double *p, size_t off) { return ((cdouble *)p)[off]; }

// Geometry tables: per location a [33][ncol] block (16 gcov, 16 gcon, gdet), column
// index col = i*SN2 + j.  All lanes of a wave share one column in 3D runs.
HD void load_geom(const double *__restrict__ tab, size_t ncol, size_t col, Geom &g)
{
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            g.gcov[a][b] = ldc(tab, (size_t)(a * 4 + b) * ncol + col);
            g.gcon[a][b] = ldc(tab, (size_t)(16 + a * 4 + b) * ncol + col);
        }
    g.g = ldc(tab, (size_t)32 * ncol + col);
    double d[GEO_DERIVED];
#pragma unroll
    for (int k = 0; k < GEO_DERIVED; k++) d[k] = ldc(tab, (size_t)(GEO_BASE + k) * ncol + col);
    geom_set_derived(g, d);
}

// Column-major copy of the tables for wave-uniform kernels: all values of one
// (i,j) column are contiguous, so the scalar unit fetches them with wide loads.
// CENT block (108 doubles): gcov[16] gcon[16] gdet rhoflr uuflr pad conn[64]
// derived[7] pad; face blocks (40 doubles): gcov[16] gcon[16] gdet derived[7].
constexpr int CENT_STRIDE = 108, CENT_CONN = 36, CENT_DERIVED = 100, FACE_STRIDE = 40, FACE_DERIVED = GEO_BASE;

template <int DERIVED>
HD void load_geom_block(const double *__restrict__ blk, Geom &g)
{
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            g.gcov[a][b] = ldc(blk, a * 4 + b);
            g.gcon[a][b] = ldc(blk, 16 + a * 4 + b);
        }
    g.g = ldc(blk, 32);
    double d[GEO_DERIVED];
#pragma unroll
    for (int k = 0; k < GEO_DERIVED; k++) d[k] = ldc(blk, DERIVED + k);
    geom_set_derived(g, d);
}

// Wave-uniform column index: N3 % 64 == 0 maps a wave onto one (i,j) column.
template <bool UNI>
HD size_t uni_col(size_t col)
{
    if (UNI) return (size_t)__builtin_amdgcn_readfirstlane((int)col);
    return col;
}

// metric.c:210-229
HD void lower(const double *ucon, const Geom &geom, double *ucov)
{
#pragma unroll
    for (int a = 0; a < 4; a++)
        ucov[a] = geom.gcov[a][0] * ucon[0] + geom.gcov[a][1] * ucon[1] + geom.gcov[a][2] * ucon[2] +
                  geom.gcov[a][3] * ucon[3];
}
// metric.c:232-255
HD void raise(const double *ucov, const Geom &geom, double *ucon)
{
#pragma unroll
    for (int a = 0; a < 4; a++)
        ucon[a] = geom.gcon[a][0] * ucov[0] + geom.gcon[a][1] * ucov[1] + geom.gcon[a][2] * ucov[2] +
                  geom.gcon[a][3] * ucov[3];
}

// phys.c:300-324
HD int gamma_calc(const double *pr, const Geom &geom, double *gamma)
{
    double qsq = geom.gcov[1][1] * pr[U1] * pr[U1] + geom.gcov[2][2] * pr[U2] * pr[U2] +
                 geom.gcov[3][3] * pr[U3] * pr[U3] +
                 2. * (geom.gcov[1][2] * pr[U1] * pr[U2] + geom.gcov[1][3] * pr[U1] * pr[U3] +
                       geom.gcov[2][3] * pr[U2] * pr[U3]);
    if (qsq < 0.) {
        if (fabs(qsq) > 1.E-10) {
            *gamma = 1.;
            return 1;
        }
        qsq = 1.E-10;
    }
    *gamma = sqrt(1. + qsq);
    return 0;
}

// phys.c:232-252; fatal counts where the reference calls fail()
// alpha, beta and 1/alpha come from the geometry table (same expressions).
HD void ucon_calc(const double *pr, const Geom &geom, double *ucon, int &fatal)
{
    const double alpha = geom.alpha, ralpha = geom.ralpha;
    double gamma;
    if (gamma_calc(pr, geom, &gamma)) fatal++;
    ucon[0] = div_r(gamma, alpha, ralpha);
#pragma unroll
    for (int j = 1; j < 4; j++) ucon[j] = pr[U1 + j - 1] - div_r(gamma * geom.beta[j - 1], alpha, ralpha);
}

// phys.c:108-117
HD void bcon_calc(const double *pr, const double *ucon, const double *ucov, double *bcon)
{
    bcon[0] = pr[B1] * ucov[1] + pr[B2] * ucov[2] + pr[B3] * ucov[3];
    const double r0 = 1. / ucon[0];
#pragma unroll
    for (int j = 1; j < 4; j++) bcon[j] = div_r(pr[B1 - 1 + j] + bcon[0] * ucon[j], ucon[0], r0);
}

// phys.c:219-229
HD void get_state(const double *pr, const Geom &geom, State &q, int &fatal)
{
    ucon_calc(pr, geom, q.ucon, fatal);
    lower(q.ucon, geom, q.ucov);
    bcon_calc(pr, q.ucon, q.ucov, q.bcon);
    lower(q.bcon, geom, q.bcov);
}

// The Gamma-law EOS policy of the physics templates below (p = (gam - 1) u,
// phys.c:150); the tabulated policy lives with the table (harm_eos5d.hpp).
struct EosGamma {
    static constexpr bool TAB = false;
    double gam;
    HD double pres(double, double u) const { return (gam - 1.) * u; }
};

// phys.c:122-161: one row of the MHD stress tensor for a given pressure P.
HD void mhd_calc_P(const double *pr, int dir, const State &q, double P, double *mhd)
{
    double r = pr[RHO], u = pr[UU];
    double w = P + r + u;
    double bsq = dot4(q.bcon, q.bcov);
    double eta = w + bsq;
    double ptot = P + 0.5 * bsq;
#pragma unroll
    for (int j = 0; j < 4; j++)
        mhd[j] = eta * q.ucon[dir] * q.ucov[j] + ptot * ((dir == j) ? 1. : 0.) - q.bcon[dir] * q.bcov[j];
}

// phys.c:143-161
HD void mhd_calc(const double *pr, int dir, const State &q, double gam, double *mhd)
{
    mhd_calc_P(pr, dir, q, (gam - 1.) * pr[UU], mhd);
}

// phys.c:59-96 for a given pressure P
HD void primtoflux_P(const double *pr, const State &q, int dir, const Geom &geom, double P, double *flux)
{
    double mhd[4];
    flux[RHO] = pr[RHO] * q.ucon[dir];
    mhd_calc_P(pr, dir, q, P, mhd);
    flux[UU] = mhd[0] + flux[RHO];
    flux[U1] = mhd[1];
    flux[U2] = mhd[2];
    flux[U3] = mhd[3];
    flux[B1] = q.bcon[1] * q.ucon[dir] - q.bcon[dir] * q.ucon[1];
    flux[B2] = q.bcon[2] * q.ucon[dir] - q.bcon[dir] * q.ucon[2];
    flux[B3] = q.bcon[3] * q.ucon[dir] - q.bcon[dir] * q.ucon[3];
#pragma unroll
    for (int k = 0; k < NPR; k++) flux[k] *= geom.g;
}

// phys.c:59-96
template <class E>
HD void primtoflux(const double *pr, const State &q, int dir, const Geom &geom, const E &eos, double *flux)
{
    primtoflux_P(pr, q, dir, geom, eos.pres(pr[RHO], pr[UU]), flux);
}
HD void primtoflux(const double *pr, const State &q, int dir, const Geom &geom, double gam, double *flux)
{
    primtoflux_P(pr, q, dir, geom, (gam - 1.) * pr[UU], flux);
}

// phys.c:337-476; returns how often the reference would fail().  Gamma-law:
// cs2 = gam (gam - 1) u / (rho + gam u) and fail() outside [0, 1].  Tabulated
// (phys.c:359-378, 398-406): cs2 from the EOS derivatives (E::cs2), clipped
// to [SMALL, 1] without fail().
template <class E>
HD int vchar(const double *pr, const State &q, const Geom &geom, int js, const E &eos, double *vmax, double *vmin)
{
    int nfail = 0;
    double bsq = dot4(q.bcon, q.bcov);
    double rho = pr[RHO], u = pr[UU];
    double cms2;
    if constexpr (E::TAB) {
        const double p = eos.pres(rho, u);
        double EF = rho + u + p;
        double EE = bsq + EF;
        double va2 = bsq / EE;
        double cs2 = eos.cs2(rho, u, p);
        cms2 = cs2 + va2 - cs2 * va2;
        if (cms2 < SMALL) cms2 = SMALL;
        if (cms2 > 1.) cms2 = 1.;
    } else {
        const double gam = eos.gam;
        double EF = rho + gam * u;
        double EE = bsq + EF;
        double va2 = bsq / EE;
        double cs2 = gam * (gam - 1.) * u / EF;
        cms2 = cs2 + va2 - cs2 * va2;
        if (cms2 < 0.) {
            nfail++;
            cms2 = SMALL;
        }
        if (cms2 > 1.) {
            nfail++;
            cms2 = 1.;
        }
    }
    // The reference raises the unit covectors A = e_js, B = e_0 and contracts
    // them; for finite metric and velocity those contractions are exactly the
    // entries picked here (a zero's sign can differ, and every use of vmax and
    // vmin passes through fabs).
    double Asq = geom.gcon[js][js], Bsq = geom.gcon[0][0];
    double Au = q.ucon[js], Bu = q.ucon[0], AB = geom.gcon[0][js];
    double Au2 = Au * Au, Bu2 = Bu * Bu, AuBu = Au * Bu;
    double A = Bu2 - (Bsq + Bu2) * cms2;
    double B = 2. * (AuBu - (AB + AuBu) * cms2);
    double C = Au2 - (Asq + Au2) * cms2;
    double discr = B * B - 4. * A * C;
    if ((discr < 0.0) && (discr > -1.e-10))
        discr = 0.0;
    else if (discr < -1.e-10) {
        nfail++;
        discr = 0.;
    }
    discr = sqrt(discr);
    double vp = -(-B + discr) / (2. * A);
    double vm = -(-B - discr) / (2. * A);
    if (vp > vm) {
        *vmax = vp;
        *vmin = vm;
    } else {
        *vmax = vm;
        *vmin = vp;
    }
    return nfail;
}
HD int vchar(const double *pr, const State &q, const Geom &geom, int js, double gam, double *vmax, double *vmin)
{
    return vchar(pr, q, geom, js, EosGamma{gam}, vmax, vmin);
}

// interp.c:49-93
HD double slope_lim(int lim, double y1, double y2, double y3)
{
    double Dqm, Dqp, Dqc, s;
    if (lim == 0) {
        Dqm = 2. * (y2 - y1);
        Dqp = 2. * (y3 - y2);
        Dqc = 0.5 * (y3 - y1);
        s = Dqm * Dqp;
        if (s <= 0.) return 0.;
        if (fabs(Dqm) < fabs(Dqp) && fabs(Dqm) < fabs(Dqc)) return Dqm;
        if (fabs(Dqp) < fabs(Dqc)) return Dqp;
        return Dqc;
    } else if (lim == 1) {
        Dqm = (y2 - y1);
        Dqp = (y3 - y2);
        s = Dqm * Dqp;
        if (s <= 0.) return 0.;
        return 2. * s / (Dqm + Dqp);
    }
    Dqm = (y2 - y1);
    Dqp = (y3 - y2);
    s = Dqm * Dqp;
    if (s <= 0.) return 0.;
    if (fabs(Dqm) < fabs(Dqp)) return Dqm;
    return Dqp;
}

// Slope-limiter codes: the reference's MC, VANL, MINM (decs.h:214-216) and the
// piecewise-parabolic option, which the reference does not have.
constexpr int LIM_MC = 0, LIM_PPM = 3;

// Piecewise-parabolic face values of one zone (Colella & Woodward 1984,
// J. Comput. Phys. 54, 174, eqs. 1.6 and 1.10) on the MC slopes of slope_lim:
// y0 the zone's value, ym/yp its neighbours, dqm/dq0/dqp the three zones' MC
// slopes.  Interface a(m+1/2) = (y(m) + y(m+1))/2 - (dq(m+1) - dq(m))/6 (the
// same expression from both zones of an interface, so both start from the same
// bits); a local extremum flattens the zone; a parabola that overshoots beyond
// one face has the other face's value moved to 3 y0 - 2 (overshooting face).
HD void ppm_para(double ym, double y0, double yp, double dqm, double dq0, double dqp, double &l, double &r)
{
    double ql = 0.5 * (ym + y0) - (1. / 6.) * (dq0 - dqm);
    double qr = 0.5 * (y0 + yp) - (1. / 6.) * (dqp - dq0);
    const double qd = qr - ql, qe = 6. * (y0 - 0.5 * (ql + qr));
    if ((qr - y0) * (y0 - ql) <= 0.) {
        ql = y0;
        qr = y0;
    } else if (qd * (qd - qe) < 0.) {
        ql = 3. * y0 - 2. * qr;
    } else if (qd * (qd + qe) < 0.) {
        qr = 3. * y0 - 2. * ql;
    }
    l = ql;
    r = qr;
}

// One face of step_ch.c:429-475: reconstructed states -> numerical flux and ctop.
template <class E>
HD void face_flux(const double *p_l, const double *p_r, const Geom &geom, int dir, const E &eos, double hllf,
                  double laxf, double *F, double &ctop, int &fatal)
{
    double F_l[NPR], F_r[NPR], U_l[NPR], U_r[NPR];
    double cmax_l, cmin_l, cmax_r, cmin_r;
    {
        // the left state is finished before the right one is formed
        State sl;
        get_state(p_l, geom, sl, fatal);
        const double P_l = eos.pres(p_l[RHO], p_l[UU]);
        primtoflux_P(p_l, sl, dir, geom, P_l, F_l);
        primtoflux_P(p_l, sl, 0, geom, P_l, U_l);
        fatal += vchar(p_l, sl, geom, dir, eos, &cmax_l, &cmin_l);
    }
    {
        State sr;
        get_state(p_r, geom, sr, fatal);
        const double P_r = eos.pres(p_r[RHO], p_r[UU]);
        primtoflux_P(p_r, sr, dir, geom, P_r, F_r);
        primtoflux_P(p_r, sr, 0, geom, P_r, U_r);
        fatal += vchar(p_r, sr, geom, dir, eos, &cmax_r, &cmin_r);
    }
    double cmax = fabs(dmax(dmax(0., cmax_l), cmax_r));
    double cmin = fabs(dmax(dmax(0., -cmin_l), -cmin_r));
    ctop = dmax(cmax, cmin);
    const double den = cmax + cmin + SMALL, rden = 1. / den;
#pragma unroll
    for (int k = 0; k < NPR; k++)
        F[k] = hllf * div_r(cmax * F_l[k] + cmin * F_r[k] - cmax * cmin * (U_r[k] - U_l[k]), den, rden) +
               laxf * (0.5 * (F_l[k] + F_r[k] - ctop * (U_r[k] - U_l[k])));
}
HD void face_flux(const double *p_l, const double *p_r, const Geom &geom, int dir, double gam, double hllf,
                  double laxf, double *F, double &ctop, int &fatal)
{
    face_flux(p_l, p_r, geom, dir, EosGamma{gam}, hllf, laxf, F, ctop, fatal);
}

// phys.c:165-207: dU = g * sum_jk T^j_k Gamma^k_{mu j}; conn is [64][ncol] (k,mu,j order)
// Each stress-tensor row is formed just before it is contracted; the sums run
// in the reference's order (j outer, k inner), so the result is unchanged while
// only one row is live.  The state at ph is returned for reuse.
// conn[((k*4+mu)*4+j)*cstride + col] = Gamma^k_{mu j}: cstride = ncol for the
// variable-major table, 1 (with col = 0 and conn at the column block) for the
// column-major one.
template <class E>
HD void source(const double *ph, const Geom &geom, const double *__restrict__ conn, size_t ncol, size_t col,
               const E &eos, double *dU, State &q, int &fatal)
{
    get_state(ph, geom, q, fatal);
    const double P = eos.pres(ph[RHO], ph[UU]);
#pragma unroll
    for (int k = 0; k < NPR; k++) dU[k] = 0.;
    // mhd_calc_P's row j with u^j, b^j picked by selects (an index into the
    // state would move it to scratch memory in the loop below)
    const double bsq = dot4(q.bcon, q.bcov);
    const double eta = (P + ph[RHO] + ph[UU]) + bsq, ptot = P + 0.5 * bsq;
    // Not unrolled: the 16 connection coefficients of row j are fetched in
    // iteration j, so only one row occupies scalar registers at a time.
#pragma unroll 1
    for (int j = 0; j < 4; j++) {
        const double uj = (j == 0) ? q.ucon[0] : (j == 1) ? q.ucon[1] : (j == 2) ? q.ucon[2] : q.ucon[3];
        const double bj = (j == 0) ? q.bcon[0] : (j == 1) ? q.bcon[1] : (j == 2) ? q.bcon[2] : q.bcon[3];
        double mhd[4];
#pragma unroll
        for (int k = 0; k < 4; k++) mhd[k] = eta * uj * q.ucov[k] + ptot * ((j == k) ? 1. : 0.) - bj * q.bcov[k];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            dU[UU] += mhd[k] * ldc(conn, (size_t)((k * 4 + 0) * 4 + j) * ncol + col);
            dU[U1] += mhd[k] * ldc(conn, (size_t)((k * 4 + 1) * 4 + j) * ncol + col);
            dU[U2] += mhd[k] * ldc(conn, (size_t)((k * 4 + 2) * 4 + j) * ncol + col);
            dU[U3] += mhd[k] * ldc(conn, (size_t)((k * 4 + 3) * 4 + j) * ncol + col);
        }
    }
#pragma unroll
    for (int k = 0; k < NPR; k++) dU[k] *= geom.g;
}
HD void source(const double *ph, const Geom &geom, const double *__restrict__ conn, size_t ncol, size_t col,
               double gam, double *dU, State &q, int &fatal)
{
    source(ph, geom, conn, ncol, col, EosGamma{gam}, dU, q, fatal);
}

// ---------------------------------------------------------------------------
// Conserved -> primitive inversion (utoprim_2d.c, utoprim_1dvsq2fix1.c,
// utoprim_1dfix1.c, u2p_util.c).  The reference's file-scope solver globals are
// the per-thread struct U2P.
// ---------------------------------------------------------------------------
struct U2P {
    double Bsq, QdotBsq, Qtsq, Qdotn, D, K_atm, gam;
    double rgam;  // 1 / gam, correctly rounded (for div_r in the 2D solver)
    double W_for_gnr2, rho_for_gnr2;
};

// u2p_util.c:204-256 + utoprim_2d.c:240-266
HD void u2p_scalars(U2P &s, const double *U, const Geom &geom, double *Bcon, double *Qcon, double *ncon,
                    double &QdotB)
{
    double Bcov[4], Qcov[4], ncov[4];
    Bcon[0] = 0.;
#pragma unroll
    for (int i = 1; i < 4; i++) Bcon[i] = U[B1 + i - 1];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        Bcov[i] = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) Bcov[i] += geom.gcov[i][j] * Bcon[j];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) Qcov[i] = U[1 + i];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        Qcon[i] = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) Qcon[i] += geom.gcon[i][j] * Qcov[j];
    }
    s.Bsq = 0.;
#pragma unroll
    for (int i = 1; i < 4; i++) s.Bsq += Bcon[i] * Bcov[i];
    QdotB = 0.;
#pragma unroll
    for (int i = 0; i < 4; i++) QdotB += Qcov[i] * Bcon[i];
    s.QdotBsq = QdotB * QdotB;
    double lapse = geom.lapse;
    ncov[0] = -lapse;
    ncov[1] = ncov[2] = ncov[3] = 0.;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        ncon[i] = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) ncon[i] += geom.gcon[i][j] * ncov[j];
    }
    s.Qdotn = Qcon[0] * ncov[0];
    double Qsq = 0.;
#pragma unroll
    for (int i = 0; i < 4; i++) Qsq += Qcov[i] * Qcon[i];
    s.Qtsq = Qsq + s.Qdotn * s.Qdotn;
    s.D = U[RHO];
}

// utoprim_2d.c:374-383
HD double vsq_calc(const U2P &s, double W)
{
    double Wsq = W * W;
    double Xsq = (s.Bsq + W) * (s.Bsq + W);
    return ((Wsq * s.Qtsq + s.QdotBsq * (s.Bsq + 2. * W)) / (Wsq * Xsq));
}

// utoprim_2d.c:562-624 with the Gamma-law EOS of utoprim_2d.c:645-678
HD void func_vsq(const U2P &s, double W, double vsq, double &dx0, double &dx1, double &f, double &df)
{
    const double GAMMA = s.gam;
    double Wsq = W * W;
    double gtmp = 1. - vsq;
    double p_tmp = div_r((GAMMA - 1.) * (W * gtmp - s.D * sqrt(gtmp)), GAMMA, s.rgam);
    double dPdW = div_r((GAMMA - 1.) * (1. - vsq), GAMMA, s.rgam);
    double dPdvsq = div_r((GAMMA - 1.) * (0.5 * s.D / sqrt(1. - vsq) - W), GAMMA, s.rgam);
    double t2 = -0.5 * s.Bsq + dPdvsq;
    double t3 = s.Bsq + W;
    double t4 = t3 * t3;
    double t9 = 1. / Wsq;
    double t11 = s.Qtsq - vsq * t4 + s.QdotBsq * (s.Bsq + 2.0 * W) * t9;
    double t16 = s.QdotBsq * t9;
    double t18 = -s.Qdotn - 0.5 * s.Bsq * (1.0 + vsq) + 0.5 * t16 - W + p_tmp;
    double t21 = 1. / t3;
    double t23 = 1. / W;
    double t24 = t16 * t23;
    double t25 = -1.0 + dPdW - t24;
    double t35 = t25 * t3 + (s.Bsq - 2.0 * dPdvsq) * (s.QdotBsq + vsq * Wsq * W) * t9 * t23;
    double t36 = 1. / t35;
    dx0 = -(t2 * t11 + t4 * t18) * t21 * t36;
    double t40 = (vsq + t24) * t3;
    dx1 = -(-t25 * t11 - 2.0 * t40 * t18) * t21 * t36;
    df = -t11 * t11 - t18 * t18;
    f = -0.5 * df;
}

// utoprim_2d.c:442-541 (validate_x: utoprim_2d.c:417-431), as a state that
// advances one Newton iteration per gnr_step, so a solve can pause after any
// iteration and resume with the same bits.
struct Gnr2 {
    double x0, x1, f, df, errx;
    int n_iter, i_extra, doing_extra, k_conv;
    bool keep;
};

HD void gnr_start(Gnr2 &g, double x0, double x1)
{
    g.x0 = x0;
    g.x1 = x1;
    g.f = 1.;
    g.df = 1.;
    g.errx = 1.;
    g.n_iter = 0;
    g.i_extra = 0;
    g.doing_extra = 0;
    g.k_conv = -1;
    g.keep = true;
}

HD void gnr_step(const U2P &s, Gnr2 &g)
{
    double dx0, dx1;
    func_vsq(s, g.x0, g.x1, dx0, dx1, g.f, g.df);
    double xo0 = g.x0;
    g.x0 += dx0;
    g.x1 += dx1;
    g.errx = (g.x0 == 0.) ? fabs(dx0) : fabs(dx0 / g.x0);
    g.errx += (g.x1 == 0.) ? fabs(dx1) : fabs(dx1 / g.x1);
    g.x0 = fabs(g.x0);
    g.x0 = (g.x0 > W_TOO_BIG) ? xo0 : g.x0;
    g.x1 = (g.x1 < 0.) ? 0. : g.x1;
    g.x1 = (g.x1 > 1.) ? (1. - 1.e-15) : g.x1;
    if ((fabs(g.errx) <= NEWT_TOL) && (g.doing_extra == 0)) {
        g.doing_extra = 1;
        g.k_conv = g.n_iter + 1;
    }
    if (g.doing_extra == 1) g.i_extra++;
    if (((fabs(g.errx) <= NEWT_TOL) && (g.doing_extra == 0)) || (g.i_extra > EXTRA_NEWT_ITER) ||
        (g.n_iter >= (MAX_NEWT_ITER - 1)))
        g.keep = false;
    g.n_iter++;
}

HD int gnr_end(const Gnr2 &g, int &n_iter_out, int &k_conv_out)
{
    n_iter_out = g.n_iter;
    k_conv_out = (g.k_conv < 0) ? g.n_iter : g.k_conv;
    if (!isfinite(g.f) || !isfinite(g.df)) return 2;
    if (fabs(g.errx) > MIN_NEWT_TOL) return 1;
    return 0;
}

HD int gnr_2d(const U2P &s, double &x0, double &x1, int &n_iter_out, int &k_conv_out)
{
    Gnr2 g;
    gnr_start(g, x0, x1);
    while (g.keep) gnr_step(s, g);
    x0 = g.x0;
    x1 = g.x1;
    return gnr_end(g, n_iter_out, k_conv_out);
}

// Conserved-variable transform of utoprim_2d.c:145-163 (shared by all solvers).
HD void u2p_transform(const double *U, const Geom &geom, double *U_tmp)
{
    const double alpha = geom.alpha, g = geom.g, rg = geom.rg;
    U_tmp[RHO] = div_r(alpha * U[RHO], g, rg);
    U_tmp[UU] = div_r(alpha * (U[UU] - U[RHO]), g, rg);
#pragma unroll
    for (int i = U1; i <= U3; i++) U_tmp[i] = div_r(alpha * U[i], g, rg);
#pragma unroll
    for (int i = B1; i <= B3; i++) U_tmp[i] = div_r(alpha * U[i], g, rg);
}

// Primitive recovery shared by the solvers (utoprim_2d.c:349-356).
HD void u2p_recover_vel(const U2P &s, const double *Bcon, const double *Qcon, const double *ncon, double QdotB,
                        double W, double gamma, double *prim)
{
#pragma unroll
    for (int i = 1; i < 4; i++) {
        double Qtcon = Qcon[i] + ncon[i] * s.Qdotn;
        prim[U1 + i - 1] = gamma / (W + s.Bsq) * (Qtcon + QdotB * Bcon[i] / W);
    }
}

HD double utsq_of(const Geom &geom, const double *prim)
{
    double utsq = 0.;
#pragma unroll
    for (int i = 1; i < 4; i++)
#pragma unroll
        for (int j = 1; j < 4; j++) utsq += geom.gcov[i][j] * prim[U1 + i - 1] * prim[U1 + j - 1];
    return utsq;
}

// Initial W guess of utoprim_2d.c:269-303 from the guess primitives.
HD bool guess_W(const U2P &s, const Geom &geom, const double *prim, double &W_last)
{
    double utsq = utsq_of(geom, prim);
    if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
    if (utsq < 0. || utsq > UTSQ_TOO_BIG) return false;
    double gammasq = 1. + utsq;
    double gamma = sqrt(gammasq);
    double rho0 = s.D / gamma;
    double u = prim[UU];
    double p = (s.gam - 1.) * u;
    double w = rho0 + u + p;
    W_last = w * gammasq;
    return true;
}

// Raise W until v^2 < 1 can hold (utoprim_2d.c:297-303).
HD double raise_W(const U2P &s, double W_last)
{
    int i_increase = 0;
    while (((W_last * W_last * W_last * (W_last + 2. * s.Bsq) - s.QdotBsq * (2. * W_last + s.Bsq)) <=
            W_last * W_last * (s.Qtsq - s.Bsq * s.Bsq)) &&
           (i_increase < 10)) {
        W_last *= 10.;
        i_increase++;
    }
    return W_last;
}

// utoprim_2d.c:306-356 from a starting W: Newton on (W, v^2), then recovery.
// Returns the reference's code; prim[0..4] written only on success.
// utoprim_2d.c:322-356: the primitives from the Newton result (retval, W, vsq).
HD int finish_2d(U2P &s, const double *Bcon, const double *Qcon, const double *ncon, double QdotB, int retval,
                 double W, double vsq, double *prim)
{
    if ((retval != 0) || (W == FAIL_VAL)) return retval * 100 + 1;
    if (W <= 0. || W > W_TOO_BIG) return 3;
    if (vsq >= 1.) return 4;
    double gtmp = sqrt(1. - vsq);
    double gamma = 1. / gtmp;
    double rho0 = s.D * gtmp;
    double w = W * (1. - vsq);
    double p = div_r((s.gam - 1.) * (w - rho0), s.gam, s.rgam);
    double u = w - (rho0 + p);
    if ((rho0 <= 0.) || (u <= 0.)) return 5;
    prim[RHO] = rho0;
    prim[UU] = u;
    u2p_recover_vel(s, Bcon, Qcon, ncon, QdotB, W, gamma, prim);
    return 0;
}

HD int solve_2d_from_x(U2P &s, const double *Bcon, const double *Qcon, const double *ncon, double QdotB, double x0,
                       double x1, double *prim, int &n_iter, int &k_conv)
{
    int retval = gnr_2d(s, x0, x1, n_iter, k_conv);
    return finish_2d(s, Bcon, Qcon, ncon, QdotB, retval, x0, x1, prim);
}

// utoprim_2d.c:306-307 (x1_of_x0, utoprim_2d.c:395-406): the Newton start from a W guess.
HD double x1_of_x0(const U2P &s, double W)
{
    double vsq0 = fabs(vsq_calc(s, W));
    return (vsq0 > 1.) ? (1.0 - 1.e-15) : vsq0;
}

HD int solve_2d_from_W(U2P &s, const double *U_tmp, const double *Bcon, const double *Qcon, const double *ncon,
                       double QdotB, double W_last, double *prim, int &n_iter, int &k_conv)
{
    return solve_2d_from_x(s, Bcon, Qcon, ncon, QdotB, fabs(W_last), x1_of_x0(s, W_last), prim, n_iter, k_conv);
}

// utoprim_2d.c:126-176 + 219-362.  U is HARM's conserved vector (x gdet).
HD int utoprim_2d(const double *U, const Geom &geom, double gam, double *prim, int &n_iter, int &k_conv)
{
    n_iter = 0;
    k_conv = 0;
#pragma unroll
    for (int i = B1; i <= B3; i++) prim[i] = div_r(U[i], geom.g, geom.rg);
    if (U[0] <= 0.) return -100;
    double U_tmp[NPR];
    u2p_transform(U, geom, U_tmp);
    U2P s;
    s.gam = gam;
    s.rgam = 1. / gam;
    s.K_atm = 0.;
    double Bcon[4], Qcon[4], ncon[4], QdotB;
    u2p_scalars(s, U_tmp, geom, Bcon, Qcon, ncon, QdotB);
    double W_last;
    if (!guess_W(s, geom, prim, W_last)) return 2;
    W_last = raise_W(s, W_last);
    double pt[5];
    int ret = solve_2d_from_W(s, U_tmp, Bcon, Qcon, ncon, QdotB, W_last, pt, n_iter, k_conv);
    if (ret == 0) {
#pragma unroll
        for (int i = 0; i < 5; i++) prim[i] = pt[i];
    }
    return ret;
}

// utoprim_2d with the Newton start (W0, vsq0) given instead of the guess of
// utoprim_2d.c:268-307: x = (|W0|, x1_of_x0(W0)) with vsq_from_W, else
// (|W0|, vsq0 clipped to [0, 1-1e-15]).  Same codes and writes as utoprim_2d.
HD int utoprim_2d_seeded(const double *U, const Geom &geom, double gam, double *prim, double W0, double vsq0,
                         int vsq_from_W, int &n_iter, int &k_conv)
{
    n_iter = 0;
    k_conv = 0;
#pragma unroll
    for (int i = B1; i <= B3; i++) prim[i] = div_r(U[i], geom.g, geom.rg);
    if (U[0] <= 0.) return -100;
    double U_tmp[NPR];
    u2p_transform(U, geom, U_tmp);
    U2P s;
    s.gam = gam;
    s.rgam = 1. / gam;
    s.K_atm = 0.;
    double Bcon[4], Qcon[4], ncon[4], QdotB;
    u2p_scalars(s, U_tmp, geom, Bcon, Qcon, ncon, QdotB);
    const double x1 = vsq_from_W ? x1_of_x0(s, W0) : ((vsq0 < 0.) ? 0. : ((vsq0 > 1. - 1.e-15) ? (1. - 1.e-15) : vsq0));
    double pt[5];
    int ret = solve_2d_from_x(s, Bcon, Qcon, ncon, QdotB, fabs(W0), x1, pt, n_iter, k_conv);
    if (ret == 0) {
#pragma unroll
        for (int i = 0; i < 5; i++) prim[i] = pt[i];
    }
    return ret;
}

// The learned-seed solve: seeded Newton, and on failure utoprim_2d from the
// guess primitives in prim (the previous-step state).  Counts add up.
HD int utoprim_2d_learned(const double *U, const Geom &geom, double gam, double *prim, double W0, double vsq0,
                          int vsq_from_W, int &n_iter, int &k_conv, int &fell_back)
{
    double pt[NPR];
#pragma unroll
    for (int i = 0; i < NPR; i++) pt[i] = prim[i];
    int n1, k1;
    int ret = utoprim_2d_seeded(U, geom, gam, pt, W0, vsq0, vsq_from_W, n1, k1);
    fell_back = 0;
    if (ret == 0 || ret == -100) {
#pragma unroll
        for (int i = 0; i < NPR; i++) prim[i] = pt[i];
        n_iter = n1;
        k_conv = k1;
        return ret;
    }
    fell_back = 1;
    int n2, k2;
    ret = utoprim_2d(U, geom, gam, prim, n2, k2);
    n_iter = n1 + n2;
    k_conv = n1 + k2;
    return ret;
}

// The gated learned solve, in two phases around the wave-level network pass.
// Phase 1 (u2p_gate_first): utoprim_2d's start from the guess primitives and
// its first Newton iteration.  need: the zone asks for the network's seed --
// no valid start from the guess (code 2), or a first Newton step
// |dW/W| + |dv2/v2| above tau.  Returns -100 (nothing to solve), 2 (no start)
// or 0 (the Newton state st holds the first iteration).
HD int u2p_gate_first(const double *U, const Geom &geom, double gam, double *prim, double tau, Gnr2 &st,
                      bool &need)
{
    need = false;
#pragma unroll
    for (int i = B1; i <= B3; i++) prim[i] = div_r(U[i], geom.g, geom.rg);
    if (U[0] <= 0.) return -100;
    double U_tmp[NPR];
    u2p_transform(U, geom, U_tmp);
    U2P s;
    s.gam = gam;
    s.rgam = 1. / gam;
    s.K_atm = 0.;
    double Bcon[4], Qcon[4], ncon[4], QdotB;
    u2p_scalars(s, U_tmp, geom, Bcon, Qcon, ncon, QdotB);
    double W_last;
    if (!guess_W(s, geom, prim, W_last)) {
        need = true;
        return 2;
    }
    W_last = raise_W(s, W_last);
    gnr_start(st, fabs(W_last), x1_of_x0(s, W_last));
    gnr_step(s, st);
    need = st.keep && fabs(st.errx) > tau;
    return 0;
}

// Phase 2: a zone that asked for the seed solves from (W0, vsq0) (as
// utoprim_2d_seeded); where that fails, or for every other zone, the plain
// solve resumes from its first iteration, so a zone that did not ask gets
// utoprim_2d's result bit for bit.  Iteration counts add up (the first
// iteration counts once); seeded = 1 where the seeded solve gave the result.
HD int u2p_gate_finish(const double *U, const Geom &geom, double gam, double *prim, int ph1, Gnr2 &st, bool need,
                       double W0, double vsq0, int vsq_from_W, int &n_iter, int &k_conv, int &seeded)
{
    n_iter = 0;
    k_conv = 0;
    seeded = 0;
    if (ph1 == -100) return -100;
    if (ph1 == 2 && !need) return 2;  // no start and no seed: utoprim_2d's code 2
    const int n_first = (ph1 == 0) ? 1 : 0;
    int n1 = 0, k1 = 0;
    if (need) {
        double pt[NPR];
#pragma unroll
        for (int i = 0; i < NPR; i++) pt[i] = prim[i];
        const int ret = utoprim_2d_seeded(U, geom, gam, pt, W0, vsq0, vsq_from_W, n1, k1);
        if (ret == 0) {
#pragma unroll
            for (int i = 0; i < NPR; i++) prim[i] = pt[i];
            n_iter = n_first + n1;
            k_conv = n_first + k1;
            seeded = 1;
            return 0;
        }
        if (ph1 == 2) {
            n_iter = n1;
            k_conv = n1;
            return 2;
        }
    }
    double U_tmp[NPR];
    u2p_transform(U, geom, U_tmp);
    U2P s;
    s.gam = gam;
    s.rgam = 1. / gam;
    s.K_atm = 0.;
    double Bcon[4], Qcon[4], ncon[4], QdotB;
    u2p_scalars(s, U_tmp, geom, Bcon, Qcon, ncon, QdotB);
    while (st.keep) gnr_step(s, st);
    int n2, k2;
    const int retval = gnr_end(st, n2, k2);
    double pt[5];
    const int ret = finish_2d(s, Bcon, Qcon, ncon, QdotB, retval, st.x0, st.x1, pt);
    if (ret == 0) {
#pragma unroll
        for (int i = 0; i < 5; i++) prim[i] = pt[i];
    }
    n_iter = n1 + n2;
    k_conv = n1 + k2;
    return ret;
}

// ---- Font fix #1 (utoprim_1dvsq2fix1.c) ----
// utoprim_1dvsq2fix1.c:586-614
HD double W_of_vsq_iso(const U2P &s, double vsq, double &p, double &rho, double &u)
{
    double gtmp = (1. - vsq);
    rho = s.D * sqrt(gtmp);
    p = s.K_atm * cr_pow(rho, s.gam);
    u = p / (s.gam - 1.);
    return ((rho + u + p) / gtmp);
}

// utoprim_1dvsq2fix1.c:398-519 with func_1d_gnr (:538-566) and validate_x (:372-384)
HD int gnr_1dvsq(const U2P &s, double &x)
{
    const double GAMMA = s.gam, G_ATM = s.gam;
    double f = 1., df = 1., errx = 1., W = 0., W_old;
    int n_iter = 0, i_extra = 0, doing_extra = 0;
    bool keep = true;
    while (keep) {
        double p, rho, u;
        double vsq = x;
        double Wv = W_of_vsq_iso(s, vsq, p, rho, u);
        double Wsq = Wv * Wv;
        double W3 = Wv * Wsq;
        double dWdvsq = ((GAMMA * (2. - G_ATM) * p + (GAMMA - 1.) * rho) / (2. * (GAMMA - 1.) * (1. - vsq) * (1. - vsq)));
        double fact_tmp = (s.Bsq + Wv);
        double resid = s.Qtsq - vsq * fact_tmp * fact_tmp + s.QdotBsq * (s.Bsq + 2. * Wv) / Wsq;
        double jac = -fact_tmp * (fact_tmp + 2. * dWdvsq * (vsq + s.QdotBsq / W3));
        double dx = -resid / jac;
        f = 0.5 * resid * resid;
        df = -2. * f;
        double x_old = x;
        x += dx;
        x = (x >= 1.0) ? (0.5 * (x_old + 1.)) : x;
        x = (x < -1.e-10) ? (0.5 * x_old) : x;
        x = fabs(x);
        W_old = W;
        W = W_of_vsq_iso(s, x, p, rho, u);
        errx = (W == 0.) ? fabs(W - W_old) : fabs((W - W_old) / W);
        errx += (x == 0.) ? fabs(x - x_old) : fabs((x - x_old) / x);
        if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) doing_extra = 1;
        if (doing_extra == 1) i_extra++;
        if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
            (n_iter >= (MAX_NEWT_ITER - 1)))
            keep = false;
        n_iter++;
    }
    if (!isfinite(f) || !isfinite(df)) return 2;
    if (fabs(errx) > MIN_NEWT_TOL) return 1;
    return 0;
}

// utoprim_1dvsq2fix1.c:143-190 + 242-370
HD int utoprim_1dvsq2fix1(const double *U, const Geom &geom, double gam, double K, double *prim)
{
#pragma unroll
    for (int i = B1; i <= B3; i++) prim[i] = U[i] / geom.g;
    if (U[0] <= 0.) return -100;
    double U_tmp[NPR];
    u2p_transform(U, geom, U_tmp);
    U2P s;
    s.gam = gam;
    s.K_atm = K;
    double Bcon[4], Qcon[4], ncon[4], QdotB;
    u2p_scalars(s, U_tmp, geom, Bcon, Qcon, ncon, QdotB);
    double utsq = utsq_of(geom, prim);
    if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
    if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
    double gammasq = 1. + utsq;
    double x = 1. - 1. / gammasq;
    int retval = gnr_1dvsq(s, x);
    if (retval != 0) return retval * 100 + 1;
    double vsq = x;
    if ((vsq >= 1.) || (vsq < 0.)) return 4;
    double p, rho0, u;
    double W = W_of_vsq_iso(s, vsq, p, rho0, u);
    double gtmp = sqrt(1. - vsq);
    double gamma = 1. / gtmp;
    if ((rho0 <= 0.) || (u <= 0.)) return 5;
    prim[RHO] = rho0;
    prim[UU] = u;
    u2p_recover_vel(s, Bcon, Qcon, ncon, QdotB, W, gamma, prim);
    return 0;
}

// ---- Font fix #2 (utoprim_1dfix1.c) ----
// gnr2 (utoprim_1dfix1.c:572-660) on func_gnr2_rho (:814-838); the tolerance
// branch there never sets doing_extra, so it stops at |drho/rho| <= 1e-15.
HD int gnr2_rho(const U2P &s, double &x)
{
    const double GAMMA = s.gam;
    double f = 1., df = 1., errx = 1.;
    int n_iter = 0;
    bool keep = true;
    while (keep) {
        double A = s.D * s.D;
        double B0 = A * GAMMA * s.K_atm;
        double B = B0 / (GAMMA - 1.);
        double rho = x;
        double W = s.W_for_gnr2;
        double C = cr_pow(rho, GAMMA - 1.);
        double resid = rho * W - A - B * C;
        double jac = W - B0 * C / rho;
        double dx = -resid / jac;
        f = 0.5 * resid * resid;
        df = -2. * f;
        x += dx;
        errx = 0.;
        errx += (x == 0.) ? fabs(dx) : fabs(dx / x);
        errx /= 1. * 1;
        x = fabs(x);
        if ((fabs(errx) <= NEWT_TOL2) || (n_iter >= (MAX_NEWT_ITER - 1))) keep = false;
        n_iter++;
    }
    if (!isfinite(f) || !isfinite(df) || !isfinite(x)) return 2;
    if (fabs(errx) > MIN_NEWT_TOL) return 1;
    return 0;
}

// utoprim_1dfix1.c:695-765 (func_1d_orig1)
HD void func_1d_orig1(U2P &s, double W, double &dx, double &f, double &df)
{
    const double GAMMA = s.gam, Bsq = s.Bsq, QdotBsq = s.QdotBsq, Qtsq = s.Qtsq, K_atm = s.K_atm;
    s.W_for_gnr2 = W;
    double rho_g, x_rho;
    rho_g = x_rho = s.rho_for_gnr2;
    int ntries = 0, retval;
    while ((retval = gnr2_rho(s, x_rho)) && (ntries++ < 10)) {
        rho_g *= 10.;
        x_rho = rho_g;
    }
    double rho = s.rho_for_gnr2 = x_rho;
    double Dc = s.D;
    double t1 = Dc * Dc;
    double t2 = QdotBsq * t1;
    double t3 = t2 * Bsq;
    double t5 = Bsq * Bsq;
    double t8 = t1 * Bsq;
    double t10 = t1 * W;
    double t21 = W * W;
    double t23 = rho * rho;
    double t26 = 1. / t1;
    double resid = (t3 + (2.0 * t2 + ((Qtsq - t5) * t1 + (-2.0 * t8 - t10) * W) * W) * W +
                    (t5 + (2.0 * Bsq + W) * W) * t21 * t23) *
                   t26 / t21;
    double t29 = t1 * t1;
    double t30 = QdotBsq * t29;
    double t32 = GAMMA * K_atm;
    double t33 = cr_pow(rho, 1.0 * GAMMA);
    double t34 = t32 * t33;
    double t38 = t23 * t33;
    double t51 = GAMMA * t1 * K_atm * t33;
    double t67 = t21 * W;
    double jac = -2.0 *
                 (t30 * Bsq * t34 +
                  (t30 * t34 +
                   ((-t38 * Bsq * t32 + Bsq * GAMMA * t1 * K_atm * t33) * t1 + (-t38 * GAMMA * K_atm + t51) * t1 * W) *
                       t21) *
                      W +
                  ((-t3 + (-t2 + (-t8 - t10) * t21) * W) * W + (-t5 - Bsq * W) * t67 * t23) * t23) *
                 t26 / (t51 - W * t23) / t67;
    dx = -resid / jac;
    f = 0.5 * resid * resid;
    df = -2. * f;
}

// utoprim_1dfix1.c:470-560
HD int gnr_1dfix(U2P &s, double &x)
{
    double f = 1., df = 1., errx = 1.;
    int n_iter = 0, i_extra = 0, doing_extra = 0;
    bool keep = true;
    while (keep) {
        double dx;
        func_1d_orig1(s, x, dx, f, df);
        x += dx;
        int i_increase = 0;
        while (((x * x * x * (x + 2. * s.Bsq) - s.QdotBsq * (2. * x + s.Bsq)) <= x * x * (s.Qtsq - s.Bsq * s.Bsq)) &&
               (i_increase < 10)) {
            x -= (1. * i_increase) * dx / 10.;
            i_increase++;
        }
        errx = (x == 0.) ? fabs(dx) : fabs(dx / x);
        x = fabs(x);
        if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) doing_extra = 1;
        if (doing_extra == 1) i_extra++;
        if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
            (n_iter >= (MAX_NEWT_ITER - 1)))
            keep = false;
        n_iter++;
    }
    if (!isfinite(f) || !isfinite(df) || !isfinite(x)) return 2;
    if (fabs(errx) > MIN_NEWT_TOL) return 1;
    return 0;
}

// utoprim_1dfix1.c:137-200 + 254-406
HD int utoprim_1dfix1(const double *U, const Geom &geom, double gam, double K, double *prim)
{
#pragma unroll
    for (int i = B1; i <= B3; i++) prim[i] = U[i] / geom.g;
    if (U[0] <= 0.) return -100;
    double U_tmp[NPR];
    u2p_transform(U, geom, U_tmp);
    U2P s;
    s.gam = gam;
    s.K_atm = K;
    double Bcon[4], Qcon[4], ncon[4], QdotB;
    u2p_scalars(s, U_tmp, geom, Bcon, Qcon, ncon, QdotB);
    double utsq = utsq_of(geom, prim);
    if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
    if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
    double gammasq = 1. + utsq;
    double gamma = sqrt(gammasq);
    double rho0 = s.D / gamma;
    double p = s.K_atm * cr_pow(rho0, gam);
    double u = p / (gam - 1.);
    double w = rho0 + u + p;
    double W_last = w * gammasq;
    W_last = raise_W(s, W_last);
    s.W_for_gnr2 = W_last;
    s.rho_for_gnr2 = rho0;
    double x = W_last;
    int retval = gnr_1dfix(s, x);
    double W = x;
    if ((retval != 0) || (W == FAIL_VAL)) return retval * 100 + 1;
    if (W <= 0. || W > W_TOO_BIG) return 3;
    double vsq = vsq_calc(s, W);
    if (vsq >= 1.) return 4;
    double gtmp = sqrt(1. - vsq);
    gamma = 1. / gtmp;
    rho0 = s.D * gtmp;
    p = s.K_atm * cr_pow(rho0, gam);
    u = p / (gam - 1.);
    if ((rho0 <= 0.) || (u <= 0.)) return 5;
    prim[RHO] = rho0;
    prim[UU] = u;
    u2p_recover_vel(s, Bcon, Qcon, ncon, QdotB, W, gamma, prim);
    return 0;
}

// ---------------------------------------------------------------------------
// Floors and Lorentz-factor limit: fixup.c:142-383 (DRIFT_FLOOR=1).
// rhoflr0/uuflr0 are the radius-only floors of get_rho_u_floor (fixup.c:117-140),
// tabulated per (i,j) on the host.  Returns the flags of what happened.
// ---------------------------------------------------------------------------
enum { FIX_FLOOR = 1, FIX_GAMMA_FAIL = 2, FIX_GAMMA_CAP = 4 };

// phys.c:255-271
HD void ucon_to_utcon(const double *ucon, const Geom &geom, double *utcon)
{
    double alpha = 1. / sqrt(-geom.gcon[0][0]);
    double beta[4];
#pragma unroll
    for (int j = 1; j < 4; j++) beta[j] = geom.gcon[0][j] * alpha * alpha;
    double gamma = alpha * ucon[0];
    utcon[0] = 0;
#pragma unroll
    for (int j = 1; j < 4; j++) utcon[j] = ucon[j] + gamma * beta[j] / alpha;
}

// phys.c:273-297
HD double ut_calc_3vel(const double *vcon, const Geom &geom)
{
    double AA = geom.gcov[0][0];
    double BB = 2. * (geom.gcov[0][1] * vcon[1] + geom.gcov[0][2] * vcon[2] + geom.gcov[0][3] * vcon[3]);
    double CC = geom.gcov[1][1] * vcon[1] * vcon[1] + geom.gcov[2][2] * vcon[2] * vcon[2] +
                geom.gcov[3][3] * vcon[3] * vcon[3] +
                2. * (geom.gcov[1][2] * vcon[1] * vcon[2] + geom.gcov[1][3] * vcon[1] * vcon[3] +
                      geom.gcov[2][3] * vcon[2] * vcon[3]);
    double DD = -1. / (AA + BB + CC);
    double one_over_alpha_sq = -geom.gcon[0][0];
    if (DD < one_over_alpha_sq) DD = one_over_alpha_sq;
    return sqrt(DD);
}

HD int fixup1zone(double *pv, const Geom &geom, double rhoflr0, double uuflr0, double gam, int &fatal)
{
    double pv_prefloor[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) pv_prefloor[k] = pv[k];
    double rhoflr = rhoflr0, uuflr = uuflr0;
    State q;
    get_state(pv, geom, q, fatal);
    double bsq = dot4(q.bcon, q.bcov);
    if (rhoflr < bsq / BSQORHOMAX) rhoflr = bsq / BSQORHOMAX;
    if (uuflr < bsq / BSQOUMAX) uuflr = bsq / BSQOUMAX;
    if (rhoflr < pv[UU] / UORHOMAX) rhoflr = pv[UU] / UORHOMAX;
    int dofloor = 0, out = 0;
    if (pv[RHO] < rhoflr) {
        pv[RHO] = rhoflr;
        dofloor = 1;
    }
    if (pv[UU] < uuflr) {
        pv[UU] = uuflr;
        dofloor = 1;
    }
    if (dofloor) out |= FIX_FLOOR;
    double trans;
    if (dofloor && (trans = 10. * bsq / dmin(pv[RHO], pv[UU]) - 1.) > 0.) {
        if (trans > 1.) trans = 1.;
        double betapar = -q.bcon[0] / ((bsq + SMALL) * q.ucon[0]);
        double betasq = betapar * betapar * bsq;
        double betasqmax = 1. - 1. / (GAMMAMAX * GAMMAMAX);
        if (betasq > betasqmax) betasq = betasqmax;
        double gamma = 1. / sqrt(1 - betasq);
        double ucondr[4], Bcon[4], Bcov[4], vcon[4], ucon[4], utcon[4];
#pragma unroll
        for (int k = 0; k < 4; k++) ucondr[k] = gamma * (q.ucon[k] + betapar * q.bcon[k]);
        Bcon[0] = 0.;
#pragma unroll
        for (int k = 1; k < 4; k++) Bcon[k] = pv[B1 - 1 + k];
        lower(Bcon, geom, Bcov);
        double udotB = dot4(q.ucon, Bcov);
        double Bsq = dot4(Bcon, Bcov);
        double B = sqrt(Bsq);
        double wold = pv_prefloor[RHO] + pv_prefloor[UU] * gam;
        double QdotB = udotB * wold * q.ucon[0];
        double wnew = pv[RHO] + pv[UU] * gam;
        double x = 2. * QdotB / (B * wnew * ucondr[0] + SMALL);
        double vpar = x / (ucondr[0] * (1. + sqrt(1. + x * x)));
        double one_over_ucondr_t = 1. / ucondr[0];
        vcon[0] = 1.;
#pragma unroll
        for (int k = 1; k < 4; k++) vcon[k] = vpar * Bcon[k] / (B + SMALL) + ucondr[k] * one_over_ucondr_t;
        double ut = ut_calc_3vel(vcon, geom);
#pragma unroll
        for (int k = 0; k < 4; k++) ucon[k] = ut * vcon[k];
        ucon_to_utcon(ucon, geom, utcon);
#pragma unroll
        for (int k = 1; k < 4; k++) pv[k + UU] = utcon[k] * trans + pv_prefloor[k + UU] * (1. - trans);
    }
    double gamma;
    if (gamma_calc(pv, geom, &gamma)) {
        out |= FIX_GAMMA_FAIL;
    } else if (gamma > GAMMAMAX) {
        double f = sqrt((GAMMAMAX * GAMMAMAX - 1.) / (gamma * gamma - 1.));
        pv[U1] *= f;
        pv[U2] *= f;
        pv[U3] *= f;
        out |= FIX_GAMMA_CAP;
    }
    return out;
}

// bounds.c:333-378: remove inflow through the radial boundaries.
HD void inflow_check(double *pr, const Geom &geom, int type, int &fatal)
{
    double ucon[4];
    ucon_calc(pr, geom, ucon, fatal);
    if (((ucon[1] > 0.) && (type == 0)) || ((ucon[1] < 0.) && (type == 1))) {
        double gamma;
        if (gamma_calc(pr, geom, &gamma)) fatal++;
        pr[U1] /= gamma;
        pr[U2] /= gamma;
        pr[U3] /= gamma;
        double alpha = 1. / sqrt(-geom.gcon[0][0]);
        double beta1 = geom.gcon[0][1] * alpha * alpha;
        pr[U1] = beta1 / alpha;
        double vsq = 0.;
#pragma unroll
        for (int j = 1; j < 4; j++)
#pragma unroll
            for (int k = 1; k < 4; k++) vsq += geom.gcov[j][k] * pr[U1 + j - 1] * pr[U1 + k - 1];
        if (fabs(vsq) < 1.e-13) vsq = 1.e-13;
        if (vsq >= 1.) vsq = 1. - 1. / (GAMMAMAX * GAMMAMAX);
        gamma = 1. / sqrt(1. - vsq);
        pr[U1] *= gamma;
        pr[U2] *= gamma;
        pr[U3] *= gamma;
    }
}

}  // namespace harm
