// harm_nse.hpp -- one point of the cooling-table generator: lepton and
// nucleon densities, the chemical (n <-> p) equilibrium, neutrino cooling
// rates and optical depths, and the total EOS, as __host__ __device__
// functions over harm_fermi.hpp.
//
// Follows number_den.f90 (fxnn, fxnp, lepton, neta, peta, eeta),
// mnewtf.f90 (mnewt with ludcmp/lubksb, here for the one-unknown systems the
// chain solves), neu_rate.f90 (pair, urca, brem, plasmon), accretion.f90
// (opt_dep), chem_eq.f90 (chem_eq, eqnucl, nuclear, funcnb, neutrino,
// rtbis_funcnb) and abundance.f90 (the table point).  Integer powers x**n
// are evaluated as the oracle build of the reference (amdflang -O2) compiles
// them: left-to-right products up to the 8th power, binary powering above
// (gfortran's powi uses binary powering throughout, so a gfortran build can
// differ in the last bit).  Undecorated literals keep their single-precision
// values (0.23, 13.8, 0.557, 0.3, 1./3.).
#pragma once
#include "harm_fermi.hpp"

namespace harm {

// 40-point Gauss-Legendre positive nodes and weights (gauss_legendre,
// gauss_integral.f90:118-219), and the 20-point Laguerre rule of FermiRules
// (gauss_laguerre, gauss_integral.f90:26-113).
struct NseRules {
    FermiRules f;
    double lx40[20], lw40[20];
};

HD double ipow(double x, int n)
{
    if (n <= 8) {
        double r = x;
        for (int i = 1; i < n; i++) r = r * x;
        return r;
    }
    double res = 1.0, sq = x;
    bool have = false;
    while (n) {
        if (n & 1) {
            res = have ? res * sq : sq;
            have = true;
        }
        n >>= 1;
        if (n) sq = sq * sq;
    }
    return res;
}

constexpr double NSE_SQRT2 = 1.4142135623730951;  // dsqrt(2.0d0)

// number_den.f90: densities of a Fermi species [fm^-3] at (eta, kT)
HD_CALL double fd_density(double mass, double xkt, double eta, const FermiRules &R)
{
    const double beta = xkt / mass, h = mass / FD_HC;
    const double c = NSE_SQRT2 / (FD_PI * FD_PI) * ipow(h, 3) * pow(beta, 1.5);
    double f1, f2, a, b;
    dfermi(0.5, eta, beta, R, f1, a, b);
    dfermi(1.5, eta, beta, R, f2, a, b);
    return c * (f1 + beta * f2);
}

// lepton (number_den.f90:372-395): e- and e+ densities; nl = n_e- - n_e+
HD_CALL void lepton(double xkt, double xetae, const FermiRules &R, double &nl, double &xnee, double &xnep)
{
    const double betae = xkt / FD_XME;
    xnee = fd_density(FD_XME, xkt, xetae, R);
    const double etapos = -xetae - (double)2.0f / betae;
    xnep = fd_density(FD_XME, xkt, etapos, R);
    nl = xnee - xnep;
}

// mnewt (mnewtf.f90:97-151) for one unknown: ludcmp of the 1x1 Jacobian is
// singular when it is zero (or NaN), and then the step is fvec itself.
template <class F>
HD double mnewt1(int ntrial, double x, double tolx, double tolf, F usrfun)
{
    for (int k = 0; k < ntrial; k++) {
        double fjac, fvec;
        usrfun(x, fjac, fvec);
        const double errf = 0. + fabs(fvec);
        if (errf <= tolf) return x;
        if (fabs(fjac) > 0.) fvec = fvec / fjac;
        const double errx = 0. + fabs(fvec);
        x = x + fvec;
        if (errx <= tolx) return x;
    }
    return x;
}

// usrfun_neta / usrfun_peta / usrfun_eeta (number_den.f90): central
// difference of the density over a step |x| 1e-8 (at least 1e-50)
template <class D>
HD void fd_usrfun(double x, double target, D dens, double &fjac, double &fvec)
{
    double xacc = fabs(x) * 1.0e-8;
    if (xacc < 1.0e-50) xacc = 1.0e-50;
    const double n0 = dens(x), np_ = dens(x + xacc), nm = dens(x - xacc);
    const double f = target - n0, fp = target - np_, fm = target - nm;
    fjac = (fp - fm) / (2.0 * xacc);
    fvec = -f;
}

HD_CALL double eta_of_density(double mass, double xkt, double n, const FermiRules &R)  // neta, peta
{
    double tol = 1.0e-8 * n;
    if (tol < 1.0e-50) tol = 1.0e-50;
    return mnewt1(200, -10.0, 1.0e-20, tol, [&](double x, double &fj, double &fv) {
        fd_usrfun(x, n, [&](double e) { return fd_density(mass, xkt, e, R); }, fj, fv);
    });
}

// eeta (number_den.f90:303-323): electron degeneracy for a net lepton density
HD_CALL double eeta(double xkt, double xne, const FermiRules &R)
{
    double tol = 1.0e-8 * xne;
    if (tol < 1.0e-50) tol = 1.0e-50;
    double e = mnewt1(200, 0.510998910 / xkt, 1.0e-20, tol, [&](double x, double &fj, double &fv) {
        fd_usrfun(x, xne,
                  [&](double eta) {
                      double nl, a, b;
                      lepton(xkt, eta, R, nl, a, b);
                      return nl;
                  },
                  fj, fv);
    });
    if (e > 0.0) e = 0.0;
    return e;
}

// pair (neu_rate.f90:18-96): e+e- annihilation emissivities
HD_CALL void pair(double xkt, double xetae, const FermiRules &R, double &ev, double &ev1, double &ev2, double &ev3)
{
    constexpr double QC = 1.023e23;
    const double betae = xkt / 0.510998910;
    double cv = 2.0 * 0.23 + 0.5, ca = 0.5;
    const double cv1e = cv * cv + ca * ca, cv2e = cv * cv - ca * ca;
    cv = 2.0 * 0.23 - 0.5;
    ca = -0.5;
    const double cv1m = cv * cv + ca * ca, cv2m = cv * cv - ca * ca;
    const double cv1t = cv1m, cv2t = cv2m;
    double f12, f32, f52, f72, a, b;
    auto moments = [&](double eta, double &m_1, double &m0, double &m1, double &m2) {
        dfermi(0.5, eta, betae, R, f12, a, b);
        dfermi(1.5, eta, betae, R, f32, a, b);
        dfermi(2.5, eta, betae, R, f52, a, b);
        dfermi(3.5, eta, betae, R, f72, a, b);
        m_1 = f12;
        m0 = f12 + betae * f32;
        m1 = f12 + 2.0 * betae * f32 + betae * betae * f52;
        m2 = f12 + 3.0 * betae * f32 + 3.0 * betae * betae * f52 + ipow(betae, 3) * f72;
    };
    double u_1, u0, u1, u2, fai_1, fai0, fai1, fai2;
    moments(xetae, u_1, u0, u1, u2);
    moments(-xetae - 2.0 / betae, fai_1, fai0, fai1, fai2);
    const double pre = QC / (36.0 * FD_PI) * (2.0 / ipow(FD_PI, 4)) * ipow(betae, 3);
    const double A = 8.0 * (fai1 * u2 + fai2 * u1) - 2.0 * (fai_1 * u2 + fai2 * u_1) + 7.0 * (fai0 * u1 + fai1 * u0) +
                     5.0 * (fai0 * u_1 + fai_1 * u0);
    const double B = fai0 * (u1 + u_1) + (fai_1 + fai1) * u0;
    ev1 = pre * (cv1e * A + 9.0 * cv2e * B);
    ev2 = pre * (cv1m * A + 9.0 * cv2m * B);
    ev3 = pre * (cv1t * A + 9.0 * cv2t * B);
    ev = ev1 + ev2 + ev3;
}

// The URCA-type integrands (neu_rate.f90 func1-3, chem_eq.f90 func1k..func3r):
// x sqrt(x^2 - 1/beta^2) (s (q - x) or s (x +- q))^P times an occupation and a
// blocking factor.  kind selects the reference function.
enum { NSE_F1 = 0, NSE_F2, NSE_F3, NSE_F1K, NSE_F1R, NSE_F2K, NSE_F2R, NSE_F3K, NSE_F3R };

template <int K>
HD double urca_integrand(double x, double xkt, double xetae, double b1)
{
    const double betae = xkt / FD_XME, q = (FD_XMN - FD_XMP) / xkt;
    const double ib = 1.0 / betae;
    double xe, f2, base;
    int pw;
    switch (K) {
    case NSE_F1:  xe = xetae - (x - ib); f2 = fabs(1.0 - b1 / (exp(q - x) + 1.0)); base = q - x; pw = 3; break;
    case NSE_F2:  xe = -xetae + (x - ib); f2 = fabs(1.0 - b1 / (exp(x - q) + 1.0)); base = x - q; pw = 3; break;
    case NSE_F3:  xe = (x - ib) - xetae; f2 = fabs(1.0 - b1 / (exp(q + x) + 1.0)); base = x + q; pw = 3; break;
    case NSE_F1K: xe = xetae - (x - ib); f2 = fabs(1.0 - b1 / (exp(q - x) + 1.0)); base = q - x; pw = 2; break;
    case NSE_F1R: xe = (x - ib) - xetae; f2 = fabs(b1 / (exp(q - x) + 1.0)); base = q - x; pw = 2; break;
    case NSE_F2K: xe = -xetae + (x - ib); f2 = fabs(1.0 - b1 / (exp(x - q) + 1.0)); base = x - q; pw = 2; break;
    case NSE_F2R: xe = +xetae - (x - ib); f2 = fabs(b1 / (exp(x - q) + 1.0)); base = x - q; pw = 2; break;
    case NSE_F3K: xe = -xetae + (x - ib); f2 = fabs(1.0 - b1 / (exp(x + q) + 1.0)); base = x + q; pw = 2; break;
    default:      xe = xetae - (x - ib); f2 = fabs(b1 / (exp(x + q) + 1.0)); base = x + q; pw = 2; break;  // F3R
    }
    const double f1 = 1.0 / (exp(xe) + 1.0);
    return x * sqrt(x * x - 1.0 / (betae * betae)) * ipow(base, pw) * f1 * f2;
}

template <int K>
HD double gauss_legendre40(const NseRules &R, double a, double b, double xkt, double xetae, double b1)
{
    const double center = 0.5 * (a + b), hlfrun = 0.5 * (b - a);
    double r = 0.0;
    for (int j = 0; j < 20; j++) {
        const double f1 = urca_integrand<K>(center + hlfrun * R.lx40[j], xkt, xetae, b1);
        const double f2 = urca_integrand<K>(center - hlfrun * R.lx40[j], xkt, xetae, b1);
        r = r + (f1 + f2) * R.lw40[j];
    }
    return r * hlfrun;
}

template <int K>
HD double gauss_laguerre20(const NseRules &R, double a, double b, double xkt, double xetae, double b1)
{
    double r = 0.0;
    for (int j = 0; j < 20; j++) r = r + urca_integrand<K>(a + b * R.f.gx[j], xkt, xetae, b1) * R.f.gw[j];
    return r * b;
}

// urca (neu_rate.f90:101-146); b1 is set to zero as the reference does
HD_CALL void urca(double xkt, double xnn, double xnp, double xetae, double xetaep, const NseRules &R, double &evep_k,
             double &even_k, double &evn_k)
{
    const double b1 = 0.0, t11 = xkt * 1.16059e-1, t6 = ipow(t11, 6);
    double a = FD_XME / xkt;
    const double b = (FD_XMN - FD_XMP) / xkt;
    evn_k = 1.14473926e37 * t6 * xnn * gauss_legendre40<NSE_F1>(R, a, b, xkt, xetae, b1);
    a = (FD_XMN - FD_XMP) / xkt;
    evep_k = 1.14473926e37 * t6 * xnp * gauss_laguerre20<NSE_F2>(R, a, 1.0, xkt, xetae, b1);
    a = FD_XME / xkt;
    even_k = 1.14473926e37 * t6 * xnn * gauss_laguerre20<NSE_F3>(R, a, 1.0, xkt, xetaep, b1);
}

// brem, plasmon (neu_rate.f90:208-259)
HD double brem(double nb, double xkt)
{
    const double rho10 = nb * 1.66e5, t11 = xkt * 1.16059e-1;
    return 3.35e27 * pow(rho10, 2.0) * pow(t11, 5.5);
}

HD double plasmon(double xkt, double xetae)
{
    const double t11 = xkt * 1.16059e-1, etae = xetae + 0.510998910 / xkt;
    const double gp = 5.565e-2 * sqrt((FD_PI * FD_PI + 3.0 * (etae * etae)) / 3.0);
    return 1.5e32 * ipow(t11, 9) * ipow(gp, 6) * exp(-gp) * (1.0 + gp) * (2.0 + (gp * gp) / (1.0 + gp));
}

// opt_dep (accretion.f90:63-98): absorption and scattering optical depths
HD_CALL void opt_dep(double hd, double xkt, double nnucl, double xnn, double xnp, double xetae, double xetaep,
                const NseRules &R, double &ta1, double &ta2, double &ts)
{
    const double t11 = xkt * 1.16059e-1;
    double ev, ev1, ev2, ev3, evep_k, even_k, evn_k;
    pair(xkt, xetae, R.f, ev, ev1, ev2, ev3);
    urca(xkt, xnn, xnp, xetae, xetaep, R, evep_k, even_k, evn_k);
    const double evbrem = brem(nnucl, xkt), evplasmon = plasmon(xkt, xetae);
    const double qe = ev1 + evep_k + even_k + evn_k + evbrem / 3.0 + evplasmon;
    const double qo = ev2 + evbrem / 3.0;
    const double xl = 5.67051e39 * 4.0 * 7.0 / 8.0 * ipow(t11, 4);
    ta1 = qe * hd / xl;
    ta2 = qo * hd / xl;
    const double cv = (double)(0.5f + 2.0f * 0.23f), alph = 1.25, xme = 0.510998910;
    const double csn = (1.0 + 5.0 * (alph * alph)) / 24.0;
    const double csp = (4.0 * ((cv - 1.0) * (cv - 1.0)) + 5.0 * (alph * alph)) / 24.0;
    const double r = xkt / xme;
    const double tsn = csn * 1.76e-5 * xnn * hd * (double)13.8f * (r * r);
    const double tsp = csp * 1.76e-5 * xnp * hd * (double)13.8f * (r * r);
    ts = tsn + tsp;
}

// The neutrino trapping loop shared by funcnb and abundance (chem_eq.f90 and
// abundance.f90:173-183): ten passes of eos_neu, h_dep, opt_dep, b_neu.
HD_CALL void trap_loop(double xkt, double p1, double e1, double HdFac, double nb, double nnucl, double xnn, double xnp,
                  double xetae, double xetaep, const NseRules &R, double &b1, double &b2, double &ta1, double &ta2,
                  double &ts)
{
    b1 = 0.0;
    b2 = 0.0;
    for (int i = 0; i < 10; i++) {
        double eneu, pneu, sneu;
        eos_neu(xkt, b1, b2, eneu, pneu, sneu);
        const double p = p1 + pneu, e = e1 + eneu;
        const double hd = h_dep(HdFac, nb, p, e);
        opt_dep(hd, xkt, nnucl, xnn, xnp, xetae, xetaep, R, ta1, ta2, ts);
        b_neu(ta1, ta2, ts, b1, b2);
    }
}

struct NuclCommon {  // common /nucl/ (chem_eq.f90)
    double xnnucl, xkt, xne0, HdFac, nb0;
};

// nuclear (chem_eq.f90): proton and neutron degeneracies for a proton fraction
HD_CALL void nuclear(double xkt, double xnb, double xp, const FermiRules &R, double &xetan, double &xetap)
{
    double xnp = xnb * xp, xnn = xnb - xnp;
    if (xnp <= 0.0) xnp = 0.0;
    xetap = eta_of_density(FD_XMP, xkt, xnp, R);
    if (xnn <= 0.0) xnn = 0.0;
    xetan = eta_of_density(FD_XMN, xkt, xnn, R);
}

// neutrino (chem_eq.f90): relative imbalance of the n <-> p rates
HD_CALL double neutrino(double b1, double xkt, double xnn, double xnp, double xetae, const NseRules &R)
{
    const double xetaep = -xetae - 2.0 * FD_XME / xkt;
    double a = FD_XME / xkt;
    const double b = (FD_XMN - FD_XMP) / xkt, t11 = xkt * 1.16059e-1, t5 = ipow(t11, 5);
    if (xnn < 0.0) xnn = 0.0;
    if (xnp < 0.0) xnp = 0.0;
    double s1 = gauss_legendre40<NSE_F1K>(R, a, b, xkt, xetae, b1);
    double s1r = gauss_legendre40<NSE_F1R>(R, a, b, xkt, xetae, b1);
    const double evn_k = 2.03813046e35 * t5 * xnn * (s1 + s1r);
    a = (FD_XMN - FD_XMP) / xkt;
    s1 = gauss_laguerre20<NSE_F2K>(R, a, 1.0, xkt, xetae, b1);
    s1r = gauss_laguerre20<NSE_F2R>(R, a, 1.0, xkt, xetae, b1);
    const double evep_k = 2.03813046e35 * t5 * xnp * (s1 + s1r);
    a = FD_XME / xkt;
    s1 = gauss_laguerre20<NSE_F3K>(R, a, 1.0, xkt, xetaep, b1);
    s1r = gauss_laguerre20<NSE_F3R>(R, a, 1.0, xkt, xetaep, b1);
    const double even_k = 2.03813046e35 * t5 * xnn * (s1 + s1r);
    if (fabs(evep_k) < 1.0e-80) return (evep_k - even_k - evn_k) / (fabs(evep_k) + 1.0e-80);
    return (evep_k - even_k - evn_k) / fabs(evep_k);
}

// funcnb (chem_eq.f90): the rate imbalance at proton fraction xp
HD_CALL double funcnb(double xp, const NuclCommon &c, const NseRules &R)
{
    const double xkt = c.xkt, betae = xkt / FD_XME;
    double xetan, xetap;
    nuclear(xkt, c.xnnucl, xp, R.f, xetan, xetap);
    const double nhe = c.xne0 / (double)2.0f;
    const double xnp = c.xnnucl * xp, xnn = c.xnnucl - xnp, xne = c.xnnucl * xp + c.xne0;
    const double xetae = eeta(xkt, xne, R.f);
    const double xetaep = -xetae - (double)2.0f / betae;
    double nl, xnee, xnep;
    lepton(xkt, xetae, R.f, nl, xnee, xnep);
    double e, p, s, er, pr, sr, ehe, phe, she;
    eos_npe(xkt, xetan, xetap, xetae, xnn, xnp, xnee, xnep, R.f, e, p, s);
    eos_rad(xkt, er, pr, sr);
    eos_he(xkt, nhe, ehe, phe, she);
    const double p1 = p + pr + phe, e1 = e + er + ehe;
    double b1, b2, ta1, ta2, ts;
    trap_loop(xkt, p1, e1, c.HdFac, c.nb0, c.xnnucl, xnn, xnp, xetae, xetaep, R, b1, b2, ta1, ta2, ts);
    return neutrino(b1, xkt, xnn, xnp, xetae, R);
}

// eqnucl / chem_eq (chem_eq.f90): the equilibrium proton fraction by Newton on
// funcnb (rtbis_funcnb), then the densities and degeneracies
HD_CALL void chem_eq(double xnb, double xkt, double hdf, const NseRules &R, double &xnn, double &xnp, double &xnee,
                double &xnep, double &xetan, double &xetap, double &xetae)
{
    const double xnuc0 = xnuc(xnb, xkt);
    NuclCommon c;
    c.nb0 = xnb;
    c.xnnucl = xnb * xnuc0;
    c.xne0 = fabs(1.0 - xnuc0) * xnb / 2.0;
    c.HdFac = hdf;
    c.xkt = xkt;
    const double X1 = 0.0, X2 = 1.0;
    double tol = 1.0e-8 * c.xnnucl;
    if (tol < 1.0e-50) tol = 1.0e-50;
    double xp = mnewt1(200, (X2 - X1) / (double)2.0f, 1.0e-20, tol, [&](double x, double &fj, double &fv) {
        fd_usrfun(x, 0.0, [&](double y) { return funcnb(y, c, R); }, fj, fv);
    });
    if (xp < X1) xp = X1;
    if (xp > X2) xp = X2;
    nuclear(xkt, c.xnnucl, xp, R.f, xetan, xetap);
    xnp = c.xnnucl * xp;
    xnn = c.xnnucl - xnp;
    const double xne = xnp + c.xne0;
    xetae = eeta(xkt, xne, R.f);
    double nl;
    lepton(xkt, xetae, R.f, nl, xnee, xnep);
    if (xnep < 1.0e-60) xnep = 0.0;
}

// abundance (abundance.f90:6-243): one table point.  rho [g cm^-3], T [K];
// e, p [erg cm^-3], Qnu [erg cm^-2 s^-1], tau, and the pair densities [fm^-3].
HD_CALL void abundance(double rho, double T, double HdFac, const NseRules &R, double &e, double &p, double &Qnu,
                  double &tau, double &xnee, double &xnep)
{
    const double xnb = rho / 1.66e15, xkt = T / 1.16059e10;
    const double betae = xkt / FD_XME;
    const double xnuc0 = xnuc(xnb, xkt), nnuc = xnb * xnuc0, nhe = (xnb - nnuc) / 4.0;
    double xnn, xnp, xetan, xetap, xetae;
    chem_eq(xnb, xkt, HdFac, R, xnn, xnp, xnee, xnep, xetan, xetap, xetae);
    const double xetaep = -xetae - 2.0 / betae;
    double nl;
    lepton(xkt, xetae, R.f, nl, xnee, xnep);
    double enucl, pnucl, snucl, erad, prad, srad, ehe, phe, she;
    eos_npe(xkt, xetan, xetap, xetae, xnn, xnp, xnee, xnep, R.f, enucl, pnucl, snucl);
    eos_rad(xkt, erad, prad, srad);
    eos_he(xkt, nhe, ehe, phe, she);
    const double p1 = pnucl + prad + phe, e1 = enucl + erad + ehe;
    double b1, b2, ta1, ta2, ts;
    trap_loop(xkt, p1, e1, HdFac, xnb, nnuc, xnn, xnp, xetae, xetaep, R, b1, b2, ta1, ta2, ts);
    double eneu, pneu, sneu;
    eos_neu(xkt, b1, b2, eneu, pneu, sneu);
    const double c557 = (double)0.557f, c3 = (double)0.3f;
    Qnu = 7.0 / 8.0 * FD_SIGMASB * ipow(T, 4) / (double)0.75f *
          (1. / (0.5 * ta1 + 0.5 * ts + c557 + c3 / ta1) + 1. / (0.5 * ta2 + 0.5 * ts + c557 + c3 / ta2));
    tau = ts + ta1 + ta2;
    p = (pnucl + prad + pneu + phe) * 1.6021772e33;
    e = (enucl + erad + eneu + ehe) * 1.6021772e33;
}

}  // namespace harm
