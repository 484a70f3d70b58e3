// harm_fermi.hpp -- the gas physics of the tabulated-EOS generator (the
// reference's Fortran chain behind its cooling table): generalized
// Fermi-Dirac integrals and the per-species equations of state, as __host__
// __device__ functions so that a table point is one GPU thread.
//
// Follows gauss_integral.f90:241-440, 516-596 (dfermi: Aparicio's breakpoint
// scheme, ApJS 117, 632 (1998), with 20-point Gauss-Legendre and
// Gauss-Laguerre rules), eos.f90:31-183 (eos_npe, eos_rad, eos_neu, eos_he)
// and accretion.f90:25-136 (xnuc, h_dep, b_neu, c_neu).
//
// const_mod.f90 declares pi, HC, sigmasb, ... as real(8) parameters from
// default-real (single precision) literals, so their values are the float
// roundings of the printed digits; the same holds for undecorated literals
// such as 2.0/3.0 inside the routines.  They are kept so here.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#ifndef HD
#define HD __host__ __device__ inline
#endif
// Heavy routines keep their own register allocation on the device: inlined
// into one table-point kernel they exceed the register file and spill.
#define HD_CALL __host__ __device__ __attribute__((noinline)) inline

namespace harm {

constexpr double FD_PI = (double)3.1415926535897932f;  // const_mod.f90:17
constexpr double FD_HC = (double)197.3269718e0f;       // const_mod.f90:38, MeV fm
constexpr double FD_SIGMASB = (double)5.6e-5f;        // const_mod.f90:26
constexpr double FD_XME = 0.510998910, FD_XMP = 938.272013, FD_XMN = 939.56536;  // MeV

// 20-point rules: Gauss-Legendre positive nodes and weights on (-1, 1), and
// Gauss-Laguerre nodes with the weights times exp(node) (the integrand carries
// the weight function), both in increasing node order.
struct FermiRules {
    double lx[10], lw[10], gx[20], gw[20];
    int wave;  // device: 1 = the FD_GROUP lanes of a point share each dfermi (dfermi_group)
};

// fdfunc1 (gauss_integral.f90:354-393)
HD void fd_func1(double x, double dk, double eta, double theta, double &fd, double &fdeta, double &fdtheta)
{
    const double xdk = pow(x, dk), xdkp1 = x * xdk;
    const double dxst = sqrt(1.0 + 0.5 * x * theta);
    if ((x - eta) < 1.0e2) {
        const double factor = exp(x - eta), denom = factor + 1.0;
        fd = xdk * dxst / denom;
        fdeta = fd * factor / denom;
        const double denom2 = 4.0 * dxst * denom;
        fdtheta = xdkp1 / denom2;
    } else {
        const double factor = exp(eta - x);
        fd = xdk * dxst * factor;
        fdeta = fd;
        const double denom2 = 4.0 * dxst;
        fdtheta = xdkp1 / denom2 * factor;
    }
}

// fdfunc2 (gauss_integral.f90:398-438): the same after z^2 = x
HD void fd_func2(double x, double dk, double eta, double theta, double &fd, double &fdeta, double &fdtheta)
{
    const double xsq = x * x;
    const double xdk = pow(x, 2.0 * dk + 1.0), xdkp1 = xsq * xdk;
    const double dxst = sqrt(1.0 + 0.5 * xsq * theta);
    if ((xsq - eta) < 1.0e2) {
        const double factor = exp(xsq - eta), denom = factor + 1.0;
        fd = 2.0 * xdk * dxst / denom;
        fdeta = fd * factor / denom;
        const double denom2 = 4.0 * dxst * denom;
        fdtheta = 2.0 * xdkp1 / denom2;
    } else {
        const double factor = exp(eta - xsq);
        fd = 2.0 * xdk * dxst * factor;
        fdeta = fd;
        const double denom2 = 4.0 * dxst;
        fdtheta = 2.0 * xdkp1 / denom2 * factor;
    }
}

// dqleg020 (gauss_integral.f90:516-596) over (a, b)
template <int F>
HD void fd_leg20(const FermiRules &R, double a, double b, double dk, double eta, double theta, double &res,
                 double &dres, double &ddres)
{
    const double center = 0.5 * (a + b), hlfrun = 0.5 * (b - a);
    res = dres = ddres = 0.0;
    for (int j = 0; j < 10; j++) {
        const double x1 = center + hlfrun * R.lx[j], x2 = center - hlfrun * R.lx[j];
        double f1, d1, dd1, f2, d2, dd2;
        if (F == 1) {
            fd_func1(x1, dk, eta, theta, f1, d1, dd1);
            fd_func1(x2, dk, eta, theta, f2, d2, dd2);
        } else {
            fd_func2(x1, dk, eta, theta, f1, d1, dd1);
            fd_func2(x2, dk, eta, theta, f2, d2, dd2);
        }
        res = res + (f1 + f2) * R.lw[j];
        dres = dres + (d1 + d2) * R.lw[j];
        ddres = ddres + (dd1 + dd2) * R.lw[j];
    }
    res = res * hlfrun;
    dres = dres * hlfrun;
    ddres = ddres * hlfrun;
}

// dqlag020 (gauss_integral.f90): exp(-(x-a)/b) weighted rule on (a, inf)
HD void fd_lag20(const FermiRules &R, double a, double b, double dk, double eta, double theta, double &res,
                 double &dres, double &ddres)
{
    res = dres = ddres = 0.0;
    for (int j = 0; j < 20; j++) {
        double f, d, dd;
        fd_func1(a + b * R.gx[j], dk, eta, theta, f, d, dd);
        res = res + f * R.gw[j];
        dres = dres + d * R.gw[j];
        ddres = ddres + dd * R.gw[j];
    }
    res = res * b;
    dres = dres * b;
    ddres = ddres * b;
}

// Lanes per cooling-table point in the group kernel: 8 measured fastest for a
// 200x200 table (22 s; 16: 26 s, 4: 30 s, 32: 37 s, 64: 64 s).
constexpr int FD_GROUP = 8;

// dfermi's four quadratures with a group of G lanes (G = 4 ... 64, the
// groups aligned in the wavefront) on one (dk, eta, theta): its 50 integrand
// evaluations -- e = 0..29 the Gauss-Legendre node pairs of the three finite
// subintervals (e / 10, e % 10), e = 30..49 the Gauss-Laguerre nodes -- go to
// lane e % G of the group, each term formed as the serial loop forms it; then
// every lane of the group adds each rule's terms in the serial loop's order,
// so the sums are the serial ones bit for bit.  Every lane of the group must
// call it with the same values.
template <int G>
__device__ inline void dfermi_group(const FermiRules &R, double dk, double eta, double theta, double s1, double s2,
                                    double s3, double s12, double &fd, double &fdeta, double &fdtheta)
{
    constexpr int NE = 50, NS = (NE + G - 1) / G;
    const int lane = (int)(threadIdx.x & 63u), gl = lane % G, gbase = lane - gl;
    double t0[NS], t1[NS], t2[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const int e = gl + k * G;
        t0[k] = t1[k] = t2[k] = 0.0;
        if (e < 30) {
            const int q = e / 10, j = e - 10 * q;
            const double a = (q == 0) ? 0.0 : (q == 1) ? s1 : s2;
            const double b = (q == 0) ? s12 : (q == 1) ? s2 : s3;
            const double center = 0.5 * (a + b), hlfrun = 0.5 * (b - a);
            const double x1 = center + hlfrun * R.lx[j], x2 = center - hlfrun * R.lx[j];
            double f1, d1, dd1, f2, d2, dd2;
            if (q == 0) {
                fd_func2(x1, dk, eta, theta, f1, d1, dd1);
                fd_func2(x2, dk, eta, theta, f2, d2, dd2);
            } else {
                fd_func1(x1, dk, eta, theta, f1, d1, dd1);
                fd_func1(x2, dk, eta, theta, f2, d2, dd2);
            }
            t0[k] = (f1 + f2) * R.lw[j];
            t1[k] = (d1 + d2) * R.lw[j];
            t2[k] = (dd1 + dd2) * R.lw[j];
        } else if (e < NE) {
            const int j = e - 30;
            double f, d, dd;
            fd_func1(s3 + 1.0 * R.gx[j], dk, eta, theta, f, d, dd);
            t0[k] = f * R.gw[j];
            t1[k] = d * R.gw[j];
            t2[k] = dd * R.gw[j];
        }
    }
    double r[4], dr[4], ddr[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int base = (q < 3) ? 10 * q : 30, m = (q < 3) ? 10 : 20;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
        for (int j = 0; j < m; j++) {
            const int e = base + j, src = gbase + e % G;
            a0 = a0 + __shfl(t0[e / G], src, 64);
            a1 = a1 + __shfl(t1[e / G], src, 64);
            a2 = a2 + __shfl(t2[e / G], src, 64);
        }
        r[q] = a0;
        dr[q] = a1;
        ddr[q] = a2;
    }
    // fd_leg20's closing products by its half-run, fd_lag20's by b = 1
    const double h0 = 0.5 * (s12 - 0.0), h1 = 0.5 * (s2 - s1), h2 = 0.5 * (s3 - s2);
    const double r1 = r[0] * h0, r2 = r[1] * h1, r3 = r[2] * h2, r4 = r[3] * 1.0;
    const double dr1 = dr[0] * h0, dr2 = dr[1] * h1, dr3 = dr[2] * h2, dr4 = dr[3] * 1.0;
    const double ddr1 = ddr[0] * h0, ddr2 = ddr[1] * h1, ddr3 = ddr[2] * h2, ddr4 = ddr[3] * 1.0;
    fd = r1 + r2 + r3 + r4;
    fdeta = dr1 + dr2 + dr3 + dr4;
    fdtheta = ddr1 + ddr2 + ddr3 + ddr4;
}

// dfermi (gauss_integral.f90:241-349): F_k(eta, theta) = int_0^inf x^k
// sqrt(1 + x theta / 2) / (exp(x - eta) + 1) dx and its eta and theta
// derivatives, over Aparicio's four subintervals.  The breakpoint fit
// parameters are the published ones.
HD_CALL void dfermi(double dk, double eta, double theta, const FermiRules &R, double &fd, double &fdeta, double &fdtheta)
{
    constexpr double d = 3.3609, sg = 9.1186e-2, a1 = 6.7774, b1 = 1.1418, c1 = 2.9826, a2 = 3.7601,
                     b2 = 9.3719e-2, c2 = 2.1063e-2, d2 = 3.1084e1, e2 = 1.0056, a3 = 7.5669, b3 = 1.1695,
                     c3 = 7.5416e-1, d3 = 6.6558, e3 = -1.2819e-1;
    const double eta1 = sg * (eta - d);
    const double xi = (eta1 <= 5.0e1) ? log(1.0 + exp(eta1)) / sg : eta - d;
    const double xi2 = xi * xi;
    const double x1 = (a1 + b1 * xi + c1 * xi2) / (1.0 + c1 * xi);
    const double x2 = (a2 + b2 * xi + c2 * d2 * xi2) / (1.0 + e2 * xi + c2 * xi2);
    const double x3 = (a3 + b3 * xi + c3 * d3 * xi2) / (1.0 + e3 * xi + c3 * xi2);
    const double s1 = x1 - x2, s2 = x1, s3 = x1 + x3, s12 = sqrt(s1);
#if defined(__HIP_DEVICE_COMPILE__)
    if (R.wave) {
        dfermi_group<FD_GROUP>(R, dk, eta, theta, s1, s2, s3, s12, fd, fdeta, fdtheta);
        return;
    }
#endif
    double r1, dr1, ddr1, r2, dr2, ddr2, r3, dr3, ddr3, r4, dr4, ddr4;
    fd_leg20<2>(R, 0.0, s12, dk, eta, theta, r1, dr1, ddr1);
    fd_leg20<1>(R, s1, s2, dk, eta, theta, r2, dr2, ddr2);
    fd_leg20<1>(R, s2, s3, dk, eta, theta, r3, dr3, ddr3);
    fd_lag20(R, s3, 1.0, dk, eta, theta, r4, dr4, ddr4);
    fd = r1 + r2 + r3 + r4;
    fdeta = dr1 + dr2 + dr3 + dr4;
    fdtheta = ddr1 + ddr2 + ddr3 + ddr4;
}

// One Fermi species of eos_npe (eos.f90:55-100): pressure and energy from
// F_{3/2} and F_{5/2} at (eta, beta); entropy from them and the density.
HD void fd_species(double eta, double beta, double cp, double ce, double xkt, double n, const FermiRules &R,
                   double &e, double &p, double &s)
{
    double f2, f3, de, dt;
    dfermi(1.5, eta, beta, R, f2, de, dt);
    dfermi(2.5, eta, beta, R, f3, de, dt);
    p = cp * (f2 + 0.5 * beta * f3);
    e = ce * (f2 + beta * f3);
    s = (e + p) / xkt - n * (eta + 1.0 / beta);
    if (s <= 0.0) s = 0.0;
}

// eos_npe (eos.f90:31-113): free neutrons, protons and e-/e+ pairs.
// xkt [MeV]; eta's are chemical potentials over kT without rest mass; n's in
// fm^-3.  e, p in MeV fm^-3, s in k_B fm^-3.
HD_CALL void eos_npe(double xkt, double xetan, double xetap, double xetae, double xnn, double xnp, double xnee, double xnep,
                const FermiRules &R, double &e, double &p, double &s)
{
    const double betae = xkt / FD_XME, betap = xkt / FD_XMP, betan = xkt / FD_XMN;
    constexpr double c23 = (double)(2.0f / 3.0f), c32 = (double)(3.0f / 2.0f);
    const double pi2 = FD_PI * FD_PI;
    const double he = FD_XME / FD_HC, hp = FD_XMP / FD_HC, hn = FD_XMN / FD_HC;
    const double cpe = c23 * sqrt(2.0) / pi2 * (he * he * he) * FD_XME * pow(betae, 2.5);
    const double cpp = c23 * sqrt(2.0) / pi2 * (hp * hp * hp) * FD_XMP * pow(betap, 2.5);
    const double cpn = c23 * sqrt(2.0) / pi2 * (hn * hn * hn) * FD_XMN * pow(betan, 2.5);
    const double cee = c32 * cpe, cep = c32 * cpp, cen = c32 * cpn;
    double eee, pee, see, eep = 0.0, pep = 0.0, sep = 0.0, ep, pp, sp, en, pn, sn;
    fd_species(xetae, betae, cpe, cee, xkt, xnee, R, eee, pee, see);
    if (xnee > 0.0) {
        const double xetaep = -xetae - 2.0 / betae;
        fd_species(xetaep, betae, cpe, cee, xkt, xnep, R, eep, pep, sep);
    }
    fd_species(xetap, betap, cpp, cep, xkt, xnp, R, ep, pp, sp);
    fd_species(xetan, betan, cpn, cen, xkt, xnn, R, en, pn, sn);
    p = pee + pep + pp + pn;
    e = eee + eep + ep + en;
    s = see + sep + sp + sn;
}

// eos_rad (eos.f90:122-132): black-body radiation
HD void eos_rad(double xkt, double &e, double &p, double &s)
{
    const double h = xkt / FD_HC;
    p = FD_PI * FD_PI / (double)15.0f * xkt * (h * h * h) / (double)3.0f;
    e = 3.0 * p;
    s = 4.0 * p / xkt;
}

// eos_neu (eos.f90:143-153): trapped neutrinos with blocking factors b1, b2
HD void eos_neu(double xkt, double b1, double b2, double &e, double &p, double &s)
{
    const double h = xkt / FD_HC;
    p = 7.0 / 8.0 * (b1 + b2) * FD_PI * FD_PI / (double)15.0f * xkt * (h * h * h) / (double)3.0f;
    e = 3.0 * p;
    s = 4.0 * p / xkt;
}

// eos_he (eos.f90:164-181): ideal helium gas of number density nhe [fm^-3]
HD void eos_he(double xkt, double nhe, double &e, double &p, double &s)
{
    const double mhe = 3755.6740;
    if (nhe <= 0.0) {
        e = p = s = 0.0;
        return;
    }
    e = 1.5 * nhe * xkt;
    p = nhe * xkt;
    s = nhe * (2.5 + 1.5 * log(mhe * xkt / (FD_HC * FD_HC) / (double)2.0f / FD_PI) - log(nhe));
    if (s <= 0.0) e = p = s = 0.0;
}

// xnuc (accretion.f90:25-38): free-nucleon mass fraction
HD double xnuc(double nb, double xkt)
{
    const double rho10 = nb * 1.66e5, t11 = xkt * 1.16059e-1;
    double x = 295.5 / pow(rho10, 0.75) * pow(t11, 1.125) * exp(-0.8209 / t11);
    if (x > 1.0) x = 1.0;
    if (x < 0.0) x = 0.0;
    return x;
}

// h_dep (accretion.f90:43-58): disk height from the sound speed, with p and e
// in MeV fm^-3 (converted to erg cm^-3)
HD double h_dep(double HdFac, double nb, double p, double e)
{
    const double CC = 2.99792458e10, rho = nb * 1.66e15;
    const double pc = p * 1.6021772e33, ec = e * 1.6021772e33;
    double cs = CC * sqrt(pc / (rho * CC * CC + ec));
    if (cs > CC) cs = CC;
    return HdFac * cs;
}

// b_neu (accretion.f90:103-116): neutrino blocking factors
HD void b_neu(double ta1, double ta2, double ts, double &b1, double &b2)
{
    constexpr double third = (double)(1.0f / 3.0f);  // 1./3. is a default-real quotient
    const double te = ta1 + ts, tt = ta2 + ts, r3 = sqrt(3.0);
    b1 = (te / 2.0 + 1.0 / r3) / (te / 2.0 + 1.0 / r3 + third / ta1);
    b2 = (tt / 2.0 + 1.0 / r3) / (tt / 2.0 + 1.0 / r3 + third / ta2);
    if (ta1 <= 0.0 || ta2 <= 0.0) b1 = b2 = 0.0;
}

}  // namespace harm
