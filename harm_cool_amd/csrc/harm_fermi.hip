// harm_fermi.hip -- device batches of the tabulated-EOS generator's gas
// physics (harm_fermi.hpp): generalized Fermi-Dirac integrals and the
// per-species equations of state, one element per thread.
//
// The 20-point Gauss-Legendre and Gauss-Laguerre rules are computed here once
// (Newton iteration on P_20 and L_20 in long double, then rounded), not read
// from tables: the reference's rules were evaluated in 100-digit arithmetic
// and rounded to double, which the long-double roots reproduce.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/harm_cool.h"
#include "harm_fermi.hpp"
#include "harm_host.hpp"
#include "harm_nse.hpp"

using harm::FermiRules;
using harm::NseRules;

namespace {

constexpr int TPB_F = 256;

#define FCHK(x)                                                                              \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) return harm_host::set_error(std::string(#x) + ": " + hipGetErrorString(e_), 2); \
    } while (0)

// positive nodes (increasing) and weights of the n-point Gauss-Legendre rule
void legendre(int n, long double *x, long double *w)
{
    for (int i = 1; i <= n / 2; i++) {
        long double z = cosl(3.14159265358979323846264338327950288L * (i - 0.25L) / (n + 0.5L)), z1, pp;
        for (int it = 0; it < 100; it++) {
            long double p1 = 1.0L, p2 = 0.0L;
            for (int j = 1; j <= n; j++) {
                const long double p3 = p2;
                p2 = p1;
                p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0L);
            z1 = z;
            z = z1 - p1 / pp;
            if (fabsl(z - z1) <= 1e-21L) break;
        }
        long double p1 = 1.0L, p2 = 0.0L;
        for (int j = 1; j <= n; j++) {
            const long double p3 = p2;
            p2 = p1;
            p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
        }
        pp = n * (z * p1 - p2) / (z * z - 1.0L);
        x[n / 2 - i] = z;  // increasing order
        w[n / 2 - i] = 2.0L / ((1.0L - z * z) * pp * pp);
    }
}

void laguerre20(long double *x, long double *w)
{
    const int n = 20;
    long double z = 0.0L;
    for (int i = 0; i < n; i++) {
        // starting values of the usual Laguerre-root recurrence
        if (i == 0)
            z = 3.0L / (1.0L + 2.4L * n);
        else if (i == 1)
            z += 15.0L / (1.0L + 2.5L * n);
        else
            z += ((1.0L + 2.55L * (i - 1)) / (1.9L * (i - 1))) * (z - x[i - 2]);
        for (int it = 0; it < 200; it++) {
            long double p1 = 1.0L, p2 = 0.0L;
            for (int j = 1; j <= n; j++) {
                const long double p3 = p2;
                p2 = p1;
                p1 = ((2.0L * j - 1.0L - z) * p2 - (j - 1.0L) * p3) / j;
            }
            const long double pp = n * (p1 - p2) / z;
            const long double z1 = z;
            z = z1 - p1 / pp;
            if (fabsl(z - z1) <= 1e-21L * fabsl(z)) break;
        }
        x[i] = z;
        // w = x / ((n+1)^2 L_{n+1}(x)^2)
        long double p1 = 1.0L, p2 = 0.0L;
        for (int j = 1; j <= n + 1; j++) {
            const long double p3 = p2;
            p2 = p1;
            p1 = ((2.0L * j - 1.0L - z) * p2 - (j - 1.0L) * p3) / j;
        }
        w[i] = z / ((n + 1.0L) * (n + 1.0L) * p1 * p1);
    }
}

const NseRules &nse_rules();

const FermiRules &rules() { return nse_rules().f; }

const NseRules &nse_rules()
{
    static NseRules N;
    static std::once_flag once;
    std::call_once(once, [] {
        FermiRules &R = N.f;
        long double lx[10], lw[10], gx[20], gw[20], l40x[20], l40w[20];
        legendre(20, lx, lw);
        legendre(40, l40x, l40w);
        laguerre20(gx, gw);
        for (int i = 0; i < 10; i++) {
            R.lx[i] = (double)lx[i];
            R.lw[i] = (double)lw[i];
        }
        for (int i = 0; i < 20; i++) {
            R.gx[i] = (double)gx[i];
            R.gw[i] = (double)(gw[i] * expl(gx[i]));
            N.lx40[i] = (double)l40x[i];
            N.lw40[i] = (double)l40w[i];
        }
        R.wave = 0;
    });
    return N;
}

// The rules on the device, copied once per device before the first launch
// (kernels read them in place rather than copying a by-value argument to
// their stack for the noinline routines that take them by reference).
__constant__ NseRules c_rules;
__constant__ NseRules c_rules_group;  // the same rules for the lane-group-per-point kernel

constexpr int GEN_G = harm::FD_GROUP;  // lanes per point of k_abundance_group

int upload_rules()
{
    static bool done[64] = {};
    static std::mutex mu;
    int dev = 0;
    FCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if (dev < 0 || dev >= 64) return harm_host::set_error("device index out of range", 3);
    if (!done[dev]) {
        const NseRules &N = nse_rules();
        FCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_rules), &N, sizeof(NseRules)));
        NseRules W = N;
        W.f.wave = 1;
        FCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_rules_group), &W, sizeof(NseRules)));
        done[dev] = true;
    }
    return 0;
}

__global__ void __launch_bounds__(TPB_F) k_fermi(const double *__restrict__ dk,
                                                 const double *__restrict__ eta, const double *__restrict__ theta,
                                                 double *__restrict__ fd, double *__restrict__ fdeta,
                                                 double *__restrict__ fdtheta, size_t n)
{
    const size_t i = (size_t)blockIdx.x * TPB_F + threadIdx.x;
    if (i >= n) return;
    double a, b, c;
    harm::dfermi(dk[i], eta[i], theta[i], c_rules.f, a, b, c);
    fd[i] = a;
    if (fdeta) fdeta[i] = b;
    if (fdtheta) fdtheta[i] = c;
}

// in: [nin][n], out: [3][n] (e, p, s)
__global__ void __launch_bounds__(TPB_F) k_eos_species(int kind, const double *__restrict__ in,
                                                       double *__restrict__ out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * TPB_F + threadIdx.x;
    if (i >= n) return;
    auto I = [&](int k) { return in[(size_t)k * n + i]; };
    const FermiRules &R = c_rules.f;
    double e = 0., p = 0., s = 0.;
    switch (kind) {
    case HARM_EOS_NPE:
        harm::eos_npe(I(0), I(1), I(2), I(3), I(4), I(5), I(6), I(7), R, e, p, s);
        break;
    case HARM_EOS_RAD:
        harm::eos_rad(I(0), e, p, s);
        break;
    case HARM_EOS_NEU:
        harm::eos_neu(I(0), I(1), I(2), e, p, s);
        break;
    case HARM_EOS_HE:
        harm::eos_he(I(0), I(1), e, p, s);
        break;
    }
    out[i] = e;
    out[n + i] = p;
    out[2 * n + i] = s;
}

// One cooling-table point per thread (abundance.f90:6-243); out [6][n]: e, p,
// Qnu, tau, n_e-, n_e+.
__global__ void __launch_bounds__(64) k_abundance(const double *__restrict__ rho,
                                                  const double *__restrict__ T, const double *__restrict__ hdfac,
                                                  double *__restrict__ out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    double v[6];
    harm::abundance(rho[i], T[i], hdfac[i], c_rules, v[0], v[1], v[2], v[3], v[4], v[5]);
    for (int k = 0; k < 6; k++) out[(size_t)k * n + i] = v[k];
}

// One cooling-table point per group of GEN_G lanes (64 / GEN_G points per
// wavefront): every lane of a group runs its point's chain (the same values,
// so the same branches) and the group shares each dfermi's quadrature nodes
// (dfermi_group), which shortens the chain of the slowest points -- the ones
// whose Newton solves run long -- and spreads the points over the chip.
__global__ void __launch_bounds__(64) k_abundance_group(const double *__restrict__ rho,
                                                        const double *__restrict__ T,
                                                        const double *__restrict__ hdfac, double *__restrict__ out,
                                                        size_t n)
{
    const size_t i = (size_t)blockIdx.x * (64 / GEN_G) + threadIdx.x / GEN_G;
    if (i >= n) return;  // whole groups leave together
    double v[6];
    harm::abundance(rho[i], T[i], hdfac[i], c_rules_group, v[0], v[1], v[2], v[3], v[4], v[5]);
    if (threadIdx.x % GEN_G == 0)
        for (int k = 0; k < 6; k++) out[(size_t)k * n + i] = v[k];
}

unsigned nb(size_t n) { return (unsigned)((n + TPB_F - 1) / TPB_F); }

}  // namespace

extern "C" {

int harm_fermi_rules(double *out)
{
    if (!out) return harm_host::set_error("null argument", 1);
    const NseRules &N = nse_rules();
    const FermiRules &R = N.f;
    std::copy(R.lx, R.lx + 10, out);
    std::copy(R.lw, R.lw + 10, out + 10);
    std::copy(R.gx, R.gx + 20, out + 20);
    std::copy(R.gw, R.gw + 20, out + 40);
    std::copy(N.lx40, N.lx40 + 20, out + 60);
    std::copy(N.lw40, N.lw40 + 20, out + 80);
    return 0;
}

int harm_fermi_batch(const double *dk, const double *eta, const double *theta, double *fd, double *fdeta,
                     double *fdtheta, size_t n, void *stream)
{
    if (!n) return 0;
    if (!dk || !eta || !theta || !fd) return harm_host::set_error("null argument", 1);
    hipStream_t s = (hipStream_t)stream;
    int rc = upload_rules();
    if (rc) return rc;
    hipLaunchKernelGGL(k_fermi, dim3(nb(n)), dim3(TPB_F), 0, s, dk, eta, theta, fd, fdeta, fdtheta, n);
    FCHK(hipGetLastError());
    FCHK(hipStreamSynchronize(s));
    return 0;
}

int harm_eos_point_batch(const double *rho, const double *T, const double *hdfac, double *out, size_t n,
                         void *stream)
{
    if (!n) return 0;
    if (!rho || !T || !hdfac || !out) return harm_host::set_error("null argument", 1);
    hipStream_t s = (hipStream_t)stream;
    int rc = upload_rules();
    if (rc) return rc;
    const size_t nblk = (n + 64 / GEN_G - 1) / (64 / GEN_G);
    if (nblk > 0x7fffffffu) return harm_host::set_error("harm_eos_point_batch: too many points for one call", 3);
    hipLaunchKernelGGL(k_abundance_group, dim3((unsigned)nblk), dim3(64), 0, s, rho, T, hdfac, out, n);
    FCHK(hipGetLastError());
    FCHK(hipStreamSynchronize(s));
    return 0;
}

int harm_eos_point_batch_thread(const double *rho, const double *T, const double *hdfac, double *out, size_t n,
                                void *stream)
{
    if (!n) return 0;
    if (!rho || !T || !hdfac || !out) return harm_host::set_error("null argument", 1);
    hipStream_t s = (hipStream_t)stream;
    int rc = upload_rules();
    if (rc) return rc;
    hipLaunchKernelGGL(k_abundance, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, rho, T, hdfac, out, n);
    FCHK(hipGetLastError());
    FCHK(hipStreamSynchronize(s));
    return 0;
}

// The step's tabulated EOS from this generator (the reference's COOL_EOS build
// fills its per-radius cooling tables with the same abundance() points,
// cooleos.c:908-973, calc_cache_val): e and p of every (rho, T) grid point at
// one disk-height factor, in code units (cooleos.c:114-135: U_UNIT = RHO_UNIT
// c^2, the u of pressure_rho0_u_CoolEOS is e / U_UNIT).
int harm_eos_table_fermi(const harm_eos_spec *s, double hdfac, double *rho, double *T, double *e, double *p,
                         double *raw)
{
    if (!s || !rho || !T || !e || !p) return harm_host::set_error("null argument", 1);
    if (s->nr < 2 || s->nt < 2 || !(s->rho_unit > 0.) || !(s->lrho_hi > s->lrho_lo) || !(s->lT_hi > s->lT_lo) ||
        !(hdfac > 0.))
        return harm_host::set_error("bad EOS table spec (nr, nt >= 2, increasing ranges, hdfac > 0)", 4);
    const double c = 2.99792458e10;  // cm/s, cooleos.c:87
    const double u_unit = s->rho_unit * c * c;
    const size_t nr = (size_t)s->nr, nt = (size_t)s->nt, n = nr * nt;
    std::vector<double> in(3 * n), out(6 * n);
    for (size_t i = 0; i < nr; i++) rho[i] = std::pow(10., s->lrho_lo + (s->lrho_hi - s->lrho_lo) * i / (nr - 1));
    for (size_t j = 0; j < nt; j++) T[j] = std::pow(10., s->lT_lo + (s->lT_hi - s->lT_lo) * j / (nt - 1));
    for (size_t i = 0; i < nr; i++)
        for (size_t j = 0; j < nt; j++) {
            in[i * nt + j] = rho[i];
            in[n + i * nt + j] = T[j];
            in[2 * n + i * nt + j] = hdfac;
        }
    double *d = nullptr;
    FCHK(hipMalloc((void **)&d, 9 * n * sizeof(double)));
    int rc = 0;
    if (hipMemcpy(d, in.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = harm_host::set_error("harm_eos_table_fermi: copy to the device failed", 2);
    if (!rc) rc = harm_eos_point_batch(d, d + n, d + 2 * n, d + 3 * n, n, nullptr);
    if (!rc && hipMemcpy(out.data(), d + 3 * n, 6 * n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = harm_host::set_error("harm_eos_table_fermi: copy from the device failed", 2);
    (void)hipFree(d);
    if (rc) return rc;
    if (raw) std::copy(out.begin(), out.end(), raw);
    for (size_t i = 0; i < nr; i++) {
        for (size_t j = 0; j < nt; j++) {
            const double ev = out[i * nt + j], pv = out[n + i * nt + j];
            if (!std::isfinite(ev) || !std::isfinite(pv) || !(ev > 0.) || !(pv > 0.))
                return harm_host::set_error("harm_eos_table_fermi: no positive finite e, p at rho " +
                                                std::to_string(rho[i]) + " T " + std::to_string(T[j]),
                                            5);
            // the table's inversion p(rho, u) needs e strictly increasing in T
            if (j && !(ev > out[i * nt + j - 1]))
                return harm_host::set_error("harm_eos_table_fermi: e does not grow with T at rho " +
                                                std::to_string(rho[i]) + " T " + std::to_string(T[j]),
                                            5);
            e[i * nt + j] = ev / u_unit;
            p[i * nt + j] = pv / u_unit;
        }
        rho[i] /= s->rho_unit;
    }
    return 0;
}

int harm_eos_species_batch(int kind, const double *in, double *out, size_t n, void *stream)
{
    if (kind < HARM_EOS_NPE || kind > HARM_EOS_HE) return harm_host::set_error("harm_eos_species_batch: unknown kind", 3);
    if (!n) return 0;
    if (!in || !out) return harm_host::set_error("null argument", 1);
    hipStream_t s = (hipStream_t)stream;
    int rc = upload_rules();
    if (rc) return rc;
    hipLaunchKernelGGL(k_eos_species, dim3(nb(n)), dim3(TPB_F), 0, s, kind, in, out, n);
    FCHK(hipGetLastError());
    FCHK(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"
