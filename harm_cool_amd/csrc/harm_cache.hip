// harm_cache.hip -- the cooling-table cache: its file format (cache4d.c:152-326),
// its axis generator (cache4d.c:329-388) and batched lookups on the device,
// one query point per thread (harm_cache.hpp).
//
// File layout (the reference's, byte for byte): the packed 229-byte header,
// the x, y and z axes (double), the values [z][x][y][d] (double) and one
// computed flag per grid point (char; int in "AGACACHE_1.0" files).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/harm_cool.h"
#include "harm_cache.hpp"
#include "harm_host.hpp"

static_assert(sizeof(harm_cache_desc) == 229, "the cache header is the reference's packed 229-byte record");

struct harm_cache {
    harm_cache_desc desc;
    int whole;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr, *values = nullptr;
    uint8_t *counted = nullptr;
    int device = 0;
};

namespace {

constexpr int TPB_C = 256;

#define CCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return harm_host::set_error(std::string(#x) + ": " + hipGetErrorString(e_), 2); \
    } while (0)

size_t npoints(const harm_cache_desc &d) { return (size_t)d.x * d.y * d.z; }

int check_desc(const harm_cache_desc &d)
{
    if (d.x < 1 || d.y < 1 || d.z < 1 || d.d < 1)
        return harm_host::set_error("cache: every axis and the value count must be at least 1", 3);
    if ((double)d.x * d.y * d.z * d.d * 8.0 > 64.0 * (1ull << 30))
        return harm_host::set_error("cache: larger than 64 GiB", 3);
    return 0;
}

bool is_v1(const harm_cache_desc &d) { return std::strncmp(d.ver, "AGACACHE_1.0", 13) == 0; }

__global__ void __launch_bounds__(TPB_C) k_cache_lookup(harm::CacheView c, int mode, uint64_t mask,
                                                        const double *__restrict__ x, const double *__restrict__ y,
                                                        const double *__restrict__ z, const int *__restrict__ zindex,
                                                        double *__restrict__ out, double *__restrict__ dery,
                                                        int *__restrict__ status, size_t n)
{
    const size_t i = (size_t)blockIdx.x * TPB_C + threadIdx.x;
    if (i >= n) return;
    const int zi = zindex ? zindex[i] : -1;
    const double zv = z ? z[i] : 0.0;
    const int s = harm::cache_lookup(c, x[i], y[i], zv, zi, mode, mask, out + i, dery ? dery + i : nullptr, n);
    if (status) status[i] = s;
}

__global__ void __launch_bounds__(TPB_C) k_cache_solve_T(harm::CacheView c, const double *__restrict__ rho,
                                                         const double *__restrict__ e,
                                                         const int *__restrict__ zindex, double *__restrict__ T,
                                                         double lo, double hi, int ntrial, double tolf,
                                                         double *__restrict__ berr, int *__restrict__ ret, size_t n)
{
    const size_t i = (size_t)blockIdx.x * TPB_C + threadIdx.x;
    if (i >= n) return;
    double t, b;
    const int r = harm::cache_solve_T(c, rho[i], e[i], zindex ? zindex[i] : 0, T[i], lo, hi, ntrial, tolf, t, b);
    T[i] = t;
    if (berr) berr[i] = b;
    if (ret) ret[i] = r;
}

}  // namespace

extern "C" {

int harm_cache_axis(int type, int n, double vmin, double vmax, double rat, const double *cust, double *out)
{
    if (n < 1 || !out) return harm_host::set_error("harm_cache_axis: n >= 1 and an output are required", 1);
    double delta;
    switch (type) {
    case HARM_AXIS_LOG:
        delta = (log10(vmax) - log10(vmin)) / (n - 1);
        for (int r = 0; r < n; r++) out[r] = pow(10.0, log10(vmin) + delta * r);
        break;
    case HARM_AXIS_LINEAR:
        delta = (vmax - vmin) / (n - 1);
        for (int r = 0; r < n; r++) out[r] = vmin + delta * r;
        break;
    case HARM_AXIS_NONUNIFORM:
        delta = (vmax - vmin) * (rat - 1.0) / (pow(rat, n) - 1.0) * rat;
        out[0] = vmin;
        for (int r = 1; r < n; r++) {
            out[r] = out[r - 1] + delta;
            delta *= rat;
        }
        break;
    case HARM_AXIS_CUSTOM:
        if (n == 1)
            out[0] = vmin;
        else if (!cust)
            return harm_host::set_error("harm_cache_axis: a custom axis needs its values", 1);
        else
            std::memcpy(out, cust, n * sizeof(double));
        break;
    case HARM_AXIS_ROLLING:
        out[n / 2] = (vmin + vmax) / 2.0;
        for (int r = n / 2 + 1; r < n; r++) out[r] = out[r - 1] * (1.0 + rat);
        for (int r = n / 2 - 1; r >= 0; r--) out[r] = out[r + 1] * (1.0 - rat);
        break;
    default:
        return harm_host::set_error("harm_cache_axis: unknown axis type", 3);
    }
    return 0;
}

int harm_cache_peek(const char *path, harm_cache_desc *desc)
{
    if (!path || !desc) return harm_host::set_error("null argument", 1);
    FILE *f = std::fopen(path, "rb");
    if (!f) return harm_host::set_error(std::string("cannot open ") + path, 4);
    const size_t got = std::fread(desc, sizeof(harm_cache_desc), 1, f);
    std::fclose(f);
    if (got != 1) return harm_host::set_error(std::string("short cache header in ") + path, 4);
    return check_desc(*desc);
}

int harm_cache_read(const char *path, harm_cache_desc *desc, double *vx, double *vy, double *vz, double *values,
                    uint8_t *counted)
{
    if (!vx || !vy || !vz || !values || !counted) return harm_host::set_error("null argument", 1);
    int rc = harm_cache_peek(path, desc);
    if (rc) return rc;
    std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path, "rb"), std::fclose);
    if (!f) return harm_host::set_error(std::string("cannot open ") + path, 4);
    const size_t np = npoints(*desc);
    bool ok = std::fseek(f.get(), sizeof(harm_cache_desc), SEEK_SET) == 0 &&
              std::fread(vx, sizeof(double), desc->x, f.get()) == (size_t)desc->x &&
              std::fread(vy, sizeof(double), desc->y, f.get()) == (size_t)desc->y &&
              std::fread(vz, sizeof(double), desc->z, f.get()) == (size_t)desc->z &&
              std::fread(values, sizeof(double), np * desc->d, f.get()) == np * desc->d;
    if (ok) {
        if (is_v1(*desc)) {  // one int per flag (cache4d.c:209-222)
            std::vector<int> tmp(np);
            ok = std::fread(tmp.data(), sizeof(int), np, f.get()) == np;
            for (size_t i = 0; ok && i < np; i++) counted[i] = (uint8_t)tmp[i];
        } else {
            ok = std::fread(counted, 1, np, f.get()) == np;
        }
    }
    if (!ok) return harm_host::set_error(std::string("short cache file ") + path, 4);
    return 0;
}

int harm_cache_write(const char *path, const harm_cache_desc *desc, const double *vx, const double *vy,
                     const double *vz, const double *values, const uint8_t *counted)
{
    if (!path || !desc || !vx || !vy || !vz || !values || !counted) return harm_host::set_error("null argument", 1);
    int rc = check_desc(*desc);
    if (rc) return rc;
    std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path, "wb"), std::fclose);
    if (!f) return harm_host::set_error(std::string("cannot create ") + path, 4);
    const size_t np = npoints(*desc);
    // the flags are written as chars whatever the header's version (cache4d.c:301-305)
    const bool ok = std::fwrite(desc, sizeof(harm_cache_desc), 1, f.get()) == 1 &&
                    std::fwrite(vx, sizeof(double), desc->x, f.get()) == (size_t)desc->x &&
                    std::fwrite(vy, sizeof(double), desc->y, f.get()) == (size_t)desc->y &&
                    std::fwrite(vz, sizeof(double), desc->z, f.get()) == (size_t)desc->z &&
                    std::fwrite(values, sizeof(double), np * desc->d, f.get()) == np * desc->d &&
                    std::fwrite(counted, 1, np, f.get()) == np;
    if (!ok || std::fflush(f.get()) != 0) return harm_host::set_error(std::string("write failed: ") + path, 4);
    return 0;
}

void harm_cache_free(harm_cache *c)
{
    if (!c) return;
    int prev = 0;
    if (hipGetDevice(&prev) == hipSuccess && hipSetDevice(c->device) == hipSuccess) {
        (void)hipFree(c->vx);
        (void)hipFree(c->vy);
        (void)hipFree(c->vz);
        (void)hipFree(c->values);
        (void)hipFree(c->counted);
        (void)hipSetDevice(prev);
    }
    delete c;
}

static int cache_upload(harm_cache *c, const double *vx, const double *vy, const double *vz, const double *values,
                        const uint8_t *counted)
{
    const harm_cache_desc &d = c->desc;
    const size_t np = npoints(d);
    CCHK(hipGetDevice(&c->device));
    CCHK(hipMalloc(&c->vx, d.x * sizeof(double)));
    CCHK(hipMalloc(&c->vy, d.y * sizeof(double)));
    CCHK(hipMalloc(&c->vz, d.z * sizeof(double)));
    CCHK(hipMalloc(&c->values, np * d.d * sizeof(double)));
    CCHK(hipMalloc(&c->counted, np));
    CCHK(hipMemcpy(c->vx, vx, d.x * sizeof(double), hipMemcpyHostToDevice));
    CCHK(hipMemcpy(c->vy, vy, d.y * sizeof(double), hipMemcpyHostToDevice));
    CCHK(hipMemcpy(c->vz, vz, d.z * sizeof(double), hipMemcpyHostToDevice));
    CCHK(hipMemcpy(c->values, values, np * d.d * sizeof(double), hipMemcpyHostToDevice));
    CCHK(hipMemcpy(c->counted, counted, np, hipMemcpyHostToDevice));
    return 0;
}

harm_cache *harm_cache_create(const harm_cache_desc *desc, const double *vx, const double *vy, const double *vz,
                              const double *values, const uint8_t *counted)
{
    if (!desc || !vx || !vy || !vz || !values || !counted) {
        harm_host::set_error("null argument", 1);
        return nullptr;
    }
    if (check_desc(*desc)) return nullptr;
    const harm_cache_desc &d = *desc;
    if (d.mode != 0) {
        harm_host::set_error("harm_cache_create: only a constant-grid cache (mode 0) is looked up by grid", 3);
        return nullptr;
    }
    if (d.x < 3 || d.y < 3 || d.d > 64 || (d.ztype == HARM_AXIS_ROLLING && d.z < 3)) {
        harm_host::set_error("harm_cache_create: x and y need at least 3 points, a rolling z axis 3, and d <= 64", 3);
        return nullptr;
    }
    // Akima's fit needs increasing abscissae at least 1e-12 apart (akima.c:39-43)
    const double *ax[3] = {vx, vy, vz};
    const int an[3] = {d.x, d.y, d.z};
    for (int a = 0; a < 3; a++)
        for (int i = 1; i < an[a]; i++)
            if (!(ax[a][i] - ax[a][i - 1] >= 1e-12)) {
                harm_host::set_error("harm_cache_create: an axis is not increasing by at least 1e-12", 3);
                return nullptr;
            }
    auto *c = new harm_cache;
    c->desc = d;
    const size_t np = npoints(d);
    size_t ones = 0;
    for (size_t i = 0; i < np; i++) ones += counted[i] == 1;
    // a file that says it is whole is taken at its word, as the reference does
    c->whole = d.wholecachecalculated || ones == np;
    if (cache_upload(c, vx, vy, vz, values, counted)) {
        harm_cache_free(c);
        return nullptr;
    }
    return c;
}

harm_cache *harm_cache_open(const char *path)
{
    harm_cache_desc d;
    if (harm_cache_peek(path, &d)) return nullptr;
    const size_t np = npoints(d);
    std::vector<double> vx(d.x), vy(d.y), vz(d.z), values(np * d.d);
    std::vector<uint8_t> counted(np);
    if (harm_cache_read(path, &d, vx.data(), vy.data(), vz.data(), values.data(), counted.data())) return nullptr;
    return harm_cache_create(&d, vx.data(), vy.data(), vz.data(), values.data(), counted.data());
}

}  // extern "C"

// The device view of a cache and its header, for the step's cache EOS.
int harm_cache_view_of(const harm_cache *c, harm::CacheView &v, harm_cache_desc &d, int &device)
{
    if (!c) return harm_host::set_error("null cache", 1);
    v = harm::CacheView{c->desc.x,     c->desc.y,  c->desc.z,  c->desc.d,  c->desc.xtype, c->desc.ytype, c->desc.ztype,
                        c->whole,      c->vx,      c->vy,      c->vz,      c->values,     c->counted};
    d = c->desc;
    device = c->device;
    return 0;
}

extern "C" {

int harm_cache_info(const harm_cache *c, harm_cache_desc *desc, int *whole)
{
    if (!c) return harm_host::set_error("null cache", 1);
    if (desc) *desc = c->desc;
    if (whole) *whole = c->whole;
    return 0;
}

int harm_cache_lookup(const harm_cache *c, int mode, const int *dimmask, const double *x, const double *y,
                      const double *z, const int *zindex, double *out, double *dery, int *status, size_t n,
                      void *stream)
{
    if (!c) return harm_host::set_error("null cache", 1);
    if (mode != HARM_INTERP_LIN && mode != HARM_INTERP_SPLINE)
        return harm_host::set_error("harm_cache_lookup: mode is HARM_INTERP_LIN or HARM_INTERP_SPLINE", 3);
    if (!n) return 0;
    if (!x || !y || !out || (!z && !zindex)) return harm_host::set_error("harm_cache_lookup: x, y, out and z or a z index are required", 1);
    uint64_t mask = 0;
    for (int k = 0; k < c->desc.d; k++)
        if (!dimmask || dimmask[k]) mask |= 1ull << k;
    harm::CacheView v{c->desc.x, c->desc.y, c->desc.z, c->desc.d, c->desc.xtype, c->desc.ytype, c->desc.ztype,
                      c->whole, c->vx, c->vy, c->vz, c->values, c->counted};
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_cache_lookup, dim3((unsigned)((n + TPB_C - 1) / TPB_C)), dim3(TPB_C), 0, s, v, mode, mask, x,
                       y, z, zindex, out, dery, status, n);
    CCHK(hipGetLastError());
    CCHK(hipStreamSynchronize(s));
    return 0;
}

int harm_cache_solve_T(const harm_cache *c, const double *rho, const double *e, const int *zindex, double *T,
                       double Tlo, double Thi, int ntrial, double tolf, double *berr, int *ret, size_t n, void *stream)
{
    if (!c) return harm_host::set_error("null cache", 1);
    if (!n) return 0;
    if (!rho || !e || !T) return harm_host::set_error("harm_cache_solve_T: rho, e and T are required", 1);
    if (ntrial < 1) return harm_host::set_error("harm_cache_solve_T: ntrial >= 1", 3);
    if (Tlo == 0. && Thi == 0.) {
        // the reference's bounds TMin (1 + 1e-6), TMax (1 - 1e-6) (cooleos.c:368-369)
        // from the y axis the cache holds (a rolling axis' header range is not it)
        double ends[2];
        CCHK(hipMemcpy(&ends[0], c->vy, sizeof(double), hipMemcpyDeviceToHost));
        CCHK(hipMemcpy(&ends[1], c->vy + (c->desc.y - 1), sizeof(double), hipMemcpyDeviceToHost));
        Tlo = ends[0] * (1. + 1.e-6);
        Thi = ends[1] * (1. - 1.e-6);
    }
    if (!(Tlo < Thi)) return harm_host::set_error("harm_cache_solve_T: needs Tlo < Thi", 3);
    harm::CacheView v{c->desc.x, c->desc.y, c->desc.z, c->desc.d, c->desc.xtype, c->desc.ytype, c->desc.ztype,
                      c->whole, c->vx, c->vy, c->vz, c->values, c->counted};
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_cache_solve_T, dim3((unsigned)((n + TPB_C - 1) / TPB_C)), dim3(TPB_C), 0, s, v, rho, e,
                       zindex, T, Tlo, Thi, ntrial, tolf, berr, ret, n);
    CCHK(hipGetLastError());
    CCHK(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"
