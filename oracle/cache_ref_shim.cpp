/*
 * cache_ref_shim.cpp -- TEST INFRASTRUCTURE.  extern "C" entry points into the
 * reference's own cooling-table cache (src/cache4d.c, src/akima.c, compiled
 * where they lie in the reference tree into oracle/_ref/libcacheref.so).
 *
 * Used only to pin oracle/cache_oracle.c and the cache file layout: a Cache4d
 * opened on a complete file (every point computed) interpolates without
 * starting its worker processes; one built on a file name that does not exist
 * only lays out its axes.  Its destructor saves nothing unless a point was
 * computed, which never happens here.
 */
#include <cstring>
#include <ctime>  // cache4d.h uses time_t without including it

#include "akima.h"
#include "cache4d.h"
#include "mnewt.h"

static int no_calc(double, double, double, double *) { return 1; }

extern "C" {

int cref_desc_size(void) { return (int)sizeof(CacheDesc); }

/* A Cache4d on an existing complete cache file (custom axes, if any, given). */
void *cref_open(const char *path, double *xcust, double *ycust, double *zcust)
{
    CacheDesc d;
    std::memset(&d, 0, sizeof d);
    d.xcust = xcust;
    d.ycust = ycust;
    d.zcust = zcust;
    return new Cache4d(&d, no_calc, path);
}

/* The axes a new cache lays out for a header (path must not exist). */
int cref_axes(const void *desc, const char *path, double *vx, double *vy, double *vz)
{
    CacheDesc d;
    std::memcpy(&d, desc, sizeof d);
    Cache4d *c = new Cache4d(&d, no_calc, path);
    std::memcpy(vx, c->VectorX, sizeof(double) * c->X);
    std::memcpy(vy, c->VectorY, sizeof(double) * c->Y);
    std::memcpy(vz, c->VectorZ, sizeof(double) * c->Z);
    delete c;
    return 0;
}

int cref_get(void *c, double x, double y, double z, int spline, double *ret, double *der, int zindic, int *dimmask)
{
    return ((Cache4d *)c)->getvalues(x, y, z, spline ? SPLINE : LIN, ret, der, zindic, dimmask);
}

/* What load() read: the header, axes, values and flags (arrays sized from it). */
int cref_arrays(void *cv, void *desc, double *vx, double *vy, double *vz, double *values, unsigned char *counted)
{
    Cache4d *c = (Cache4d *)cv;
    const size_t np = (size_t)c->X * c->Y * c->Z;
    std::memcpy(desc, &c->cdesc, sizeof(CacheDesc));
    if (vx) {
        std::memcpy(vx, c->VectorX, sizeof(double) * c->X);
        std::memcpy(vy, c->VectorY, sizeof(double) * c->Y);
        std::memcpy(vz, c->VectorZ, sizeof(double) * c->Z);
        std::memcpy(values, c->MValues, sizeof(double) * np * c->D);
        std::memcpy(counted, c->MCounted, np);
    }
    return 0;
}

void cref_close(void *c) { delete (Cache4d *)c; }

void cref_spline_akima(int ndim, int *dimmask, int ndata, double *t, double *y, double tval, double *out)
{
    spline_akima(ndim, dimmask, ndata, t, y, tval, out);
}

void cref_spline_akima_der(int ndim, int *dimmask, int ndata, double *t, double *y, double tval, double *out,
                           double *der)
{
    spline_akima_der(ndim, dimmask, ndata, t, y, tval, out, der);
}

void cref_linear(int ndim, int ndata, double *t, double *y, double tval, double *out)
{
    spline_linear_val2(ndim, ndata, t, y, tval, out);
}

struct SolveArgs {
    Cache4d *c;
    double rho, e;
    int zidx;
};

/* The function cooleos.c hands mnewtc<1> (usrfun_mnewt_rho_e with
 * mnewt_mini_rho_e_der, cooleos.c:228-262, 296-306), over a Cache4d. */
static void rho_e_fun(double x[2], double a[2][2], double b[2], void *par)
{
    SolveArgs *s = (SolveArgs *)par;
    double val[64], der[64];
    int mask[64] = {1};
    s->c->getvalues(s->rho, x[1], 0.0, SPLINE, val, der, s->zidx, mask);
    a[1][1] = -der[0] / s->e;
    b[1] = -((s->e - val[0]) / s->e);
}

/* One temperature solve by the reference's own mnewtc<1> (mnewt.h:104-205). */
int cref_solve_T(void *c, double rho, double e, int zidx, double T0, double lo, double hi, int ntrial, double tolf,
                 double *T, double *berr)
{
    SolveArgs s{(Cache4d *)c, rho, e, zidx};
    double x[2] = {0.0, T0}, mn[2] = {0.0, lo}, mx[2] = {0.0, hi};
    int r = mnewtc<1>(ntrial, x, tolf, mn, mx, rho_e_fun, &s, berr, 1, nullptr);
    *T = x[1];
    return r;
}

}  // extern "C"
