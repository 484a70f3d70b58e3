/*
 * cache_oracle.c -- TEST INFRASTRUCTURE (the oracle; never linked into the
 * product).  Plain-C restatement of the cooling-table cache lookup for the CPU:
 *
 *   orc_akima      interpolate_akima(), akima.c:9-160: secants, two points
 *                  extrapolated on each side (akima.c:56-82), node slopes by
 *                  Akima's weights (akima.c:84-104), a cubic per interval
 *                  (akima.c:105-124) and the forward-only interval search
 *                  over the output points (akima.c:125-158)
 *   orc_akima_der  spline_akima_der(), akima.c:196-231 (centred difference,
 *                  step t * 1e-6)
 *   orc_lin        spline_linear_val2(), akima.c:300-370, with dvec_bracket,
 *                  akima.c:234-298
 *   orc_cache_get  Cache4d::getvalues(), cache4d.c:863-1134, CONSTGRID mode,
 *                  with calcbox(), cache4d.c:408-440
 *
 * Like the reference, the spline fits coefficients for every interval of its
 * window (the device code forms only the interval it evaluates).  Pinned
 * against the reference's own akima.c and cache4d.c compiled from where they
 * lie (cache_ref_shim.cpp).  Statuses: 1 done, 0 outside the grid, 2 outside a
 * rolling axis' inner range, 3 points not yet computed, 4 bad z index.
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>

#define MAXW 8 /* window points */

static int bracket1(int n, const double *x, double t)
{
    for (int i = 2; i <= n - 1; i++)
        if (t < x[i - 1]) return i - 1;
    return n - 1;
}

int orc_akima(int n, const double *x, const double *y, int nt, const double *t, double *out)
{
    if (n < 3 || n > MAXW) return -2;
    double m[MAXW + 2], slope[MAXW], c1[MAXW], c2[MAXW];
    for (int k = 0; k + 1 < n; k++) {
        double h = x[k + 1] - x[k];
        if (h < 1e-12) return -1;
        m[k] = (y[k + 1] - y[k]) / h;
    }
    /* two extrapolated points left of x[0] */
    double xa = x[0] + x[1] - x[2];
    double xb = xa + x[0] - x[1];
    double ya = (x[0] - xa) * (m[1] - 2.0 * m[0]) + y[0];
    double gl1 = (y[0] - ya) / (x[0] - xa);
    double yb = (xa - xb) * (m[0] - 2.0 * gl1) + ya;
    double gl0 = (ya - yb) / (xa - xb);
    /* and two right of x[n-1] */
    double xc = x[n - 1] + x[n - 2] - x[n - 3];
    double xd = xc + x[n - 1] - x[n - 2];
    double yc = (2.0 * m[n - 2] - m[n - 3]) * (xc - x[n - 1]) + y[n - 1];
    m[n - 1] = (yc - y[n - 1]) / (xc - x[n - 1]);
    double yd = (2.0 * m[n - 1] - m[n - 2]) * (xd - xc) + yc;
    m[n] = (yd - yc) / (xd - xc);
    /* Akima's slope at node k from the secants k-2 .. k+1 */
    for (int k = 0; k < n; k++) {
        double s0 = k >= 2 ? m[k - 2] : (k == 1 ? gl1 : gl0);
        double s1 = k >= 1 ? m[k - 1] : gl1;
        double s2 = m[k], s3 = m[k + 1];
        double w0 = s3 - s2, w1 = s1 - s0;
        if (w0 < 0.0) w0 = -1.0 * w0;
        if (w1 < 0.0) w1 = -1.0 * w1;
        if ((w0 + w1) < 1e-9)
            slope[k] = 0.5 * (s1 + s2);
        else
            slope[k] = (w0 * s1 + w1 * s2) / (w0 + w1);
    }
    /* cubic coefficients of every interval */
    for (int k = 0; k + 1 < n; k++) {
        double h = x[k + 1] - x[k];
        double sec = (y[k + 1] - y[k]) / h;
        c1[k] = (3.0 * sec - 2.0 * slope[k] - slope[k + 1]) / h;
        c2[k] = (slope[k] + slope[k + 1] - 2.0 * sec) / (h * h);
    }
    int i = -1;
    for (int q = 0; q < nt; q++) {
        while (t[q] > x[i + 1] && i < n - 2) i++;
        if (i < 0) i = 0;
        double d = t[q] - x[i];
        out[q] = y[i] + slope[i] * d + c1[i] * d * d + c2[i] * d * d * d;
    }
    return 0;
}

int orc_akima_der(int n, const double *x, const double *y, double t, double *val, double *der)
{
    const double h = t * 1.0E-6;
    double ts[3] = {t - h, t, t + h}, v[3];
    int rc = orc_akima(n, x, y, 3, ts, v);
    if (rc) return rc;
    *val = v[1];
    *der = (v[2] - v[0]) / (2.0 * h);
    return 0;
}

double orc_lin(int n, const double *x, const double *y, double t)
{
    int l = bracket1(n, x, t); /* interval [l-1, l] */
    double r = 1.0 / (x[l] - x[l - 1]);
    return y[l - 1] + (t - x[l - 1]) * ((y[l] - y[l - 1]) * r);
}

/* calcbox(), cache4d.c:408-440; the linear window stops at the last point */
static void window(int n, const double *v, double t, int spline, int *lo, int *hi)
{
    int l = bracket1(n, v, t);
    if (!spline) {
        *lo = l - 1;
        *hi = l + 1 > n - 1 ? n - 1 : l + 1;
    } else {
        *lo = l - 2 > 0 ? l - 2 : 0;
        *hi = l + 1 < n ? l + 1 : n - 1;
    }
}

/* out[k*ostride], der[k*ostride] for k in [0, d) */
int orc_cache_get(int X, int Y, int Z, int D, const int *types, int whole, const double *vx, const double *vy,
                  const double *vz, const double *values, const uint8_t *counted, double x, double y, double z,
                  int zindex, int spline, const int *dimmask, double *out, double *der, size_t ostride)
{
    const double *ax[3] = {vx, vy, vz};
    const int nn[3] = {X, Y, Z};
    const double tv[3] = {x, y, z};
    int rolling = 0;
    for (int a = 0; a < 3; a++) {
        const double *v = ax[a];
        int n = nn[a];
        if (types[a] == 4) {
            if (tv[a] < v[1] || tv[a] > v[n - 2]) rolling = 1;
        } else if ((a < 2 || zindex == -1) && (tv[a] < v[0] || tv[a] > v[n - 1])) {
            return 0;
        }
    }
    if (rolling) return 2;
    if (zindex < -1 || zindex >= Z || (zindex == -1 && Z < (spline ? 3 : 2))) return 4;
    int lo[3], hi[3];
    window(X, vx, x, spline, &lo[0], &hi[0]);
    window(Y, vy, y, spline, &lo[1], &hi[1]);
    if (zindex >= 0)
        lo[2] = hi[2] = zindex;
    else
        window(Z, vz, z, spline, &lo[2], &hi[2]);
    const size_t XY = (size_t)X * Y;
    if (!whole)
        for (int iz = lo[2]; iz <= hi[2]; iz++)
            for (int ix = lo[0]; ix <= hi[0]; ix++)
                for (int iy = lo[1]; iy <= hi[1]; iy++) {
                    uint8_t f = counted[iz * XY + (size_t)ix * Y + iy];
                    if (f == 0) return 3;
                    if (f != 1) return 0;
                }
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    double ycol[MAXW * 64], xv[MAXW * 64], xd[MAXW * 64], zv[MAXW * 64], dy[64];
    /* along y for every (z, x) row of the window, all components */
    for (int a = 0; a < nz; a++) {
        for (int b = 0; b < nx; b++) {
            const double *row = values + (((size_t)(lo[2] + a) * XY + (size_t)(lo[0] + b) * Y + lo[1]) * D);
            for (int k = 0; k < D; k++) {
                if (spline && dimmask && !dimmask[k]) continue;
                for (int q = 0; q < ny; q++) ycol[q] = row[(size_t)q * D + k];
                if (!spline)
                    xv[b * D + k] = orc_lin(ny, vy + lo[1], ycol, y);
                else if (der)
                    orc_akima_der(ny, vy + lo[1], ycol, y, &xv[b * D + k], &xd[b * D + k]);
                else
                    orc_akima(ny, vy + lo[1], ycol, 1, &y, &xv[b * D + k]);
            }
        }
        /* then along x */
        for (int k = 0; k < D; k++) {
            if (spline && dimmask && !dimmask[k]) continue;
            double col[MAXW], dcol[MAXW];
            for (int b = 0; b < nx; b++) {
                col[b] = xv[b * D + k];
                dcol[b] = xd[b * D + k];
            }
            if (!spline) {
                zv[a * D + k] = orc_lin(nx, vx + lo[0], col, x);
            } else {
                orc_akima(nx, vx + lo[0], col, 1, &x, &zv[a * D + k]);
                if (der) orc_akima(nx, vx + lo[0], dcol, 1, &x, &dy[k]);
            }
        }
    }
    for (int k = 0; k < D; k++) {
        if (spline && dimmask && !dimmask[k]) continue;
        double v = zv[k];
        if (zindex == -1) {
            double col[MAXW];
            for (int a = 0; a < nz; a++) col[a] = zv[a * D + k];
            if (!spline)
                v = orc_lin(nz, vz + lo[2], col, z);
            else
                orc_akima(nz, vz + lo[2], col, 1, &z, &v);
        }
        out[k * ostride] = v;
        if (der && spline) der[k * ostride] = dy[k];
    }
    return 1;
}

/*
 * orc_cache_solve_T -- the temperature solve of cooleos.c:365-404 and 545-577:
 * mnewtc<1> (mnewt.h:104-205) over F = (e - e(rho, T)) / e with the Jacobian
 * -de/dT / e (cooleos.c:228-262, 296-306), every trial run as the reference
 * runs it (a rejected step leaves T where it is for the next trial).
 */
int orc_cache_solve_T(int X, int Y, int Z, int D, const int *types, int whole, const double *vx, const double *vy,
                      const double *vz, const double *values, const uint8_t *counted, double rho, double e,
                      int zindex, double T0, double lo, double hi, int ntrial, double tolf, double *T, double *berr)
{
    int mask[64] = {0};
    mask[0] = 1;
    double x = T0, best = T0, besterr = 1.0E100;
    int ret = 0;
    for (int k = 1; k <= ntrial; k++) {
        double val[64], der[64];
        if (orc_cache_get(X, Y, Z, D, types, whole, vx, vy, vz, values, counted, rho, x, 0.0, zindex, 1, mask, val,
                          der, 1) != 1) {
            *T = x;
            *berr = besterr;
            return -1;
        }
        double alpha = -der[0] / e;
        double bet = -((e - val[0]) / e);
        double errf = 0.0;
        errf += bet < 0.0 ? -bet : bet;
        if (errf <= tolf) {
            *T = x;
            *berr = errf;
            return 1;
        }
        if (errf < besterr) {
            if (k > 1) ret = 2;
            besterr = errf;
            best = x;
        }
        double big = alpha < 0.0 ? -alpha : alpha;
        if (big > 0.0) bet = bet / alpha; /* the 1x1 LU solve; big == 0 (or NaN) is singular */
        double nv = x + bet;
        if (nv >= lo && nv <= hi && nv != x) x = nv;
    }
    *T = best;
    *berr = besterr;
    return ret;
}
