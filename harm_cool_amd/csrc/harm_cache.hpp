// harm_cache.hpp -- lookups in the cooling-table cache: a grid of D values at
// every point of a 3D (x, y, z) = (rho, T, H_d factor) grid, stored
// [z][x][y][D] (cache4d.c:152-260). One lookup is cache4d.c's getvalues() in
// CONSTGRID mode (cache4d.c:863-1134): a window of grid points around the
// query along each axis, interpolated along y, then x, then z, either
// piecewise-linearly (akima.c:300-370) or with Akima's local spline (H. Akima,
// J. ACM 17(4), 589-602, 1970) fitted to the window alone, with the
// two-point extrapolation at each end and the slope rule of akima.c:9-160.
//
// Every operation is +, -, * or / in the reference's order, so with
// contraction off the host and the device give the reference's bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace harm {

#define HC_HD __host__ __device__ inline

enum CacheStatus : int {
    CACHE_OUT_OF_RANGE = 0,  // getvalues() returns 0 (cache4d.c:896-929)
    CACHE_OK = 1,
    CACHE_NEEDS_RECENTRE = 2,  // outside a rolling axis' inner range: the reference moves the grid
    CACHE_MISSING = 3,         // the window holds points not yet computed
    CACHE_BAD_ZINDEX = 4,      // z index outside [0, Z)
};

enum { CACHE_LIN = 0, CACHE_SPLINE = 1 };
enum { AXIS_LOG = 0, AXIS_LINEAR, AXIS_CUSTOM, AXIS_NONUNIFORM, AXIS_ROLLING };

// 1-based index `left` of the bracket [x[left-1], x[left]] of xval in the
// increasing x[0..n): the first i in 2..n-1 with xval < x[i-1] gives i-1,
// none gives n-1 (akima.c:234-298).  The reference scans i upwards; as x
// increases, xval < x[i-1] holds from that first i on, so bisection finds the
// same i in log2(n) reads (a NaN never compares less: n-1 either way).
HC_HD int cache_bracket(int n, const double *x, double xval)
{
    int lo = 2, hi = n;  // the first i in [lo, hi) with xval < x[i-1], hi = none
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (xval < x[mid - 1])
            hi = mid;
        else
            lo = mid + 1;
    }
    return (lo <= n - 1) ? lo - 1 : n - 1;
}

// Window [lo, hi] of grid indices a lookup reads along one axis
// (cache4d.c:408-440). The linear rule's window is [left-1, left+1]; its third
// point is never used except at xval == x[n-1], where the reference reads one
// point past the axis: the window stops at n-1 here and that point takes the
// last interval.
HC_HD void cache_window(int n, const double *x, double xval, int mode, int &lo, int &hi)
{
    const int left = cache_bracket(n, x, xval);
    if (mode == CACHE_LIN) {
        lo = left - 1;
        hi = left + 1 < n - 1 ? left + 1 : n - 1;
    } else {
        lo = left - 2 < 0 ? 0 : left - 2;
        hi = left + 1 >= n ? n - 1 : left + 1;
    }
}

// Piecewise-linear value at t from m points (akima.c:300-370).
HC_HD double cache_lin(int m, const double *xs, const double *ys, double t)
{
    const int left = cache_bracket(m, xs, t);  // 1-based
    const double T1 = 1.0 / (xs[left] - xs[left - 1]);
    const double T2 = t - xs[left - 1];
    const double yp = (ys[left] - ys[left - 1]) * T1;
    return ys[left - 1] + T2 * yp;
}

// Akima's spline through m (3 or 4) points, evaluated at nt increasing-search
// points ts (akima.c:9-160). Returns false when two abscissae are closer than
// 1e-12 (the reference then leaves the output unwritten).
HC_HD bool cache_akima(int m, const double *xs, const double *ys, int nt, const double *ts, double *out)
{
    // secants, with two extrapolated on each side: S = g0, g1, m_0 .. m_{m-2}, m_{m-1}, m_m
    double S[7];
    for (int k = 0; k < m - 1; k++) {
        const double dx = xs[k + 1] - xs[k];
        if (dx < 1e-12) return false;
        S[k + 2] = (ys[k + 1] - ys[k]) / dx;
    }
    {
        const double x0 = xs[0], x1 = xs[1], y0 = ys[0];
        const double xl1 = x0 + x1 - xs[2];
        const double xl0 = xl1 + x0 - x1;
        const double yl1 = (x0 - xl1) * (S[3] - 2.0 * S[2]) + y0;
        const double g1 = (y0 - yl1) / (x0 - xl1);
        const double yl0 = (xl1 - xl0) * (S[2] - 2.0 * g1) + yl1;
        S[0] = (yl1 - yl0) / (xl1 - xl0);
        S[1] = g1;
    }
    {
        const double xa = xs[m - 2], xb = xs[m - 1], yb = ys[m - 1];
        const double xr2 = xb + xa - xs[m - 3];
        const double xr3 = xr2 + xb - xa;
        const double yr2 = (2.0 * S[m] - S[m - 1]) * (xr2 - xb) + yb;
        S[m + 1] = (yr2 - yb) / (xr2 - xb);
        const double yr3 = (2.0 * S[m + 1] - S[m]) * (xr3 - xr2) + yr2;
        S[m + 2] = (yr3 - yr2) / (xr3 - xr2);
    }
    // node slopes from the four secants around each node
    double sl[4];
    for (int k = 0; k < m; k++) {
        const double a = S[k], b = S[k + 1], c = S[k + 2], d = S[k + 3];
        double d0 = d - c, d1 = b - a;
        if (d0 < 0.0) d0 *= -1.0;
        if (d1 < 0.0) d1 *= -1.0;
        sl[k] = (d0 + d1) < 1e-9 ? 0.5 * (b + c) : (d0 * b + d1 * c) / (d0 + d1);
    }
    // the interval search continues from one output point to the next
    int i = -1;
    for (int q = 0; q < nt; q++) {
        const double t = ts[q];
        while (t > xs[i + 1] && i < m - 2) i++;
        if (i < 0) i = 0;
        const double d0 = xs[i + 1] - xs[i];
        const double t2 = (ys[i + 1] - ys[i]) / d0;
        const double g0 = sl[i], g1 = sl[i + 1];
        const double p1 = (3.0 * t2 - 2.0 * g0 - g1) / d0;
        const double p2 = (g0 + g1 - 2.0 * t2) / (d0 * d0);
        const double t1 = t - xs[i];
        out[q] = ys[i] + g0 * t1 + p1 * t1 * t1 + p2 * t1 * t1 * t1;
    }
    return true;
}

// Value and centred difference (step t * 1e-6) of the spline at t
// (akima.c:196-231).
HC_HD bool cache_akima_der(int m, const double *xs, const double *ys, double t, double &y, double &dy)
{
    const double h = t * 1.0E-6;
    const double ts[3] = {t - h, t, t + h};
    double v[3];
    if (!cache_akima(m, xs, ys, 3, ts, v)) return false;
    y = v[1];
    dy = (v[2] - v[0]) / (2.0 * h);
    return true;
}

// cache_akima with the window size M (3 or 4) and the number of output points
// NT fixed at compile time: the same operations in the same order, the
// interval search unrolled (it advances i one interval at a time, as the
// loop it replaces does), so every array stays in registers.
template <int M, int NT>
HC_HD bool cache_akima_m(const double *xs, const double *ys, const double *ts, double *out)
{
    double S[M + 3];
#pragma unroll
    for (int k = 0; k < M - 1; k++) {
        const double dx = xs[k + 1] - xs[k];
        if (dx < 1e-12) return false;
        S[k + 2] = (ys[k + 1] - ys[k]) / dx;
    }
    {
        const double x0 = xs[0], x1 = xs[1], y0 = ys[0];
        const double xl1 = x0 + x1 - xs[2];
        const double xl0 = xl1 + x0 - x1;
        const double yl1 = (x0 - xl1) * (S[3] - 2.0 * S[2]) + y0;
        const double g1 = (y0 - yl1) / (x0 - xl1);
        const double yl0 = (xl1 - xl0) * (S[2] - 2.0 * g1) + yl1;
        S[0] = (yl1 - yl0) / (xl1 - xl0);
        S[1] = g1;
    }
    {
        const double xa = xs[M - 2], xb = xs[M - 1], yb = ys[M - 1];
        const double xr2 = xb + xa - xs[M - 3];
        const double xr3 = xr2 + xb - xa;
        const double yr2 = (2.0 * S[M] - S[M - 1]) * (xr2 - xb) + yb;
        S[M + 1] = (yr2 - yb) / (xr2 - xb);
        const double yr3 = (2.0 * S[M + 1] - S[M]) * (xr3 - xr2) + yr2;
        S[M + 2] = (yr3 - yr2) / (xr3 - xr2);
    }
    double sl[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
        const double a = S[k], b = S[k + 1], c = S[k + 2], d = S[k + 3];
        double d0 = d - c, d1 = b - a;
        if (d0 < 0.0) d0 *= -1.0;
        if (d1 < 0.0) d1 *= -1.0;
        sl[k] = (d0 + d1) < 1e-9 ? 0.5 * (b + c) : (d0 * b + d1 * c) / (d0 + d1);
    }
    int i = -1;
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const double t = ts[q];
#pragma unroll
        for (int s = 0; s <= M - 2; s++)
            if (i == s - 1 && t > xs[s]) i = s;
        if (i < 0) i = 0;
        double xi = xs[0], xi1 = xs[1], yi = ys[0], yi1 = ys[1], g0 = sl[0], g1 = sl[1];
#pragma unroll
        for (int s = 1; s <= M - 2; s++)
            if (i == s) {
                xi = xs[s];
                xi1 = xs[s + 1];
                yi = ys[s];
                yi1 = ys[s + 1];
                g0 = sl[s];
                g1 = sl[s + 1];
            }
        const double d0 = xi1 - xi;
        const double t2 = (yi1 - yi) / d0;
        const double p1 = (3.0 * t2 - 2.0 * g0 - g1) / d0;
        const double p2 = (g0 + g1 - 2.0 * t2) / (d0 * d0);
        const double t1 = t - xi;
        out[q] = yi + g0 * t1 + p1 * t1 * t1 + p2 * t1 * t1 * t1;
    }
    return true;
}

// The per-interval numbers cache_akima_m's evaluation uses, for the M-1
// intervals of a window: co[4 i ..] = (y_i, g_i, p1_i, p2_i), formed by the
// same operations.  False as cache_akima_m when two abscissae are too close.
template <int M>
HC_HD bool cache_akima_coefs(const double *xs, const double *ys, double *co)
{
    double S[M + 3];
#pragma unroll
    for (int k = 0; k < M - 1; k++) {
        const double dx = xs[k + 1] - xs[k];
        if (dx < 1e-12) return false;
        S[k + 2] = (ys[k + 1] - ys[k]) / dx;
    }
    {
        const double x0 = xs[0], x1 = xs[1], y0 = ys[0];
        const double xl1 = x0 + x1 - xs[2];
        const double xl0 = xl1 + x0 - x1;
        const double yl1 = (x0 - xl1) * (S[3] - 2.0 * S[2]) + y0;
        const double g1 = (y0 - yl1) / (x0 - xl1);
        const double yl0 = (xl1 - xl0) * (S[2] - 2.0 * g1) + yl1;
        S[0] = (yl1 - yl0) / (xl1 - xl0);
        S[1] = g1;
    }
    {
        const double xa = xs[M - 2], xb = xs[M - 1], yb = ys[M - 1];
        const double xr2 = xb + xa - xs[M - 3];
        const double xr3 = xr2 + xb - xa;
        const double yr2 = (2.0 * S[M] - S[M - 1]) * (xr2 - xb) + yb;
        S[M + 1] = (yr2 - yb) / (xr2 - xb);
        const double yr3 = (2.0 * S[M + 1] - S[M]) * (xr3 - xr2) + yr2;
        S[M + 2] = (yr3 - yr2) / (xr3 - xr2);
    }
    double sl[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
        const double a = S[k], b = S[k + 1], c = S[k + 2], d = S[k + 3];
        double d0 = d - c, d1 = b - a;
        if (d0 < 0.0) d0 *= -1.0;
        if (d1 < 0.0) d1 *= -1.0;
        sl[k] = (d0 + d1) < 1e-9 ? 0.5 * (b + c) : (d0 * b + d1 * c) / (d0 + d1);
    }
#pragma unroll
    for (int i = 0; i < M - 1; i++) {
        const double d0 = xs[i + 1] - xs[i];
        const double t2 = (ys[i + 1] - ys[i]) / d0;
        co[4 * i + 0] = ys[i];
        co[4 * i + 1] = sl[i];
        co[4 * i + 2] = (3.0 * t2 - 2.0 * sl[i] - sl[i + 1]) / d0;
        co[4 * i + 3] = (sl[i] + sl[i + 1] - 2.0 * t2) / (d0 * d0);
    }
    return true;
}

// The first three abscissae of a window (every window has at least three
// points), read together and handed to the device code that follows as
// register values (an empty asm takes them): their reads go out at once ahead
// of the first comparison, and a pick among them stays a select.
struct CacheAbsc {
    double a0, a1, a2;
};

HC_HD CacheAbsc cache_absc(const double *ys)
{
    CacheAbsc w{ys[0], ys[1], ys[2]};
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(w.a0), "+v"(w.a1), "+v"(w.a2));
#endif
    return w;
}

HC_HD double cache_absc_at(const CacheAbsc &w, int s) { return s == 0 ? w.a0 : (s == 1 ? w.a1 : w.a2); }

// cache_akima_m's output points from those numbers: the same interval search
// and the same cubic.  The search's abscissae are read together, ahead of it.
template <int M, int NT>
HC_HD void cache_akima_eval(const double *xs, const double *co, const double *ts, double *out)
{
    const CacheAbsc xv = cache_absc(xs);
    int i = -1;
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const double t = ts[q];
#pragma unroll
        for (int s = 0; s <= M - 2; s++)
            if ((i == s - 1) & (t > cache_absc_at(xv, s))) i = s;
        if (i < 0) i = 0;
        const double *ci = co + 4 * i;
        const double t1 = t - cache_absc_at(xv, i);
        out[q] = ci[0] + ci[1] * t1 + ci[2] * t1 * t1 + ci[3] * t1 * t1 * t1;
    }
}

struct CacheView {
    int X, Y, Z, D;
    int xtype, ytype, ztype;
    int whole;                    // every point computed: the window check is skipped
    const double *vx, *vy, *vz;   // axes
    const double *values;         // [Z][X][Y][D]
    const uint8_t *counted;       // [Z][X][Y], 1 = computed
};

HC_HD int cache_axis_check(int type, int n, const double *v, double t, bool checked)
{
    if (type == AXIS_ROLLING) return (t < v[1] || t > v[n - 2]) ? CACHE_NEEDS_RECENTRE : CACHE_OK;
    if (checked && (t < v[0] || t > v[n - 1])) return CACHE_OUT_OF_RANGE;
    return CACHE_OK;
}

// One getvalues() call (cache4d.c:863-1134, CONSTGRID). zindex >= 0 selects a
// z plane (no z interpolation), -1 interpolates in z. out[k * ostride] and,
// with dery, dery[k * ostride] receive the masked components (the reference's
// dimmask; the linear rule fills every component and no derivative). The
// derivative along y is interpolated in x only and comes from the last z plane
// of the window, as the reference's does.
HC_HD int cache_lookup(const CacheView &c, double x, double y, double z, int zindex, int mode, uint64_t mask,
                       double *out, double *dery, size_t ostride)
{
    int rolling = 0;
    int s = cache_axis_check(c.xtype, c.X, c.vx, x, true);
    if (s == CACHE_OUT_OF_RANGE) return s;
    rolling |= s == CACHE_NEEDS_RECENTRE;
    s = cache_axis_check(c.ytype, c.Y, c.vy, y, true);
    if (s == CACHE_OUT_OF_RANGE) return s;
    rolling |= s == CACHE_NEEDS_RECENTRE;
    s = cache_axis_check(c.ztype, c.Z, c.vz, z, zindex == -1);
    if (s == CACHE_OUT_OF_RANGE) return s;
    rolling |= s == CACHE_NEEDS_RECENTRE;
    if (rolling) return CACHE_NEEDS_RECENTRE;
    if (zindex < -1 || zindex >= c.Z) return CACHE_BAD_ZINDEX;
    // a window needs 3 points for the spline and 2 for the linear rule (x and y
    // are checked when the cache is built)
    if (zindex == -1 && c.Z < (mode == CACHE_LIN ? 2 : 3)) return CACHE_BAD_ZINDEX;

    int x0, x1, y0, y1, z0, z1;
    cache_window(c.X, c.vx, x, mode, x0, x1);
    cache_window(c.Y, c.vy, y, mode, y0, y1);
    if (zindex != -1)
        z0 = z1 = zindex;
    else
        cache_window(c.Z, c.vz, z, mode, z0, z1);
    const long XY = (long)c.X * c.Y;
    if (!c.whole) {
        for (int iz = z0; iz <= z1; iz++)
            for (int ix = x0; ix <= x1; ix++)
                for (int iy = y0; iy <= y1; iy++) {
                    const uint8_t f = c.counted[iz * XY + (long)ix * c.Y + iy];
                    if (f == 0) return CACHE_MISSING;
                    if (f != 1) return CACHE_OUT_OF_RANGE;  // a broken entry (cache4d.c:1018-1021)
                }
    }
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1, nz = z1 - z0 + 1;
    const double *xs = c.vx + x0, *ys = c.vy + y0, *zs = c.vz + z0;
    for (int k = 0; k < c.D; k++) {
        const bool lin = mode == CACHE_LIN;
        if (!lin && !((mask >> k) & 1)) continue;
        double zt[4], dlast = 0.0;
        bool ok = true;
        for (int a = 0; a < nz; a++) {
            double xt[4], xd[4];
            for (int b = 0; b < nx; b++) {
                const double *row = c.values + ((size_t)(z0 + a) * XY + (size_t)(x0 + b) * c.Y + y0) * c.D + k;
                double yv[4];
                for (int q = 0; q < ny; q++) yv[q] = row[(size_t)q * c.D];
                if (lin)
                    xt[b] = cache_lin(ny, ys, yv, y);
                else if (dery)
                    ok &= cache_akima_der(ny, ys, yv, y, xt[b], xd[b]);
                else
                    ok &= cache_akima(ny, ys, yv, 1, &y, &xt[b]);
            }
            if (lin) {
                zt[a] = cache_lin(nx, xs, xt, x);
            } else {
                ok &= cache_akima(nx, xs, xt, 1, &x, &zt[a]);
                if (dery) ok &= cache_akima(nx, xs, xd, 1, &x, &dlast);
            }
        }
        if (!ok) return CACHE_OUT_OF_RANGE;
        double v = zt[0];
        if (zindex == -1) {
            if (lin)
                v = cache_lin(nz, zs, zt, z);
            else if (!cache_akima(nz, zs, zt, 1, &z, &v))
                return CACHE_OUT_OF_RANGE;
        }
        out[k * ostride] = v;
        if (dery && !lin) dery[k * ostride] = dlast;
    }
    return CACHE_OK;
}

// The spline value (and with DER the derivative along y, cache_akima_der's
// centred difference) of component k over an NX x NY window at (x0, y0) of z
// plane zp: cache_lookup's operations for one component.
template <int NX, int NY, bool DER>
HC_HD bool cache_spline_win(const CacheView &c, int x0, int y0, int zp, int k, double x, double y, double &v,
                            double &dv)
{
    const size_t XY = (size_t)c.X * c.Y;
    const double *xs = c.vx + x0, *ys = c.vy + y0;
    double xt[NX], xd[NX];
    bool ok = true;
#pragma unroll
    for (int b = 0; b < NX; b++) {
        const double *row = c.values + ((size_t)zp * XY + (size_t)(x0 + b) * c.Y + y0) * c.D + k;
        double yv[NY];
#pragma unroll
        for (int q = 0; q < NY; q++) yv[q] = row[(size_t)q * c.D];
        if (DER) {
            const double h = y * 1.0E-6;
            const double ts[3] = {y - h, y, y + h};
            double o3[3];
            ok &= cache_akima_m<NY, 3>(ys, yv, ts, o3);
            xt[b] = o3[1];
            xd[b] = (o3[2] - o3[0]) / (2.0 * h);
        } else {
            ok &= cache_akima_m<NY, 1>(ys, yv, &y, &xt[b]);
        }
    }
    ok &= cache_akima_m<NX, 1>(xs, xt, &x, &v);
    if (DER) ok &= cache_akima_m<NX, 1>(xs, xd, &x, &dv);
    return ok;
}

// Coefficient table of the splines along y of one component of one z plane:
// for x row ix and bracket left = 1..Y-1 (the y window of a query in that
// bracket, cache_window) the window's intervals' numbers (cache_akima_coefs),
// [X][Y][3][4] doubles (left 0 unused; a 3-point window fills 2 intervals).
constexpr int CACHE_YCO = 12;

HC_HD void cache_ycoef_one(const CacheView &c, int zp, int k, int ix, int left, double *co)
{
    const int lo = left - 2 < 0 ? 0 : left - 2;
    const int hi = left + 1 >= c.Y ? c.Y - 1 : left + 1;
    const int m = hi - lo + 1;
    const double *row = c.values + ((size_t)zp * c.X * c.Y + (size_t)ix * c.Y + lo) * c.D + k;
    double yv[4];
    for (int q = 0; q < m; q++) yv[q] = row[(size_t)q * c.D];
    if (m == 4)
        cache_akima_coefs<4>(c.vy + lo, yv, co);
    else
        cache_akima_coefs<3>(c.vy + lo, yv, co);
}

// cache_spline_win with the y splines read from a coefficient table.
template <int NX, int NY, bool DER>
HC_HD void cache_spline_win_co(const CacheView &c, const double *yco, int x0, int y0, int left, double x, double y,
                               double &v, double &dv, bool &ok)
{
    const double *xs = c.vx + x0, *ys = c.vy + y0;
    double xt[NX], xd[NX];
    if (DER) {
        // cache_akima_eval's intervals of the three points, shared by every row
        // (the window's abscissae are the same); they almost always coincide, and
        // then each row reads one interval's numbers
        const double h = y * 1.0E-6;
        const double ts[3] = {y - h, y, y + h};
        const CacheAbsc yv = cache_absc(ys);
        int iq[3], i = -1;
#pragma unroll
        for (int q = 0; q < 3; q++) {
#pragma unroll
            for (int s = 0; s <= NY - 2; s++)
                if ((i == s - 1) & (ts[q] > cache_absc_at(yv, s))) i = s;
            if (i < 0) i = 0;
            iq[q] = i;
        }
        if (iq[0] == iq[2]) {
            const int i0 = iq[0];
            const double y0v = cache_absc_at(yv, i0);
            const double t1[3] = {ts[0] - y0v, ts[1] - y0v, ts[2] - y0v};
#pragma unroll
            for (int b = 0; b < NX; b++) {
                const double *ci = yco + ((size_t)(x0 + b) * c.Y + left) * CACHE_YCO + 4 * i0;
                const double a0 = ci[0], a1 = ci[1], a2 = ci[2], a3 = ci[3];
                double o3[3];
#pragma unroll
                for (int q = 0; q < 3; q++)
                    o3[q] = a0 + a1 * t1[q] + a2 * t1[q] * t1[q] + a3 * t1[q] * t1[q] * t1[q];
                xt[b] = o3[1];
                xd[b] = (o3[2] - o3[0]) / (2.0 * h);
            }
        } else {
#pragma unroll
            for (int b = 0; b < NX; b++) {
                const double *co = yco + ((size_t)(x0 + b) * c.Y + left) * CACHE_YCO;
                double o3[3];
                cache_akima_eval<NY, 3>(ys, co, ts, o3);
                xt[b] = o3[1];
                xd[b] = (o3[2] - o3[0]) / (2.0 * h);
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < NX; b++) {
            const double *co = yco + ((size_t)(x0 + b) * c.Y + left) * CACHE_YCO;
            cache_akima_eval<NY, 1>(ys, co, &y, &xt[b]);
        }
    }
    ok &= cache_akima_m<NX, 1>(xs, xt, &x, &v);
    if (DER) ok &= cache_akima_m<NX, 1>(xs, xd, &x, &dv);
}

// cache_lookup for one component k of a z plane (zindex >= 0) in spline mode:
// the same status and bits, with the window's size fixed per call.
HC_HD int cache_lookup(const CacheView &c, double x, double y, double z, int zindex, int mode, uint64_t mask,
                       double *out, double *dery, size_t ostride);

// With yco (the coefficient table of component k, plane zindex, of a cache
// whose axes were checked when it was built) the y splines are read from it.
template <bool DER>
HC_HD int cache_lookup1(const CacheView &c, double x, double y, int zindex, int k, double &v, double &dv,
                        const double *yco = nullptr)
{
    if (zindex == -1) {
        // interpolation in z (at z = 0): the general lookup
        double o[64], d[64];
        const int st = cache_lookup(c, x, y, 0.0, -1, CACHE_SPLINE, 1ull << k, o, DER ? d : nullptr, 1);
        if (st == CACHE_OK) {
            v = o[k];
            if (DER) dv = d[k];
        }
        return st;
    }
    int rolling = 0;
    int s = cache_axis_check(c.xtype, c.X, c.vx, x, true);
    if (s == CACHE_OUT_OF_RANGE) return s;
    rolling |= s == CACHE_NEEDS_RECENTRE;
    s = cache_axis_check(c.ytype, c.Y, c.vy, y, true);
    if (s == CACHE_OUT_OF_RANGE) return s;
    rolling |= s == CACHE_NEEDS_RECENTRE;
    s = cache_axis_check(c.ztype, c.Z, c.vz, 0.0, false);
    rolling |= s == CACHE_NEEDS_RECENTRE;
    if (rolling) return CACHE_NEEDS_RECENTRE;
    if (zindex < 0 || zindex >= c.Z) return CACHE_BAD_ZINDEX;
    int x0, x1, y0, y1;
    cache_window(c.X, c.vx, x, CACHE_SPLINE, x0, x1);
    cache_window(c.Y, c.vy, y, CACHE_SPLINE, y0, y1);
    if (!c.whole) {
        const long XY = (long)c.X * c.Y;
        for (int ix = x0; ix <= x1; ix++)
            for (int iy = y0; iy <= y1; iy++) {
                const uint8_t f = c.counted[zindex * XY + (long)ix * c.Y + iy];
                if (f == 0) return CACHE_MISSING;
                if (f != 1) return CACHE_OUT_OF_RANGE;
            }
    }
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1;
    bool ok = true;
    if (yco) {
        const int left = cache_bracket(c.Y, c.vy, y);
        if (nx == 4 && ny == 4)
            cache_spline_win_co<4, 4, DER>(c, yco, x0, y0, left, x, y, v, dv, ok);
        else if (nx == 4)
            cache_spline_win_co<4, 3, DER>(c, yco, x0, y0, left, x, y, v, dv, ok);
        else if (ny == 4)
            cache_spline_win_co<3, 4, DER>(c, yco, x0, y0, left, x, y, v, dv, ok);
        else
            cache_spline_win_co<3, 3, DER>(c, yco, x0, y0, left, x, y, v, dv, ok);
        return ok ? CACHE_OK : CACHE_OUT_OF_RANGE;
    }
    if (nx == 4 && ny == 4)
        ok = cache_spline_win<4, 4, DER>(c, x0, y0, zindex, k, x, y, v, dv);
    else if (nx == 4)
        ok = cache_spline_win<4, 3, DER>(c, x0, y0, zindex, k, x, y, v, dv);
    else if (ny == 4)
        ok = cache_spline_win<3, 4, DER>(c, x0, y0, zindex, k, x, y, v, dv);
    else
        ok = cache_spline_win<3, 3, DER>(c, x0, y0, zindex, k, x, y, v, dv);
    return ok ? CACHE_OK : CACHE_OUT_OF_RANGE;
}

// Temperature at which the cache's energy density (component 0) equals e, by
// the bounded Newton iteration the reference's COOL_EOS pressure and cooling
// routines run (mnewt.h:104-205 mnewtc<1> with cooleos.c:228-262/296-306 as
// its function: F = (e - e(rho, T)) / e, dF/dT = -de/dT / e from the spline's
// centred difference; cooleos.c:365-404, 545-577). A step is taken when it
// stays inside [lo, hi] and moves T; otherwise every later trial would repeat
// this one, so the loop stops with the result those trials give. Returns the
// reference's code (1 converged, errf <= tolf, T the last iterate; 2 or 0 not
// converged, T the iterate of least error) or -1 when a lookup fails (the
// reference reads values it did not get).
HC_HD int cache_solve_T(const CacheView &c, double rho, double e, int zindex, double T0, double lo, double hi,
                        int ntrial, double tolf, double &T, double &berr, const double *yco = nullptr)
{
    double x = T0, best = T0, besterr = 1.0E100;
    int ret = 0;
    for (int k = 1; k <= ntrial; k++) {
        double val = 0.0, der = 0.0;
        if (cache_lookup1<true>(c, rho, x, zindex, 0, val, der, yco) != CACHE_OK) {
            T = x;
            berr = besterr;
            return -1;
        }
        const double F = (e - val) / e;
        const double D = -der / e;
        double bet = -F;
        const double errf = 0.0 + fabs(bet);
        if (errf <= tolf) {
            T = x;
            berr = errf;
            return 1;
        }
        if (errf < besterr) {
            if (k > 1) ret = 2;
            besterr = errf;
            best = x;
        }
        // 1x1 LU solve: a zero (or NaN) derivative is a singular matrix, and the
        // step stays -F (mnewt.h:46-62, 149-150)
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

}  // namespace harm
