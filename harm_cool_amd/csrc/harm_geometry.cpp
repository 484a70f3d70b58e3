// harm_geometry.cpp -- host-side setup for libharmcool.so: metric tables in code
// coordinates, per-column floors, Font-fix isentrope constants and the
// axisymmetric part of the Fishbone-Moncrief torus.
//
// These run once per context on the host (the reference computes them once in
// set_grid()/init(), main.c:413-508, init.c:70-472) and are uploaded to HBM.  The
// metric uses the reference's finite-difference Jacobian and connection
// (metric.c:158,161-207,373-399), which analytic derivatives would not reproduce
// bitwise, so every expression keeps the reference's operation order.
//
// Where the reference computes sin and cos of one angle in one function, its g++
// -O2 build calls glibc's sincos(), which differs from a separate cos() by 1 ulp
// for ~0.1% of angles; those sites call sincos() here as well.
#include <cmath>
#include <cstring>
#include <vector>

#include "harm_host.hpp"

namespace harm_host {

namespace {

constexpr double DELTA = 1.e-5;          // metric.c:158
constexpr double SINGSMALL = 1.E-20;     // decs.h:171
constexpr double RHOMIN = 1.e-6, UUMIN = 1.e-8, RHOMINLIMIT = 1.e-20, UUMINLIMIT = 1.e-20, POWRHO = 2.0;
constexpr double SMALL = 1.e-20;

typedef double M4[4][4];

// coord.c:61-112 (gammie_coords + polar-axis guard of bl_coord_vec)
void bl_coord_vec(const harm_params &P, const double *X, double *V)
{
    double theexp = X[1];
    if (P.Rbr >= 0.) {
        double x1br = log(P.Rbr - P.R0);
        if (X[1] > x1br) theexp += 1.0 * pow(X[1] - x1br, 4.0);
    }
    V[0] = X[0];
    V[1] = exp(theexp) + P.R0;
    V[2] = M_PI_2 * (1. + X[2]) + ((1. - P.hslope) / 2.) * sin(M_PI * (1. + X[2]));
    V[3] = X[3];
    if (fabs(V[2]) < SINGSMALL) V[2] = (V[2] >= 0) ? SINGSMALL : -SINGSMALL;
    if (fabs(M_PI - V[2]) < SINGSMALL) V[2] = (V[2] >= M_PI) ? M_PI + SINGSMALL : M_PI - SINGSMALL;
}

// metric.c:373-399 (BL=1): central differences
void dxdxp_func(const harm_params &P, const double *X, M4 dxdxp)
{
    for (int k = 0; k < 4; k++) {
        double Xh[4], Xl[4], Vh[4], Vl[4];
        for (int l = 0; l < 4; l++) Xh[l] = Xl[l] = X[l];
        Xh[k] += DELTA;
        Xl[k] -= DELTA;
        bl_coord_vec(P, Xh, Vh);
        bl_coord_vec(P, Xl, Vl);
        for (int j = 0; j < 4; j++) dxdxp[j][k] = (Vh[j] - Vl[j]) / (Xh[k] - Xl[k]);
    }
}

// coord.c:128-199
void gcov_func(const harm_params &P, const double *X, M4 gcovp)
{
    const double a = P.a;
    double V[4];
    M4 gcov, dxdxp;
    for (int j = 0; j < 4; j++)
        for (int k = 0; k < 4; k++) gcov[j][k] = gcovp[j][k] = 0.;
    bl_coord_vec(P, X, V);
    double r = V[1], th = V[2];
    double sth, cth;
    sincos(th, &sth, &cth);
    double s2 = sth * sth;
    double rho2 = r * r + a * a * cth * cth;
    dxdxp_func(P, X, dxdxp);
    gcov[0][0] = (-1. + 2. * r / rho2);
    gcov[0][1] = (2. * r / rho2);
    gcov[0][3] = (-2. * a * r * s2 / rho2);
    gcov[1][0] = gcov[0][1];
    gcov[1][1] = (1. + 2. * r / rho2);
    gcov[1][3] = (-a * s2 * (1. + 2. * r / rho2));
    gcov[2][2] = rho2;
    gcov[3][0] = gcov[0][3];
    gcov[3][1] = gcov[1][3];
    gcov[3][3] = s2 * (rho2 + a * a * s2 * (1. + 2. * r / rho2));
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            gcovp[i][j] = 0;
            for (int k = 0; k < 4; k++)
                for (int l = 0; l < 4; l++) gcovp[i][j] += gcov[k][l] * dxdxp[k][i] * dxdxp[l][j];
        }
}

// lu.c:60-324 (Crout LU with implicit partial pivoting)
int lu_decompose(M4 A, int permute[4])
{
    double row_norm[4], absmax, maxtemp;
    int max_row = 0;
    const int n = 4;
    for (int i = 0; i < n; i++) {
        absmax = 0.;
        for (int j = 0; j < n; j++) {
            maxtemp = fabs(A[i][j]);
            if (maxtemp > absmax) absmax = maxtemp;
        }
        if (absmax == 0.) return 1;
        row_norm[i] = 1. / absmax;
    }
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < j; i++)
            for (int k = 0; k < i; k++) A[i][j] -= A[i][k] * A[k][j];
        absmax = 0.0;
        for (int i = j; i < n; i++) {
            for (int k = 0; k < j; k++) A[i][j] -= A[i][k] * A[k][j];
            maxtemp = fabs(A[i][j]) * row_norm[i];
            if (maxtemp >= absmax) {
                absmax = maxtemp;
                max_row = i;
            }
        }
        if (max_row != j) {
            if ((j == (n - 2)) && (A[j][j + 1] == 0.)) {
                max_row = j;
            } else {
                for (int k = 0; k < n; k++) {
                    maxtemp = A[j][k];
                    A[j][k] = A[max_row][k];
                    A[max_row][k] = maxtemp;
                }
                row_norm[max_row] = row_norm[j];
            }
        }
        permute[j] = max_row;
        if (A[j][j] == 0.) A[j][j] = 1.e-30;
        if (j != (n - 1)) {
            maxtemp = 1. / A[j][j];
            for (int i = (j + 1); i < n; i++) A[i][j] *= maxtemp;
        }
    }
    return 0;
}

void lu_substitution(M4 A, double B[4], const int permute[4])
{
    const int n = 4;
    for (int i = 0; i < n; i++) {
        double tmpvar = B[permute[i]];
        B[permute[i]] = B[i];
        for (int j = (i - 1); j >= 0; j--) tmpvar -= A[i][j] * B[j];
        B[i] = tmpvar;
    }
    for (int i = (n - 1); i >= 0; i--) {
        for (int j = (i + 1); j < n; j++) B[i] -= A[i][j] * B[j];
        B[i] /= A[i][i];
    }
}

int invert_matrix(const M4 Am, M4 Aminv)
{
    M4 t;
    int permute[4];
    std::memcpy(t, Am, sizeof(M4));
    if (lu_decompose(t, permute) != 0) return 1;
    for (int i = 0; i < 4; i++) {
        double d[4] = {0., 0., 0., 0.};
        d[i] = 1.;
        lu_substitution(t, d, permute);
        for (int j = 0; j < 4; j++) Aminv[j][i] = d[j];
    }
    return 0;
}

// metric.c:118-138
double gdet_func(const M4 gcov, int &bad)
{
    M4 t;
    int permute[4];
    std::memcpy(t, gcov, sizeof(M4));
    if (lu_decompose(t, permute) != 0) bad = 1;
    double detg = 1.;
    for (int j = 0; j < 4; j++) detg *= t[j][j];
    return sqrt(fabs(detg));
}

// metric.c:161-207
void conn_func(const harm_params &P, const double *X, const M4 gcon, double conn[4][4][4])
{
    double tmp[4][4][4];
    for (int k = 0; k < 4; k++) {
        double Xh[4], Xl[4];
        M4 gh, gl;
        for (int l = 0; l < 4; l++) Xh[l] = Xl[l] = X[l];
        Xh[k] += DELTA;
        Xl[k] -= DELTA;
        gcov_func(P, Xh, gh);
        gcov_func(P, Xl, gl);
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) conn[i][j][k] = (gh[i][j] - gl[i][j]) / (Xh[k] - Xl[k]);
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) tmp[i][j][k] = 0.5 * (conn[j][i][k] + conn[k][i][j] - conn[k][j][i]);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) {
                conn[i][j][k] = 0.;
                for (int l = 0; l < 4; l++) conn[i][j][k] += gcon[i][l] * tmp[l][j][k];
            }
}

// metric.c:62-115 for the whole domain on one device (h is the 0-based phi plane)
void coord(const Grid &g, int h, int i, int j, int loc, double *X)
{
    i -= 2;
    j -= 2;
    X[0] = 1.;
    const bool n3gt2 = g.N3 > 2;
    if (loc == 0) {  // FACE1
        X[1] = g.startx[1] + i * g.dx[1];
        X[2] = g.startx[2] + (j + 0.5) * g.dx[2];
        X[3] = n3gt2 ? g.startx[3] + (h + 0.5) * g.dx[3] : g.startx[3] + h * g.dx[3];
    } else if (loc == 1) {  // FACE2
        X[1] = g.startx[1] + (i + 0.5) * g.dx[1];
        X[2] = g.startx[2] + j * g.dx[2];
        X[3] = n3gt2 ? g.startx[3] + (h + 0.5) * g.dx[3] : g.startx[3] + h * g.dx[3];
    } else if (loc == 4) {  // FACE3
        X[1] = g.startx[1] + (i + 0.5) * g.dx[1];
        X[2] = g.startx[2] + (j + 0.5) * g.dx[2];
        X[3] = g.startx[3] + h * g.dx[3];
    } else if (loc == 3) {  // CENT
        X[1] = g.startx[1] + (i + 0.5) * g.dx[1];
        X[2] = g.startx[2] + (j + 0.5) * g.dx[2];
        X[3] = n3gt2 ? g.startx[3] + (h + 0.5) * g.dx[3] : g.startx[3] + h * g.dx[3];
    } else {  // CORN
        X[1] = g.startx[1] + i * g.dx[1];
        X[2] = g.startx[2] + j * g.dx[2];
        X[3] = g.startx[3] + h * g.dx[3];
    }
}

// metric.c:340-360
void bl_gcov_func(double a, double r, double th, M4 gcov)
{
    for (int j = 0; j < 4; j++)
        for (int k = 0; k < 4; k++) gcov[j][k] = 0.;
    double sth, cth;
    sincos(th, &sth, &cth);
    sth = fabs(sth);
    double s2 = sth * sth;
    double a2 = a * a;
    double r2 = r * r;
    double DD = 1. - 2. / r + a2 / r2;
    double mu = 1. + a2 * cth * cth / r2;
    gcov[0][0] = -(1. - 2. / (r * mu));
    gcov[0][3] = -2. * a * s2 / (r * mu);
    gcov[3][0] = gcov[0][3];
    gcov[1][1] = mu / DD;
    gcov[2][2] = r2 * mu;
    gcov[3][3] = r2 * sth * sth * (1. + a2 / r2 + 2. * a2 * s2 / (r2 * r * mu));
}

// init.c:557-567
double lfish_calc(double a, double r)
{
    return (((pow(a, 2) - 2. * a * sqrt(r) + pow(r, 2)) *
             ((-2. * a * r * (pow(a, 2) - 2. * a * sqrt(r) + pow(r, 2))) / sqrt(2. * a * sqrt(r) + (-3. + r) * r) +
              ((a + (-2. + r) * sqrt(r)) * (pow(r, 3) + pow(a, 2) * (2. + r))) /
                  sqrt(1 + (2. * a) / pow(r, 1.5) - 3. / r))) /
            (pow(r, 3) * sqrt(2. * a * sqrt(r) + (-3. + r) * r) * (pow(a, 2) + (-2. + r) * r)));
}

}  // namespace

int build_geometry(const harm_params &P, Grid &g, GeomTables &t)
{
    // coord.c:202-262 (set_points)
    double x1max = log(P.Rout - P.R0);
    if (!(P.Rbr < 0. || P.Rout < P.Rbr)) {
        double x1max0 = 1., x1br = log(P.Rbr - P.R0);
        int iter;
        x1max = 2.;
        for (iter = 0; iter < 200; iter++) {
            if (fabs((x1max - x1max0) / x1max) < 1e-20) break;
            x1max0 = x1max;
            double dxmax = (pow((log(P.Rout - P.R0) - x1max0) / 1.0, 1. / 4.0) + x1br) - x1max0;
            x1max = x1max0 + 0.5 * dxmax;
            if (x1max > log(P.Rout - P.R0)) x1max = log(P.Rout - P.R0);
        }
        if (iter == 200) return 1;
    }
    g.startx[0] = 0.;
    g.startx[1] = log(P.Rin - P.R0);
    g.startx[2] = -1.;
    g.startx[3] = 0.;
    g.dx[0] = 0.;
    g.dx[1] = (x1max - g.startx[1]) / g.N1;
    g.dx[2] = 2. / g.N2;
    g.dx[3] = 2. * M_PI / g.N3;

    const size_t ncol = (size_t)g.SN1 * g.SN2;
    for (int l = 0; l < 5; l++) t.geom[l].assign(33 * ncol, 0.);
    t.conn.assign(64 * ncol, 0.);
    t.coords.assign(4 * ncol, 0.);
    t.floors.assign(2 * ncol, 0.);
    t.katm.assign(g.SN1, 0.);
    int bad = 0;
    const int locs[5] = {3, 2, 0, 1, 4};  // CENT, CORN, FACE1, FACE2, FACE3 (main.c:434-490)
    for (int i = 0; i < g.SN1; i++)
        for (int j = 0; j < g.SN2; j++) {
            const size_t col = (size_t)i * g.SN2 + j;
            for (int li = 0; li < 5; li++) {
                const int loc = locs[li];
                double X[4];
                M4 gc, gn;
                coord(g, 0, i, j, loc, X);
                gcov_func(P, X, gc);
                double gd = gdet_func(gc, bad);
                if (invert_matrix(gc, gn)) bad = 1;
                std::vector<double> &T = t.geom[loc];
                for (int a = 0; a < 4; a++)
                    for (int b = 0; b < 4; b++) {
                        T[(size_t)(a * 4 + b) * ncol + col] = gc[a][b];
                        T[(size_t)(16 + a * 4 + b) * ncol + col] = gn[a][b];
                    }
                T[(size_t)32 * ncol + col] = gd;
                if (loc == 3) {
                    double cn[4][4][4], V[4];
                    conn_func(P, X, gn, cn);
                    for (int a = 0; a < 4; a++)
                        for (int b = 0; b < 4; b++)
                            for (int c = 0; c < 4; c++) t.conn[(size_t)((a * 4 + b) * 4 + c) * ncol + col] = cn[a][b][c];
                    bl_coord_vec(P, X, V);
                    for (int k = 0; k < 4; k++) t.coords[(size_t)k * ncol + col] = V[k];
                    // fixup.c:117-140 (BL branch): radius-only floors
                    double rhoscal = pow(V[1], -POWRHO);
                    double uuscal = pow(rhoscal, P.gam);
                    double rhoflr = RHOMIN * rhoscal;
                    double uuflr = UUMIN * uuscal;
                    if (rhoflr < RHOMINLIMIT) rhoflr = RHOMINLIMIT;
                    if (uuflr < UUMINLIMIT) uuflr = UUMINLIMIT;
                    t.floors[col] = rhoflr;
                    t.floors[ncol + col] = uuflr;
                }
            }
        }
    // fixup.c:555-585 (set_Katm): the floor state of the j=0 ghost row has zero
    // velocity and field, so fixup1zone reduces to rho = rhoflr, u = uuflr there.
    for (int i = g.is; i <= g.ie; i++) {
        const size_t col = (size_t)i * g.SN2 + 0;
        t.katm[i] = (P.gam - 1.) * t.floors[ncol + col] / pow(t.floors[col], P.gam);
    }
    return bad;
}

// init.c:180-291 for one (i,j) column: the torus is axisymmetric, so everything but
// the seeded perturbation of u is a function of (i,j).  Output per column:
// [0] rho, [1] u (unperturbed), [2..4] relative 4-velocity, [5] 1 inside the torus.
int build_torus2d(const harm_params &P, const Grid &g, const GeomTables &t, std::vector<double> &tor, double &rhomax,
                  double &umax)
{
    const size_t ncol = (size_t)g.SN1 * g.SN2;
    const double a = P.a, gam = P.gam, rin = P.rin, kappa = P.kappa;
    const double l = lfish_calc(a, P.rmax);
    tor.assign(6 * ncol, 0.);
    rhomax = 0.;
    umax = 0.;
    for (int i = g.is; i <= g.ie; i++)
        for (int j = g.js; j <= g.je; j++) {
            const size_t col = (size_t)i * g.SN2 + j;
            double X[4], V[4];
            coord(g, 0, i, j, 3, X);
            bl_coord_vec(P, X, V);
            double r = V[1], th = V[2];
            double sth, cth;
            sincos(th, &sth, &cth);
            double DD = r * r - 2. * r + a * a;
            double AA = (r * r + a * a) * (r * r + a * a) - DD * a * a * sth * sth;
            double SS = r * r + a * a * cth * cth;
            double thin = M_PI / 2.;
            double sthin, cthin;
            sincos(thin, &sthin, &cthin);
            double DDin = rin * rin - 2. * rin + a * a;
            double AAin = (rin * rin + a * a) * (rin * rin + a * a) - DDin * a * a * sthin * sthin;
            double SSin = rin * rin + a * a * cthin * cthin;
            double lnh;
            if (r >= rin) {
                lnh = 0.5 * log((1. + sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * sth * AA * sth))) / (SS * DD / AA)) -
                      0.5 * sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * AA * sth * sth)) - 2. * a * r * l / AA -
                      (0.5 * log((1. + sqrt(1. + 4. * (l * l * SSin * SSin) * DDin / (AAin * AAin * sthin * sthin))) /
                                 (SSin * DDin / AAin)) -
                       0.5 * sqrt(1. + 4. * (l * l * SSin * SSin) * DDin / (AAin * AAin * sthin * sthin)) -
                       2. * a * rin * l / AAin);
            } else
                lnh = 1.;
            if (lnh < SMALL || r < rin) {
                tor[0 * ncol + col] = 1.e-7 * RHOMIN;
                tor[1 * ncol + col] = 1.e-7 * UUMIN;
                tor[5 * ncol + col] = 0.;
                continue;
            }
            double hm1 = exp(lnh) - 1.;
            double rho = pow(hm1 * (gam - 1.) / (kappa * gam), 1. / (gam - 1.));
            double u = kappa * pow(rho, gam) / (gam - 1.);
            double expm2chi = SS * SS * DD / (AA * AA * sth * sth);
            double up1 = sqrt((-1. + sqrt(1. + 4. * l * l * expm2chi)) / 2.);
            double up = 2. * a * r * sqrt(1. + up1 * up1) / sqrt(AA * SS * DD) + sqrt(SS / AA) * up1 / sth;
            if (rho > rhomax) rhomax = rho;
            if (u > umax && r >= rin) umax = u;

            // init.c:479-555 (coord_transform): BL 4-velocity -> KS -> MKS relative 4-velocity
            double ucon[4], tmp[4], uconp[4], utconp[4];
            M4 blg, trans, dxdxp, dxpdx;
            {
                double tb = th;
                if (tb < 0) tb *= -1.;
                if (tb > M_PI) tb = 2. * M_PI - tb;
                bl_gcov_func(a, r, tb, blg);
            }
            X[0] = 1;
            ucon[1] = 0.;
            ucon[2] = 0.;
            ucon[3] = up;
            double AAq = blg[0][0];
            double BBq = 2. * (blg[0][1] * ucon[1] + blg[0][2] * ucon[2] + blg[0][3] * ucon[3]);
            double CCq = 1. + blg[1][1] * ucon[1] * ucon[1] + blg[2][2] * ucon[2] * ucon[2] +
                         blg[3][3] * ucon[3] * ucon[3] +
                         2. * (blg[1][2] * ucon[1] * ucon[2] + blg[1][3] * ucon[1] * ucon[3] + blg[2][3] * ucon[2] * ucon[3]);
            double discr = BBq * BBq - 4. * AAq * CCq;
            ucon[0] = (-BBq - sqrt(discr)) / (2. * AAq);
            for (int jj = 0; jj < 4; jj++)
                for (int kk = 0; kk < 4; kk++) trans[jj][kk] = 0.;
            for (int jj = 0; jj < 4; jj++) trans[jj][jj] = 1.;
            trans[0][1] = 2. * r / (r * r - 2. * r + a * a);
            trans[3][1] = a / (r * r - 2. * r + a * a);
            for (int jj = 0; jj < 4; jj++) tmp[jj] = 0.;
            for (int jj = 0; jj < 4; jj++)
                for (int kk = 0; kk < 4; kk++) tmp[jj] += trans[jj][kk] * ucon[kk];
            for (int jj = 0; jj < 4; jj++) ucon[jj] = tmp[jj];
            dxdxp_func(P, X, dxdxp);
            invert_matrix(dxdxp, dxpdx);
            for (int ii = 0; ii < 4; ii++) {
                uconp[ii] = 0;
                for (int jj = 0; jj < 4; jj++) uconp[ii] += dxpdx[ii][jj] * ucon[jj];
            }
            // phys.c:255-271 (ucon_to_utcon) with the CENT metric
            const std::vector<double> &T = t.geom[3];
            double gcon00 = T[(size_t)16 * ncol + col];
            double alpha = 1. / sqrt(-gcon00);
            double beta[4];
            for (int jj = 1; jj < 4; jj++) beta[jj] = T[(size_t)(16 + jj) * ncol + col] * alpha * alpha;
            double gamma = alpha * uconp[0];
            utconp[0] = 0;
            for (int jj = 1; jj < 4; jj++) utconp[jj] = uconp[jj] + gamma * beta[jj] / alpha;

            tor[0 * ncol + col] = rho;
            tor[1 * ncol + col] = u;
            tor[2 * ncol + col] = utconp[1];
            tor[3 * ncol + col] = utconp[2];
            tor[4 * ncol + col] = utconp[3];
            tor[5 * ncol + col] = 1.;
        }
    return 0;
}

}  // namespace harm_host
