/*
 * harm_oracle.c -- CPU oracle for the HARM-COOL per-timestep GRMHD update.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this file's library, and only as the checker.
 *
 * Plain-C restatement of the reference (agnieszkajaniuk/HARM_COOL, src/):
 * each function cites the file:line it follows.  Arithmetic expressions are
 * kept in the reference's operation order because the parity bar is bitwise
 * (FP64, no FMA contraction: build with -ffp-contract=off).
 *
 * Differences from the reference that do not change results:
 *   - single process, whole domain: the MPI halo exchanges of step_ch.c,
 *     fixup.c and bounds.c become periodic phi indexing and in-place copies;
 *   - fail() (diag.c:381-397) and exit() paths count into ORC_FATAL and the
 *     step returns a nonzero status instead of terminating the process;
 *   - file-scope solver globals (utoprim_1dfix1.c:78, u2p_util.c:93,285)
 *     become per-call structs, so zone loops may run on OpenMP threads.
 *
 * Parity pins (tests/test_oracle_*.py): Utoprim_2d / Font-fix / primtoU_g /
 * geometry / slope limiter / KISS against the compiled reference subset in
 * oracle/_ref; whole-step composition against closed forms and invariants
 * (divB, mass budget, reflection symmetry, stationary torus) and against the
 * reference's measured config-1 Newton-iteration and Font-fix counts.
 */
#define _GNU_SOURCE
#include <sys/stat.h>

#include "harm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NPR 8
#define NDIM 4
#define NPG 5
#define RHO 0
#define UU 1
#define U1 2
#define U2 3
#define U3 4
#define B1 5
#define B2 6
#define B3 7
#define TT 0
#define RR 1
#define TH 2
#define PH 3
#define FACE1 0
#define FACE2 1
#define CORN 2
#define CENT 3
#define FACE3 4
#define MC 0
#define VANL 1
#define MINM 2
#define PPM 3 /* not in the reference (interp.c:49-93 has MC, VANL, MINM only) */

/* decs.h:147-171 */
#define RHOMIN (1.e-6)
#define UUMIN (1.e-8)
#define RHOMINLIMIT (1.e-20)
#define UUMINLIMIT (1.e-20)
#define POWRHO (2.0)
#define BSQORHOMAX (50.)
#define BSQOUMAX (2500.)
#define UORHOMAX (50.)
#define SMALL (1.e-20)
#define GAMMAMAX (50.)
#define SAFE (1.3)
#define SINGSMALL (1.E-20)
/* metric.c:158 */
#define DELTA 1.e-5
/* u2p_defs.h:79-96 */
#define MAX_NEWT_ITER 30
#define NEWT_TOL 1.0e-10
#define MIN_NEWT_TOL 1.0e-10
#define EXTRA_NEWT_ITER 2
#define NEWT_TOL2 1.0e-15
#define NUMEPS5 (2.2204460492503131e-16) /* NUMEPSILON, u2p_defs.h:96 */
#define TOL_RESID5 (1.e-4)               /* TOL_RESID, utoprim_5d.c:55 */
#define W_TOO_BIG 1.e20
#define UTSQ_TOO_BIG 1.e20
#define FAIL_VAL 1.e30

#define MY_MIN(a, b) (((a) < (b)) ? (a) : (b))
#define MY_MAX(a, b) (((a) > (b)) ? (a) : (b))
#define dot(a, b) (a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3])
#define delta(i, j) (((i) == (j)) ? 1. : 0.)

/* The reference is built with g++ -O2, which fuses sin(x) and cos(x) of one
 * argument in one function into glibc's sincos(); glibc 2.35's sincos differs
 * from separate cos() by 1 ulp for ~0.1% of arguments.  Where the reference's
 * compiled code computes both (coord.c:147-154, metric.c:345-347, init.c:188-197)
 * the restatement calls sincos explicitly so it does not depend on the
 * compiler's choice. */
static inline void sincos_ref(double x, double *s, double *c)
{
    sincos(x, s, c);
}

typedef struct {
    double gcon[NDIM][NDIM];
    double gcov[NDIM][NDIM];
    double g;
} geom_t;

typedef struct {
    double ucon[NDIM], ucov[NDIM], bcon[NDIM], bcov[NDIM];
} state_t;

struct orc_ctx {
    orc_params P;
    int N1, N2, N3, SN1, SN2;
    int istart, istop, jstart, jstop;
    int polefix;
    double dx[NDIM], startx[NDIM];
    double t, dt;
    double *p, *ph, *dq, *F1, *F2, *F3, *e1, *e2, *e3;
    int8_t *pflag;
    double *gcov, *gcon, *gdet, *conn, *coords;
    double *Katm;
    int64_t cnt[ORC_NCOUNTERS];
    double mass_in, mass_out, mass_start;
    int flux_valid; /* F1 holds the fluxes of a completed full step (diag_flux) */
    /* tracers (tracers.c): slot k = (h N1 + i) N2 + j, X[3][n], state per slot */
    int tr_on;
    long long tr_n, tr_nev, tr_cap;
    double *tr_X;
    int8_t *tr_state;
    orc_tracer_rec *tr_ev;
    double tr_Rin, tr_Rout, tr_thmin, tr_thmax, tr_rho_unit, tr_t_unit;
    const orc_eos *eos; /* tabulated EOS of the step (NULL: Gamma-law) */
    int eos_refd;
    /* the cooling-table cache EOS of the step (cooleos.c), when ceos_on */
    int ceos_on;
    orc_ceos ceos;
    double *lastT; /* the zones' last temperatures (cooleos.c:67-69), [zi] */
};

/* The EOS of the step being computed: set from the context on entry to
 * orc_step / orc_debug_fluxes and read by the physics functions (one step at
 * a time per process; OpenMP threads only read it). */
static const orc_eos *g_step_eos = NULL;
static int g_step_refd = 0;

/* Planted mistakes for the tests that show a pin catches them: built with
 * -DORC_MUTANT=n, site MUT(n) takes its wrong branch (tests/oracle_lib.py
 * build_mutant).  The default build has none. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define MUT(n) (ORC_MUTANT == (n))

/* The cache EOS of the step being computed (the context, or NULL), and per
 * OpenMP thread the zone InitCache_CoolEOS selected: its last temperature and
 * its radial index's cache values. */
static const orc_ctx *g_step_ceos = NULL;
static inline size_t zi(const orc_ctx *o, int h, int i, int j);
static _Thread_local double *t_lastT = NULL;
static _Thread_local const double *t_cvals = NULL;

/* cooleos.c:72, 86-95 */
#define CEOS_HH 1.0E-6
static const double CEOS_KBOL = 1.3e-16;
static const double CEOS_CL = 2.99792458e10;
static const double CEOS_MAV = (1.67262163783e-24 + 1.67492729e-24 + 2. * 9.10956e-28 + 6.6465e-24) / 5.;

/* InitCache_CoolEOS (cooleos.c:770-807): the zone's last temperature, or, when
 * it was never set, its j-1 neighbour's (the loops run j upwards), and the
 * cache of radial index i. */
static void ceos_init(const orc_ctx *o, int h, int i, int j)
{
    double *L = o->lastT;
    t_lastT = &L[zi(o, h, i, j)];
    if (*t_lastT == 0.0 && !MUT(16)) {
        if (j > o->jstart && (*t_lastT = L[zi(o, h, i, j - 1)]) != 0.0) {
        }
    }
    const orc_ceos *c = &o->ceos;
    t_cvals = c->values + (size_t)(c->n == 1 ? 0 : i) * c->X * c->Y * c->D;
}

/* pressure_rho0_u_CoolEOS (cooleos.c:314-444), MAXLOS 1 (one solve). */
static double ceos_pres(double rho0, double u)
{
    const orc_ceos *c = &g_step_ceos->ceos;
    static const int types[3] = {0, 0, 2};
    const double zv = 0.0;
    double rho = rho0 * c->rho_unit;
    double e = u * c->u_unit;
    if (*t_lastT == 0.0) {
        const double corrfact = 1.0;
        *t_lastT = (c->gam - 1.) * u / rho0 / CEOS_KBOL * CEOS_MAV * CEOS_CL * CEOS_CL * corrfact;
        if (*t_lastT < c->T_min * (1.0 + CEOS_HH)) *t_lastT = c->T_min * (1.0 + CEOS_HH);
    }
    if (rho < c->rho_min) return (c->gam - 1.) * u;
    if (rho > c->rho_max) return (c->gam - 1.) * u;
    if (e < 0.0) return 0.0;
    double minx = c->T_min * (1.0 + CEOS_HH), maxx = c->T_max * (1.0 - CEOS_HH);
    double x = *t_lastT, berr;
    orc_cache_solve_T(c->X, c->Y, 1, c->D, types, 1, c->vx, c->vy, &zv, t_cvals, NULL, rho, e, 0, x, minx, maxx, 300,
                      1.0E-6, &x, &berr);
    if (berr > 1.0E-6) return (c->gam - 1.) * u;
    double val[64];
    int mask[64] = {0};
    mask[1] = 1;
    if (MUT(18)) mask[0] = 1;
    orc_cache_get(c->X, c->Y, 1, c->D, types, 1, c->vx, c->vy, &zv, t_cvals, NULL, rho, x, 0.0, 0, 1, mask, val, NULL,
                  1);
    *t_lastT = x;
    return val[MUT(18) ? 0 : 1] / c->p_unit;
}

/* dpdu_calc_CoolEOS (cooleos.c:667-701) */
static double ceos_dpdu(double rho0, double u)
{
    double h = u * CEOS_HH / g_step_ceos->ceos.u_unit;
    double uP = u + h;
    double uM = u - h;
    double ret;
    if (uM > 0) {
        double pP = ceos_pres(rho0, uP);
        double pM = ceos_pres(rho0, uM);
        ret = (pP - pM) / (2.0 * h);
    } else {
        uM = u;
        double pP = ceos_pres(rho0, uP);
        double pM = ceos_pres(rho0, uM);
        ret = (pP - pM) / (h);
    }
    return ret;
}

/* dpdrho_calc_CoolEOS (cooleos.c:704-741) */
static double ceos_dpdrho(double rho0, double u)
{
    double h = rho0 * CEOS_HH / g_step_ceos->ceos.rho_unit;
    double rhoP = rho0 + h;
    double rhoM = rho0 - h;
    double ret;
    if (rhoM > 0) {
        double pP = ceos_pres(rhoP, u);
        double pM = ceos_pres(rhoM, u);
        ret = (pP - pM) / (2.0 * h);
    } else {
        rhoM = rho0;
        double pP = ceos_pres(rhoP, u);
        double pM = ceos_pres(rhoM, u);
        ret = (pP - pM) / (h);
    }
    return ret;
}

/* mhd_calc_CoolEOS's pressure (phys.c:122-141) or the Gamma-law one. */
static double step_pres(double gam, double r, double u)
{
    double p;
    if (g_step_ceos) return ceos_pres(r, u);
    if (!g_step_eos) return (gam - 1.) * u;
    orc_eos_table_p(g_step_eos, r, u, &p, NULL, NULL, NULL);
    return p;
}

/* ------------------------------------------------------------------------ */
/* indexing                                                                  */
/* ------------------------------------------------------------------------ */
static inline size_t zi(const orc_ctx *o, int h, int i, int j)
{
    return ((size_t)h * o->SN1 + (size_t)i) * o->SN2 + (size_t)j;
}
#define PV(o, arr, h, i, j) (&(arr)[zi(o, h, i, j) * NPR])

static inline int hwrap(const orc_ctx *o, int h)
{
    return (h + o->N3) % o->N3;
}
static inline double *GCOV(const orc_ctx *o, int i, int j, int loc)
{
    return &o->gcov[(((size_t)i * o->SN2 + j) * NPG + loc) * 16];
}
static inline double *GCON(const orc_ctx *o, int i, int j, int loc)
{
    return &o->gcon[(((size_t)i * o->SN2 + j) * NPG + loc) * 16];
}
static inline double GDET(const orc_ctx *o, int i, int j, int loc)
{
    return o->gdet[((size_t)i * o->SN2 + j) * NPG + loc];
}
static inline const double *CONN(const orc_ctx *o, int i, int j)
{
    return &o->conn[((size_t)i * o->SN2 + j) * 64];
}

/* metric.c:258-274 */
static void get_geometry(const orc_ctx *o, int i, int j, int loc, geom_t *g)
{
    memcpy(g->gcov, GCOV(o, i, j, loc), 16 * sizeof(double));
    memcpy(g->gcon, GCON(o, i, j, loc), 16 * sizeof(double));
    g->g = GDET(o, i, j, loc);
}

/* ------------------------------------------------------------------------ */
/* lu.c:60-324                                                               */
/* ------------------------------------------------------------------------ */
static int LU_decompose(double A[NDIM][NDIM], int permute[])
{
    const double absmin = 1.e-30;
    double row_norm[NDIM];
    double absmax, maxtemp;
    int i, j, k, max_row;
    int n = NDIM;

    max_row = 0;
    for (i = 0; i < n; i++) {
        absmax = 0.;
        for (j = 0; j < n; j++) {
            maxtemp = fabs(A[i][j]);
            if (maxtemp > absmax) absmax = maxtemp;
        }
        if (absmax == 0.) return 1;
        row_norm[i] = 1. / absmax;
    }
    for (j = 0; j < n; j++) {
        for (i = 0; i < j; i++)
            for (k = 0; k < i; k++) A[i][j] -= A[i][k] * A[k][j];
        absmax = 0.0;
        for (i = j; i < n; i++) {
            for (k = 0; k < j; k++) A[i][j] -= A[i][k] * A[k][j];
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
                for (k = 0; k < n; k++) {
                    maxtemp = A[j][k];
                    A[j][k] = A[max_row][k];
                    A[max_row][k] = maxtemp;
                }
                row_norm[max_row] = row_norm[j];
            }
        }
        permute[j] = max_row;
        if (A[j][j] == 0.) A[j][j] = absmin;
        if (j != (n - 1)) {
            maxtemp = 1. / A[j][j];
            for (i = (j + 1); i < n; i++) A[i][j] *= maxtemp;
        }
    }
    return 0;
}

static void LU_substitution(double A[NDIM][NDIM], double B[], int permute[])
{
    int i, j;
    int n = NDIM;
    double tmpvar;
    for (i = 0; i < n; i++) {
        tmpvar = B[permute[i]];
        B[permute[i]] = B[i];
        for (j = (i - 1); j >= 0; j--) tmpvar -= A[i][j] * B[j];
        B[i] = tmpvar;
    }
    for (i = (n - 1); i >= 0; i--) {
        for (j = (i + 1); j < n; j++) B[i] -= A[i][j] * B[j];
        B[i] /= A[i][i];
    }
}

static int invert_matrix(double Am[NDIM][NDIM], double Aminv[NDIM][NDIM])
{
    int i, j, n = NDIM, permute[NDIM];
    double dxm[NDIM], Amtmp[NDIM][NDIM];
    for (j = 0; j < NDIM; j++)
        for (i = 0; i < NDIM; i++) Amtmp[j][i] = Am[j][i];
    if (LU_decompose(Amtmp, permute) != 0) return 1;
    for (i = 0; i < n; i++) {
        for (j = 0; j < n; j++) dxm[j] = 0.;
        dxm[i] = 1.;
        LU_substitution(Amtmp, dxm, permute);
        for (j = 0; j < n; j++) Aminv[j][i] = dxm[j];
    }
    return 0;
}

/* metric.c:118-138 */
static double gdet_func(double gcov[NDIM][NDIM], int *bad)
{
    int j, k, permute[NDIM];
    double gcovtmp[NDIM][NDIM], detg;
    for (j = 0; j < NDIM; j++)
        for (k = 0; k < NDIM; k++) gcovtmp[j][k] = gcov[j][k];
    if (LU_decompose(gcovtmp, permute) != 0) {
        if (bad) *bad = 1;
    }
    detg = 1.;
    for (j = 0; j < NDIM; j++) detg *= gcovtmp[j][j];
    return sqrt(fabs(detg));
}

/* ------------------------------------------------------------------------ */
/* coordinates and metric                                                    */
/* ------------------------------------------------------------------------ */
/* coord.c:61-81 (gammie_coords) + coord.c:84-112 (bl_coord_vec) */
static void bl_coord_vec(const orc_params *P, const double *X, double *V)
{
    double theexp = X[RR];
    if (P->Rbr >= 0.) {
        double x1br = log(P->Rbr - P->R0);
        if (X[RR] > x1br) {
            const double npow2 = 4.0;
            const double cpow2 = 1.0;
            theexp += cpow2 * pow(X[RR] - x1br, npow2);
        }
    }
    V[TT] = X[TT];
    V[RR] = exp(theexp) + P->R0;
    V[TH] = M_PI_2 * (1. + X[TH]) + ((1. - P->hslope) / 2.) * sin(M_PI * (1. + X[TH]));
    V[PH] = X[PH];
    if (fabs(V[TH]) < SINGSMALL) {
        if (V[TH] >= 0) V[TH] = SINGSMALL;
        if (V[TH] < 0) V[TH] = -SINGSMALL;
    }
    if (fabs(M_PI - V[TH]) < SINGSMALL) {
        if (V[TH] >= M_PI) V[TH] = M_PI + SINGSMALL;
        if (V[TH] < M_PI) V[TH] = M_PI - SINGSMALL;
    }
}

/* metric.c:373-399 (BL=1 branch): central differences with step DELTA */
static void dxdxp_func(const orc_params *P, const double *X, double dxdxp[NDIM][NDIM])
{
    int j, k, l;
    double Xh[NDIM], Xl[NDIM], Vh[NDIM], Vl[NDIM];
    for (k = 0; k < NDIM; k++) {
        for (l = 0; l < NDIM; l++) Xh[l] = X[l];
        for (l = 0; l < NDIM; l++) Xl[l] = X[l];
        Xh[k] += DELTA;
        Xl[k] -= DELTA;
        bl_coord_vec(P, Xh, Vh);
        bl_coord_vec(P, Xl, Vl);
        for (j = 0; j < NDIM; j++) dxdxp[j][k] = (Vh[j] - Vl[j]) / (Xh[k] - Xl[k]);
    }
}

/* coord.c:128-199: Kerr-Schild metric transformed to code coordinates */
static void gcov_func(const orc_params *P, const double *X, double gcovp[NDIM][NDIM])
{
    int i, j, k, l;
    double sth, cth, s2, rho2, r, th;
    double gcov[NDIM][NDIM], dxdxp[NDIM][NDIM], V[NDIM];
    const double a = P->a;

    for (j = 0; j < NDIM; j++)
        for (k = 0; k < NDIM; k++) gcov[j][k] = 0.;
    for (j = 0; j < NDIM; j++)
        for (k = 0; k < NDIM; k++) gcovp[j][k] = 0.;
    bl_coord_vec(P, X, V);
    r = V[RR];
    th = V[TH];
    sincos_ref(th, &sth, &cth);
    s2 = sth * sth;
    rho2 = r * r + a * a * cth * cth;
    dxdxp_func(P, X, dxdxp);

    gcov[TT][TT] = (-1. + 2. * r / rho2);
    gcov[TT][1] = (2. * r / rho2);
    gcov[TT][3] = (-2. * a * r * s2 / rho2);
    gcov[1][TT] = gcov[TT][1];
    gcov[1][1] = (1. + 2. * r / rho2);
    gcov[1][3] = (-a * s2 * (1. + 2. * r / rho2));
    gcov[2][2] = rho2;
    gcov[3][TT] = gcov[TT][3];
    gcov[3][1] = gcov[1][3];
    gcov[3][3] = s2 * (rho2 + a * a * s2 * (1. + 2. * r / rho2));

    for (i = 0; i < NDIM; i++)
        for (j = 0; j < NDIM; j++) {
            gcovp[i][j] = 0;
            for (k = 0; k < NDIM; k++)
                for (l = 0; l < NDIM; l++) gcovp[i][j] += gcov[k][l] * dxdxp[k][i] * dxdxp[l][j];
        }
}

/* metric.c:161-207: connection by central differences of gcov */
static void conn_func(const orc_params *P, const double *X, double gcon[NDIM][NDIM],
                      double conn[NDIM][NDIM][NDIM])
{
    int i, j, k, l;
    double tmp[NDIM][NDIM][NDIM], Xh[NDIM], Xl[NDIM], gh[NDIM][NDIM], gl[NDIM][NDIM];
    for (k = 0; k < NDIM; k++) {
        for (l = 0; l < NDIM; l++) Xh[l] = X[l];
        for (l = 0; l < NDIM; l++) Xl[l] = X[l];
        Xh[k] += DELTA;
        Xl[k] -= DELTA;
        gcov_func(P, Xh, gh);
        gcov_func(P, Xl, gl);
        for (i = 0; i < NDIM; i++)
            for (j = 0; j < NDIM; j++) conn[i][j][k] = (gh[i][j] - gl[i][j]) / (Xh[k] - Xl[k]);
    }
    for (i = 0; i < NDIM; i++)
        for (j = 0; j < NDIM; j++)
            for (k = 0; k < NDIM; k++)
                tmp[i][j][k] = 0.5 * (conn[j][i][k] + conn[k][i][j] - conn[k][j][i]);
    for (i = 0; i < NDIM; i++)
        for (j = 0; j < NDIM; j++)
            for (k = 0; k < NDIM; k++) {
                conn[i][j][k] = 0.;
                for (l = 0; l < NDIM; l++) conn[i][j][k] += gcon[i][l] * tmp[l][j][k];
            }
}

/* metric.c:62-115 with the single-rank offsets; h is the 0-based phi plane */
static void coord(const orc_ctx *o, int h, int i, int j, int loc, double *X)
{
    const int N3 = o->N3;
    i -= 2;
    j -= 2;
    X[TT] = 1.;
    if (loc == FACE1) {
        X[RR] = o->startx[RR] + i * o->dx[RR];
        X[TH] = o->startx[TH] + (j + 0.5) * o->dx[TH];
        X[PH] = (N3 > 2) ? o->startx[PH] + (h + 0.5) * o->dx[PH] : o->startx[PH] + h * o->dx[PH];
    } else if (loc == FACE2) {
        X[RR] = o->startx[RR] + (i + 0.5) * o->dx[RR];
        X[TH] = o->startx[TH] + j * o->dx[TH];
        X[PH] = (N3 > 2) ? o->startx[PH] + (h + 0.5) * o->dx[PH] : o->startx[PH] + h * o->dx[PH];
    } else if (loc == FACE3) {
        X[RR] = o->startx[RR] + (i + 0.5) * o->dx[RR];
        X[TH] = o->startx[TH] + (j + 0.5) * o->dx[TH];
        X[PH] = o->startx[PH] + h * o->dx[PH];
    } else if (loc == CENT) {
        X[RR] = o->startx[RR] + (i + 0.5) * o->dx[RR];
        X[TH] = o->startx[TH] + (j + 0.5) * o->dx[TH];
        X[PH] = (N3 > 2) ? o->startx[PH] + (h + 0.5) * o->dx[PH] : o->startx[PH] + h * o->dx[PH];
    } else {
        X[RR] = o->startx[RR] + i * o->dx[RR];
        X[TH] = o->startx[TH] + j * o->dx[TH];
        X[PH] = o->startx[PH] + h * o->dx[PH];
    }
}

/* coord.c:202-262 (Rbr < 0 branch and hyperexponential root find) */
static int set_points(orc_ctx *o)
{
    const orc_params *P = &o->P;
    double x1max = log(P->Rout - P->R0);
    if (!(P->Rbr < 0. || P->Rout < P->Rbr)) {
        const double RELACC = 1e-20;
        const int ITERMAX = 200;
        const double npow2 = 4.0, cpow2 = 1.0;
        double x1max0 = 1.;
        int iter;
        double x1br = log(P->Rbr - P->R0);
        x1max = 2.;
        for (iter = 0; iter < ITERMAX; iter++) {
            if (fabs((x1max - x1max0) / x1max) < RELACC) break;
            x1max0 = x1max;
            double dxmax = (pow((log(P->Rout - P->R0) - x1max0) / cpow2, 1. / npow2) + x1br) - x1max0;
            double dampingfactor = 0.5;
            x1max = x1max0 + dampingfactor * dxmax;
            if (x1max > log(P->Rout - P->R0)) x1max = log(P->Rout - P->R0);
        }
        if (iter == ITERMAX) return 1;
    }
    o->startx[RR] = log(P->Rin - P->R0);
    o->startx[TH] = -1.;
    o->startx[PH] = 0.;
    o->dx[RR] = (x1max - o->startx[RR]) / o->N1;
    o->dx[TH] = 2. / o->N2;
    o->dx[PH] = 2. * M_PI / o->N3;
    return 0;
}

/* main.c:413-508: metric tables at CENT, CORN, FACE1..3 for every (i,j) */
static int set_grid(orc_ctx *o)
{
    int i, j, k, bad = 0;
    double X[NDIM], V[NDIM];
    if (set_points(o)) return 1;
    const int h = 0;
    for (i = 0; i < o->SN1; ++i)
        for (j = 0; j < o->SN2; ++j) {
            double(*gc)[NDIM];
            coord(o, h, i, j, CENT, X);
            gc = (double(*)[NDIM])GCOV(o, i, j, CENT);
            gcov_func(&o->P, X, gc);
            o->gdet[((size_t)i * o->SN2 + j) * NPG + CENT] = gdet_func(gc, &bad);
            invert_matrix(gc, (double(*)[NDIM])GCON(o, i, j, CENT));
            conn_func(&o->P, X, (double(*)[NDIM])GCON(o, i, j, CENT),
                      (double(*)[NDIM][NDIM])CONN(o, i, j));
            bl_coord_vec(&o->P, X, V);
            for (k = 0; k < NDIM; k++) o->coords[((size_t)i * o->SN2 + j) * NDIM + k] = V[k];
            const int locs[4] = {CORN, FACE1, FACE2, FACE3};
            for (int l = 0; l < 4; l++) {
                coord(o, h, i, j, locs[l], X);
                gc = (double(*)[NDIM])GCOV(o, i, j, locs[l]);
                gcov_func(&o->P, X, gc);
                o->gdet[((size_t)i * o->SN2 + j) * NPG + locs[l]] = gdet_func(gc, &bad);
                invert_matrix(gc, (double(*)[NDIM])GCON(o, i, j, locs[l]));
            }
        }
    return bad;
}

/* ------------------------------------------------------------------------ */
/* phys.c                                                                    */
/* ------------------------------------------------------------------------ */
/* metric.c:210-229 */
static void lower(const double *ucon, const geom_t *geom, double *ucov)
{
    ucov[0] = geom->gcov[0][0] * ucon[0] + geom->gcov[0][1] * ucon[1] + geom->gcov[0][2] * ucon[2] +
              geom->gcov[0][3] * ucon[3];
    ucov[1] = geom->gcov[1][0] * ucon[0] + geom->gcov[1][1] * ucon[1] + geom->gcov[1][2] * ucon[2] +
              geom->gcov[1][3] * ucon[3];
    ucov[2] = geom->gcov[2][0] * ucon[0] + geom->gcov[2][1] * ucon[1] + geom->gcov[2][2] * ucon[2] +
              geom->gcov[2][3] * ucon[3];
    ucov[3] = geom->gcov[3][0] * ucon[0] + geom->gcov[3][1] * ucon[1] + geom->gcov[3][2] * ucon[2] +
              geom->gcov[3][3] * ucon[3];
}

/* metric.c:232-255 */
static void raise_(const double *ucov, const geom_t *geom, double *ucon)
{
    ucon[0] = geom->gcon[0][0] * ucov[0] + geom->gcon[0][1] * ucov[1] + geom->gcon[0][2] * ucov[2] +
              geom->gcon[0][3] * ucov[3];
    ucon[1] = geom->gcon[1][0] * ucov[0] + geom->gcon[1][1] * ucov[1] + geom->gcon[1][2] * ucov[2] +
              geom->gcon[1][3] * ucov[3];
    ucon[2] = geom->gcon[2][0] * ucov[0] + geom->gcon[2][1] * ucov[1] + geom->gcon[2][2] * ucov[2] +
              geom->gcon[2][3] * ucov[3];
    ucon[3] = geom->gcon[3][0] * ucov[0] + geom->gcon[3][1] * ucov[1] + geom->gcon[3][2] * ucov[2] +
              geom->gcon[3][3] * ucov[3];
}

/* phys.c:300-324 */
static int gamma_calc(const double *pr, const geom_t *geom, double *gamma)
{
    double qsq;
    qsq = geom->gcov[1][1] * pr[U1] * pr[U1] + geom->gcov[2][2] * pr[U2] * pr[U2] +
          geom->gcov[3][3] * pr[U3] * pr[U3] +
          2. * (geom->gcov[1][2] * pr[U1] * pr[U2] + geom->gcov[1][3] * pr[U1] * pr[U3] +
                geom->gcov[2][3] * pr[U2] * pr[U3]);
    if (qsq < 0.) {
        if (fabs(qsq) > 1.E-10) {
            *gamma = 1.;
            return 1;
        } else
            qsq = 1.E-10;
    }
    *gamma = sqrt(1. + qsq);
    return 0;
}

/* phys.c:232-252; *fatal incremented where the reference calls fail() */
static void ucon_calc(const double *pr, const geom_t *geom, double *ucon, int64_t *fatal)
{
    double alpha, gamma, beta[NDIM];
    int j;
    alpha = 1. / sqrt(-geom->gcon[TT][TT]);
    for (j = 1; j < NDIM; j++) beta[j] = geom->gcon[TT][j] * alpha * alpha;
    if (gamma_calc(pr, geom, &gamma)) (*fatal)++;
    ucon[TT] = gamma / alpha;
    for (j = 1; j < NDIM; j++) ucon[j] = pr[U1 + j - 1] - gamma * beta[j] / alpha;
}

/* phys.c:108-117 */
static void bcon_calc(const double *pr, const double *ucon, const double *ucov, double *bcon)
{
    int j;
    bcon[TT] = pr[B1] * ucov[1] + pr[B2] * ucov[2] + pr[B3] * ucov[3];
    for (j = 1; j < 4; j++) bcon[j] = (pr[B1 - 1 + j] + bcon[TT] * ucon[j]) / ucon[TT];
}

/* phys.c:219-229 */
static void get_state(const double *pr, const geom_t *geom, state_t *q, int64_t *fatal)
{
    ucon_calc(pr, geom, q->ucon, fatal);
    lower(q->ucon, geom, q->ucov);
    bcon_calc(pr, q->ucon, q->ucov, q->bcon);
    lower(q->bcon, geom, q->bcov);
}

/* phys.c:143-161 */
static void mhd_calc(const double *pr, int dir, const state_t *q, double gam, double *mhd)
{
    int j;
    double r, u, P, w, bsq, eta, ptot;
    r = pr[RHO];
    u = pr[UU];
    P = step_pres(gam, r, u);
    w = P + r + u;
    bsq = dot(q->bcon, q->bcov);
    eta = w + bsq;
    ptot = P + 0.5 * bsq;
    for (j = 0; j < NDIM; j++)
        mhd[j] = eta * q->ucon[dir] * q->ucov[j] + ptot * delta(dir, j) - q->bcon[dir] * q->bcov[j];
}

/* phys.c:59-96 */
static void primtoflux(const double *pr, const state_t *q, int dir, const geom_t *geom,
                       double gam, double *flux)
{
    int k;
    double mhd[NDIM];
    flux[RHO] = pr[RHO] * q->ucon[dir];
    mhd_calc(pr, dir, q, gam, mhd);
    flux[UU] = mhd[0] + flux[RHO];
    flux[U1] = mhd[1];
    flux[U2] = mhd[2];
    flux[U3] = mhd[3];
    flux[B1] = q->bcon[1] * q->ucon[dir] - q->bcon[dir] * q->ucon[1];
    flux[B2] = q->bcon[2] * q->ucon[dir] - q->bcon[dir] * q->ucon[2];
    flux[B3] = q->bcon[3] * q->ucon[dir] - q->bcon[dir] * q->ucon[3];
    for (k = 0; k < NPR; k++) flux[k] *= geom->g;
}

/* phys.c:165-207 */
static void source(const double *ph, const geom_t *geom, const double *conn, double gam,
                   double *dU, int64_t *fatal)
{
    double mhd[NDIM][NDIM];
    int j, k;
    state_t q;
    get_state(ph, geom, &q, fatal);
    mhd_calc(ph, 0, &q, gam, mhd[0]);
    mhd_calc(ph, 1, &q, gam, mhd[1]);
    mhd_calc(ph, 2, &q, gam, mhd[2]);
    mhd_calc(ph, 3, &q, gam, mhd[3]);
    for (k = 0; k < NPR; k++) dU[k] = 0.;
    for (j = 0; j < NDIM; j++)
        for (k = 0; k < NDIM; k++) {
            dU[UU] += mhd[j][k] * conn[(k * 4 + 0) * 4 + j];
            dU[U1] += mhd[j][k] * conn[(k * 4 + 1) * 4 + j];
            dU[U2] += mhd[j][k] * conn[(k * 4 + 2) * 4 + j];
            dU[U3] += mhd[j][k] * conn[(k * 4 + 3) * 4 + j];
        }
    for (k = 0; k < NPR; k++) dU[k] *= geom->g;
}

/* phys.c:337-476 (Gamma-law branch); returns the fail() count it would make */
static int vchar(const double *pr, const state_t *q, const geom_t *geom, int js, double gam,
                 double *vmax, double *vmin)
{
    double discr, vp, vm, bsq, EE, EF, va2, cs2, cms2, rho, u;
    double Acov[NDIM], Bcov[NDIM], Acon[NDIM], Bcon[NDIM];
    double Asq, Bsq, Au, Bu, AB, Au2, Bu2, AuBu, A, B, C;
    int j, nfail = 0;

    for (j = 0; j < NDIM; j++) Acov[j] = 0.;
    Acov[js] = 1.;
    raise_(Acov, geom, Acon);
    for (j = 0; j < NDIM; j++) Bcov[j] = 0.;
    Bcov[TT] = 1.;
    raise_(Bcov, geom, Bcon);

    bsq = dot(q->bcon, q->bcov);
    rho = pr[RHO];
    u = pr[UU];
    if (g_step_eos || g_step_ceos) {
        /* phys.c:359-378, 398-406 */
        double p = step_pres(gam, rho, u), pp, dpdrho, dpdu, e, enthalpy, dpde;
        EF = rho + u + p;
        EE = bsq + EF;
        va2 = bsq / EE;
        if (g_step_ceos) {
            e = u / rho;
            enthalpy = 1.0 + e + p / rho;
            dpdrho = ceos_dpdrho(rho, u);
            dpde = ceos_dpdu(rho, u) * rho;
            cs2 = (dpdrho + p / rho / rho * dpde) / enthalpy;
            goto have_cs2;
        }
        orc_eos_table_p(g_step_eos, rho, u, &pp, &dpdrho, &dpdu, NULL);
        if (g_step_refd) dpdrho = dpdu = 0.;
        e = u / rho;
        enthalpy = 1.0 + e + p / rho;
        dpde = dpdu * rho;
        cs2 = (dpdrho + p / rho / rho * dpde) / enthalpy;
    have_cs2:
        cms2 = cs2 + va2 - cs2 * va2;
        if (cms2 < SMALL) cms2 = SMALL;
        if (cms2 > 1.) cms2 = 1.;
    } else {
        EF = rho + gam * u;
        EE = bsq + EF;
        va2 = bsq / EE;
        cs2 = gam * (gam - 1.) * u / EF;
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
    Asq = dot(Acon, Acov);
    Bsq = dot(Bcon, Bcov);
    Au = dot(Acov, q->ucon);
    Bu = dot(Bcov, q->ucon);
    AB = dot(Acon, Bcov);
    Au2 = Au * Au;
    Bu2 = Bu * Bu;
    AuBu = Au * Bu;
    A = Bu2 - (Bsq + Bu2) * cms2;
    B = 2. * (AuBu - (AB + AuBu) * cms2);
    C = Au2 - (Asq + Au2) * cms2;
    discr = B * B - 4. * A * C;
    if ((discr < 0.0) && (discr > -1.e-10))
        discr = 0.0;
    else if (discr < -1.e-10) {
        nfail++;
        discr = 0.;
    }
    discr = sqrt(discr);
    vp = -(-B + discr) / (2. * A);
    vm = -(-B - discr) / (2. * A);
    if (vp > vm) {
        *vmax = vp;
        *vmin = vm;
    } else {
        *vmax = vm;
        *vmin = vp;
    }
    return nfail;
}

/* interp.c:49-93 */
static double slope_lim(int lim, double y1, double y2, double y3)
{
    double Dqm, Dqp, Dqc, s;
    if (lim == MC) {
        Dqm = 2. * (y2 - y1);
        Dqp = 2. * (y3 - y2);
        Dqc = 0.5 * (y3 - y1);
        s = Dqm * Dqp;
        if (s <= 0.) return 0.;
        if (fabs(Dqm) < fabs(Dqp) && fabs(Dqm) < fabs(Dqc)) return Dqm;
        else if (fabs(Dqp) < fabs(Dqc)) return Dqp;
        else return Dqc;
    } else if (lim == VANL) {
        Dqm = (y2 - y1);
        Dqp = (y3 - y2);
        s = Dqm * Dqp;
        if (s <= 0.) return 0.;
        return 2. * s / (Dqm + Dqp);
    } else {
        Dqm = (y2 - y1);
        Dqp = (y3 - y2);
        s = Dqm * Dqp;
        if (s <= 0.) return 0.;
        else if (fabs(Dqm) < fabs(Dqp)) return Dqm;
        else return Dqp;
    }
}

/* Piecewise-parabolic reconstruction of one zone (Colella & Woodward 1984,
 * J. Comput. Phys. 54, 174, eqs. 1.6 and 1.10), built on the MC-limited slopes
 * slope_lim() gives (interp.c:49-93; the reference has no PPM).  y0 is the
 * zone's value, ym/yp its neighbours along the direction, dqm/dq0/dqp the MC
 * slopes of the three zones.  Returns the zone's left and right face values:
 *   interface  a(m+1/2) = (y(m) + y(m+1))/2 - (dq(m+1) - dq(m))/6
 *   monotone   a local extremum flattens the zone (l = r = y0); a parabola
 *              that overshoots moves the far face value to 3 y0 - 2 (near one).
 * The expression of a shared interface is the same from both zones, so the two
 * zones start from identical bits. */
static void para(double ym, double y0, double yp, double dqm, double dq0, double dqp, double *l, double *r)
{
    double ql = 0.5 * (ym + y0) - (MUT(13) ? 0.125 : 1. / 6.) * (dq0 - dqm);
    double qr = 0.5 * (y0 + yp) - (MUT(13) ? 0.125 : 1. / 6.) * (dqp - dq0);
    const double qd = qr - ql, qe = 6. * (y0 - 0.5 * (ql + qr));
    if ((qr - y0) * (y0 - ql) <= 0.) {
        ql = y0;
        qr = y0;
    } else if (qd * (qd - qe) < 0.) {
        ql = MUT(14) ? 3. * y0 - 2. * ql : 3. * y0 - 2. * qr;
    } else if (qd * (qd + qe) < 0.) {
        qr = 3. * y0 - 2. * ql;
    }
    *l = ql;
    *r = qr;
}

double orc_para(double ym, double y0, double yp, double dqm, double dq0, double dqp, double *l, double *r)
{
    para(ym, y0, yp, dqm, dq0, dqp, l, r);
    return *r - *l;
}

/* ------------------------------------------------------------------------ */
/* u2p_util.c:84-366                                                         */
/* ------------------------------------------------------------------------ */
static void raise_g(const double vcov[NDIM], double gcon[NDIM][NDIM], double vcon[NDIM])
{
    int i, j;
    for (i = 0; i < NDIM; i++) {
        vcon[i] = 0.;
        for (j = 0; j < NDIM; j++) vcon[i] += gcon[i][j] * vcov[j];
    }
}
static void lower_g(const double vcon[NDIM], double gcov[NDIM][NDIM], double vcov[NDIM])
{
    int i, j;
    for (i = 0; i < NDIM; i++) {
        vcov[i] = 0.;
        for (j = 0; j < NDIM; j++) vcov[i] += gcov[i][j] * vcon[j];
    }
}
static void ncov_calc(double gcon[NDIM][NDIM], double ncov[NDIM])
{
    double lapse = sqrt(-1. / gcon[0][0]);
    ncov[0] = -lapse;
    for (int i = 1; i < NDIM; i++) ncov[i] = 0.;
}
static void ucon_calc_g(const double prim[NPR], double gcov[NDIM][NDIM], double gcon[NDIM][NDIM],
                        double ucon[NDIM])
{
    double u_tilde_con[4], u_tilde_sq, gamma, lapse;
    int i, j;
    u_tilde_con[0] = 0.;
    u_tilde_con[1] = prim[U1];
    u_tilde_con[2] = prim[U2];
    u_tilde_con[3] = prim[U3];
    u_tilde_sq = 0.;
    for (i = 0; i < NDIM; i++)
        for (j = 0; j < NDIM; j++) u_tilde_sq += gcov[i][j] * u_tilde_con[i] * u_tilde_con[j];
    u_tilde_sq = fabs(u_tilde_sq);
    gamma = sqrt(1. + u_tilde_sq);
    lapse = sqrt(-1. / gcon[0][0]);
    for (i = 0; i < NDIM; i++) ucon[i] = u_tilde_con[i] - lapse * gamma * gcon[0][i];
}
static void bcon_calc_g(const double prim[NPR], const double ucon[NDIM], const double ucov[NDIM],
                        const double ncov[NDIM], double bcon[NDIM])
{
    double Bcon[NDIM], u_dot_B, gamma;
    int i;
    Bcon[0] = 0.;
    for (i = 1; i < NDIM; i++) Bcon[i] = -ncov[0] * prim[B1 + i - 1];
    u_dot_B = 0.;
    for (i = 0; i < NDIM; i++) u_dot_B += ucov[i] * Bcon[i];
    gamma = -ucon[0] * ncov[0];
    for (i = 0; i < NDIM; i++) bcon[i] = (Bcon[i] + ucon[i] * u_dot_B) / gamma;
}
static void primtoU_g(const double prim[NPR], double gcov[NDIM][NDIM], double gcon[NDIM][NDIM],
                      double gdet, double gam, double U[NPR])
{
    int i;
    double ucon[NDIM], ucov[NDIM], bcon[NDIM], bcov[NDIM], ncov[NDIM];
    double rho0, gamma, n_dot_b, bsq, u, p, w, alpha;
    alpha = 1.0 / sqrt(-gcon[0][0]);
    ucon_calc_g(prim, gcov, gcon, ucon);
    lower_g(ucon, gcov, ucov);
    ncov_calc(gcon, ncov);
    gamma = -ncov[0] * ucon[0];
    bcon_calc_g(prim, ucon, ucov, ncov, bcon);
    lower_g(bcon, gcov, bcov);
    n_dot_b = 0.;
    for (i = 0; i < 4; i++) n_dot_b += ncov[i] * bcon[i];
    bsq = 0.;
    for (i = 0; i < 4; i++) bsq += bcov[i] * bcon[i];
    rho0 = prim[RHO];
    u = prim[UU];
    p = (gam - 1.) * u;
    w = rho0 + u + p;
    U[RHO] = ucon[0] * rho0;
    for (i = 0; i < 4; i++) {
        U[1 + i] = gamma * (w + bsq) * ucov[i] - (p + bsq / 2.) * ncov[i] + n_dot_b * bcov[i];
        U[1 + i] /= alpha;
    }
    U[1] = U[1] + U[RHO];
    U[B1] = prim[B1];
    U[B2] = prim[B2];
    U[B3] = prim[B3];
    for (i = 0; i < NPR; i++) U[i] *= gdet;
}

/* ------------------------------------------------------------------------ */
/* Noble 2D inversion: utoprim_2d.c:126-686                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    double Bsq, QdotBsq, Qtsq, Qdotn, D, K_atm, gam;
    double W_for_gnr2, rho_for_gnr2, W_for_gnr2_old, rho_for_gnr2_old;
} u2p_t;

/* utoprim_2d.c:374-383 */
static double vsq_calc(const u2p_t *s, double W)
{
    double Wsq, Xsq;
    Wsq = W * W;
    Xsq = (s->Bsq + W) * (s->Bsq + W);
    return ((Wsq * s->Qtsq + s->QdotBsq * (s->Bsq + 2. * W)) / (Wsq * Xsq));
}
/* utoprim_2d.c:395-406 */
static double x1_of_x0(const u2p_t *s, double x0)
{
    double dv = 1.e-15;
    double vsq = fabs(vsq_calc(s, x0));
    return ((vsq > 1.) ? (1.0 - dv) : vsq);
}
/* utoprim_2d.c:417-431 */
static void validate_x_2d(double x[2], const double x0[2])
{
    double dv = 1.e-15;
    x[0] = fabs(x[0]);
    x[0] = (x[0] > W_TOO_BIG) ? x0[0] : x[0];
    x[1] = (x[1] < 0.) ? 0. : x[1];
    x[1] = (x[1] > 1.) ? (1. - dv) : x[1];
}
/* utoprim_2d.c:562-624 + EOS utoprim_2d.c:645-678 */
static void func_vsq(const u2p_t *s, const double x[], double dx[], double resid[],
                     double jac[][2], double *f, double *df)
{
    double W, vsq, Wsq, p_tmp, dPdvsq, dPdW;
    double t11, t16, t18, t2, t21, t23, t24, t25, t3, t35, t36, t4, t40, t9, gtmp;
    const double GAMMA = s->gam;
    W = x[0];
    vsq = x[1];
    Wsq = W * W;
    gtmp = 1. - vsq;
    p_tmp = (GAMMA - 1.) * (W * gtmp - s->D * sqrt(gtmp)) / GAMMA;
    dPdW = (GAMMA - 1.) * (1. - vsq) / GAMMA;
    dPdvsq = (GAMMA - 1.) * (0.5 * s->D / sqrt(1. - vsq) - W) / GAMMA;
    t2 = -0.5 * s->Bsq + dPdvsq;
    t3 = s->Bsq + W;
    t4 = t3 * t3;
    t9 = 1. / Wsq;
    t11 = s->Qtsq - vsq * t4 + s->QdotBsq * (s->Bsq + 2.0 * W) * t9;
    t16 = s->QdotBsq * t9;
    t18 = -s->Qdotn - 0.5 * s->Bsq * (1.0 + vsq) + 0.5 * t16 - W + p_tmp;
    t21 = 1. / t3;
    t23 = 1. / W;
    t24 = t16 * t23;
    t25 = -1.0 + dPdW - t24;
    t35 = t25 * t3 + (s->Bsq - 2.0 * dPdvsq) * (s->QdotBsq + vsq * Wsq * W) * t9 * t23;
    t36 = 1. / t35;
    dx[0] = -(t2 * t11 + t4 * t18) * t21 * t36;
    t40 = (vsq + t24) * t3;
    dx[1] = -(-t25 * t11 - 2.0 * t40 * t18) * t21 * t36;
    jac[0][0] = -2.0 * t40;
    jac[0][1] = -t4;
    jac[1][0] = t25;
    jac[1][1] = t2;
    resid[0] = t11;
    resid[1] = t18;
    *df = -resid[0] * resid[0] - resid[1] * resid[1];
    *f = -0.5 * (*df);
}
/* utoprim_2d.c:442-541; errx_first (may be NULL) receives |dW/W| + |dv2/v2| of
 * the first iteration (the gate of the learned seed, no reference counterpart). */
static int gnr_2d_first(const u2p_t *s, double x[2], int *n_iter_out, int *k_conv_out, double *errx_first)
{
    double f, df, dx[2], x_old[2], resid[2], jac[2][2], errx;
    int n_iter, id, i_extra, doing_extra, keep_iterating, k_conv = -1;
    errx = 1.;
    df = f = 1.;
    i_extra = doing_extra = 0;
    for (id = 0; id < 2; id++) x_old[id] = x[id];
    n_iter = 0;
    keep_iterating = 1;
    while (keep_iterating) {
        func_vsq(s, x, dx, resid, jac, &f, &df);
        errx = 0.;
        for (id = 0; id < 2; id++) x_old[id] = x[id];
        for (id = 0; id < 2; id++) x[id] += dx[id];
        errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0] / x[0]);
        errx += (x[1] == 0.) ? fabs(dx[1]) : fabs(dx[1] / x[1]);
        validate_x_2d(x, x_old);
        if (errx_first && n_iter == 0) *errx_first = fabs(errx);
        if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) {
            doing_extra = 1;
            k_conv = n_iter + 1;
        }
        if (doing_extra == 1) i_extra++;
        if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
            (n_iter >= (MAX_NEWT_ITER - 1)))
            keep_iterating = 0;
        n_iter++;
    }
    if (n_iter_out) *n_iter_out = n_iter;
    if (k_conv_out) *k_conv_out = (k_conv < 0) ? n_iter : k_conv;
    if ((isfinite(f) == 0) || (isfinite(df) == 0)) return 2;
    if (fabs(errx) > MIN_NEWT_TOL) return 1;
    return 0;
}
static int gnr_2d(const u2p_t *s, double x[2], int *n_iter_out, int *k_conv_out)
{
    return gnr_2d_first(s, x, n_iter_out, k_conv_out, NULL);
}

/* The scalars shared by all three solvers (utoprim_2d.c:240-266). */
static void u2p_scalars(u2p_t *s, const double U[NPR], double gcov[NDIM][NDIM],
                        double gcon[NDIM][NDIM], double Bcon[NDIM], double Qcon[NDIM],
                        double ncon[NDIM], double *QdotB_out)
{
    double QdotB, Bcov[NDIM], Qcov[NDIM], ncov[NDIM], Qsq;
    int i;
    Bcon[0] = 0.;
    for (i = 1; i < 4; i++) Bcon[i] = U[B1 + i - 1];
    lower_g(Bcon, gcov, Bcov);
    for (i = 0; i < 4; i++) Qcov[i] = U[1 + i];
    raise_g(Qcov, gcon, Qcon);
    s->Bsq = 0.;
    for (i = 1; i < 4; i++) s->Bsq += Bcon[i] * Bcov[i];
    QdotB = 0.;
    for (i = 0; i < 4; i++) QdotB += Qcov[i] * Bcon[i];
    s->QdotBsq = QdotB * QdotB;
    ncov_calc(gcon, ncov);
    raise_g(ncov, gcon, ncon);
    s->Qdotn = Qcon[0] * ncov[0];
    Qsq = 0.;
    for (i = 0; i < 4; i++) Qsq += Qcov[i] * Qcon[i];
    s->Qtsq = Qsq + s->Qdotn * s->Qdotn;
    s->D = U[RHO];
    *QdotB_out = QdotB;
}

/* utoprim_2d.c:219-362 */
static int new_body_2d(u2p_t *s, const double U[NPR], double gcov[NDIM][NDIM],
                       double gcon[NDIM][NDIM], double prim[NPR], int *n_iter, int *k_conv, double *errx_first)
{
    double x_2d[2], QdotB, Bcon[NDIM], Qcon[NDIM], ncon[NDIM], Qtcon[NDIM];
    double rho0, u, p, w, gammasq, gamma, gtmp, W_last, W, utsq, vsq;
    int i, j, retval, i_increase;
    const double GAMMA = s->gam;

    for (i = B1; i <= B3; i++) prim[i] = U[i];
    u2p_scalars(s, U, gcov, gcon, Bcon, Qcon, ncon, &QdotB);

    utsq = 0.;
    for (i = 1; i < 4; i++)
        for (j = 1; j < 4; j++) utsq += gcov[i][j] * prim[U1 + i - 1] * prim[U1 + j - 1];
    if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
    if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;

    gammasq = 1. + utsq;
    gamma = sqrt(gammasq);
    rho0 = s->D / gamma;
    u = prim[UU];
    p = (GAMMA - 1.) * u;
    w = rho0 + u + p;
    W_last = w * gammasq;

    i_increase = 0;
    while (((W_last * W_last * W_last * (W_last + 2. * s->Bsq) - s->QdotBsq * (2. * W_last + s->Bsq)) <=
            W_last * W_last * (s->Qtsq - s->Bsq * s->Bsq)) &&
           (i_increase < 10)) {
        W_last *= 10.;
        i_increase++;
    }
    x_2d[0] = fabs(W_last);
    x_2d[1] = x1_of_x0(s, W_last);
    retval = gnr_2d_first(s, x_2d, n_iter, k_conv, errx_first);
    W = x_2d[0];
    vsq = x_2d[1];
    if ((retval != 0) || (W == FAIL_VAL)) return retval * 100 + 1;
    if (W <= 0. || W > W_TOO_BIG) return 3;
    if (vsq >= 1.) return 4;

    gtmp = sqrt(1. - vsq);
    gamma = 1. / gtmp;
    rho0 = s->D * gtmp;
    w = W * (1. - vsq);
    p = (GAMMA - 1.) * (w - rho0) / GAMMA;
    u = w - (rho0 + p);
    if ((rho0 <= 0.) || (u <= 0.)) return 5;

    prim[RHO] = rho0;
    prim[UU] = u;
    for (i = 1; i < 4; i++) Qtcon[i] = Qcon[i] + ncon[i] * s->Qdotn;
    for (i = 1; i < 4; i++) prim[U1 + i - 1] = gamma / (W + s->Bsq) * (Qtcon[i] + QdotB * Bcon[i] / W);
    for (i = B1; i <= B3; i++) prim[i] = U[i];
    return 0;
}

/* Driver shared by the three solvers: utoprim_2d.c:126-176,
 * utoprim_1dvsq2fix1.c:143-190, utoprim_1dfix1.c:137-200. */
static void u2p_transform(const double U[NPR], double gcon[NDIM][NDIM], double gdet,
                          const double prim[NPR], double U_tmp[NPR], double prim_tmp[NPR])
{
    int i;
    double alpha = 1.0 / sqrt(-gcon[0][0]);
    U_tmp[RHO] = alpha * U[RHO] / gdet;
    U_tmp[UU] = alpha * (U[UU] - U[RHO]) / gdet;
    for (i = U1; i <= U3; i++) U_tmp[i] = alpha * U[i] / gdet;
    for (i = B1; i <= B3; i++) U_tmp[i] = alpha * U[i] / gdet;
    for (i = 0; i < B1; i++) prim_tmp[i] = prim[i];
    for (i = B1; i <= B3; i++) prim_tmp[i] = alpha * prim[i];
}

static int Utoprim_2d_first(const double U[NPR], double gcov[NDIM][NDIM], double gcon[NDIM][NDIM], double gdet,
                            double prim[NPR], double gam, int *n_iter, int *k_conv, double *errx_first)
{
    double U_tmp[NPR], prim_tmp[NPR];
    int i, ret;
    u2p_t s;
    memset(&s, 0, sizeof(s));
    s.gam = gam;
    if (n_iter) *n_iter = 0;
    if (k_conv) *k_conv = 0;
    for (i = B1; i <= B3; i++) prim[i] = U[i] / gdet;
    if (U[0] <= 0.) return -100;
    u2p_transform(U, gcon, gdet, prim, U_tmp, prim_tmp);
    ret = new_body_2d(&s, U_tmp, gcov, gcon, prim_tmp, n_iter, k_conv, errx_first);
    if (ret == 0)
        for (i = 0; i < B1; i++) prim[i] = prim_tmp[i];
    return ret;
}
static int Utoprim_2d(const double U[NPR], double gcov[NDIM][NDIM], double gcon[NDIM][NDIM],
                      double gdet, double prim[NPR], double gam, int *n_iter, int *k_conv)
{
    return Utoprim_2d_first(U, gcov, gcon, gdet, prim, gam, n_iter, k_conv, NULL);
}

/* ---- Font fix #1: utoprim_1dvsq2fix1.c ---- */
/* utoprim_1dvsq2fix1.c:586-614 */
static double W_of_vsq_iso(const u2p_t *s, double vsq, double *p, double *rho, double *u)
{
    double gtmp = (1. - vsq);
    *rho = s->D * sqrt(gtmp);
    *p = s->K_atm * pow(*rho, s->gam);
    *u = *p / (s->gam - 1.);
    return ((*rho + *u + *p) / gtmp);
}
/* utoprim_1dvsq2fix1.c:619-624 */
static double dWdvsq_calc(const u2p_t *s, double vsq, double rho, double p)
{
    const double GAMMA = s->gam, G_ATM = s->gam;
    return ((GAMMA * (2. - G_ATM) * p + (GAMMA - 1.) * rho) / (2. * (GAMMA - 1.) * (1. - vsq) * (1. - vsq)));
}
/* utoprim_1dvsq2fix1.c:538-566 */
static void func_1d_gnr(const u2p_t *s, const double x[], double dx[], double resid[],
                        double jac[][1], double *f, double *df)
{
    double vsq, W, Wsq, W3, dWdvsq, fact_tmp, rho, p, u;
    vsq = x[0];
    W = W_of_vsq_iso(s, vsq, &p, &rho, &u);
    Wsq = W * W;
    W3 = W * Wsq;
    dWdvsq = dWdvsq_calc(s, vsq, rho, p);
    fact_tmp = (s->Bsq + W);
    resid[0] = s->Qtsq - vsq * fact_tmp * fact_tmp + s->QdotBsq * (s->Bsq + 2. * W) / Wsq;
    jac[0][0] = -fact_tmp * (fact_tmp + 2. * dWdvsq * (vsq + s->QdotBsq / W3));
    dx[0] = -resid[0] / jac[0][0];
    *f = 0.5 * resid[0] * resid[0];
    *df = -2. * (*f);
}
/* utoprim_1dvsq2fix1.c:372-384 */
static void validate_x_1d(double x[1], const double x0[1])
{
    double small = 1.e-10;
    x[0] = (x[0] >= 1.0) ? (0.5 * (x0[0] + 1.)) : x[0];
    x[0] = (x[0] < -small) ? (0.5 * x0[0]) : x[0];
    x[0] = fabs(x[0]);
}
/* utoprim_1dvsq2fix1.c:398-519 */
static int gnr_1dvsq(const u2p_t *s, double x[1])
{
    double f, df, dx[1], x_old[1], resid[1], jac[1][1], errx, W, W_old, rho, p, u;
    int n_iter, i_extra, doing_extra, keep_iterating;
    errx = 1.;
    df = f = 1.;
    i_extra = doing_extra = 0;
    x_old[0] = x[0];
    W = W_old = 0.;
    n_iter = 0;
    keep_iterating = 1;
    while (keep_iterating) {
        func_1d_gnr(s, x, dx, resid, jac, &f, &df);
        errx = 0.;
        x_old[0] = x[0];
        x[0] += dx[0];
        validate_x_1d(x, x_old);
        W_old = W;
        W = W_of_vsq_iso(s, x[0], &p, &rho, &u);
        errx = (W == 0.) ? fabs(W - W_old) : fabs((W - W_old) / W);
        errx += (x[0] == 0.) ? fabs(x[0] - x_old[0]) : fabs((x[0] - x_old[0]) / x[0]);
        if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
        if (doing_extra == 1) i_extra++;
        if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
            (n_iter >= (MAX_NEWT_ITER - 1)))
            keep_iterating = 0;
        n_iter++;
    }
    if ((isfinite(f) == 0) || (isfinite(df) == 0)) return 2;
    if (fabs(errx) > MIN_NEWT_TOL) return 1;
    return 0;
}
/* utoprim_1dvsq2fix1.c:242-370 */
static int new_body_1dvsq(u2p_t *s, const double U[NPR], double gcov[NDIM][NDIM],
                          double gcon[NDIM][NDIM], double prim[NPR])
{
    double x_1d[1], QdotB, Bcon[NDIM], Qcon[NDIM], ncon[NDIM], Qtcon[NDIM];
    double rho0, u, p, gammasq, gamma, gtmp, W, utsq, vsq;
    int i, j, retval;

    for (i = B1; i <= B3; i++) prim[i] = U[i];
    u2p_scalars(s, U, gcov, gcon, Bcon, Qcon, ncon, &QdotB);
    utsq = 0.;
    for (i = 1; i < 4; i++)
        for (j = 1; j < 4; j++) utsq += gcov[i][j] * prim[U1 + i - 1] * prim[U1 + j - 1];
    if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
    if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
    gammasq = 1. + utsq;
    x_1d[0] = 1. - 1. / gammasq;
    retval = gnr_1dvsq(s, x_1d);
    if (retval != 0) return retval * 100 + 1;
    vsq = x_1d[0];
    if ((vsq >= 1.) || (vsq < 0.)) return 4;
    W = W_of_vsq_iso(s, vsq, &p, &rho0, &u);
    gtmp = sqrt(1. - vsq);
    gamma = 1. / gtmp;
    if ((rho0 <= 0.) || (u <= 0.)) return 5;
    prim[RHO] = rho0;
    prim[UU] = u;
    for (i = 1; i < 4; i++) Qtcon[i] = Qcon[i] + ncon[i] * s->Qdotn;
    for (i = 1; i < 4; i++) prim[U1 + i - 1] = gamma / (W + s->Bsq) * (Qtcon[i] + QdotB * Bcon[i] / W);
    for (i = B1; i <= B3; i++) prim[i] = U[i];
    return 0;
}
static int Utoprim_1dvsq2fix1(const double U[NPR], double gcov[NDIM][NDIM], double gcon[NDIM][NDIM],
                              double gdet, double prim[NPR], double K, double gam)
{
    double U_tmp[NPR], prim_tmp[NPR];
    int i, ret;
    u2p_t s;
    memset(&s, 0, sizeof(s));
    s.gam = gam;
    for (i = B1; i <= B3; i++) prim[i] = U[i] / gdet;
    if (U[0] <= 0.) return -100;
    s.K_atm = K;
    u2p_transform(U, gcon, gdet, prim, U_tmp, prim_tmp);
    ret = new_body_1dvsq(&s, U_tmp, gcov, gcon, prim_tmp);
    if (ret == 0)
        for (i = 0; i < B1; i++) prim[i] = prim_tmp[i];
    return ret;
}

/* ---- Font fix #2: utoprim_1dfix1.c ---- */
/* utoprim_1dfix1.c:814-838 */
static void func_gnr2_rho(const u2p_t *s, const double x[], double dx[], double resid[],
                          double jac[][1], double *f, double *df)
{
    double A, B, C, rho, W, B0;
    const double GAMMA = s->gam;
    A = s->D * s->D;
    B0 = A * GAMMA * s->K_atm;
    B = B0 / (GAMMA - 1.);
    rho = x[0];
    W = s->W_for_gnr2;
    C = pow(rho, GAMMA - 1.);
    resid[0] = rho * W - A - B * C;
    jac[0][0] = W - B0 * C / rho;
    dx[0] = -resid[0] / jac[0][0];
    *f = 0.5 * resid[0] * resid[0];
    *df = -2. * (*f);
}
/* utoprim_1dfix1.c:572-660 (the tolerance test never sets doing_extra) */
static int gnr2(const u2p_t *s, double x[1])
{
    double f, df, dx[1], x_old[1], resid[1], jac[1][1], errx;
    int n_iter, i_extra, doing_extra, keep_iterating;
    errx = 1.;
    df = f = 1.;
    i_extra = doing_extra = 0;
    x_old[0] = x[0];
    n_iter = 0;
    keep_iterating = 1;
    while (keep_iterating) {
        func_gnr2_rho(s, x, dx, resid, jac, &f, &df);
        errx = 0.;
        x_old[0] = x[0];
        x[0] += dx[0];
        errx += (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0] / x[0]);
        errx /= 1. * 1;
        x[0] = fabs(x[0]);
        if ((fabs(errx) <= NEWT_TOL2) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 0;
        if (doing_extra == 1) i_extra++;
        if (((fabs(errx) <= NEWT_TOL2) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
            (n_iter >= (MAX_NEWT_ITER - 1)))
            keep_iterating = 0;
        n_iter++;
    }
    if ((isfinite(f) == 0) || (isfinite(df) == 0) || (isfinite(x[0]) == 0)) return 2;
    if (fabs(errx) > MIN_NEWT_TOL) return 1;
    return 0;
}
/* utoprim_1dfix1.c:695-765 (func_1d_orig1, USE_ISENTROPIC) */
static void func_1d_orig1(u2p_t *s, const double x[], double dx[], double resid[],
                          double jac[][1], double *f, double *df)
{
    int retval, ntries;
    double Dc, t1, t10, t2, t21, t23, t26, t29, t3, t30;
    double t32, t33, t34, t38, t5, t51, t67, t8, W, x_rho[1], rho, rho_g;
    const double GAMMA = s->gam;
    const double Bsq = s->Bsq, QdotBsq = s->QdotBsq, Qtsq = s->Qtsq, K_atm = s->K_atm;

    W = x[0];
    s->W_for_gnr2_old = s->W_for_gnr2;
    s->W_for_gnr2 = W;
    rho_g = x_rho[0] = s->rho_for_gnr2;
    ntries = 0;
    while ((retval = gnr2(s, x_rho)) && (ntries++ < 10)) {
        rho_g *= 10.;
        x_rho[0] = rho_g;
    }
    s->rho_for_gnr2_old = s->rho_for_gnr2;
    rho = s->rho_for_gnr2 = x_rho[0];
    Dc = s->D;
    t1 = Dc * Dc;
    t2 = QdotBsq * t1;
    t3 = t2 * Bsq;
    t5 = Bsq * Bsq;
    t8 = t1 * Bsq;
    t10 = t1 * W;
    t21 = W * W;
    t23 = rho * rho;
    t26 = 1. / t1;
    resid[0] = (t3 + (2.0 * t2 + ((Qtsq - t5) * t1 + (-2.0 * t8 - t10) * W) * W) * W +
                (t5 + (2.0 * Bsq + W) * W) * t21 * t23) *
               t26 / t21;
    t29 = t1 * t1;
    t30 = QdotBsq * t29;
    t32 = GAMMA * K_atm;
    t33 = pow(rho, 1.0 * GAMMA);
    t34 = t32 * t33;
    t38 = t23 * t33;
    t51 = GAMMA * t1 * K_atm * t33;
    t67 = t21 * W;
    jac[0][0] = -2.0 *
                (t30 * Bsq * t34 +
                 (t30 * t34 +
                  ((-t38 * Bsq * t32 + Bsq * GAMMA * t1 * K_atm * t33) * t1 +
                   (-t38 * GAMMA * K_atm + t51) * t1 * W) *
                      t21) *
                     W +
                 ((-t3 + (-t2 + (-t8 - t10) * t21) * W) * W + (-t5 - Bsq * W) * t67 * t23) * t23) *
                t26 / (t51 - W * t23) / t67;
    dx[0] = -resid[0] / jac[0][0];
    *f = 0.5 * resid[0] * resid[0];
    *df = -2. * (*f);
}
/* utoprim_1dfix1.c:470-560 */
static int gnr_1dfix(u2p_t *s, double x[1])
{
    double f, df, dx[1], x_old[1], resid[1], jac[1][1], errx;
    int n_iter, i_extra, doing_extra, keep_iterating, i_increase;
    errx = 1.;
    df = f = 1.;
    i_extra = doing_extra = 0;
    x_old[0] = x[0];
    n_iter = 0;
    keep_iterating = 1;
    while (keep_iterating) {
        func_1d_orig1(s, x, dx, resid, jac, &f, &df);
        errx = 0.;
        x_old[0] = x[0];
        x[0] += dx[0];
        i_increase = 0;
        while (((x[0] * x[0] * x[0] * (x[0] + 2. * s->Bsq) - s->QdotBsq * (2. * x[0] + s->Bsq)) <=
                x[0] * x[0] * (s->Qtsq - s->Bsq * s->Bsq)) &&
               (i_increase < 10)) {
            x[0] -= (1. * i_increase) * dx[0] / 10.;
            i_increase++;
        }
        errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0] / x[0]);
        x[0] = fabs(x[0]);
        if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
        if (doing_extra == 1) i_extra++;
        if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
            (n_iter >= (MAX_NEWT_ITER - 1)))
            keep_iterating = 0;
        n_iter++;
    }
    if ((isfinite(f) == 0) || (isfinite(df) == 0) || (isfinite(x[0]) == 0)) return 2;
    if (fabs(errx) > MIN_NEWT_TOL) return 1;
    return 0;
}
/* utoprim_1dfix1.c:254-406 */
static int new_body_1dfix(u2p_t *s, const double U[NPR], double gcov[NDIM][NDIM],
                          double gcon[NDIM][NDIM], double prim[NPR])
{
    double x_1d[1], QdotB, Bcon[NDIM], Qcon[NDIM], ncon[NDIM], Qtcon[NDIM];
    double rho0, u, p, w, gammasq, gamma, gtmp, W_last, W, utsq, vsq;
    int i, j, retval, i_increase;
    const double GAMMA = s->gam;

    for (i = B1; i <= B3; i++) prim[i] = U[i];
    u2p_scalars(s, U, gcov, gcon, Bcon, Qcon, ncon, &QdotB);
    utsq = 0.;
    for (i = 1; i < 4; i++)
        for (j = 1; j < 4; j++) utsq += gcov[i][j] * prim[U1 + i - 1] * prim[U1 + j - 1];
    if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
    if (utsq < 0. || utsq > UTSQ_TOO_BIG) return 2;
    gammasq = 1. + utsq;
    gamma = sqrt(gammasq);
    rho0 = s->D / gamma;
    p = s->K_atm * pow(rho0, GAMMA);
    u = p / (GAMMA - 1.);
    w = rho0 + u + p;
    W_last = w * gammasq;
    i_increase = 0;
    while (((W_last * W_last * W_last * (W_last + 2. * s->Bsq) - s->QdotBsq * (2. * W_last + s->Bsq)) <=
            W_last * W_last * (s->Qtsq - s->Bsq * s->Bsq)) &&
           (i_increase < 10)) {
        W_last *= 10.;
        i_increase++;
    }
    s->W_for_gnr2 = s->W_for_gnr2_old = W_last;
    s->rho_for_gnr2 = s->rho_for_gnr2_old = rho0;
    x_1d[0] = W_last;
    retval = gnr_1dfix(s, x_1d);
    W = x_1d[0];
    if ((retval != 0) || (W == FAIL_VAL)) return retval * 100 + 1;
    if (W <= 0. || W > W_TOO_BIG) return 3;
    vsq = vsq_calc(s, W);
    if (vsq >= 1.) return 4;
    gtmp = sqrt(1. - vsq);
    gamma = 1. / gtmp;
    rho0 = s->D * gtmp;
    w = W * (1. - vsq);
    p = s->K_atm * pow(rho0, GAMMA);
    u = p / (GAMMA - 1.);
    if ((rho0 <= 0.) || (u <= 0.)) return 5;
    prim[RHO] = rho0;
    prim[UU] = u;
    for (i = 1; i < 4; i++) Qtcon[i] = Qcon[i] + ncon[i] * s->Qdotn;
    for (i = 1; i < 4; i++) prim[U1 + i - 1] = gamma / (W + s->Bsq) * (Qtcon[i] + QdotB * Bcon[i] / W);
    for (i = B1; i <= B3; i++) prim[i] = U[i];
    (void)w;
    return 0;
}
static int Utoprim_1dfix1(const double U[NPR], double gcov[NDIM][NDIM], double gcon[NDIM][NDIM],
                          double gdet, double prim[NPR], double K, double gam)
{
    double U_tmp[NPR], prim_tmp[NPR];
    int i, ret;
    u2p_t s;
    memset(&s, 0, sizeof(s));
    s.gam = gam;
    for (i = B1; i <= B3; i++) prim[i] = U[i] / gdet;
    if (U[0] <= 0.) return -100;
    s.K_atm = K;
    u2p_transform(U, gcon, gdet, prim, U_tmp, prim_tmp);
    ret = new_body_1dfix(&s, U_tmp, gcov, gcon, prim_tmp);
    if (ret == 0)
        for (i = 0; i < B1; i++) prim[i] = prim_tmp[i];
    return ret;
}

/* ------------------------------------------------------------------------ */
/* fixup.c                                                                   */
/* ------------------------------------------------------------------------ */
/* phys.c:255-271 */
static void ucon_to_utcon(const double *ucon, const geom_t *geom, double *utcon)
{
    double alpha, beta[NDIM], gamma;
    int j;
    alpha = 1. / sqrt(-geom->gcon[TT][TT]);
    for (j = 1; j < NDIM; j++) beta[j] = geom->gcon[TT][j] * alpha * alpha;
    gamma = alpha * ucon[TT];
    utcon[0] = 0;
    for (j = 1; j < NDIM; j++) utcon[j] = ucon[j] + gamma * beta[j] / alpha;
}
/* phys.c:273-297 */
static void ut_calc_3vel(const double *vcon, const geom_t *geom, double *ut)
{
    double AA, BB, CC, DD, one_over_alpha_sq;
    AA = geom->gcov[TT][TT];
    BB = 2. * (geom->gcov[TT][1] * vcon[1] + geom->gcov[TT][2] * vcon[2] + geom->gcov[TT][3] * vcon[3]);
    CC = geom->gcov[1][1] * vcon[1] * vcon[1] + geom->gcov[2][2] * vcon[2] * vcon[2] +
         geom->gcov[3][3] * vcon[3] * vcon[3] +
         2. * (geom->gcov[1][2] * vcon[1] * vcon[2] + geom->gcov[1][3] * vcon[1] * vcon[3] +
               geom->gcov[2][3] * vcon[2] * vcon[3]);
    DD = -1. / (AA + BB + CC);
    one_over_alpha_sq = -geom->gcon[TT][TT];
    if (DD < one_over_alpha_sq) DD = one_over_alpha_sq;
    *ut = sqrt(DD);
}

/* fixup.c:117-140 */
static void get_rho_u_floor(double r, double gam, double *rho_floor, double *u_floor)
{
    double rhoflr, uuflr, uuscal, rhoscal;
    rhoscal = pow(r, -POWRHO);
    uuscal = pow(rhoscal, gam);
    rhoflr = RHOMIN * rhoscal;
    uuflr = UUMIN * uuscal;
    if (rhoflr < RHOMINLIMIT) rhoflr = RHOMINLIMIT;
    if (uuflr < UUMINLIMIT) uuflr = UUMINLIMIT;
    *rho_floor = rhoflr;
    *u_floor = uuflr;
}

/* fixup.c:142-383 (DRIFT_FLOOR=1); cnt may be NULL (no counting) */
static void fixup1zone(orc_ctx *o, int h, int i, int j, double pv[NPR], int64_t *cnt)
{
    const double gam = o->P.gam;
    double r, rhoflr, uuflr, bsq, f, gamma, trans;
    double pv_prefloor[NPR];
    geom_t geom;
    state_t q;
    int k, dofloor;
    double betapar, betasq, betasqmax, ucondr[NDIM], Bcov[NDIM], Bcon[NDIM];
    double udotB, Bsq, B, wold, QdotB, wnew, x, vpar, one_over_ucondr_t, vcon[NDIM], ut;
    double ucon[NDIM], utcon[NDIM];
    int64_t fatal = 0;

    for (k = 0; k < NPR; k++) pv_prefloor[k] = pv[k];
    r = o->coords[((size_t)i * o->SN2 + j) * NDIM + RR];
    get_rho_u_floor(r, gam, &rhoflr, &uuflr);
    get_geometry(o, i, j, CENT, &geom);
    get_state(pv, &geom, &q, &fatal);
    bsq = dot(q.bcon, q.bcov);
    if (rhoflr < bsq / BSQORHOMAX) rhoflr = bsq / BSQORHOMAX;
    if (uuflr < bsq / BSQOUMAX) uuflr = bsq / BSQOUMAX;
    if (rhoflr < pv[UU] / UORHOMAX) rhoflr = pv[UU] / UORHOMAX;

    dofloor = 0;
    if (pv[RHO] < rhoflr) {
        pv[RHO] = rhoflr;
        dofloor = 1;
    }
    if (pv[UU] < uuflr) {
        pv[UU] = uuflr;
        dofloor = 1;
    }
    if (dofloor && cnt) cnt[ORC_FLOOR_ZONES]++;

    if (dofloor && (trans = 10. * bsq / MY_MIN(pv[RHO], pv[UU]) - 1.) > 0.) {
        if (trans > 1.) trans = 1.;
        betapar = -q.bcon[0] / ((bsq + SMALL) * q.ucon[0]);
        betasq = betapar * betapar * bsq;
        betasqmax = 1. - 1. / (GAMMAMAX * GAMMAMAX);
        if (betasq > betasqmax) betasq = betasqmax;
        gamma = 1. / sqrt(1 - betasq);
        for (k = 0; k < NDIM; k++) ucondr[k] = gamma * (q.ucon[k] + betapar * q.bcon[k]);
        Bcon[0] = 0.;
        for (k = 1; k < NDIM; k++) Bcon[k] = pv[B1 - 1 + k];
        lower(Bcon, &geom, Bcov);
        udotB = dot(q.ucon, Bcov);
        Bsq = dot(Bcon, Bcov);
        B = sqrt(Bsq);
        wold = pv_prefloor[RHO] + pv_prefloor[UU] * gam;
        QdotB = udotB * wold * q.ucon[0];
        wnew = pv[RHO] + pv[UU] * gam;
        x = 2. * QdotB / (B * wnew * ucondr[0] + SMALL);
        vpar = x / (ucondr[0] * (1. + sqrt(1. + x * x)));
        one_over_ucondr_t = 1. / ucondr[0];
        vcon[0] = 1.;
        for (k = 1; k < NDIM; k++) vcon[k] = vpar * Bcon[k] / (B + SMALL) + ucondr[k] * one_over_ucondr_t;
        ut_calc_3vel(vcon, &geom, &ut);
        for (k = 0; k < NDIM; k++) ucon[k] = ut * vcon[k];
        ucon_to_utcon(ucon, &geom, utcon);
        for (k = 1; k < NDIM; k++) pv[k + UU] = utcon[k] * trans + pv_prefloor[k + UU] * (1. - trans);
    }

    if (gamma_calc(pv, &geom, &gamma)) {
        o->pflag[zi(o, h, i, j)] = -33;
        if (cnt) cnt[ORC_FIXUP_GAMMA]++;
    } else {
        if (gamma > GAMMAMAX) {
            f = sqrt((GAMMAMAX * GAMMAMAX - 1.) / (gamma * gamma - 1.));
            pv[U1] *= f;
            pv[U2] *= f;
            pv[U3] *= f;
            if (cnt) cnt[ORC_GAMMA_CAP]++;
        }
    }
    if (cnt) cnt[ORC_FATAL] += fatal;
}

/* fixup.c:51-60 */
static void fixup(orc_ctx *o, double *pv, int64_t *cnt)
{
    int64_t lc[ORC_NCOUNTERS];
    memset(lc, 0, sizeof(lc));
    const int N = o->N3 * (o->istop - o->istart + 1);
#pragma omp parallel for reduction(+ : lc[:ORC_NCOUNTERS]) schedule(static)
    for (int n = 0; n < N; n++) {
        int h = n / (o->istop - o->istart + 1);
        int i = o->istart + n % (o->istop - o->istart + 1);
        for (int j = o->jstart; j <= o->jstop; j++) fixup1zone(o, h, i, j, PV(o, pv, h, i, j), cnt ? lc : NULL);
    }
    if (cnt)
        for (int k = 0; k < ORC_NCOUNTERS; k++) cnt[k] += lc[k];
}

/* fixup.c:432-538 (sequential: the fallback branch reads repaired values) */
#define AVG4_1(pr, h, i, j, k) \
    (0.25 * (PV(o, pr, h, i, j + 1)[k] + PV(o, pr, h, i, j - 1)[k] + PV(o, pr, h, i - 1, j)[k] + PV(o, pr, h, i + 1, j)[k]))
#define AVG4_2(pr, h, i, j, k)                                                                          \
    (0.25 * (PV(o, pr, h, i + 1, MUT(9) ? j : j + 1)[k] + PV(o, pr, h, i + 1, j - 1)[k] +                   \
             PV(o, pr, h, i - 1, j + 1)[k] + PV(o, pr, h, i - 1, j - 1)[k]))
static void fixup_utoprim(orc_ctx *o, double *pv, int64_t *cnt)
{
    int h, i, j, k, pf[9];
    int8_t *pflag = o->pflag;
    for (h = 0; h < o->N3; h++)
        for (i = 0; i < o->SN1; i++)
            for (j = 0; j < o->SN2; j++) pflag[zi(o, h, i, j)] = !pflag[zi(o, h, i, j)];
    const int ni = o->istop - o->istart + 1, nj = o->jstop - o->jstart + 1;
    for (h = 0; h < o->N3; h++)
        for (int q = 0; q < ni * nj; q++) {
            /* ZLOOP order (h, i, j); mutant 8 walks (h, j, i) */
            i = o->istart + (MUT(8) ? q % ni : q / nj);
            j = o->jstart + (MUT(8) ? q / ni : q % nj);
            if (pflag[zi(o, h, i, j)] == 0) {
                double *z = PV(o, pv, h, i, j);
                pf[1] = pflag[zi(o, h, i - 1, j + 1)];
                pf[2] = pflag[zi(o, h, i, j + 1)];
                pf[3] = pflag[zi(o, h, i + 1, j + 1)];
                pf[8] = pflag[zi(o, h, i - 1, j)];
                pf[4] = pflag[zi(o, h, i + 1, j)];
                pf[7] = pflag[zi(o, h, i - 1, j - 1)];
                pf[6] = pflag[zi(o, h, i, j - 1)];
                pf[5] = pflag[zi(o, h, i + 1, j - 1)];
                if (pf[2] && pf[4] && pf[6] && pf[8]) {
                    for (k = 0; k < B1; k++) z[k] = AVG4_1(pv, h, i, j, k);
                } else if (pf[1] && pf[3] && pf[5] && pf[7]) {
                    for (k = 0; k < B1; k++) z[k] = AVG4_2(pv, h, i, j, k);
                } else {
                    for (k = RHO; k <= UU; k++) z[k] = 0.5 * (AVG4_1(pv, h, i, j, k) + AVG4_2(pv, h, i, j, k));
                    z[U1] = z[U2] = z[U3] = 0.;
                    cnt[ORC_FIXUP_NOSTENCIL]++;
                }
                pflag[zi(o, h, i, j)] = MUT(7) ? 1 : 0; /* fixup.c:530: stays "bad" */
                cnt[ORC_FIXUP_REPAIRS]++;
                fixup1zone(o, h, i, j, z, cnt);
            }
        }
}

/* ------------------------------------------------------------------------ */
/* bounds.c                                                                  */
/* ------------------------------------------------------------------------ */
/* bounds.c:333-378 */
static void inflow_check(orc_ctx *o, double *pr, int ii, int jj, int type, int64_t *fatal)
{
    geom_t geom;
    double ucon[NDIM], alpha, beta1, gamma, vsq;
    int j, k;
    get_geometry(o, ii, jj, CENT, &geom);
    ucon_calc(pr, &geom, ucon, fatal);
    if (((ucon[1] > 0.) && (type == 0)) || ((ucon[1] < 0.) && (type == 1))) {
        if (gamma_calc(pr, &geom, &gamma)) (*fatal)++;
        pr[U1] /= gamma;
        pr[U2] /= gamma;
        pr[U3] /= gamma;
        alpha = 1. / sqrt(-geom.gcon[0][0]);
        beta1 = geom.gcon[0][1] * alpha * alpha;
        pr[U1] = beta1 / alpha;
        vsq = 0.;
        for (j = 1; j < NDIM; j++)
            for (k = 1; k < NDIM; k++) vsq += geom.gcov[j][k] * pr[U1 + j - 1] * pr[U1 + k - 1];
        if (fabs(vsq) < 1.e-13) vsq = 1.e-13;
        if (vsq >= 1.) vsq = 1. - 1. / (GAMMAMAX * GAMMAMAX);
        gamma = 1. / sqrt(1. - vsq);
        pr[U1] *= gamma;
        pr[U2] *= gamma;
        pr[U3] *= gamma;
    }
}

/* bounds.c:54-331 with the polefix of bounds.c:381-436 */
static void bound_prim(orc_ctx *o, double *prim, int64_t *fatal)
{
    int h, i, j, k;
    int8_t *pflag = o->pflag;
    const int N3 = o->N3, is = o->istart, ie = o->istop, js = o->jstart, je = o->jstop;

    for (h = 0; h < N3; h++)
        for (j = js; j <= je; j++) {
            for (k = 0; k < NPR; k++) {
                PV(o, prim, h, is - 1, j)[k] = PV(o, prim, h, is, j)[k];
                PV(o, prim, h, is - 2, j)[k] = PV(o, prim, h, is, j)[k];
            }
            pflag[zi(o, h, is - 1, j)] = pflag[zi(o, h, is, j)];
            pflag[zi(o, h, is - 2, j)] = pflag[zi(o, h, is, j)];
        }
    for (h = 0; h < N3; h++)
        for (j = js; j <= je; j++) {
            for (k = 0; k < NPR; k++) {
                PV(o, prim, h, ie + 1, j)[k] = PV(o, prim, h, ie, j)[k];
                PV(o, prim, h, ie + 2, j)[k] = PV(o, prim, h, ie, j)[k];
            }
            pflag[zi(o, h, ie + 1, j)] = pflag[zi(o, h, ie, j)];
            pflag[zi(o, h, ie + 2, j)] = pflag[zi(o, h, ie, j)];
        }

    if (o->polefix) {
        const int jref = o->polefix;
        for (h = 0; h < N3; h++)
            for (i = is - 2; i <= ie + 2; i++)
                for (j = 0; j < jref; j++)
                    for (k = 0; k < NPR; k++) {
                        if (k == B1 || k == B2 || (N3 > 1 && k == B3)) continue;
                        const int jr = MUT(4) ? jref - 1 : jref;
                        if (k == U2) {
                            const double w = (j + 0.5) / (MUT(3) ? jref + 1. : jref + 0.5);
                            PV(o, prim, h, i, js + j)[k] = w * PV(o, prim, h, i, js + jr)[k];
                            PV(o, prim, h, i, je - j)[k] = w * PV(o, prim, h, i, je - jr)[k];
                        } else {
                            PV(o, prim, h, i, js + j)[k] = PV(o, prim, h, i, js + jr)[k];
                            PV(o, prim, h, i, je - j)[k] = PV(o, prim, h, i, je - jr)[k];
                        }
                    }
    }

    if (N3 > 1) {
        /* bounds.c:161-275: theta ghosts from the phi+pi plane */
        for (h = 0; h < N3; h++) {
            const int hop = MUT(1) ? (h + N3 / 2 + 1) % N3 : MUT(2) ? h : (h + N3 / 2) % N3;
            for (i = is - 2; i <= ie + 2; i++) {
                for (k = 0; k < NPR; k++) {
                    PV(o, prim, h, i, js - 1)[k] = PV(o, prim, hop, i, js)[k];
                    PV(o, prim, h, i, js - 2)[k] = PV(o, prim, hop, i, js + 1)[k];
                    PV(o, prim, h, i, je + 1)[k] = PV(o, prim, hop, i, je)[k];
                    PV(o, prim, h, i, je + 2)[k] = PV(o, prim, hop, i, je - 1)[k];
                }
                pflag[zi(o, h, i, js - 1)] = pflag[zi(o, h, i, js)];
                pflag[zi(o, h, i, js - 2)] = pflag[zi(o, h, i, js + 1)];
                pflag[zi(o, h, i, je + 1)] = pflag[zi(o, h, i, je)];
                pflag[zi(o, h, i, je + 2)] = pflag[zi(o, h, i, je - 1)];
            }
        }
    } else {
        for (h = 0; h < N3; h++)
            for (i = is - 2; i <= ie + 2; i++) {
                for (k = 0; k < NPR; k++) {
                    PV(o, prim, h, i, js - 1)[k] = PV(o, prim, h, i, js)[k];
                    PV(o, prim, h, i, js - 2)[k] = PV(o, prim, h, i, js + 1)[k];
                }
                pflag[zi(o, h, i, js - 1)] = pflag[zi(o, h, i, js)];
                pflag[zi(o, h, i, js - 2)] = pflag[zi(o, h, i, js + 1)];
            }
        for (h = 0; h < N3; h++)
            for (i = is - 2; i <= ie + 2; i++) {
                for (k = 0; k < NPR; k++) {
                    PV(o, prim, h, i, je + 1)[k] = PV(o, prim, h, i, je)[k];
                    PV(o, prim, h, i, je + 2)[k] = PV(o, prim, h, i, je - 1)[k];
                }
                pflag[zi(o, h, i, je + 1)] = pflag[zi(o, h, i, je)];
                pflag[zi(o, h, i, je + 2)] = pflag[zi(o, h, i, je - 1)];
            }
    }
    for (h = 0; h < N3; h++)
        for (i = is - 2; i <= ie + 2; i++) {
            for (j = js - 2; j < js; j++) {
                PV(o, prim, h, i, j)[U2] *= -1.;
                PV(o, prim, h, i, j)[B2] *= -1.;
            }
            for (j = je + 1; j <= je + 2; j++) {
                PV(o, prim, h, i, j)[U2] *= -1.;
                PV(o, prim, h, i, j)[B2] *= -1.;
            }
        }
    for (h = 0; h < N3; h++)
        for (i = is - 2; i <= is - 1; i++)
            for (j = js - 2; j <= je + 2; j++) inflow_check(o, PV(o, prim, h, i, j), i, j, 0, fatal);
    for (h = 0; h < N3; h++)
        for (i = ie + 1; i <= ie + 2; i++)
            for (j = js - 2; j <= je + 2; j++) inflow_check(o, PV(o, prim, h, i, j), i, j, 1, fatal);
}

/* ------------------------------------------------------------------------ */
/* step_ch.c                                                                 */
/* ------------------------------------------------------------------------ */
/* One face of step_ch.c:429-475 given the reconstructed states. */
static void face_flux(const double *p_l, const double *p_r, const geom_t *geom, int dir, double gam,
                      double hllf, double laxf, double *F, double *ctop_out, int64_t *fatal)
{
    double F_l[NPR], F_r[NPR], U_l[NPR], U_r[NPR];
    double cmax_l, cmax_r, cmin_l, cmin_r, cmax, cmin, ctop;
    state_t state_l, state_r;
    int k;
    get_state(p_l, geom, &state_l, fatal);
    get_state(p_r, geom, &state_r, fatal);
    primtoflux(p_l, &state_l, dir, geom, gam, F_l);
    primtoflux(p_l, &state_l, TT, geom, gam, U_l);
    *fatal += vchar(p_l, &state_l, geom, dir, gam, &cmax_l, &cmin_l);
    primtoflux(p_r, &state_r, dir, geom, gam, F_r);
    primtoflux(p_r, &state_r, TT, geom, gam, U_r);
    *fatal += vchar(p_r, &state_r, geom, dir, gam, &cmax_r, &cmin_r);
    cmax = fabs(MY_MAX(MY_MAX(0., cmax_l), cmax_r));
    cmin = fabs(MY_MAX(MY_MAX(0., -cmin_l), -cmin_r));
    ctop = MY_MAX(cmax, cmin);
    for (k = 0; k < NPR; k++)
        F[k] = hllf * ((cmax * F_l[k] + cmin * F_r[k] - cmax * cmin * (U_r[k] - U_l[k])) / (cmax + cmin + SMALL)) +
               laxf * (0.5 * (F_l[k] + F_r[k] - ctop * (U_r[k] - U_l[k])));
    *ctop_out = ctop;
}

/* PPM needs zones m-2 .. m+2 along dir; the grid holds 2 ghost layers in r
 * and theta (phi is periodic). */
static int ppm_zone(const orc_ctx *o, int dir, int i, int j)
{
    if (dir == 1) return i - 2 >= o->istart - 2 && i + 2 <= o->istop + 2 && !(MUT(15) && i == o->istart);
    if (dir == 2) return j - 2 >= o->jstart - 2 && j + 2 <= o->jstop + 2;
    return 1;
}

/* PPM face values of zone (h, i, j) along the direction (hdel, idel, jdel)
 * from pr and the MC slopes in o->dq: left into l, right into r (each may be
 * NULL). */
static void ppm_face(const orc_ctx *o, const double *pr, int h, int i, int j, int hdel, int idel, int jdel, double *l,
                     double *r)
{
    const int hm = hwrap(o, h - hdel), hp = hwrap(o, h + hdel);
    for (int k = 0; k < NPR; k++) {
        double ql, qr;
        para(PV(o, pr, hm, i - idel, j - jdel)[k], PV(o, pr, h, i, j)[k], PV(o, pr, hp, i + idel, j + jdel)[k],
             PV(o, o->dq, hm, i - idel, j - jdel)[k], PV(o, o->dq, h, i, j)[k], PV(o, o->dq, hp, i + idel, j + jdel)[k],
             &ql, &qr);
        if (l) l[k] = ql;
        if (r) r[k] = qr;
    }
}

/* step_ch.c:331-504 */
static double fluxcalc(orc_ctx *o, const double *pr, double *F, int dir, int fullstep, int64_t *cnt)
{
    int hdel, idel, jdel, face;
    const int N3 = o->N3;
    if (dir == 1) { hdel = 0; idel = 1; jdel = 0; face = FACE1; }
    else if (dir == 2) { hdel = 0; idel = 0; jdel = 1; face = FACE2; }
    else { hdel = 1; idel = 0; jdel = 0; face = FACE3; }

    const int ni = o->istop - o->istart + 3;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N3 * ni; n++) {
        int h = n / ni, i = o->istart - 1 + n % ni;
        for (int j = o->jstart - 1; j <= o->jstop + 1; j++)
            for (int k = 0; k < NPR; k++)
                PV(o, o->dq, h, i, j)[k] =
                    slope_lim(o->P.lim == PPM ? MC : o->P.lim, PV(o, pr, hwrap(o, h - hdel), i - idel, j - jdel)[k], PV(o, pr, h, i, j)[k],
                              PV(o, pr, (MUT(11) && h + hdel >= N3) ? h : hwrap(o, h + hdel), i + idel, j + jdel)[k]);
    }

    double ndt = 1.e9;
    int64_t fatal = 0;
    const int i0 = o->istart - (hdel + jdel), i1 = o->istop + 1;
    const int j0 = o->jstart - (hdel + idel), j1 = o->jstop + 1;
    const int nI = i1 - i0 + 1;
#pragma omp parallel for reduction(min : ndt) reduction(+ : fatal) schedule(static)
    for (int n = 0; n < N3 * nI; n++) {
        int h = n / nI, i = i0 + n % nI;
        for (int j = j0; j <= j1; j++) {
            double *Fz = PV(o, F, h, i, j);
            if (dir == 2 && (j == o->jstart || j == o->jstop + 1)) {
                for (int k = 0; k < NPR; k++) Fz[k] = 0.;
                continue;
            }
            double p_l[NPR], p_r[NPR], ctop;
            geom_t geom;
            const int hm = hwrap(o, h - hdel);
            for (int k = 0; k < NPR; k++) {
                p_l[k] = PV(o, pr, hm, i - idel, j - jdel)[k] + 0.5 * PV(o, o->dq, hm, i - idel, j - jdel)[k];
                p_r[k] = PV(o, pr, h, i, j)[k] - 0.5 * PV(o, o->dq, h, i, j)[k];
            }
            if (o->P.lim == PPM) {
                /* the zone left of the face gives its right value, the zone right
                 * of it its left value; a zone whose 5-zone stencil leaves the
                 * grid (the 2 ghost layers) keeps the linear MC state */
                if (ppm_zone(o, dir, i - idel, j - jdel)) ppm_face(o, pr, hm, i - idel, j - jdel, hdel, idel, jdel, NULL, p_l);
                if (ppm_zone(o, dir, i, j)) ppm_face(o, pr, h, i, j, hdel, idel, jdel, p_r, NULL);
            }
            get_geometry(o, i, j, face, &geom);
            if (g_step_ceos) ceos_init(o, h, MUT(17) ? i - idel : i, j); /* step_ch.c:439-441 */
            face_flux(p_l, p_r, &geom, dir, o->P.gam, o->P.hllf, o->P.laxf, Fz, &ctop, &fatal);
            if (fullstep && i >= o->istart && i <= o->istop && j >= o->jstart && j <= o->jstop) {
                double dtij = o->P.cour * o->dx[dir] / ctop;
                if (dtij < ndt) ndt = dtij;
            }
        }
    }
    cnt[ORC_FATAL] += fatal;
    return ndt;
}

/* coord.c:264-324 */
static void fix_flux(orc_ctx *o)
{
    int h, i, j, k;
    const int is = o->istart, ie = o->istop, js = o->jstart, je = o->jstop;
    for (h = 0; h < o->N3; h++)
        for (i = is - 1; i <= ie + 1; i++) {
            PV(o, o->F1, h, i, js - 1)[B2] = -PV(o, o->F1, h, i, js)[B2];
            if (o->N3 > 1) PV(o, o->F3, h, i, js - 1)[B2] = -PV(o, o->F3, h, i, js)[B2];
            for (k = 0; k < NPR; k++) PV(o, o->F2, h, i, js)[k] = 0.;
        }
    for (h = 0; h < o->N3; h++)
        for (i = is - 1; i <= ie + 1; i++) {
            PV(o, o->F1, h, i, je + 1)[B2] = -PV(o, o->F1, h, i, je)[B2];
            if (o->N3 > 1) PV(o, o->F3, h, i, je + 1)[B2] = -PV(o, o->F3, h, i, je)[B2];
            for (k = 0; k < NPR; k++) PV(o, o->F2, h, i, je + 1)[k] = 0.;
        }
    for (h = 0; h < o->N3; h++)
        for (j = js - 2; j <= je + 2; j++)
            if (PV(o, o->F1, h, is, j)[RHO] > 0) PV(o, o->F1, h, is, j)[RHO] = 0;
    for (h = 0; h < o->N3; h++)
        for (j = js - 2; j <= je + 2; j++)
            if (PV(o, o->F1, h, ie + 1, j)[RHO] < 0) PV(o, o->F1, h, ie + 1, j)[RHO] = 0;
}

/* step_ch.c:514-629 (N3 <= 2) and step_ch.c:726-798 (N3 > 2) */
static void flux_ct(orc_ctx *o)
{
    int h, i, j;
    const int is = o->istart, ie = o->istop, js = o->jstart, je = o->jstop, N3 = o->N3;
    double *F1 = o->F1, *F2 = o->F2, *F3 = o->F3;
#define EM(e, h, i, j) (e)[zi(o, h, i, j)]
    if (N3 <= 2) {
        double *emf = o->e3;
        for (h = 0; h < N3; h++)
            for (i = is; i <= ie + 1; i++)
                for (j = js; j <= je + 1; j++)
                    EM(emf, h, i, j) = 0.25 * (PV(o, F1, h, i, j)[B2] + PV(o, F1, h, i, j - 1)[B2] -
                                               PV(o, F2, h, i, j)[B1] - PV(o, F2, h, i - 1, j)[B1]);
        for (h = 0; h < N3; h++)
            for (i = is; i <= ie + 1; i++)
                for (j = js; j <= je; j++) {
                    PV(o, F1, h, i, j)[B1] = 0.;
                    PV(o, F1, h, i, j)[B2] = 0.5 * (EM(emf, h, i, j) + EM(emf, h, i, j + 1));
                }
        for (h = 0; h < N3; h++)
            for (i = is; i <= ie; i++)
                for (j = js; j <= je + 1; j++) {
                    PV(o, F2, h, i, j)[B1] = -0.5 * (EM(emf, h, i, j) + EM(emf, h, i + 1, j));
                    PV(o, F2, h, i, j)[B2] = 0.;
                }
        return;
    }
    double *e1 = o->e1, *e2 = o->e2, *e3 = o->e3;
    for (h = 0; h < N3; h++) {
        const int hm = MUT(10) ? (h > 0 ? h - 1 : 0) : hwrap(o, h - 1);
        const int hm1 = MUT(5) ? hwrap(o, h + 1) : hm;
        for (i = is; i <= ie + 1; i++)
            for (j = js; j <= je + 1; j++) {
                EM(e1, h, i, j) = 0.25 * ((PV(o, F2, h, i, j)[B3] + PV(o, F2, hm1, i, j)[B3]) -
                                          (PV(o, F3, h, i, j)[B2] + PV(o, F3, h, i, j - 1)[B2]));
                EM(e2, h, i, j) = 0.25 * ((PV(o, F3, h, i, j)[B1] + PV(o, F3, h, i - 1, j)[B1]) -
                                          (PV(o, F1, h, i, j)[B3] + PV(o, F1, hm, i, j)[B3]));
                EM(e3, h, i, j) = 0.25 * ((PV(o, F1, h, i, j)[B2] + PV(o, F1, h, i, j - 1)[B2]) -
                                          (PV(o, F2, h, i, j)[B1] + PV(o, F2, h, i - 1, j)[B1]));
            }
    }
    for (h = 0; h < N3; h++) {
        const int hp = hwrap(o, h + 1);
        for (i = is; i <= ie + 1; i++)
            for (j = js; j <= je; j++) {
                PV(o, F1, h, i, j)[B1] = 0.;
                PV(o, F1, h, i, j)[B2] = 0.5 * (EM(e3, h, i, j) + EM(e3, h, i, j + 1));
                PV(o, F1, h, i, j)[B3] = -0.5 * (EM(e2, h, i, j) + EM(e2, hp, i, j));
            }
        for (i = is; i <= ie; i++)
            for (j = js; j <= je + 1; j++) {
                PV(o, F2, h, i, j)[B1] = -0.5 * (EM(e3, h, i, j) + EM(e3, h, i + 1, j));
                PV(o, F2, h, i, j)[B2] = 0.;
                PV(o, F2, h, i, j)[B3] = 0.5 * (EM(e1, h, i, j) + EM(e1, hp, i, j));
            }
        for (i = is; i <= ie; i++)
            for (j = js; j <= je; j++) {
                PV(o, F3, h, i, j)[B1] = (MUT(6) ? -0.5 : 0.5) * (EM(e2, h, i, j) + EM(e2, h, i + 1, j));
                PV(o, F3, h, i, j)[B2] = -0.5 * (EM(e1, h, i, j) + EM(e1, h, i, j + 1));
                PV(o, F3, h, i, j)[B3] = 0.;
            }
    }
#undef EM
}

static int utoprim_5d_s(const double U[8], const double *gcov, const double *gcon, double gdet, double prim[8],
                        double gam, const orc_eos *eos, int ceos, int *n_iter);

/* step_ch.c:153-318 */
static double advance(orc_ctx *o, double *pi, double *pb, double Dt, double *pf, int fullstep)
{
    const int N3 = o->N3, is = o->istart, ie = o->istop, js = o->jstart, je = o->jstop;
    double ndt1, ndt2, ndt3 = 0., ndt;
    int64_t *cnt = o->cnt;

    if (pf != pi) {
        for (int h = 0; h < N3; h++)
            for (int i = is; i <= ie; i++)
                for (int j = js; j <= je; j++)
                    for (int k = 0; k < NPR; k++) PV(o, pf, h, i, j)[k] = PV(o, pi, h, i, j)[k];
    }
    ndt1 = fluxcalc(o, pb, o->F1, 1, fullstep, cnt);
    ndt2 = fluxcalc(o, pb, o->F2, 2, fullstep, cnt);
    if (N3 > 2) ndt3 = fluxcalc(o, pb, o->F3, 3, fullstep, cnt);
    fix_flux(o);
    flux_ct(o);

    int64_t lc[ORC_NCOUNTERS];
    memset(lc, 0, sizeof(lc));
    const int ni = ie - is + 1;
    const int font = o->P.font_fix;
#pragma omp parallel for reduction(+ : lc[:ORC_NCOUNTERS]) schedule(dynamic, 4)
    for (int n = 0; n < N3 * ni; n++) {
        const int h = n / ni, i = is + n % ni;
        for (int j = js; j <= je; j++) {
            geom_t geom;
            state_t q;
            double U[NPR], dU[NPR];
            int k, niter, kconv, ret;
            get_geometry(o, i, j, CENT, &geom);
            if (g_step_ceos) ceos_init(o, h, i, j); /* step_ch.c:241-243 */
            source(PV(o, pb, h, i, j), &geom, CONN(o, i, j), o->P.gam, dU, &lc[ORC_FATAL]);
            get_state(PV(o, pi, h, i, j), &geom, &q, &lc[ORC_FATAL]);
            primtoflux(PV(o, pi, h, i, j), &q, 0, &geom, o->P.gam, U);
            for (k = 0; k < NPR; k++) {
                if (N3 > 2)
                    U[k] += Dt * (-(PV(o, o->F1, h, i + 1, j)[k] - PV(o, o->F1, h, i, j)[k]) / o->dx[RR] -
                                  (PV(o, o->F2, h, i, j + 1)[k] - PV(o, o->F2, h, i, j)[k]) / o->dx[TH] -
                                  (PV(o, o->F3, (MUT(12) && h + 1 == N3) ? h : hwrap(o, h + 1), i, j)[k] -
                                   PV(o, o->F3, h, i, j)[k]) / o->dx[PH] +
                                  dU[k]);
                else
                    U[k] += Dt * (-(PV(o, o->F1, h, i + 1, j)[k] - PV(o, o->F1, h, i, j)[k]) / o->dx[RR] -
                                  (PV(o, o->F2, h, i, j + 1)[k] - PV(o, o->F2, h, i, j)[k]) / o->dx[TH] + dU[k]);
            }
            double *pz = PV(o, pf, h, i, j);
            if (g_step_eos || g_step_ceos) {
                /* step_ch.c:265-268 with COOL_EOS: Utoprim_5d, no Font fix */
                ret = utoprim_5d_s(U, &geom.gcov[0][0], &geom.gcon[0][0], geom.g, pz, o->P.gam, g_step_eos,
                                   g_step_ceos != NULL, &niter);
                kconv = niter;
                lc[ORC_C2P_CALLS]++;
                lc[ORC_NEWTON_ITERS] += niter;
                lc[ORC_NEWTON_KCONV] += kconv;
                if (ret) lc[ORC_FAIL_2D]++;
                o->pflag[zi(o, h, i, j)] = (int8_t)ret;
                continue;
            }
            ret = Utoprim_2d(U, geom.gcov, geom.gcon, geom.g, pz, o->P.gam, &niter, &kconv);
            lc[ORC_C2P_CALLS]++;
            lc[ORC_NEWTON_ITERS] += niter;
            lc[ORC_NEWTON_KCONV] += kconv;
            int8_t flag = (int8_t)ret;
            if (flag) {
                lc[ORC_FAIL_2D]++;
                if (font) {
                    const double kappa = o->Katm[i];
                    flag = (int8_t)Utoprim_1dvsq2fix1(U, geom.gcov, geom.gcon, geom.g, pz, kappa, o->P.gam);
                    if (flag) {
                        lc[ORC_FAIL_FONT1]++;
                        flag = (int8_t)Utoprim_1dfix1(U, geom.gcov, geom.gcon, geom.g, pz, kappa, o->P.gam);
                        if (flag) lc[ORC_FAIL_FONT2]++;
                    }
                }
            }
            o->pflag[zi(o, h, i, j)] = flag;
        }
    }
    for (int k = 0; k < ORC_NCOUNTERS; k++) cnt[k] += lc[k];
    o->flux_valid = fullstep ? 1 : 0;

    if (N3 > 2)
        ndt = o->P.defcon * 1. / (1. / ndt1 + 1. / ndt2 + 1. / ndt3);
    else
        ndt = o->P.defcon * 1. / (1. / ndt1 + 1. / ndt2);
    return ndt;
}

/* ------------------------------------------------------------------------ */
/* tracers.c: Lagrangian tracers                                             */
/* ------------------------------------------------------------------------ */
/* metric.c:62-115, CENT, for array indices (i, j) and 0-based plane h */
static double tr_c1(const orc_ctx *o, int i) { return o->startx[RR] + ((i - 2) + 0.5) * o->dx[RR]; }
static double tr_c2(const orc_ctx *o, int j) { return o->startx[TH] + ((j - 2) + 0.5) * o->dx[TH]; }
static double tr_c3(const orc_ctx *o, int h)
{
    return o->N3 > 2 ? o->startx[PH] + (h + 0.5) * o->dx[PH] : o->startx[PH] + h * o->dx[PH];
}

/* search_tracer, tracers.c:37-98 (linear scans) */
static int tr_search(const orc_ctx *o, const double *X, int *hh, int *ii, int *jj)
{
    int ok;
    if (X[RR] < tr_c1(o, o->istart - 1) || X[TH] < tr_c2(o, o->jstart - 1)) return 0;
    *hh = 0;
    ok = 0;
    for (int i = o->istart - 1; i <= o->istop; i++)
        if (X[RR] >= tr_c1(o, i) && X[RR] < tr_c1(o, i + 1)) {
            *ii = i;
            ok = 1;
            break;
        }
    if (!ok) return 0;
    ok = 0;
    for (int j = o->jstart - 1; j <= o->jstop; j++)
        if (X[TH] >= tr_c2(o, j) && X[TH] < tr_c2(o, j + 1)) {
            *jj = j;
            ok = 1;
            break;
        }
    if (!ok) return 0;
    if (o->N3 > 2) {
        ok = 0;
        for (int h = 0; h < o->N3; h++)
            if (X[PH] >= o->startx[PH] + h * o->dx[PH] && X[PH] < o->startx[PH] + (h + 1) * o->dx[PH]) {
                *hh = h;
                ok = 1;
                break;
            }
        if (!ok) return 0;
    }
    return 1;
}

/* ucon_fun, tracers.c:248-260 */
static void tr_ucon(const orc_ctx *o, int h, int i, int j, double *v)
{
    geom_t geom;
    int64_t fatal = 0;
    get_geometry(o, i, j, CENT, &geom);
    ucon_calc(PV(o, o->p, h, i, j), &geom, v, &fatal);
    v[RR] /= v[TT];
    v[TH] /= v[TT];
    v[PH] /= v[TT];
}

/* rhoTYe_fun, tracers.c:432-466 (Ye = -1; with a table, its temperature) */
static void tr_rhoT(const orc_ctx *o, int h, int i, int j, double *v)
{
    const double *pz = PV(o, o->p, h, i, j);
    v[0] = pz[RHO] * o->tr_rho_unit;
    if (o->eos) {
        double pp, T;
        orc_eos_table_p(o->eos, pz[RHO], pz[UU], &pp, NULL, NULL, &T);
        v[1] = T;
    } else {
        const double CL = 2.99792458e10, KBOL = 1.38e-16, mn = 1.67492729e-24, me = 9.109565e-28, mp = 1.673e-24,
                     mhe = 6.6465e-24;
        const double mav = (mp + mn + 2. * me + mhe) / 5.;
        v[1] = (o->P.gam - 1.) * pz[UU] / pz[RHO] / KBOL * mav * CL * CL;
    }
    v[2] = -1.;
}

/* linear_interpolation, tracers.c:174-245; X is [TT, RR, TH, PH] */
static void tr_interp(const orc_ctx *o, const double *X, int hh, int ii, int jj, int nv, double *val,
                      void (*fun)(const orc_ctx *, int, int, int, double *))
{
    const int wph = o->N3 > 2 ? 2 : 1;
    double XX[2][2][2][NDIM], valt[2][2][2][NDIM], tmp1[2][2][NDIM], tmp2[2][NDIM];
    for (int l3 = 0; l3 < wph; l3++)
        for (int l1 = 0; l1 < 2; l1++)
            for (int l2 = 0; l2 < 2; l2++) {
                XX[l3][l1][l2][RR] = tr_c1(o, ii + l1);
                XX[l3][l1][l2][TH] = tr_c2(o, jj + l2);
                XX[l3][l1][l2][PH] = tr_c3(o, hh + l3);
                fun(o, hwrap(o, hh + l3), ii + l1, jj + l2, valt[l3][l1][l2]);
            }
    for (int l3 = 0; l3 < wph; l3++)
        for (int l1 = 0; l1 < 2; l1++) {
            const double T1 = 1.0 / (XX[l3][l1][1][TH] - XX[l3][l1][0][TH]);
            const double T2 = X[TH] - XX[l3][l1][0][TH];
            for (int k = 0; k < nv; k++) {
                const double ypval = (valt[l3][l1][1][k] - valt[l3][l1][0][k]) * T1;
                tmp1[l3][l1][k] = valt[l3][l1][0][k] + T2 * ypval;
            }
        }
    for (int l3 = 0; l3 < wph; l3++) {
        const double T1 = 1.0 / (XX[l3][1][1][RR] - XX[l3][0][0][RR]);
        const double T2 = X[RR] - XX[l3][0][0][RR];
        for (int k = 0; k < nv; k++) {
            const double ypval = (tmp1[l3][1][k] - tmp1[l3][0][k]) * T1;
            tmp2[l3][k] = tmp1[l3][0][k] + T2 * ypval;
        }
    }
    if (wph == 1) {
        for (int k = 0; k < nv; k++) val[k] = tmp2[0][k];
        return;
    }
    const double T1 = 1.0 / (XX[1][1][1][PH] - XX[0][0][0][PH]);
    const double T2 = X[PH] - XX[0][0][0][PH];
    for (int k = 0; k < nv; k++) {
        const double ypval = (tmp2[1][k] - tmp2[0][k]) * T1;
        val[k] = tmp2[0][k] + T2 * ypval;
    }
}

/* dump_one_tracer, tracers.c:470-510 */
static void tr_record(const orc_ctx *o, long long k, const double *X, int hh, int ii, int jj, int kind,
                      orc_tracer_rec *r)
{
    double val[3], V[NDIM];
    const long long n12 = (long long)o->N1 * o->N2;
    tr_interp(o, X, hh, ii, jj, 3, val, tr_rhoT);
    bl_coord_vec(&o->P, X, V);
    double ph = V[PH];
    if (o->N3 > 2) {
        if (ph > 2. * M_PI)
            ph -= 2. * M_PI;
        else if (ph < 0.)
            ph += 2. * M_PI;
    } else {
        ph = 0.0;
    }
    r->t = o->t * o->tr_t_unit;
    r->rho = val[0];
    r->T = val[1];
    r->ye = val[2];
    r->r = V[RR];
    r->th = V[TH];
    r->ph = ph;
    r->h = (int32_t)(k / n12);
    r->i = (int32_t)((k % n12) / o->N2);
    r->j = (int32_t)(k % o->N2);
    r->kind = kind;
}

/* update_tracers, tracers.c:298-430 (one rank: phi wraps in place) */
static void tracers_update(orc_ctx *o, double Dt)
{
    const long long n = o->tr_n;
    for (long long k = 0; k < n; k++) {
        if (o->tr_state[k] != ORC_TRACER_LIVE) continue;
        double X[NDIM] = {0., o->tr_X[k], o->tr_X[n + k], o->tr_X[2 * n + k]}, ucon[NDIM], V[NDIM];
        int hh, ii, jj, kind = 0;
        if (!tr_search(o, X, &hh, &ii, &jj)) {
            o->tr_state[k] = ORC_TRACER_LOST;
            continue;
        }
        tr_interp(o, X, hh, ii, jj, NDIM, ucon, tr_ucon);
        for (int d = 1; d < NDIM; d++) X[d] += ucon[d] * Dt;
        bl_coord_vec(&o->P, X, V);
        if (V[RR] < o->tr_Rin || V[TH] < 0 || V[TH] > M_PI)
            kind = ORC_TRACER_IN;
        else if (V[RR] >= o->tr_Rout)
            kind = (V[TH] >= o->tr_thmin && V[TH] <= o->tr_thmax) ? ORC_TRACER_OUT : ORC_TRACER_JET;
        if (kind) {
            if (o->tr_nev < o->tr_cap) tr_record(o, k, X, hh, ii, jj, kind, &o->tr_ev[o->tr_nev++]);
            o->tr_state[k] = (int8_t)kind;
        } else if (o->N3 > 2) {
            if (X[PH] < 0.)
                X[PH] += 2. * M_PI;
            else if (X[PH] > 2. * M_PI)
                X[PH] -= 2. * M_PI;
        }
        o->tr_X[k] = X[RR];
        o->tr_X[n + k] = X[TH];
        o->tr_X[2 * n + k] = X[PH];
    }
}

/* init_tracers, tracers.c:109-172 */
long long orc_tracers_init(orc_ctx *o, double rho_unit, double t_unit)
{
    const long long n = (long long)o->N1 * o->N2 * o->N3;
    long long live = 0;
    free(o->tr_X);
    free(o->tr_state);
    free(o->tr_ev);
    o->tr_n = n;
    o->tr_X = (double *)calloc(3 * (size_t)n, sizeof(double));
    o->tr_state = (int8_t *)calloc((size_t)n, 1);
    o->tr_Rin = exp(o->startx[RR] - 0.5 * o->dx[RR]) + o->P.R0;
    o->tr_Rout = 0.8 * (exp(o->startx[RR] + o->N1 * o->dx[RR]) + o->P.R0);
    o->tr_thmin = 0.02 * M_PI;
    o->tr_thmax = 0.98 * M_PI;
    o->tr_rho_unit = rho_unit;
    o->tr_t_unit = t_unit;
    long long k = 0;
    for (int h = 0; h < o->N3; h++)
        for (int i = 0; i < o->N1; i++)
            for (int j = 0; j < o->N2; j++, k++) {
                double X[NDIM];
                int hh, ii, jj;
                X[TT] = 0.;
                X[RR] = o->startx[RR] + (i + 0.5) * o->dx[RR];
                X[TH] = o->startx[TH] + (j + 0.5) * o->dx[TH];
                X[PH] = o->startx[PH] + h * o->dx[PH];
                o->tr_X[k] = X[RR];
                o->tr_X[n + k] = X[TH];
                o->tr_X[2 * n + k] = X[PH];
                if (tr_search(o, X, &hh, &ii, &jj) && PV(o, o->p, hh, ii, jj)[RHO] > RHOMIN) {
                    o->tr_state[k] = ORC_TRACER_LIVE;
                    live++;
                }
            }
    o->tr_cap = live;
    o->tr_nev = 0;
    o->tr_ev = (orc_tracer_rec *)calloc(live > 0 ? (size_t)live : 1, sizeof(orc_tracer_rec));
    o->tr_on = 1;
    return live;
}

/* dump_tracers, tracers.c:512-560 */
long long orc_tracers_dump(orc_ctx *o, orc_tracer_rec *recs, long long cap)
{
    const long long n = o->tr_n;
    long long m = 0;
    g_step_eos = o->eos;
    for (long long k = 0; k < n; k++) {
        if (o->tr_state[k] != ORC_TRACER_LIVE) continue;
        const double X[NDIM] = {0., o->tr_X[k], o->tr_X[n + k], o->tr_X[2 * n + k]};
        int hh, ii, jj;
        if (!tr_search(o, X, &hh, &ii, &jj)) {
            o->tr_state[k] = ORC_TRACER_LOST;
            continue;
        }
        if (m < cap) tr_record(o, k, X, hh, ii, jj, ORC_TRACER_LIVE, &recs[m]);
        m++;
    }
    g_step_eos = NULL;
    return m;
}

/* records of the tracers that left since the previous call, in order */
long long orc_tracers_events(orc_ctx *o, orc_tracer_rec *recs, long long cap)
{
    const long long m = o->tr_nev;
    for (long long q = 0; q < m && q < cap; q++) recs[q] = o->tr_ev[q];
    o->tr_nev = 0;
    return m;
}

void orc_tracers_update(orc_ctx *o, double Dt) { tracers_update(o, Dt); }

void orc_tracers_set(orc_ctx *o, const double *X, const int8_t *state)
{
    long long live = 0;
    if (X) memcpy(o->tr_X, X, 3 * (size_t)o->tr_n * sizeof(double));
    if (state) memcpy(o->tr_state, state, (size_t)o->tr_n);
    for (long long k = 0; k < o->tr_n; k++) live += o->tr_state[k] == ORC_TRACER_LIVE;
    if (live > o->tr_cap) {
        free(o->tr_ev);
        o->tr_ev = (orc_tracer_rec *)calloc((size_t)live, sizeof(orc_tracer_rec));
        o->tr_cap = live;
    }
    o->tr_nev = 0;
}

void orc_tracers_get(const orc_ctx *o, double *X, int8_t *state)
{
    if (X) memcpy(X, o->tr_X, 3 * (size_t)o->tr_n * sizeof(double));
    if (state) memcpy(state, o->tr_state, (size_t)o->tr_n);
}

/* step_ch.c:90-140 */
void orc_set_eos(orc_ctx *o, const orc_eos *eos, int ref_derivs)
{
    o->eos = eos;
    o->eos_refd = ref_derivs ? 1 : 0;
}

int orc_set_eos_cache(orc_ctx *o, const orc_ceos *c)
{
    const size_t nz = (size_t)o->N3 * o->SN1 * o->SN2;
    if (!c) {
        o->ceos_on = 0;
        return 0;
    }
    if (c->n != 1 && c->n != o->SN1) return 2;
    if (!o->lastT && !(o->lastT = (double *)malloc(nz * sizeof(double)))) return 3;
    memset(o->lastT, 0, nz * sizeof(double)); /* calloc'ed by SetupCache_CoolEOS, cooleos.c:834 */
    o->ceos = *c;
    o->ceos_on = 1;
    return 0;
}

int orc_get_lastT(const orc_ctx *o, double *T)
{
    if (!o->lastT) return 2;
    const size_t nh = o->N3, nj = o->SN2, ni = o->SN1;
    for (size_t i = 0; i < ni; i++)
        for (size_t j = 0; j < nj; j++)
            for (size_t h = 0; h < nh; h++) T[(i * nj + j) * nh + h] = o->lastT[zi(o, (int)h, (int)i, (int)j)];
    return 0;
}

int orc_set_lastT(orc_ctx *o, const double *T)
{
    if (!o->lastT) return 2;
    const size_t nh = o->N3, nj = o->SN2, ni = o->SN1;
    for (size_t i = 0; i < ni; i++)
        for (size_t j = 0; j < nj; j++)
            for (size_t h = 0; h < nh; h++) o->lastT[zi(o, (int)h, (int)i, (int)j)] = T[(i * nj + j) * nh + h];
    return 0;
}

double orc_ceos_pressure(orc_ctx *o, int h, int i, int j, double rho0, double u)
{
    if (!o->ceos_on) return NAN;
    g_step_ceos = o;
    ceos_init(o, h, i, j);
    const double p = ceos_pres(rho0, u);
    g_step_ceos = NULL;
    return p;
}

int orc_step(orc_ctx *o)
{
    double ndt;
    g_step_eos = o->eos;
    g_step_ceos = o->ceos_on ? o : NULL;
    g_step_refd = o->eos_refd;
    int64_t fatal0 = o->cnt[ORC_FATAL];
    ndt = advance(o, o->p, o->p, 0.5 * o->dt, o->ph, 0);
    fixup(o, o->ph, o->cnt);
    bound_prim(o, o->ph, &o->cnt[ORC_FATAL]);
    fixup_utoprim(o, o->ph, o->cnt);
    bound_prim(o, o->ph, &o->cnt[ORC_FATAL]);
    ndt = advance(o, o->p, o->ph, o->dt, o->p, 1);
    fixup(o, o->p, o->cnt);
    bound_prim(o, o->p, &o->cnt[ORC_FATAL]);
    fixup_utoprim(o, o->p, o->cnt);
    bound_prim(o, o->p, &o->cnt[ORC_FATAL]);
    /* The reference exits inside fail() (diag.c:381-397), before diag_mass and
     * the clock update; here the step's state is kept and 14 returned. */
    if (o->cnt[ORC_FATAL] != fatal0) {
        g_step_eos = NULL;
        g_step_ceos = NULL;
        return 14;
    }
    /* diag_mass (diag.c:477-580) takes no pressure: the cache EOS's zone
     * selection ends with the advance */
    if (!MUT(19)) g_step_ceos = NULL;
    {
        double out[4];
        orc_diag(o, out);
        o->mass_in += out[2] * o->dt;
        o->mass_out += out[3] * o->dt;
    }
    if (o->tr_on) tracers_update(o, o->dt); /* step_ch.c:121-123 */
    g_step_eos = NULL;
    g_step_ceos = NULL;
    if (o->dt < 1.e-9) return 11;           /* step_ch.c:126-129: exit before t += dt */
    o->t += o->dt;
    if (ndt > SAFE * o->dt) ndt = SAFE * o->dt;
    o->dt = ndt;
    if (o->t + o->dt > o->P.tf) o->dt = o->P.tf - o->t;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* init.c: Fishbone-Moncrief torus                                           */
/* ------------------------------------------------------------------------ */
/* init.c:557-567 */
static double lfish_calc(double a, double r)
{
    return (((pow(a, 2) - 2. * a * sqrt(r) + pow(r, 2)) *
             ((-2. * a * r * (pow(a, 2) - 2. * a * sqrt(r) + pow(r, 2))) / sqrt(2. * a * sqrt(r) + (-3. + r) * r) +
              ((a + (-2. + r) * sqrt(r)) * (pow(r, 3) + pow(a, 2) * (2. + r))) /
                  sqrt(1 + (2. * a) / pow(r, 1.5) - 3. / r))) /
            (pow(r, 3) * sqrt(2. * a * sqrt(r) + (-3. + r) * r) * (pow(a, 2) + (-2. + r) * r)));
}

/* metric.c:340-360 */
static void bl_gcov_func(double a, double r, double th, double gcov[NDIM][NDIM])
{
    int j, k;
    double sth, cth, s2, a2, r2, DD, mu;
    for (j = 0; j < NDIM; j++)
        for (k = 0; k < NDIM; k++) gcov[j][k] = 0.;
    sincos_ref(th, &sth, &cth);
    sth = fabs(sth);
    s2 = sth * sth;
    a2 = a * a;
    r2 = r * r;
    DD = 1. - 2. / r + a2 / r2;
    mu = 1. + a2 * cth * cth / r2;
    gcov[TT][TT] = -(1. - 2. / (r * mu));
    gcov[TT][3] = -2. * a * s2 / (r * mu);
    gcov[3][TT] = gcov[TT][3];
    gcov[1][1] = mu / DD;
    gcov[2][2] = r2 * mu;
    gcov[3][3] = r2 * sth * sth * (1. + a2 / r2 + 2. * a2 * s2 / (r2 * r * mu));
}

/* init.c:479-555 */
static void coord_transform(orc_ctx *o, double *pr, int hh, int ii, int jj)
{
    double X[NDIM], V[NDIM], r, th, ucon[NDIM], uconp[NDIM], trans[NDIM][NDIM], tmp[NDIM];
    double AA, BB, CC, discr, utconp[NDIM], dxdxp[NDIM][NDIM], dxpdx[NDIM][NDIM];
    double blg[NDIM][NDIM];
    geom_t geom;
    int i, j, k;
    const double a = o->P.a;

    coord(o, hh, ii, jj, CENT, X);
    bl_coord_vec(&o->P, X, V);
    r = V[RR];
    th = V[TH];
    {
        double tb = th;
        if (tb < 0) tb *= -1.;
        if (tb > M_PI) tb = 2. * M_PI - tb;
        bl_gcov_func(a, r, tb, blg);
    }
    X[0] = 1;
    ucon[1] = pr[U1];
    ucon[2] = pr[U2];
    ucon[3] = pr[U3];
    AA = blg[TT][TT];
    BB = 2. * (blg[TT][1] * ucon[1] + blg[TT][2] * ucon[2] + blg[TT][3] * ucon[3]);
    CC = 1. + blg[1][1] * ucon[1] * ucon[1] + blg[2][2] * ucon[2] * ucon[2] + blg[3][3] * ucon[3] * ucon[3] +
         2. * (blg[1][2] * ucon[1] * ucon[2] + blg[1][3] * ucon[1] * ucon[3] + blg[2][3] * ucon[2] * ucon[3]);
    discr = BB * BB - 4. * AA * CC;
    ucon[TT] = (-BB - sqrt(discr)) / (2. * AA);
    for (j = 0; j < NDIM; j++)
        for (k = 0; k < NDIM; k++) trans[j][k] = 0.;
    for (j = 0; j < NDIM; j++) trans[j][j] = 1.;
    trans[0][1] = 2. * r / (r * r - 2. * r + a * a);
    trans[3][1] = a / (r * r - 2. * r + a * a);
    for (j = 0; j < NDIM; j++) tmp[j] = 0.;
    for (j = 0; j < NDIM; j++)
        for (k = 0; k < NDIM; k++) tmp[j] += trans[j][k] * ucon[k];
    for (j = 0; j < NDIM; j++) ucon[j] = tmp[j];
    dxdxp_func(&o->P, X, dxdxp);
    invert_matrix(dxdxp, dxpdx);
    for (i = 0; i < NDIM; i++) {
        uconp[i] = 0;
        for (j = 0; j < NDIM; j++) uconp[i] += dxpdx[i][j] * ucon[j];
    }
    get_geometry(o, ii, jj, CENT, &geom);
    ucon_to_utcon(uconp, &geom, utconp);
    pr[U1] = utconp[1];
    pr[U2] = utconp[2];
    pr[U3] = utconp[3];
}

static double bsq_calc(const double *pr, const geom_t *geom, int64_t *fatal)
{
    state_t q;
    get_state(pr, geom, &q, fatal);
    return dot(q.bcon, q.bcov);
}

/* init.c:70-472 */
int orc_init_torus(orc_ctx *o)
{
    const orc_params *P = &o->P;
    const double a = P->a, gam = P->gam, rin = P->rin, kappa = P->kappa, beta = P->beta;
    const int N3 = o->N3, is = o->istart, ie = o->istop, js = o->jstart, je = o->jstop;
    double rhomax = 0., umax = 0., l, bsq_max, beta_act, norm;
    int64_t fatal = 0;
    int h, i, j;
    double *A = o->e3;

    l = lfish_calc(a, P->rmax);
    o->t = 0.;
    o->dt = P->dt0;
    memset(o->p, 0, sizeof(double) * (size_t)N3 * o->SN1 * o->SN2 * NPR);
    memset(o->pflag, 0, (size_t)N3 * o->SN1 * o->SN2);

    for (h = 0; h < N3; h++)
        for (i = is; i <= ie; i++)
            for (j = js; j <= je; j++) {
                double X[NDIM], V[NDIM], r, th, sth, cth, DD, AA, SS, thin, sthin, cthin, DDin, AAin, SSin;
                double lnh, rho, u, ur, uh, up, hm1, expm2chi, up1;
                double *pz = PV(o, o->p, h, i, j);
                coord(o, h, i, j, CENT, X);
                bl_coord_vec(P, X, V);
                r = V[RR];
                th = V[TH];
                sincos_ref(th, &sth, &cth);
                DD = r * r - 2. * r + a * a;
                AA = (r * r + a * a) * (r * r + a * a) - DD * a * a * sth * sth;
                SS = r * r + a * a * cth * cth;
                thin = M_PI / 2.;
                sincos_ref(thin, &sthin, &cthin);
                DDin = rin * rin - 2. * rin + a * a;
                AAin = (rin * rin + a * a) * (rin * rin + a * a) - DDin * a * a * sthin * sthin;
                SSin = rin * rin + a * a * cthin * cthin;
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
                    rho = 1.e-7 * RHOMIN;
                    u = 1.e-7 * UUMIN;
                    pz[RHO] = rho;
                    pz[UU] = u;
                    pz[U1] = 0.;
                    pz[U2] = 0.;
                    pz[U3] = 0.;
                } else {
                    hm1 = exp(lnh) - 1.;
                    rho = pow(hm1 * (gam - 1.) / (kappa * gam), 1. / (gam - 1.));
                    u = kappa * pow(rho, gam) / (gam - 1.);
                    ur = 0.;
                    uh = 0.;
                    expm2chi = SS * SS * DD / (AA * AA * sth * sth);
                    up1 = sqrt((-1. + sqrt(1. + 4. * l * l * expm2chi)) / 2.);
                    up = 2. * a * r * sqrt(1. + up1 * up1) / sqrt(AA * SS * DD) + sqrt(SS / AA) * up1 / sth;
                    pz[RHO] = rho;
                    if (rho > rhomax) rhomax = rho;
                    {
                        const unsigned int ii = (unsigned int)(i - 2);
                        const unsigned int jj = (unsigned int)(j - 2);
                        const unsigned int hh = (N3 == 1) ? 0u : (unsigned int)h;
                        const unsigned int off = hh * ii * jj + ii * jj + jj;
                        double rnd = orc_kiss_uniform(P->seed[0] + off, P->seed[1] + off, P->seed[2] + off,
                                                      P->seed[3] + off);
                        pz[UU] = u * (1. + P->pert_amp * (rnd - 0.5));
                    }
                    if (u > umax && r >= rin) umax = u;
                    pz[U1] = ur;
                    pz[U2] = uh;
                    pz[U3] = up;
                    coord_transform(o, pz, h, i, j);
                }
                pz[B1] = 0.;
                pz[B2] = 0.;
                pz[B3] = 0.;
            }

    for (h = 0; h < N3; h++)
        for (i = is; i <= ie; i++)
            for (j = js; j <= je; j++) {
                PV(o, o->p, h, i, j)[RHO] /= rhomax;
                PV(o, o->p, h, i, j)[UU] /= rhomax;
            }
    umax /= rhomax;
    rhomax = 1.;
    fixup(o, o->p, NULL);
    bound_prim(o, o->p, &fatal);

    for (h = 0; h < N3; h++)
        for (i = is; i <= ie + 1; i++)
            for (j = js; j <= je + 1; j++) A[zi(o, h, i, j)] = 0.;
    for (h = 0; h < N3; h++)
        for (i = is; i <= ie + 1; i++)
            for (j = js; j <= je + 1; j++) {
                double rho_av = 0.25 * (PV(o, o->p, h, i, j)[RHO] + PV(o, o->p, h, i - 1, j)[RHO] +
                                        PV(o, o->p, h, i, j - 1)[RHO] + PV(o, o->p, h, i - 1, j - 1)[RHO]);
                double q = rho_av / rhomax - 0.2;
                if (q > 0.) A[zi(o, h, i, j)] = q;
            }
    bsq_max = 0.;
    for (h = 0; h < N3; h++)
        for (i = is; i <= ie; i++)
            for (j = js; j <= je; j++) {
                geom_t geom;
                double *pz = PV(o, o->p, h, i, j);
                get_geometry(o, i, j, CENT, &geom);
                pz[B1] = -(A[zi(o, h, i, j)] - A[zi(o, h, i, j + 1)] + A[zi(o, h, i + 1, j)] - A[zi(o, h, i + 1, j + 1)]) /
                         (2. * o->dx[TH] * geom.g);
                pz[B2] = (A[zi(o, h, i, j)] + A[zi(o, h, i, j + 1)] - A[zi(o, h, i + 1, j)] - A[zi(o, h, i + 1, j + 1)]) /
                         (2. * o->dx[RR] * geom.g);
                pz[B3] = 0.;
                double bsq_ij = bsq_calc(pz, &geom, &fatal);
                if (bsq_ij > bsq_max) bsq_max = bsq_ij;
            }
    beta_act = (gam - 1.) * umax / (0.5 * bsq_max);
    norm = sqrt(beta_act / beta);
    bsq_max = 0.;
    for (h = 0; h < N3; h++)
        for (i = is; i <= ie; i++)
            for (j = js; j <= je; j++) {
                geom_t geom;
                double *pz = PV(o, o->p, h, i, j);
                pz[B1] *= norm;
                pz[B2] *= norm;
                get_geometry(o, i, j, CENT, &geom);
                double bsq_ij = bsq_calc(pz, &geom, &fatal);
                if (bsq_ij > bsq_max) bsq_max = bsq_ij;
            }
    fixup(o, o->p, NULL);
    bound_prim(o, o->p, &fatal);

    /* fixup.c:555-585 (set_Katm): floor values at the j=0 ghost row */
    for (i = is; i <= ie; i++) {
        double prim[NPR];
        for (int k = 0; k < NPR; k++) prim[k] = 0.;
        prim[RHO] = prim[UU] = -1.;
        fixup1zone(o, 0, i, 0, prim, NULL);
        o->Katm[i] = (gam - 1.) * prim[UU] / pow(prim[RHO], gam);
    }
    for (h = 0; h < N3; h++)
        for (i = is; i <= ie; i++)
            for (j = js; j <= je; j++)
                if (isnan(PV(o, o->p, h, i, j)[UU])) return 2;
    o->mass_in = o->mass_out = 0.;
    /* main.c:173-176 */
    o->mass_start = 0.;
    for (h = 0; h < N3; h++)
        for (i = is; i <= ie; i++)
            for (j = js; j <= je; j++)
                o->mass_start += PV(o, o->p, h, i, j)[RHO] * GDET(o, i, j, CENT) * o->dx[RR] * o->dx[TH] * o->dx[PH];
    o->flux_valid = 0;
    return fatal ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* diagnostics: diag.c:156-209, diag.c:477-580                               */
/* ------------------------------------------------------------------------ */
/* |div B| of the flux-CT corner at zone (h, i, j) (diag.c:160-209). */
static double divb_at(const orc_ctx *o, int h, int i, int j)
{
    const int N3 = o->N3;
#define PB(h, i, j, k) PV(o, o->p, h, i, j)[k]
#define GD(i, j) GDET(o, i, j, CENT)
    double divb;
    if (N3 > 2) {
        const int hm = hwrap(o, h - 1);
        divb = fabs(0.25 *
                        (PB(h, i, j, B1) * GD(i, j) + PB(hm, i, j, B1) * GD(i, j) +
                         PB(h, i, j - 1, B1) * GD(i, j - 1) + PB(hm, i, j - 1, B1) * GD(i, j - 1) -
                         PB(h, i - 1, j, B1) * GD(i - 1, j) - PB(hm, i - 1, j, B1) * GD(i - 1, j) -
                         PB(h, i - 1, j - 1, B1) * GD(i - 1, j - 1) - PB(hm, i - 1, j - 1, B1) * GD(i - 1, j - 1)) /
                        o->dx[RR] +
                    0.25 *
                        (PB(h, i, j, B2) * GD(i, j) + PB(hm, i, j, B2) * GD(i, j) +
                         PB(h, i - 1, j, B2) * GD(i - 1, j) + PB(hm, i - 1, j, B2) * GD(i - 1, j) -
                         PB(h, i, j - 1, B2) * GD(i, j - 1) - PB(hm, i, j - 1, B2) * GD(i, j - 1) -
                         PB(h, i - 1, j - 1, B2) * GD(i - 1, j - 1) - PB(hm, i - 1, j - 1, B2) * GD(i - 1, j - 1)) /
                        o->dx[TH] +
                    0.25 *
                        (PB(h, i, j, B3) * GD(i, j) + PB(h, i - 1, j, B3) * GD(i - 1, j) +
                         PB(h, i, j - 1, B3) * GD(i, j - 1) + PB(h, i - 1, j - 1, B3) * GD(i - 1, j - 1) -
                         PB(hm, i, j, B3) * GD(i, j) - PB(hm, i - 1, j, B3) * GD(i - 1, j) -
                         PB(hm, i, j - 1, B3) * GD(i, j - 1) - PB(hm, i - 1, j - 1, B3) * GD(i - 1, j - 1)) /
                        o->dx[PH]);
    } else {
        divb = fabs(0.5 *
                        (PB(h, i, j, B1) * GD(i, j) + PB(h, i, j - 1, B1) * GD(i, j - 1) -
                         PB(h, i - 1, j, B1) * GD(i - 1, j) - PB(h, i - 1, j - 1, B1) * GD(i - 1, j - 1)) /
                        o->dx[RR] +
                    0.5 *
                        (PB(h, i, j, B2) * GD(i, j) + PB(h, i - 1, j, B2) * GD(i - 1, j) -
                         PB(h, i, j - 1, B2) * GD(i, j - 1) - PB(h, i - 1, j - 1, B2) * GD(i - 1, j - 1)) /
                        o->dx[TH]);
    }
    return divb;
#undef PB
#undef GD
}

double orc_divb_max(const orc_ctx *o)
{
    double divbmax = 0.;
    for (int h = 0; h < o->N3; h++)
        for (int i = o->istart; i <= o->istop; i++)
            for (int j = o->jstart; j <= o->jstop; j++) {
                const double divb = divb_at(o, h, i, j);
                if (divb > divbmax) divbmax = divb;
            }
    return divbmax;
}

void orc_diag(const orc_ctx *o, double *out)
{
    const int N3 = o->N3;
    const double dV = o->dx[RR] * o->dx[TH] * o->dx[PH]; /* main.c:427 */
    double mass = 0., e = 0., mdot_in = 0., mdot_out = 0.;
    int64_t fatal = 0;
    for (int h = 0; h < N3; h++)
        for (int i = o->istart; i <= o->istop; i++)
            for (int j = o->jstart; j <= o->jstop; j++) {
                geom_t geom;
                state_t q;
                double U[NPR];
                const double *pz = PV(o, o->p, h, i, j);
                mass += pz[RHO] * GDET(o, i, j, CENT) * o->dx[RR] * o->dx[TH] * o->dx[PH];
                get_geometry(o, i, j, CENT, &geom);
                get_state(pz, &geom, &q, &fatal);
                primtoflux(pz, &q, 0, &geom, o->P.gam, U);
                e += U[UU] * dV;
            }
    for (int h = 0; h < N3; h++)
        for (int j = o->jstart; j <= o->jstop; j++) {
            geom_t geom;
            double ucon[NDIM];
            const double *pz = PV(o, o->p, h, o->istart - 1, j);
            get_geometry(o, o->istart - 1, j, CENT, &geom);
            ucon_calc(pz, &geom, ucon, &fatal);
            mdot_in += pz[RHO] * ucon[1] * geom.g * o->dx[TH] * o->dx[PH];
            pz = PV(o, o->p, h, o->istop + 1, j);
            get_geometry(o, o->istop + 1, j, CENT, &geom);
            ucon_calc(pz, &geom, ucon, &fatal);
            mdot_out -= pz[RHO] * ucon[1] * geom.g * o->dx[TH] * o->dx[PH];
        }
    out[0] = mass;
    out[1] = e;
    out[2] = mdot_in;
    out[3] = mdot_out;
}

/* ------------------------------------------------------------------------ */
/* context and exports                                                       */
/* ------------------------------------------------------------------------ */
/* Fluxes of one half-step after fix_flux and flux_ct, canonical layout. */
/* diag() log record (diag.c:59-326): conserved totals and |div B| with its
 * zone (ZLOOP order, first maximum), diag_flux of the last full step
 * (diag.c:437-465) and diag_mass of the current state (diag.c:477-580), each
 * summed in the reference's loop order.  The neutrino luminosities need the
 * cooling table's Qnu, Lambda and H_d, which the tabulated EOS here does not
 * carry: lnu, lnu_rho and lnu2 are 0, and with a tabulated EOS lnu_rho_2 holds
 * the COOL_EOS build's density weight sum gdet rho (diag.c:203). */
void orc_diag_log(orc_ctx *o, orc_diag_rec *r)
{
    const int N3 = o->N3, is = o->istart, ie = o->istop, js = o->jstart, je = o->jstop;
    const double dV = o->dx[RR] * o->dx[TH] * o->dx[PH]; /* main.c:427 */
    int64_t fatal = 0;
    memset(r, 0, sizeof(*r));
    r->imax[0] = r->imax[1] = r->imax[2] = -1;
    g_step_eos = o->eos; /* primtoU with the step's EOS (phys.c:122-141) */
    g_step_refd = o->eos_refd;
    for (int h = 0; h < N3; h++)
        for (int i = is; i <= ie; i++)
            for (int j = js; j <= je; j++) {
                geom_t geom;
                state_t q;
                double U[NPR];
                const double *pz = PV(o, o->p, h, i, j);
                get_geometry(o, i, j, CENT, &geom);
                get_state(pz, &geom, &q, &fatal);
                primtoflux(pz, &q, 0, &geom, o->P.gam, U);
                r->rmed += U[RHO] * dV;
                r->pp += U[U3] * dV;
                r->e += U[UU] * dV;
                const double divb = divb_at(o, h, i, j);
                if (divb > r->divbmax) {
                    r->imax[0] = i - is;
                    r->imax[1] = j - js;
                    r->imax[2] = h;
                    r->divbmax = divb;
                }
                r->mass_curr += pz[RHO] * GDET(o, i, j, CENT) * o->dx[RR] * o->dx[TH] * o->dx[PH];
                if (o->eos) r->lnu_rho_2 += GDET(o, i, j, CENT) * pz[RHO]; /* diag.c:203 */
            }
    g_step_eos = NULL;
    if (o->flux_valid)
        for (int h = 0; h < N3; h++)
            for (int j = js; j <= je; j++) {
                const double *F = PV(o, o->F1, h, is, j);
                r->mdot += F[RHO] * o->dx[TH] * o->dx[PH];
                r->edot -= (F[UU] - F[RHO]) * o->dx[TH] * o->dx[PH];
                r->ldot += F[U3] * o->dx[TH] * o->dx[PH];
            }
    for (int h = 0; h < N3; h++)
        for (int j = js; j <= je; j++) {
            geom_t geom;
            double ucon[NDIM];
            const double *pz = PV(o, o->p, h, is - 1, j);
            get_geometry(o, is - 1, j, CENT, &geom);
            ucon_calc(pz, &geom, ucon, &fatal);
            r->mdot_in += pz[RHO] * ucon[1] * geom.g * o->dx[TH] * o->dx[PH];
            pz = PV(o, o->p, h, ie + 1, j);
            get_geometry(o, ie + 1, j, CENT, &geom);
            ucon_calc(pz, &geom, ucon, &fatal);
            r->mdot_out -= pz[RHO] * ucon[1] * geom.g * o->dx[TH] * o->dx[PH];
        }
    r->t = o->t;
    r->mass_start = o->mass_start;
    r->mass_in = o->mass_in;
    r->mass_out = o->mass_out;
    r->flux_valid = o->flux_valid;
}

int orc_debug_fluxes(orc_ctx *o, int fullstep, double *F1, double *F2, double *F3, double *ndt)
{
    int64_t cnt[ORC_NCOUNTERS];
    g_step_eos = o->eos;
    g_step_ceos = o->ceos_on ? o : NULL;
    g_step_refd = o->eos_refd;
    memset(cnt, 0, sizeof(cnt));
    ndt[0] = fluxcalc(o, o->p, o->F1, 1, fullstep, cnt);
    ndt[1] = fluxcalc(o, o->p, o->F2, 2, fullstep, cnt);
    ndt[2] = (o->N3 > 2) ? fluxcalc(o, o->p, o->F3, 3, fullstep, cnt) : 1.e9;
    fix_flux(o);
    flux_ct(o);
    o->flux_valid = 0;
    const size_t nh = o->N3, nj = o->SN2, ni = o->SN1;
    double *dst[3] = {F1, F2, F3};
    const double *src[3] = {o->F1, o->F2, o->F3};
    for (int d = 0; d < 3; d++) {
        if (!dst[d]) continue;
        for (size_t i = 0; i < ni; i++)
            for (size_t j = 0; j < nj; j++)
                for (size_t h = 0; h < nh; h++)
                    for (int k = 0; k < NPR; k++)
                        dst[d][(((size_t)k * ni + i) * nj + j) * nh + h] = PV(o, src[d], (int)h, (int)i, (int)j)[k];
    }
    g_step_eos = NULL;
    g_step_ceos = NULL;
    return 0;
}

/* fixup_utoprim + bound_prim on the current state and flags (fixup.c:432-538). */
int orc_debug_fixup_utoprim(orc_ctx *o)
{
    int64_t *cnt = o->cnt;
    fixup_utoprim(o, o->p, cnt);
    bound_prim(o, o->p, &cnt[ORC_FATAL]);
    return 0;
}

/* One primitive array between the oracle's zone-major layout and the canonical
 * one (8 x (N1+4) x (N2+4) x N3); to_canon selects the direction. */
static void canon_copy(orc_ctx *o, double *arr, double *canon, int to_canon)
{
    const size_t nh = o->N3, nj = o->SN2, ni = o->SN1;
    for (size_t i = 0; i < ni; i++)
        for (size_t j = 0; j < nj; j++)
            for (size_t h = 0; h < nh; h++) {
                double *z = PV(o, arr, (int)h, (int)i, (int)j);
                for (int k = 0; k < NPR; k++) {
                    double *c = &canon[(((size_t)k * ni + i) * nj + j) * nh + h];
                    if (to_canon)
                        *c = z[k];
                    else
                        z[k] = *c;
                }
            }
}

/* advance() then fixup(), the first two calls of each half of step_ch
 * (step_ch.c:101-103 and 112-114), without the bound_prim / fixup_utoprim that
 * follow them.  fullstep 0: the predictor p -> ph with dt/2 (p unchanged; ph is
 * returned in the canonical layout).  fullstep 1: the corrector from p and the
 * given ph (canonical layout, ghosts included) -> p with dt. */
int orc_debug_advance(orc_ctx *o, int fullstep, double *ph)
{
    g_step_eos = o->eos;
    g_step_ceos = o->ceos_on ? o : NULL;
    g_step_refd = o->eos_refd;
    if (fullstep) {
        canon_copy(o, o->ph, ph, 0);
        advance(o, o->p, o->ph, o->dt, o->p, 1);
        fixup(o, o->p, o->cnt);
    } else {
        advance(o, o->p, o->p, 0.5 * o->dt, o->ph, 0);
        fixup(o, o->ph, o->cnt);
        canon_copy(o, o->ph, ph, 1);
    }
    o->flux_valid = 0;
    g_step_eos = NULL;
    g_step_ceos = NULL;
    return 0;
}

/* fixup() alone on the current state (fixup.c:51-60): floors, the drift-frame
 * velocity blend and the gamma cap per active zone; flag -33 where gamma fails. */
int orc_debug_fixup(orc_ctx *o)
{
    fixup(o, o->p, o->cnt);
    return 0;
}

/* bound_prim on the current state (bounds.c:54-331). */
int orc_debug_bound(orc_ctx *o)
{
    bound_prim(o, o->p, &o->cnt[ORC_FATAL]);
    return 0;
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

orc_ctx *orc_create(const orc_params *p)
{
    if (p->n1 < 4 || p->n2 < 8 || p->n3 < 1) return NULL;
    if (p->n3 == 2 || (p->n3 > 1 && (p->n3 % 2) != 0)) return NULL;
    orc_ctx *o = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    if (!o) return NULL;
    o->P = *p;
    if (o->P.Rin <= 0.) o->P.Rin = 0.98 * (1. + sqrt(1. - o->P.a * o->P.a));
    o->N1 = p->n1;
    o->N2 = p->n2;
    o->N3 = p->n3;
    o->SN1 = p->n1 + 4;
    o->SN2 = p->n2 + 4;
    o->istart = 2;
    o->istop = p->n1 + 1;
    o->jstart = 2;
    o->jstop = p->n2 + 1;
    o->polefix = (p->n3 > 1 && 2 < p->n2 / 2) ? 2 : 0;
    const size_t nz = (size_t)o->N3 * o->SN1 * o->SN2;
    const size_t ng = (size_t)o->SN1 * o->SN2;
    o->p = calloc(nz * NPR, sizeof(double));
    o->ph = calloc(nz * NPR, sizeof(double));
    o->dq = calloc(nz * NPR, sizeof(double));
    o->F1 = calloc(nz * NPR, sizeof(double));
    o->F2 = calloc(nz * NPR, sizeof(double));
    o->F3 = calloc(nz * NPR, sizeof(double));
    o->e1 = calloc(nz, sizeof(double));
    o->e2 = calloc(nz, sizeof(double));
    o->e3 = calloc(nz, sizeof(double));
    o->pflag = calloc(nz, 1);
    o->gcov = calloc(ng * NPG * 16, sizeof(double));
    o->gcon = calloc(ng * NPG * 16, sizeof(double));
    o->gdet = calloc(ng * NPG, sizeof(double));
    o->conn = calloc(ng * 64, sizeof(double));
    o->coords = calloc(ng * NDIM, sizeof(double));
    o->Katm = calloc(o->SN1, sizeof(double));
    if (!o->p || !o->ph || !o->dq || !o->F1 || !o->F2 || !o->F3 || !o->e1 || !o->e2 || !o->e3 || !o->pflag ||
        !o->gcov || !o->gcon || !o->gdet || !o->conn || !o->coords || !o->Katm || set_grid(o)) {
        orc_destroy(o);
        return NULL;
    }
    o->dt = p->dt0;
    return o;
}

void orc_destroy(orc_ctx *o)
{
    if (!o) return;
    free(o->p); free(o->ph); free(o->dq); free(o->F1); free(o->F2); free(o->F3);
    free(o->e1); free(o->e2); free(o->e3); free(o->pflag);
    free(o->gcov); free(o->gcon); free(o->gdet); free(o->conn); free(o->coords); free(o->Katm);
    free(o->tr_X); free(o->tr_state); free(o->tr_ev); free(o->lastT);
    free(o);
}

int orc_get_state(const orc_ctx *o, double *prim, int8_t *pflag)
{
    const size_t nh = o->N3, nj = o->SN2, ni = o->SN1;
    for (size_t i = 0; i < ni; i++)
        for (size_t j = 0; j < nj; j++)
            for (size_t h = 0; h < nh; h++) {
                const double *z = PV(o, o->p, (int)h, (int)i, (int)j);
                if (prim)
                    for (int k = 0; k < NPR; k++) prim[(((size_t)k * ni + i) * nj + j) * nh + h] = z[k];
                if (pflag) pflag[(i * nj + j) * nh + h] = o->pflag[zi(o, (int)h, (int)i, (int)j)];
            }
    return 0;
}

int orc_set_state(orc_ctx *o, const double *prim, const int8_t *pflag)
{
    const size_t nh = o->N3, nj = o->SN2, ni = o->SN1;
    for (size_t i = 0; i < ni; i++)
        for (size_t j = 0; j < nj; j++)
            for (size_t h = 0; h < nh; h++) {
                double *z = PV(o, o->p, (int)h, (int)i, (int)j);
                if (prim)
                    for (int k = 0; k < NPR; k++) z[k] = prim[(((size_t)k * ni + i) * nj + j) * nh + h];
                if (pflag) o->pflag[zi(o, (int)h, (int)i, (int)j)] = pflag[(i * nj + j) * nh + h];
            }
    o->flux_valid = 0;
    return 0;
}

void orc_set_mass_budget(orc_ctx *o, double mass_start, double mass_in, double mass_out)
{
    o->mass_start = mass_start;
    o->mass_in = mass_in;
    o->mass_out = mass_out;
}

void orc_get_time(const orc_ctx *o, double *t, double *dt)
{
    *t = o->t;
    *dt = o->dt;
}
void orc_set_time(orc_ctx *o, double t, double dt)
{
    o->t = t;
    o->dt = dt;
}
void orc_get_counters(const orc_ctx *o, int64_t *c)
{
    memcpy(c, o->cnt, sizeof(o->cnt));
}
void orc_get_grid(const orc_ctx *o, double *startx, double *dx)
{
    memcpy(startx, o->startx, sizeof(o->startx));
    memcpy(dx, o->dx, sizeof(o->dx));
}
double orc_get_katm(const orc_ctx *o, int i)
{
    return (i >= 0 && i < o->SN1) ? o->Katm[i] : 0.;
}
void orc_get_geometry(const orc_ctx *o, double *gcov, double *gcon, double *gdet, double *conn, double *coords)
{
    const size_t ng = (size_t)o->SN1 * o->SN2;
    if (gcov) memcpy(gcov, o->gcov, ng * NPG * 16 * sizeof(double));
    if (gcon) memcpy(gcon, o->gcon, ng * NPG * 16 * sizeof(double));
    if (gdet) memcpy(gdet, o->gdet, ng * NPG * sizeof(double));
    if (conn) memcpy(conn, o->conn, ng * 64 * sizeof(double));
    if (coords) memcpy(coords, o->coords, ng * NDIM * sizeof(double));
}

int orc_utoprim_2d(const double U[8], const double *gcov, const double *gcon, double gdet, double prim[8],
                   double gam, int *n_iter, int *k_conv)
{
    double gc[NDIM][NDIM], gn[NDIM][NDIM];
    memcpy(gc, gcov, sizeof(gc));
    memcpy(gn, gcon, sizeof(gn));
    return Utoprim_2d(U, gc, gn, gdet, prim, gam, n_iter, k_conv);
}
int orc_utoprim_1dvsq2fix1(const double U[8], const double *gcov, const double *gcon, double gdet, double prim[8],
                           double K, double gam)
{
    double gc[NDIM][NDIM], gn[NDIM][NDIM];
    memcpy(gc, gcov, sizeof(gc));
    memcpy(gn, gcon, sizeof(gn));
    return Utoprim_1dvsq2fix1(U, gc, gn, gdet, prim, K, gam);
}
int orc_utoprim_1dfix1(const double U[8], const double *gcov, const double *gcon, double gdet, double prim[8],
                       double K, double gam)
{
    double gc[NDIM][NDIM], gn[NDIM][NDIM];
    memcpy(gc, gcov, sizeof(gc));
    memcpy(gn, gcon, sizeof(gn));
    return Utoprim_1dfix1(U, gc, gn, gdet, prim, K, gam);
}
void orc_primtoU_g(const double prim[8], const double *gcov, const double *gcon, double gdet, double gam,
                   double U[8])
{
    double gc[NDIM][NDIM], gn[NDIM][NDIM];
    memcpy(gc, gcov, sizeof(gc));
    memcpy(gn, gcon, sizeof(gn));
    primtoU_g(prim, gc, gn, gdet, gam, U);
}
void orc_primtoU(const double prim[8], const double *gcov, const double *gcon, double gdet, double gam, double U[8])
{
    geom_t geom;
    state_t q;
    int64_t fatal = 0;
    memcpy(geom.gcov, gcov, sizeof(geom.gcov));
    memcpy(geom.gcon, gcon, sizeof(geom.gcon));
    geom.g = gdet;
    get_state(prim, &geom, &q, &fatal);
    primtoflux(prim, &q, 0, &geom, gam, U);
}
void orc_gcov_func(const orc_params *p, const double X[4], double *gcov)
{
    gcov_func(p, X, (double(*)[NDIM])gcov);
}
void orc_dxdxp_func(const orc_params *p, const double X[4], double *dxdxp)
{
    dxdxp_func(p, X, (double(*)[NDIM])dxdxp);
}
void orc_bl_coord_vec(const orc_params *p, const double X[4], double *V)
{
    bl_coord_vec(p, X, V);
}
void orc_conn_func(const orc_params *p, const double X[4], const double *gcon, double *conn)
{
    double gn[NDIM][NDIM];
    memcpy(gn, gcon, sizeof(gn));
    conn_func(p, X, gn, (double(*)[NDIM][NDIM])conn);
}
double orc_gdet_func(const double *gcov)
{
    double gc[NDIM][NDIM];
    memcpy(gc, gcov, sizeof(gc));
    return gdet_func(gc, NULL);
}
int orc_invert_matrix(const double *A, double *Ainv)
{
    double a[NDIM][NDIM];
    memcpy(a, A, sizeof(a));
    return invert_matrix(a, (double(*)[NDIM])Ainv);
}
double orc_slope_lim(int lim, double y1, double y2, double y3)
{
    return slope_lim(lim, y1, y2, y3);
}
/* random.c:44-55: one KISS draw mapped to [0,1] by UniformRandomRange(0,1) */
double orc_kiss_uniform(uint32_t z, uint32_t w, uint32_t jsr, uint32_t jcong)
{
    uint32_t mwc, cong, shr3, kiss;
    z = 36969u * (z & 65535u) + (z >> 16);
    w = 18000u * (w & 65535u) + (w >> 16);
    mwc = (z << 16) + w;
    cong = jcong = 69069u * jcong + 1234567u;
    jsr ^= (jsr << 17);
    jsr ^= (jsr >> 13);
    jsr ^= (jsr << 5);
    shr3 = jsr;
    kiss = (mwc ^ cong) + shr3;
    double u = ((double)(kiss)) / 0xFFFFFFFF;
    return 0.0 + (1.0 - 0.0) * u;
}
void orc_face_flux(const double *p_l, const double *p_r, const double *gcov, const double *gcon, double gdet, int dir,
                   double gam, double hllf, double laxf, double *F, double *ctop)
{
    geom_t geom;
    int64_t fatal = 0;
    memcpy(geom.gcov, gcov, sizeof(geom.gcov));
    memcpy(geom.gcon, gcon, sizeof(geom.gcon));
    geom.g = gdet;
    face_flux(p_l, p_r, &geom, dir, gam, hllf, laxf, F, ctop, &fatal);
}
int orc_vchar(const double *pr, const double *gcov, const double *gcon, double gdet, int dir, double gam,
              double *cmax, double *cmin)
{
    geom_t geom;
    state_t q;
    int64_t fatal = 0;
    memcpy(geom.gcov, gcov, sizeof(geom.gcov));
    memcpy(geom.gcon, gcon, sizeof(geom.gcon));
    geom.g = gdet;
    get_state(pr, &geom, &q, &fatal);
    return vchar(pr, &q, &geom, dir, gam, cmax, cmin) + (int)fatal;
}

/* ------------------------------------------------------------------------ */
/* Learned Newton seed (no reference counterpart; its seed replaces the      */
/* guess block utoprim_2d.c:268-307 and nothing else).                       */
/* ------------------------------------------------------------------------ */
/* The eight raw inputs of the seed network, from the scalars of
 * utoprim_2d.c:240-266 and the lapse of utoprim_2d.c:145.  All ratios are
 * taken to D so that the features do not change when U is rescaled. */
int orc_u2p_features(const double U[8], const double *gcov, const double *gcon, double gdet, double feat[8])
{
    double gc[NDIM][NDIM], gn[NDIM][NDIM], U_tmp[NPR], prim_tmp[NPR], prim0[NPR] = {0};
    double Bcon[NDIM], Qcon[NDIM], ncon[NDIM], QdotB, lapse, D, E;
    u2p_t s;
    int i;
    if (U[0] <= 0.) {
        for (i = 0; i < 8; i++) feat[i] = 0.;
        return -100;
    }
    memcpy(gc, gcov, sizeof(gc));
    memcpy(gn, gcon, sizeof(gn));
    memset(&s, 0, sizeof(s));
    u2p_transform(U, gn, gdet, prim0, U_tmp, prim_tmp);
    u2p_scalars(&s, U_tmp, gc, gn, Bcon, Qcon, ncon, &QdotB);
    lapse = 1.0 / sqrt(-gn[0][0]);
    D = s.D;
    E = -s.Qdotn;
    feat[0] = log(D);
    feat[1] = asinh(E / D);
    feat[2] = asinh(s.Qtsq / (D * D));
    feat[3] = asinh(s.Bsq / D);
    feat[4] = asinh(s.QdotBsq / (D * D * D));
    feat[5] = lapse;
    feat[6] = asinh((E - D - 0.5 * s.Bsq) / D);
    feat[7] = asinh(s.Bsq / E);
    return 0;
}

/* Noble 2D inversion (utoprim_2d.c:126-176,219-362) started from the pair
 * (W0, vsq0) instead of the guess that utoprim_2d.c:268-307 builds from prim.
 * x = (|W0|, vsq): with vsq_from_W, vsq = x1_of_x0(W0) as utoprim_2d.c:306-307
 * sets it from a W guess; otherwise vsq0 clipped to [0, 1-1e-15].  Return codes, prim[B] and
 * the success-only write of prim[0..4] follow Utoprim_2d. */
int orc_utoprim_2d_seeded(const double U[8], const double *gcov, const double *gcon, double gdet, double prim[8],
                          double gam, double W0, double vsq0, int vsq_from_W, int *n_iter, int *k_conv)
{
    double gc[NDIM][NDIM], gn[NDIM][NDIM], U_tmp[NPR], prim_tmp[NPR], x_2d[2];
    double Bcon[NDIM], Qcon[NDIM], ncon[NDIM], Qtcon[NDIM], QdotB;
    double W, vsq, gtmp, gamma, rho0, w, p, u;
    const double GAMMA = gam;
    u2p_t s;
    int i, retval;
    memcpy(gc, gcov, sizeof(gc));
    memcpy(gn, gcon, sizeof(gn));
    memset(&s, 0, sizeof(s));
    s.gam = gam;
    if (n_iter) *n_iter = 0;
    if (k_conv) *k_conv = 0;
    for (i = B1; i <= B3; i++) prim[i] = U[i] / gdet;
    if (U[0] <= 0.) return -100;
    u2p_transform(U, gn, gdet, prim, U_tmp, prim_tmp);
    u2p_scalars(&s, U_tmp, gc, gn, Bcon, Qcon, ncon, &QdotB);
    x_2d[0] = fabs(W0);
    if (vsq_from_W)
        x_2d[1] = x1_of_x0(&s, W0);
    else
        x_2d[1] = (vsq0 < 0.) ? 0. : ((vsq0 > 1. - 1.e-15) ? (1. - 1.e-15) : vsq0);
    retval = gnr_2d(&s, x_2d, n_iter, k_conv);
    W = x_2d[0];
    vsq = x_2d[1];
    if ((retval != 0) || (W == FAIL_VAL)) return retval * 100 + 1;
    if (W <= 0. || W > W_TOO_BIG) return 3;
    if (vsq >= 1.) return 4;
    gtmp = sqrt(1. - vsq);
    gamma = 1. / gtmp;
    rho0 = s.D * gtmp;
    w = W * (1. - vsq);
    p = (GAMMA - 1.) * (w - rho0) / GAMMA;
    u = w - (rho0 + p);
    if ((rho0 <= 0.) || (u <= 0.)) return 5;
    prim[RHO] = rho0;
    prim[UU] = u;
    for (i = 1; i < 4; i++) Qtcon[i] = Qcon[i] + ncon[i] * s.Qdotn;
    for (i = 1; i < 4; i++) prim[U1 + i - 1] = gamma / (W + s.Bsq) * (Qtcon[i] + QdotB * Bcon[i] / W);
    return 0;
}

/* The learned-seed solve: seeded Newton, then, when it fails, Utoprim_2d from
 * the previous-step primitives in prim.  n_iter/k_conv add both attempts. */
int orc_utoprim_2d_learned(const double U[8], const double *gcov, const double *gcon, double gdet,
                           double prim[8], double gam, double W0, double vsq0, int vsq_from_W, int *n_iter,
                           int *k_conv, int *fell_back)
{
    double pt[8];
    int n1 = 0, k1 = 0, n2 = 0, k2 = 0, ret, i;
    for (i = 0; i < 8; i++) pt[i] = prim[i];
    ret = orc_utoprim_2d_seeded(U, gcov, gcon, gdet, pt, gam, W0, vsq0, vsq_from_W, &n1, &k1);
    *fell_back = 0;
    if (ret == 0 || ret == -100) {
        for (i = 0; i < 8; i++) prim[i] = pt[i];
        *n_iter = n1;
        *k_conv = k1;
        return ret;
    }
    *fell_back = 1;
    ret = orc_utoprim_2d(U, gcov, gcon, gdet, prim, gam, &n2, &k2);
    *n_iter = n1 + n2;
    *k_conv = n1 + k2;
    return ret;
}

/* The gated learned solve: Utoprim_2d from the previous-step primitives in
 * prim, except for a zone that has no valid start from them (code 2) or whose
 * first Newton step |dW/W| + |dv2/v2| exceeds tau: that zone solves from the
 * seed (W0, vsq0), and where the seeded solve fails it takes the plain result.
 * Counts: the first plain iteration counts once for a seeded zone; a failed
 * seeded attempt adds its iterations to the plain ones. */
int orc_utoprim_2d_gated(const double U[8], const double *gcov, const double *gcon, double gdet, double prim[8],
                         double gam, double W0, double vsq0, int vsq_from_W, double tau, int *n_iter, int *k_conv,
                         int *seeded)
{
    double gc[NDIM][NDIM], gn[NDIM][NDIM], pp[8], pt[8], errx1 = 0.;
    int n0 = 0, k0 = 0, n1 = 0, k1 = 0, r0, r1, i, need;
    memcpy(gc, gcov, sizeof(gc));
    memcpy(gn, gcon, sizeof(gn));
    for (i = 0; i < 8; i++) pp[i] = pt[i] = prim[i];
    r0 = Utoprim_2d_first(U, gc, gn, gdet, pp, gam, &n0, &k0, &errx1);
    *seeded = 0;
    need = (r0 == 2 && n0 == 0) || (r0 != -100 && n0 > 0 && errx1 > tau);
    if (need) {
        r1 = orc_utoprim_2d_seeded(U, gcov, gcon, gdet, pt, gam, W0, vsq0, vsq_from_W, &n1, &k1);
        if (r1 == 0) {
            const int first = (n0 > 0) ? 1 : 0;
            for (i = 0; i < 8; i++) prim[i] = pt[i];
            *n_iter = first + n1;
            *k_conv = first + k1;
            *seeded = 1;
            return 0;
        }
        if (n0 == 0) {
            for (i = 0; i < 8; i++) prim[i] = pp[i];
            *n_iter = n1;
            *k_conv = n1;
            return r0;
        }
    }
    for (i = 0; i < 8; i++) prim[i] = pp[i];
    *n_iter = n1 + n0;
    *k_conv = n1 + k0;
    return r0;
}

/* ------------------------------------------------------------------------ */
/* Tabulated equation of state and the 5D inversion (utoprim_5d.c).          */
/* ------------------------------------------------------------------------ */
/* The table: rho[nr] and T[nt] increasing (log-spaced in practice), e and p
 * [nr][nt] in code units (e = internal energy density u).  p(rho, u) is the
 * piecewise-bilinear interpolant in (rho, T) inverted along T at fixed rho:
 * E_j(rho) = e[k][j] + a (e[k+1][j] - e[k][j]) with a the rho weight of cell k,
 * j such that E_j <= u <= E_{j+1}, b = (u - E_j)/(E_{j+1} - E_j),
 * p = P_j + b (P_{j+1} - P_j).  No transcendental function is involved, so the
 * device evaluates it with identical arithmetic.  Outside the table the
 * Gamma-law pressure is used, as the reference does outside its table
 * (cooleos.c:342-366).  Returns 1 when the Gamma-law branch was taken. */
static int cell_of(const double *x, int n, double v)
{
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (x[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

int orc_eos_table_p(const orc_eos *t, double rho, double u, double *p, double *dpdrho, double *dpdu, double *T)
{
    const int nr = t->nr, nt = t->nt;
    int k, lo, hi;
    double dr, a, Elo, Ehi, E0, E1, P0, P1, b, dE, dP, sE0, sE1, sP0, sP1;
    if (!(rho >= t->rho[0] && rho <= t->rho[nr - 1])) goto gamma_law;
    k = cell_of(t->rho, nr, rho);
    dr = t->rho[k + 1] - t->rho[k];
    a = (rho - t->rho[k]) / dr;
#define EAT(j) (t->e[k * nt + (j)] + a * (t->e[(k + 1) * nt + (j)] - t->e[k * nt + (j)]))
#define PAT(j) (t->p[k * nt + (j)] + a * (t->p[(k + 1) * nt + (j)] - t->p[k * nt + (j)]))
    Elo = EAT(0);
    Ehi = EAT(nt - 1);
    if (!(u >= Elo && u <= Ehi)) goto gamma_law;
    lo = 0;
    hi = nt - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (EAT(mid) <= u)
            lo = mid;
        else
            hi = mid;
    }
    E0 = EAT(lo);
    E1 = EAT(lo + 1);
    P0 = PAT(lo);
    P1 = PAT(lo + 1);
    dE = E1 - E0;
    dP = P1 - P0;
    b = (u - E0) / dE;
    *p = P0 + b * dP;
    if (dpdu) *dpdu = dP / dE;
    if (dpdrho) {
        sE0 = (t->e[(k + 1) * nt + lo] - t->e[k * nt + lo]) / dr;
        sE1 = (t->e[(k + 1) * nt + lo + 1] - t->e[k * nt + lo + 1]) / dr;
        sP0 = (t->p[(k + 1) * nt + lo] - t->p[k * nt + lo]) / dr;
        sP1 = (t->p[(k + 1) * nt + lo + 1] - t->p[k * nt + lo + 1]) / dr;
        *dpdrho = sP0 + (-sE0 * dP * dE + (u - E0) * ((sP1 - sP0) * dE - dP * (sE1 - sE0))) / (dE * dE);
    }
    if (T) *T = t->T[lo] + b * (t->T[lo + 1] - t->T[lo]);
#undef EAT
#undef PAT
    return 0;
gamma_law:
    *p = (t->gam - 1.) * u;
    if (dpdrho) *dpdrho = 0.;
    if (dpdu) *dpdu = t->gam - 1.;
    if (T) *T = 0.;
    return 1;
}

/* EOS seen by the 5D solver: t == NULL is the Gamma-law pressure_rho0_u
 * (utoprim_2d.c:645-650 / u2p_util.c). */
static double eos5_p(const orc_eos *t, double gam, double rho, double u)
{
    double p;
    if (!t) return (gam - 1.) * u;
    orc_eos_table_p(t, rho, u, &p, NULL, NULL, NULL);
    return p;
}

typedef struct {
    double U_target[NPR];
    double gcov[NDIM][NDIM], gcon[NDIM][NDIM], gdet, gam;
    const orc_eos *eos;
    int ceos; /* the step's cache EOS (the zone InitCache_CoolEOS selected) */
} u5_t;

static double eos5_ps(const u5_t *s, double rho, double u)
{
    if (s->ceos) return ceos_pres(rho, u);
    return eos5_p(s->eos, s->gam, rho, u);
}

/* primtoU_g (u2p_util.c:84-145) with the EOS of the solver. */
static void primtoU_g5(const u5_t *s, const double prim[NPR], double U[NPR])
{
    int i;
    double ucon[NDIM], ucov[NDIM], bcon[NDIM], bcov[NDIM], ncov[NDIM];
    double rho0, gamma, n_dot_b, bsq, u, p, w, alpha;
    double gc[NDIM][NDIM], gn[NDIM][NDIM];
    memcpy(gc, s->gcov, sizeof(gc));
    memcpy(gn, s->gcon, sizeof(gn));
    alpha = 1.0 / sqrt(-gn[0][0]);
    ucon_calc_g(prim, gc, gn, ucon);
    lower_g(ucon, gc, ucov);
    ncov_calc(gn, ncov);
    gamma = -ncov[0] * ucon[0];
    bcon_calc_g(prim, ucon, ucov, ncov, bcon);
    lower_g(bcon, gc, bcov);
    n_dot_b = 0.;
    for (i = 0; i < 4; i++) n_dot_b += ncov[i] * bcon[i];
    bsq = 0.;
    for (i = 0; i < 4; i++) bsq += bcov[i] * bcon[i];
    rho0 = prim[RHO];
    u = prim[UU];
    p = eos5_ps(s, rho0, u);
    w = rho0 + u + p;
    U[RHO] = ucon[0] * rho0;
    for (i = 0; i < 4; i++) {
        U[1 + i] = gamma * (w + bsq) * ucov[i] - (p + bsq / 2.) * ncov[i] + n_dot_b * bcov[i];
        U[1 + i] /= alpha;
    }
    U[1] = U[1] + U[RHO];
    U[B1] = prim[B1];
    U[B2] = prim[B2];
    U[B3] = prim[B3];
    for (i = 0; i < NPR; i++) U[i] *= s->gdet;
}

/* gamma_calc_g (u2p_util.c:319-341); the reference leaves gamma unset on
 * failure, here it is sqrt(1 + |utsq|). */
static int gamma_calc_g5(const double *pr, double gcov[NDIM][NDIM], double *gamma)
{
    double utsq = gcov[1][1] * pr[U1] * pr[U1] + gcov[2][2] * pr[U2] * pr[U2] + gcov[3][3] * pr[U3] * pr[U3] +
                  2. * (gcov[1][2] * pr[U1] * pr[U2] + gcov[1][3] * pr[U1] * pr[U3] + gcov[2][3] * pr[U2] * pr[U3]);
    int bad = 0;
    if (utsq < 0.0) {
        if (fabs(utsq) > 1.E-10) bad = 1;
        utsq = fabs(utsq);
    }
    *gamma = sqrt(1. + utsq);
    return bad;
}

#define N5 5
/* dudp_calc_g (utoprim_5d.c:409-529) and its helpers (utoprim_5d.c:531-640). */
static void dudp_calc_g5(const u5_t *s, const double *pr, double Am[N5][N5])
{
    double dutdui[NDIM], dbtdui[NDIM], dbsqdui[NDIM], dbiduj[NDIM][NDIM], dgdvi[NDIM], duidvj[NDIM][NDIM];
    double dudduu[NDIM][NDIM], dbdiduj[NDIM][NDIM], tempA[N5][N5];
    double bcon[NDIM], bcov[NDIM], ucov[NDIM], ucon[NDIM], ncov[NDIM], B[NDIM], Bcov[NDIM];
    double eta, bsq, gamma, alpha, betacon[NDIM];
    double gc[NDIM][NDIM], gn[NDIM][NDIM];
    int i, j, k, l;
    memcpy(gc, s->gcov, sizeof(gc));
    memcpy(gn, s->gcon, sizeof(gn));
    for (j = 0; j < N5; j++)
        for (k = 0; k < N5; k++) tempA[j][k] = Am[j][k] = 0.;
    ucon_calc_g(pr, gc, gn, ucon);
    lower_g(ucon, gc, ucov);
    ncov_calc(gn, ncov);
    bcon_calc_g(pr, ucon, ucov, ncov, bcon);
    lower_g(bcon, gc, bcov);
    bsq = bcon[0] * bcov[0] + bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
    for (j = 1; j < NDIM; j++) dutdui[j] = -ucov[j] / ucov[0];
    B[0] = 0.;
    for (j = 1; j < NDIM; j++) B[j] = pr[j + B1 - 1];
    lower_g(B, gc, Bcov);
    for (j = 1; j < NDIM; j++) dbtdui[j] = Bcov[j] + Bcov[0] * dutdui[j];
    for (j = 0; j < NDIM; j++)
        for (k = 0; k < NDIM; k++) dbiduj[j][k] = 0.;
    for (j = 1; j < NDIM; j++)
        for (k = 1; k < NDIM; k++) dbiduj[j][k] = -bcon[j] * dutdui[k] / ucon[0] + ucon[j] * dbtdui[k] / ucon[0];
    for (j = 1; j < NDIM; j++) dbiduj[j][j] += bcon[0] / ucon[0];
    for (j = 1; j < NDIM; j++) dbiduj[0][j] = dbtdui[j];
    for (j = 0; j < NDIM; j++) dbsqdui[j] = 0.;
    for (j = 0; j < NDIM; j++)
        for (k = 0; k < NDIM; k++) dbsqdui[j] += 2. * bcov[k] * dbiduj[k][j];
    for (j = 0; j < NDIM; j++)
        for (k = 1; k < NDIM; k++) dudduu[j][k] = gc[j][k] + gc[j][0] * dutdui[k];
    for (j = 0; j < NDIM; j++)
        for (k = 1; k < NDIM; k++) dbdiduj[j][k] = 0.0;
    for (j = 0; j < NDIM; j++)
        for (k = 1; k < NDIM; k++)
            for (l = 0; l < NDIM; l++) dbdiduj[j][k] += gc[j][l] * dbiduj[l][k];
    if (s->eos || s->ceos) {
        double p, dpdrho, dpdu;
        if (s->ceos) {
            /* utoprim_5d.c:454-481: eta's pressure, then dp/drho, then dp/du */
            p = ceos_pres(pr[RHO], pr[UU]);
            dpdrho = ceos_dpdrho(pr[RHO], pr[UU]);
            dpdu = ceos_dpdu(pr[RHO], pr[UU]);
        } else {
            orc_eos_table_p(s->eos, pr[RHO], pr[UU], &p, &dpdrho, &dpdu, NULL);
        }
        eta = pr[RHO] + pr[UU] + p + bsq;
        Am[RHO][RHO] = ucon[0];
        for (j = 0; j < NDIM; j++) Am[UU + j][RHO] = (1.0 + dpdrho) * ucon[0] * ucov[j] + (j == 0 ? 1. : 0.) * dpdrho;
        Am[RHO][UU] = 0.0;
        for (j = 0; j < NDIM; j++) Am[UU + j][UU] = (1.0 + dpdu) * ucon[0] * ucov[j] + (j == 0 ? 1. : 0.) * dpdu;
    } else {
        const double GAMMA = s->gam;
        eta = pr[RHO] + GAMMA * pr[UU] + bsq;
        Am[RHO][RHO] = ucon[0];
        for (j = 0; j < NDIM; j++) Am[UU + j][RHO] = ucon[0] * ucov[j];
        Am[RHO][UU] = 0.0;
        for (j = 0; j < NDIM; j++)
            Am[UU + j][UU] = GAMMA * ucon[0] * ucov[j] + (j == 0 ? 1. : 0.) * (GAMMA - 1.0);
    }
    for (j = 1; j < NDIM; j++) Am[RHO][U1 + j - 1] = pr[RHO] * dutdui[j];
    for (j = 0; j < NDIM; j++)
        for (k = 1; k < NDIM; k++)
            Am[UU + j][U1 + k - 1] = dbsqdui[k] * ucon[0] * ucov[j] +
                                    eta * (dutdui[k] * ucov[j] + ucon[0] * dudduu[j][k]) +
                                    (j == 0 ? 1. : 0.) * 0.5 * dbsqdui[k] -
                                    (dbtdui[k] * bcov[j] + bcon[0] * dbdiduj[j][k]);
    for (k = 0; k < N5; k++) Am[UU][k] += Am[RHO][k];
    gamma_calc_g5(pr, gc, &gamma);
    for (j = 1; j < NDIM; j++) dgdvi[j] = 0.0;
    for (j = 1; j < NDIM; j++)
        for (k = 1; k < NDIM; k++) dgdvi[j] += 1.0 / gamma * gc[j][k] * pr[U1 + k - 1];
    alpha = 1.0 / sqrt(-gn[0][0]);
    for (j = 1; j < NDIM; j++) betacon[j] = gn[0][j] * alpha * alpha;
    for (j = 1; j < NDIM; j++) duidvj[0][j] = dgdvi[j] / alpha;
    for (j = 1; j < NDIM; j++)
        for (k = 1; k < NDIM; k++) duidvj[j][k] = (j == k ? 1. : 0.) - betacon[j] / alpha * dgdvi[k];
    for (i = 0; i < N5; i++)
        for (j = RHO; j <= UU; j++) tempA[i][j] = Am[i][j];
    for (i = 0; i < N5; i++)
        for (j = 1; j < NDIM; j++)
            for (l = 1; l <= 3; l++) tempA[RHO + i][U1 + j - 1] += Am[RHO + i][U1 + l - 1] * duidvj[l][j];
    for (i = 0; i < N5; i++)
        for (j = 0; j < N5; j++) Am[i][j] = tempA[i][j];
    for (j = 0; j < N5; j++)
        for (k = 0; k < N5; k++) Am[j][k] *= s->gdet;
}

/* LU_decompose / LU_substitution (utoprim_5d.c:672-874): Crout with implicit
 * partial pivoting. */
static int lu5_decompose(double A[N5][N5], int permute[N5])
{
    const double absmin = 1e-30;
    double row_norm[N5], absmax, maxtemp;
    int i, j, k, max_row = 0, n = N5;
    for (i = 0; i < n; i++) {
        absmax = 0.;
        for (j = 0; j < n; j++) {
            maxtemp = fabs(A[i][j]);
            if (maxtemp > absmax) absmax = maxtemp;
        }
        if (absmax == 0.) return 1;
        row_norm[i] = 1. / absmax;
    }
    for (j = 0; j < n; j++) {
        for (i = 0; i < j; i++)
            for (k = 0; k < i; k++) A[i][j] -= A[i][k] * A[k][j];
        absmax = 0.0;
        for (i = j; i < n; i++) {
            for (k = 0; k < j; k++) A[i][j] -= A[i][k] * A[k][j];
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
                for (k = 0; k < n; k++) {
                    maxtemp = A[j][k];
                    A[j][k] = A[max_row][k];
                    A[max_row][k] = maxtemp;
                }
                row_norm[max_row] = row_norm[j];
            }
        }
        permute[j] = max_row;
        if (A[j][j] == 0.) A[j][j] = absmin;
        if (j != (n - 1)) {
            maxtemp = 1. / A[j][j];
            for (i = (j + 1); i < n; i++) A[i][j] *= maxtemp;
        }
    }
    return 0;
}
static void lu5_substitution(double A[N5][N5], double B[N5], const int permute[N5])
{
    int i, j, n = N5;
    double tmpvar;
    for (i = 0; i < n; i++) {
        tmpvar = B[permute[i]];
        B[permute[i]] = B[i];
        for (j = (i - 1); j >= 0; j--) tmpvar -= A[i][j] * B[j];
        B[i] = tmpvar;
    }
    for (i = (n - 1); i >= 0; i--) {
        for (j = (i + 1); j < n; j++) B[i] -= A[i][j] * B[j];
        B[i] /= A[i][i];
    }
}

/* func_5d (utoprim_5d.c:250-359), NORMMETHOD 1. */
static void func_5d(const u5_t *s, const double prguess[], double dxm[], double *f, double *df)
{
    double pr[NPR], U_curr[NPR], Am[N5][N5], beta[N5], norm, residnorm, resid;
    int i, j, k, permute[N5], numnormterms;
    for (i = 0; i < N5; i++) permute[i] = 0;
    for (i = 0; i < B1; i++) pr[i] = prguess[i];
    for (i = B1; i <= B3; i++) pr[i] = s->U_target[i] / s->gdet;
    primtoU_g5(s, pr, U_curr);
    dudp_calc_g5(s, pr, Am);
    norm = 0.0;
    numnormterms = 0;
    for (j = 0; j < N5; j++)
        for (k = 0; k < N5; k++)
            if (Am[j][k] > NUMEPS5) {
                norm += fabs(Am[j][k]);
                numnormterms++;
            }
    norm = ((double)(numnormterms)) / (norm);
    for (j = 0; j < N5; j++)
        for (k = 0; k < N5; k++) Am[j][k] *= (norm);
    *df = 0.;
    for (k = 0; k < N5; k++) {
        resid = U_curr[k] - s->U_target[k];
        beta[k] = -resid * (norm);
        residnorm = 0.5 * (fabs(U_curr[k]) + fabs(s->U_target[k]));
        residnorm = (residnorm == 0.) ? 1. : 1. / residnorm;
        *df -= residnorm * resid * resid;
    }
    *f = -0.5 * (*df);
    for (i = 0; i < N5; i++) dxm[i] = beta[i];
    lu5_decompose(Am, permute);
    lu5_substitution(Am, dxm, permute);
}

/* calc_errx_5d (utoprim_5d.c:378-400). */
static void calc_errx_5d(const u5_t *s, double *W, double *vsq, const double x[])
{
    int id, jd;
    double gsqtmp;
    *vsq = 0.;
    for (id = 1; id < 4; id++)
        for (jd = 1; jd < 4; jd++) *vsq += s->gcov[id][jd] * x[U1 + id - 1] * x[U1 + jd - 1];
    gsqtmp = 1. + *vsq;
    *vsq /= gsqtmp;
    if (s->eos || s->ceos)
        *W = gsqtmp * (x[UU] + x[RHO] + eos5_ps(s, x[RHO], x[UU]));
    else
        *W = gsqtmp * (s->gam * x[UU] + x[RHO]);
}

/* general_newton_raphson (utoprim_5d.c:890-983). */
static int gnr_5d(const u5_t *s, double x[N5], int *n_iter_out)
{
    double f = 1., df = 1., dx[N5], errx = 1., W = 0., W_old, vsq = 0.;
    int n_iter = 0, id, i_extra = 0, doing_extra = 0, keep = 1;
    while (keep) {
        func_5d(s, x, dx, &f, &df);
        for (id = 0; id < N5; id++) x[id] += dx[id];
        W_old = W;
        calc_errx_5d(s, &W, &vsq, x);
        errx = (W == 0.) ? fabs(W - W_old) : fabs((W - W_old) / W);
        if ((fabs(errx) <= NEWT_TOL) && (f <= TOL_RESID5) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0))
            doing_extra = 1;
        if (doing_extra == 1) i_extra++;
        if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
            (n_iter >= (MAX_NEWT_ITER - 1)))
            keep = 0;
        n_iter++;
    }
    if (n_iter_out) *n_iter_out = n_iter;
    if ((isfinite(f) == 0) || (isfinite(df) == 0)) return 2;
    if ((fabs(errx) > MIN_NEWT_TOL) || (f > TOL_RESID5)) return 1;
    return 0;
}

/* Utoprim_5d (utoprim_5d.c:179-239): U is HARM's conserved vector (x gdet),
 * prim the guess in and the result out (B always, the rest only on success). */
static int utoprim_5d_s(const double U[8], const double *gcov, const double *gcon, double gdet, double prim[8],
                        double gam, const orc_eos *eos, int ceos, int *n_iter)
{
    u5_t s;
    double prim_tmp[NPR], gamma_tmp;
    int k, ret;
    memcpy(s.gcov, gcov, sizeof(s.gcov));
    memcpy(s.gcon, gcon, sizeof(s.gcon));
    s.gdet = gdet;
    s.gam = gam;
    s.eos = eos;
    s.ceos = ceos;
    for (k = 0; k < NPR; k++) s.U_target[k] = U[k];
    for (k = B1; k <= B3; k++) prim[k] = U[k] / gdet;
    for (k = 0; k < NPR; k++) prim_tmp[k] = prim[k];
    ret = gnr_5d(&s, prim_tmp, n_iter);
    if (ret) return 100 * ret + 1;
    if ((prim_tmp[RHO] < 0.0) || (prim_tmp[UU] < 0.0)) return 5;
    if (gamma_calc_g5(prim_tmp, s.gcov, &gamma_tmp) == 1) return 4;
    for (k = 0; k < NPR; k++) prim[k] = prim_tmp[k];
    return 0;
}

int orc_utoprim_5d(const double U[8], const double *gcov, const double *gcon, double gdet, double prim[8], double gam,
                   const orc_eos *eos, int *n_iter)
{
    return utoprim_5d_s(U, gcov, gcon, gdet, prim, gam, eos, 0, n_iter);
}

/* primtoU_g with the solver's EOS (for round trips and tests). */
void orc_primtoU_eos(const double prim[8], const double *gcov, const double *gcon, double gdet, double gam,
                     const orc_eos *eos, double U[8])
{
    u5_t s;
    memcpy(s.gcov, gcov, sizeof(s.gcov));
    memcpy(s.gcon, gcon, sizeof(s.gcon));
    s.gdet = gdet;
    s.gam = gam;
    s.eos = eos;
    s.ceos = 0;
    primtoU_g5(&s, prim, U);
}

/* dudp_calc_g's Jacobian (5x5, row-major) for tests against finite differences. */
void orc_dudp_5d(const double prim[8], const double *gcov, const double *gcon, double gdet, double gam,
                 const orc_eos *eos, double *Am)
{
    u5_t s;
    double A[N5][N5];
    memcpy(s.gcov, gcov, sizeof(s.gcov));
    memcpy(s.gcon, gcon, sizeof(s.gcon));
    s.gdet = gdet;
    s.gam = gam;
    s.eos = eos;
    s.ceos = 0;
    dudp_calc_g5(&s, prim, A);
    memcpy(Am, A, sizeof(A));
}

/* ------------------------------------------------------------------------ */
/* Restart file, restart.c:73-383 (and cooleos.c:996-1030).                 */
/* ------------------------------------------------------------------------ */
#define ORC_FMT_DBL_OUT "%32.22e" /* decs.h:175 */
#define ORC_FMT_INT_OUT "%10d"    /* decs.h:176 */

static void restart_write_tracers(FILE *fp, const orc_ctx *o);

int orc_restart_write(const char *path, const orc_restart_hdr *h, const double *prim, const double *temp)
{
    return orc_restart_write_tr(path, h, prim, temp, NULL);
}

int orc_restart_write_tr(const char *path, const orc_restart_hdr *h, const double *prim, const double *temp,
                         const orc_ctx *o)
{
    const int sn1 = h->N1 + 4, sn2 = h->N2 + 4, sn3 = h->N3;
    FILE *fp = fopen(path, "wt");
    int hh, i, j, k;
    if (!fp) return 2;
    /* restart.c:105-138 */
    fprintf(fp, ORC_FMT_INT_OUT, h->N1);
    fprintf(fp, ORC_FMT_INT_OUT, h->N2);
    fprintf(fp, ORC_FMT_INT_OUT, h->N3);
    fprintf(fp, ORC_FMT_DBL_OUT, h->t);
    fprintf(fp, ORC_FMT_DBL_OUT, h->tf);
    fprintf(fp, ORC_FMT_INT_OUT, h->nstep);
    fprintf(fp, ORC_FMT_DBL_OUT, h->a);
    fprintf(fp, ORC_FMT_DBL_OUT, h->gam);
    fprintf(fp, ORC_FMT_DBL_OUT, h->cour);
    fprintf(fp, ORC_FMT_DBL_OUT, h->DTd);
    fprintf(fp, ORC_FMT_DBL_OUT, h->DTl);
    fprintf(fp, ORC_FMT_DBL_OUT, h->DTi);
    fprintf(fp, ORC_FMT_INT_OUT, h->DTr);
    fprintf(fp, ORC_FMT_INT_OUT, h->dump_cnt);
    fprintf(fp, ORC_FMT_INT_OUT, h->image_cnt);
    fprintf(fp, ORC_FMT_INT_OUT, h->rdump_cnt);
    fprintf(fp, ORC_FMT_DBL_OUT, h->dt);
    fprintf(fp, ORC_FMT_INT_OUT, h->lim);
    fprintf(fp, ORC_FMT_INT_OUT, h->failed);
    fprintf(fp, ORC_FMT_DBL_OUT, h->Rin);
    fprintf(fp, ORC_FMT_DBL_OUT, h->Rout);
    fprintf(fp, ORC_FMT_DBL_OUT, h->hslope);
    fprintf(fp, ORC_FMT_DBL_OUT, h->R0);
    fprintf(fp, ORC_FMT_DBL_OUT, h->mass_start);
    fprintf(fp, ORC_FMT_DBL_OUT, h->mass_in);
    fprintf(fp, ORC_FMT_DBL_OUT, h->mass_out);
    fprintf(fp, ORC_FMT_DBL_OUT, h->tdump);
    fprintf(fp, ORC_FMT_DBL_OUT, h->timage);
    fprintf(fp, ORC_FMT_DBL_OUT, h->tlog);
    fprintf(fp, ORC_FMT_DBL_OUT, h->defcon);
    fprintf(fp, ORC_FMT_INT_OUT, h->nfailed);
    fprintf(fp, ORC_FMT_DBL_OUT, h->Rbr);
    fprintf(fp, ORC_FMT_DBL_OUT, h->cyl_R0);
    fprintf(fp, ORC_FMT_DBL_OUT, h->cyl_Th0);
    fprintf(fp, "\n");
    /* restart.c:144-147: p[h][i][j][k] */
    for (hh = 0; hh < sn3; ++hh)
        for (i = 0; i < sn1; ++i)
            for (j = 0; j < sn2; ++j) {
                for (k = 0; k < 8; k++) fprintf(fp, ORC_FMT_DBL_OUT, prim[(((size_t)k * sn1 + i) * sn2 + j) * sn3 + hh]);
                fprintf(fp, "\n");
            }
    if (h->ener_size >= 0) fprintf(fp, ORC_FMT_INT_OUT, h->ener_size); /* restart.c:150-157 */
    if (o && o->tr_on) restart_write_tracers(fp, o);                   /* restart.c:159-161 */
    if (temp)                                                          /* cooleos.c:1002-1005 */
        for (hh = 0; hh < sn3; ++hh)
            for (i = 0; i < sn1; ++i)
                for (j = 0; j < sn2; ++j) fprintf(fp, ORC_FMT_DBL_OUT, temp[((size_t)i * sn2 + j) * sn3 + hh]);
    return fclose(fp) == 0 ? 0 : 2;
}

int orc_restart_read(const char *path, orc_restart_hdr *h, double *prim, double *temp, int read_ener)
{
    FILE *fp = fopen(path, "r");
    int hh, i, j, k, n = 0, sn1, sn2, sn3;
    if (!fp) return 2;
    memset(h, 0, sizeof(*h));
    h->ener_size = -1;
    /* restart.c:317-367 */
    n += fscanf(fp, "%d", &h->N1);
    n += fscanf(fp, "%d", &h->N2);
    n += fscanf(fp, "%d", &h->N3);
    n += fscanf(fp, "%lf", &h->t);
    n += fscanf(fp, "%lf", &h->tf);
    n += fscanf(fp, "%d", &h->nstep);
    n += fscanf(fp, "%lf", &h->a);
    n += fscanf(fp, "%lf", &h->gam);
    n += fscanf(fp, "%lf", &h->cour);
    n += fscanf(fp, "%lf", &h->DTd);
    n += fscanf(fp, "%lf", &h->DTl);
    n += fscanf(fp, "%lf", &h->DTi);
    n += fscanf(fp, "%d", &h->DTr);
    n += fscanf(fp, "%d", &h->dump_cnt);
    n += fscanf(fp, "%d", &h->image_cnt);
    n += fscanf(fp, "%d", &h->rdump_cnt);
    n += fscanf(fp, "%lf", &h->dt);
    n += fscanf(fp, "%d", &h->lim);
    n += fscanf(fp, "%d", &h->failed);
    n += fscanf(fp, "%lf", &h->Rin);
    n += fscanf(fp, "%lf", &h->Rout);
    n += fscanf(fp, "%lf", &h->hslope);
    n += fscanf(fp, "%lf", &h->R0);
    n += fscanf(fp, "%lf", &h->mass_start);
    n += fscanf(fp, "%lf", &h->mass_in);
    n += fscanf(fp, "%lf", &h->mass_out);
    n += fscanf(fp, "%lf", &h->tdump);
    n += fscanf(fp, "%lf", &h->timage);
    n += fscanf(fp, "%lf", &h->tlog);
    n += fscanf(fp, "%lf", &h->defcon);
    n += fscanf(fp, "%d", &h->nfailed);
    n += fscanf(fp, "%lf", &h->Rbr);
    n += fscanf(fp, "%lf", &h->cyl_R0);
    n += fscanf(fp, "%lf", &h->cyl_Th0);
    if (n != 34) {
        fclose(fp);
        return 3;
    }
    sn1 = h->N1 + 4;
    sn2 = h->N2 + 4;
    sn3 = h->N3;
    if (prim) { /* restart.c:372 */
        for (hh = 0; hh < sn3; ++hh)
            for (i = 0; i < sn1; ++i)
                for (j = 0; j < sn2; ++j)
                    for (k = 0; k < 8; k++)
                        if (fscanf(fp, "%lf", &prim[(((size_t)k * sn1 + i) * sn2 + j) * sn3 + hh]) != 1) {
                            fclose(fp);
                            return 3;
                        }
        if (read_ener && fscanf(fp, "%d", &h->ener_size) != 1) { /* restart.c:375-380 */
            fclose(fp);
            return 3;
        }
        if (temp) /* cooleos.c:1024-1027 */
            for (hh = 0; hh < sn3; ++hh)
                for (i = 0; i < sn1; ++i)
                    for (j = 0; j < sn2; ++j)
                        if (fscanf(fp, "%lf", &temp[((size_t)i * sn2 + j) * sn3 + hh]) != 1) {
                            fclose(fp);
                            return 3;
                        }
    }
    fclose(fp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* dump(), dump.c:1710-1990 (HDFDUMP = 0), one rank                          */
/* ------------------------------------------------------------------------ */
int orc_dump_write(orc_ctx *o, const char *path, const orc_restart_hdr *hin)
{
    const int N3 = o->N3;
    orc_restart_hdr h = *hin;
    FILE *fp = fopen(path, "wt");
    int i, j, hh, k;
    int64_t fatal = 0;
    if (!fp) return 2;
    h.N1 = o->N1;
    h.N2 = o->N2;
    h.N3 = o->N3;
    h.t = o->t;
    h.dt = o->dt;
    h.tf = o->P.tf;
    h.a = o->P.a;
    h.gam = o->P.gam;
    h.cour = o->P.cour;
    h.lim = o->P.lim;
    h.Rin = o->P.Rin;
    h.Rout = o->P.Rout;
    h.hslope = o->P.hslope;
    h.R0 = o->P.R0;
    /* dump.c:1745-1790 */
    fprintf(fp, ORC_FMT_DBL_OUT, h.t);
    fprintf(fp, ORC_FMT_INT_OUT, h.N1);
    if (N3 == 1) fprintf(fp, ORC_FMT_INT_OUT, h.N2);
    if (N3 > 1) fprintf(fp, ORC_FMT_INT_OUT, h.N3);
    fprintf(fp, ORC_FMT_DBL_OUT, o->startx[RR]);
    if (N3 == 1) fprintf(fp, ORC_FMT_DBL_OUT, o->startx[TH]);
    if (N3 > 1) fprintf(fp, ORC_FMT_DBL_OUT, o->startx[PH]);
    fprintf(fp, ORC_FMT_DBL_OUT, o->dx[RR]);
    if (N3 == 1) fprintf(fp, ORC_FMT_DBL_OUT, o->dx[TH]);
    if (N3 > 1) fprintf(fp, ORC_FMT_DBL_OUT, o->dx[PH]);
    fprintf(fp, ORC_FMT_DBL_OUT, h.tf);
    fprintf(fp, ORC_FMT_INT_OUT, h.nstep);
    fprintf(fp, ORC_FMT_DBL_OUT, h.a);
    fprintf(fp, ORC_FMT_DBL_OUT, h.gam);
    fprintf(fp, ORC_FMT_DBL_OUT, h.cour);
    fprintf(fp, ORC_FMT_DBL_OUT, h.DTd);
    fprintf(fp, ORC_FMT_DBL_OUT, h.DTl);
    fprintf(fp, ORC_FMT_DBL_OUT, h.DTi);
    fprintf(fp, ORC_FMT_INT_OUT, h.DTr);
    fprintf(fp, ORC_FMT_INT_OUT, h.dump_cnt);
    fprintf(fp, ORC_FMT_INT_OUT, h.image_cnt);
    fprintf(fp, ORC_FMT_INT_OUT, h.rdump_cnt);
    fprintf(fp, ORC_FMT_DBL_OUT, h.dt);
    fprintf(fp, ORC_FMT_INT_OUT, h.lim);
    fprintf(fp, ORC_FMT_INT_OUT, h.failed);
    fprintf(fp, ORC_FMT_DBL_OUT, h.Rin);
    fprintf(fp, ORC_FMT_DBL_OUT, h.Rout);
    fprintf(fp, ORC_FMT_DBL_OUT, h.hslope);
    fprintf(fp, ORC_FMT_DBL_OUT, h.R0);
    fprintf(fp, "\n");
    g_step_eos = o->eos;
    g_step_refd = o->eos_refd;
#define P_(hq, iq, jq, kq) PV(o, o->p, hq, iq, jq)[kq]
#define G_(iq, jq) GDET(o, iq, jq, CENT)
    for (i = o->istart; i <= o->istop; i++)
        for (j = o->jstart; j <= o->jstop; j++)
            for (hh = 0; hh < N3; hh++) {
                double X[NDIM], V[NDIM], divb, vmin, vmax;
                const double *pz = PV(o, o->p, hh, i, j);
                geom_t geom;
                state_t q;
                coord(o, hh, i, j, CENT, X);
                bl_coord_vec(&o->P, X, V);
                fprintf(fp, ORC_FMT_DBL_OUT, X[1]);
                if (N3 == 1) fprintf(fp, ORC_FMT_DBL_OUT, X[2]);
                if (N3 > 1) fprintf(fp, ORC_FMT_DBL_OUT, X[3]);
                fprintf(fp, ORC_FMT_DBL_OUT, V[RR]);
                if (N3 == 1) fprintf(fp, ORC_FMT_DBL_OUT, V[TH]);
                if (N3 > 1) fprintf(fp, ORC_FMT_DBL_OUT, V[PH]);
                for (k = 0; k < NPR; k++) fprintf(fp, ORC_FMT_DBL_OUT, pz[k]);
                /* dump.c:1835-1860: the cooling values of the COOL_EOS build
                 * (only P and T exist in this table), zeros otherwise */
                {
                    double cv[7] = {0., 0., 0., 0., 0., 0., 0.};
                    if (o->eos) {
                        double pp, T;
                        orc_eos_table_p(o->eos, pz[RHO], pz[UU], &pp, NULL, NULL, &T);
                        cv[1] = pp;
                        cv[2] = T;
                    }
                    for (k = 0; k < 7; k++) fprintf(fp, ORC_FMT_DBL_OUT, cv[k]);
                }
                if (N3 == 1) {
                    if (i > 0 && j > 0 && i < o->N1 && j < o->N2)
                        divb = fabs(0.5 *
                                        (P_(hh, i, j, B1) * G_(i, j) + P_(hh, i, j - 1, B1) * G_(i, j - 1) -
                                         P_(hh, i - 1, j, B1) * G_(i - 1, j) -
                                         P_(hh, i - 1, j - 1, B1) * G_(i - 1, j - 1)) /
                                        o->dx[RR] +
                                    0.5 *
                                        (P_(hh, i, j, B2) * G_(i, j) + P_(hh, i - 1, j, B2) * G_(i - 1, j) -
                                         P_(hh, i, j - 1, B2) * G_(i, j - 1) -
                                         P_(hh, i - 1, j - 1, B2) * G_(i - 1, j - 1)) /
                                        o->dx[TH]);
                    else
                        divb = 0.;
                } else {
                    const int m = hwrap(o, hh - 1);
                    divb = fabs(0.25 *
                                    (+P_(m, i, j, B1) * G_(i, j) + P_(m, i, j - 1, B1) * G_(i, j - 1) -
                                     P_(m, i - 1, j, B1) * G_(i - 1, j) - P_(m, i - 1, j - 1, B1) * G_(i - 1, j - 1)

                                     + P_(hh, i, j, B1) * G_(i, j) + P_(hh, i, j - 1, B1) * G_(i, j - 1) -
                                     P_(hh, i - 1, j, B1) * G_(i - 1, j) -
                                     P_(hh, i - 1, j - 1, B1) * G_(i - 1, j - 1)) /
                                    o->dx[RR] +
                                0.25 *
                                    (+P_(m, i, j, B2) * G_(i, j) + P_(m, i - 1, j, B2) * G_(i - 1, j) -
                                     P_(m, i, j - 1, B2) * G_(i, j - 1) - P_(m, i - 1, j - 1, B2) * G_(i - 1, j - 1)

                                     + P_(hh, i, j, B2) * G_(i, j) + P_(hh, i - 1, j, B2) * G_(i - 1, j) -
                                     P_(hh, i, j - 1, B2) * G_(i, j - 1) -
                                     P_(hh, i - 1, j - 1, B2) * G_(i - 1, j - 1)) /
                                    o->dx[TH] +
                                0.25 *
                                    (+P_(hh, i, j, B3) * G_(i, j) + P_(hh, i - 1, j, B3) * G_(i - 1, j) -
                                     P_(m, i, j, B3) * G_(i, j) - P_(m, i - 1, j, B3) * G_(i - 1, j)

                                     + P_(hh, i, j - 1, B3) * G_(i, j) + P_(hh, i - 1, j - 1, B3) * G_(i - 1, j - 1) -
                                     P_(m, i, j - 1, B3) * G_(i, j - 1) - P_(m, i - 1, j - 1, B3) * G_(i - 1, j - 1)) /
                                    o->dx[PH]);
                }
                fprintf(fp, ORC_FMT_DBL_OUT, divb);
                if (!h.failed) {
                    double vsq = 0., gsqtmp, ps_p, ps_cs, uur;
                    get_geometry(o, i, j, CENT, &geom);
                    get_state(pz, &geom, &q, &fatal);
                    for (k = 0; k < NDIM; k++) fprintf(fp, ORC_FMT_DBL_OUT, q.ucon[k]);
                    for (k = 0; k < NDIM; k++) fprintf(fp, ORC_FMT_DBL_OUT, q.ucov[k]);
                    for (k = 0; k < NDIM; k++) fprintf(fp, ORC_FMT_DBL_OUT, q.bcon[k]);
                    for (k = 0; k < NDIM; k++) fprintf(fp, ORC_FMT_DBL_OUT, q.bcov[k]);
                    vchar(pz, &q, &geom, 1, o->P.gam, &vmax, &vmin);
                    fprintf(fp, ORC_FMT_DBL_OUT, vmin);
                    fprintf(fp, ORC_FMT_DBL_OUT, vmax);
                    vchar(pz, &q, &geom, 2, o->P.gam, &vmax, &vmin);
                    fprintf(fp, ORC_FMT_DBL_OUT, vmin);
                    fprintf(fp, ORC_FMT_DBL_OUT, vmax);
                    fprintf(fp, ORC_FMT_DBL_OUT, geom.g);
                    for (int id = 1; id < 4; id++)
                        for (int jd = 1; jd < 4; jd++) vsq += geom.gcov[id][jd] * q.ucon[id] * q.ucon[jd];
                    gsqtmp = 1. + vsq;
                    vsq /= gsqtmp;
                    fprintf(fp, ORC_FMT_DBL_OUT, sqrt(vsq));
                    ps_p = pz[UU] * (o->P.gam - 1.0);
                    ps_cs = sqrt(4.0 / 3.0 * ps_p / (pz[RHO] + 4.0 * ps_p));
                    uur = sqrt(geom.gcov[RR][RR]) * q.ucon[RR] / q.ucon[TT];
                    fprintf(fp, ORC_FMT_DBL_OUT, ps_p);
                    fprintf(fp, ORC_FMT_DBL_OUT, ps_cs);
                    fprintf(fp, ORC_FMT_DBL_OUT, uur);
                    fprintf(fp, ORC_FMT_DBL_OUT, -uur / ps_cs);
                }
                fprintf(fp, "\n");
            }
#undef P_
#undef G_
    g_step_eos = NULL;
    return fclose(fp) == 0 ? 0 : 2;
}

/* tracers.c:565-606, one rank: the live tracers in their vector order */
static void restart_write_tracers(FILE *fp, const orc_ctx *o)
{
    const long long n = o->tr_n, n12 = (long long)o->N1 * o->N2;
    long long cnt = 0, k;
    int d;
    fprintf(fp, ORC_FMT_DBL_OUT, o->tr_Rin);
    fprintf(fp, ORC_FMT_DBL_OUT, o->tr_Rout);
    fprintf(fp, ORC_FMT_DBL_OUT, o->tr_thmin);
    fprintf(fp, ORC_FMT_DBL_OUT, o->tr_thmax);
    /* tracers.c:133-145: coord() of the first and one-past-last centres */
    fprintf(fp, ORC_FMT_DBL_OUT, o->startx[RR] + (0 + 0.5) * o->dx[RR]);
    fprintf(fp, ORC_FMT_DBL_OUT, o->startx[RR] + (o->N1 + 0.5) * o->dx[RR]);
    fprintf(fp, ORC_FMT_DBL_OUT, o->N3 > 2 ? o->startx[PH] + (0 + 0.5) * o->dx[PH] : 0.0);
    fprintf(fp, ORC_FMT_DBL_OUT, o->N3 > 2 ? o->startx[PH] + (o->N3 + 0.5) * o->dx[PH] : 0.0);
    for (k = 0; k < n; k++) cnt += o->tr_state[k] == ORC_TRACER_LIVE;
    fprintf(fp, ORC_FMT_INT_OUT, (int)cnt);
    for (k = 0; k < n; k++) {
        if (o->tr_state[k] != ORC_TRACER_LIVE) continue;
        const int h = (int)(k / n12), i = (int)((k % n12) / o->N2), j = (int)(k % o->N2);
        char name[96];
        struct stat st;
        fprintf(fp, ORC_FMT_INT_OUT, i);
        fprintf(fp, ORC_FMT_INT_OUT, j);
        fprintf(fp, ORC_FMT_INT_OUT, h);
        fprintf(fp, ORC_FMT_DBL_OUT, 0.0);
        for (d = 0; d < 3; d++) fprintf(fp, ORC_FMT_DBL_OUT, o->tr_X[(size_t)d * n + k]);
        if (o->N3 > 2)
            snprintf(name, sizeof(name), "tracers/tracer_%03d_%03d_%03d.dat", i, j, h);
        else
            snprintf(name, sizeof(name), "tracers/tracer_%03d_%03d.dat", i, j);
        fprintf(fp, ORC_FMT_INT_OUT, stat(name, &st) == 0 ? (int)st.st_size : 0);
    }
}
