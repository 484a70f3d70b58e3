/*
 * harm_oracle.h -- CPU oracle for the HARM-COOL per-timestep GRMHD update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (agnieszkajaniuk/HARM_COOL, src/step_ch.c and the files it calls),
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  The product (harm_cool_amd/, libharmcool.so) never links or calls it.
 *
 * Conventions
 *   - Zone indices follow the reference: i in [0, N1+4) with active 2..N1+1,
 *     j in [0, N2+4) with active 2..N2+1 (2 ghosts each side, decs.h:403,
 *     defs.h:135).  phi (h) has N3 active planes and no ghost planes: the
 *     reference's one-plane phi halo is a periodic MPI self-exchange
 *     (step_ch.c:183-197, mpi.c:20,63), restated here as periodic indexing.
 *   - The exchange ("canonical") layout is SoA, phi fastest:
 *       prim[v][i][j][h], v in 0..7 (RHO,UU,U1,U2,U3,B1,B2,B3, decs.h:190-197),
 *       i in [0,N1+4), j in [0,N2+4), h in [0,N3).
 */
#ifndef HARM_ORACLE_H
#define HARM_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field layout as harm_params in include/harm_cool.h. */
typedef struct orc_params {
    int32_t n1, n2, n3;      /* active zones in r (X1), theta (X2), phi (X3)      */
    int32_t lim;             /* slope limiter: 0 MC, 1 van Leer, 2 minmod, 3 PPM   */
    int32_t font_fix;        /* 1: Font fix fallbacks on inversion failure         */
    int32_t pad0;
    double a, gam;           /* BH spin, adiabatic index                          */
    double rin, rmax;        /* FM torus inner edge and pressure maximum          */
    double kappa, beta;      /* polytropic constant, plasma beta normalization    */
    double Rin, Rout;        /* grid radial extent (Rin <= 0: 0.98*r_horizon)     */
    double hslope, R0, Rbr;  /* MKS theta concentration, r offset, hyperexp break */
    double cour, dt0, tf, defcon;
    double hllf, laxf;       /* relative weights of HLL and LLF fluxes            */
    double pert_amp;         /* u *= 1 + pert_amp*(U[0,1]-0.5)                    */
    uint32_t seed[4];        /* KISS base seeds                                   */
} orc_params;

enum {
    ORC_C2P_CALLS = 0, ORC_NEWTON_ITERS, ORC_NEWTON_KCONV, ORC_FAIL_2D,
    ORC_FAIL_FONT1, ORC_FAIL_FONT2, ORC_FIXUP_GAMMA, ORC_FIXUP_REPAIRS,
    ORC_FIXUP_NOSTENCIL, ORC_FLOOR_ZONES, ORC_GAMMA_CAP, ORC_FATAL,
    ORC_NCOUNTERS
};

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_params *p);
void     orc_destroy(orc_ctx *o);
void     orc_set_threads(int n);
int      orc_init_torus(orc_ctx *o);
/* One step_ch(): 0; 11 when dt < 1e-9 (state and mass budget of the step kept,
 * t not advanced, step_ch.c:126-129); 14 when a fail() event happened in the
 * step (state kept; no mass budget, tracer or clock update). */
int      orc_step(orc_ctx *o);
int      orc_get_state(const orc_ctx *o, double *prim, int8_t *pflag);
int      orc_set_state(orc_ctx *o, const double *prim, const int8_t *pflag);
void     orc_get_time(const orc_ctx *o, double *t, double *dt);
void     orc_set_time(orc_ctx *o, double t, double dt);
void     orc_get_counters(const orc_ctx *o, int64_t *c);
void     orc_get_grid(const orc_ctx *o, double *startx, double *dx);
double   orc_get_katm(const orc_ctx *o, int i);
/* Geometry tables in the reference layout: gcov/gcon [SN1][SN2][5][4][4],
 * gdet [SN1][SN2][5], conn [SN1][SN2][4][4][4], coords [SN1][SN2][4]. */
void     orc_get_geometry(const orc_ctx *o, double *gcov, double *gcon,
                          double *gdet, double *conn, double *coords);
double   orc_divb_max(const orc_ctx *o);
/* Fluxes of one half-step on the current state after fix_flux and flux_ct
 * (canonical layout, 8 x (N1+4) x (N2+4) x N3 each; F3 may be NULL in 2D);
 * ndt[3] = per-direction CFL limits (1e9 when not fullstep). */
int      orc_debug_fluxes(orc_ctx *o, int fullstep, double *F1, double *F2, double *F3, double *ndt);
/* fixup_utoprim() then bound_prim() on the current state and flags. */
int      orc_debug_fixup_utoprim(orc_ctx *o);
/* bound_prim() on the current state. */
int      orc_debug_bound(orc_ctx *o);
/* advance() then fixup() of one half of step_ch: fullstep 0 = predictor
 * p -> ph (dt/2), ph returned; fullstep 1 = corrector from p and the given ph
 * -> p (dt).  ph in the canonical layout, ghosts included. */
int      orc_debug_advance(orc_ctx *o, int fullstep, double *ph);
/* fixup() alone on the current state. */
int      orc_debug_fixup(orc_ctx *o);
/* out[0] = mass (sum rho gdet dV), out[1] = sum U[UU] dV, out[2] = mdot_in,
 * out[3] = mdot_out (diag.c:477-580, diag.c:59-230). */
void     orc_diag(const orc_ctx *o, double *out);
/* The ener.out record of diag() (diag.c:59-326), diag_flux (diag.c:437-465)
 * and diag_mass (diag.c:477-580); imax = zero-based (i, j, h) of the first
 * zone with the largest |div B|, -1 when it is 0 everywhere. */
typedef struct orc_diag_rec {
    double t, rmed, pp, e, mdot, edot, ldot, lnu, lnu_rho, lnu_rho_2, lnu2;
    double mdot_in, mdot_out, mass_start, mass_curr, mass_in, mass_out, divbmax;
    int32_t imax[3], flux_valid;
} orc_diag_rec;
void     orc_diag_log(orc_ctx *o, orc_diag_rec *r);
/* mass_start (main.c:173-176) and the accumulated mass_in / mass_out. */
void     orc_set_mass_budget(orc_ctx *o, double mass_start, double mass_in, double mass_out);
/* Lagrangian tracers (tracers.c:109-560): slot k = (h N1 + i) N2 + j of the
 * seed zone; after orc_tracers_init every orc_step advances them. */
enum { ORC_TRACER_NONE = 0, ORC_TRACER_LIVE, ORC_TRACER_IN, ORC_TRACER_OUT, ORC_TRACER_JET, ORC_TRACER_LOST };
typedef struct orc_tracer_rec {
    double t, rho, T, ye, r, th, ph;
    int32_t h, i, j, kind;
} orc_tracer_rec;
long long orc_tracers_init(orc_ctx *o, double rho_unit, double t_unit);
long long orc_tracers_dump(orc_ctx *o, orc_tracer_rec *recs, long long cap);
long long orc_tracers_events(orc_ctx *o, orc_tracer_rec *recs, long long cap);
void      orc_tracers_get(const orc_ctx *o, double *X, int8_t *state);
void      orc_tracers_set(orc_ctx *o, const double *X, const int8_t *state);
/* update_tracers(Dt) on the current state, without a step */
void      orc_tracers_update(orc_ctx *o, double Dt);

/* Single-zone primitive solvers; gcov/gcon are row-major 4x4. */
int  orc_utoprim_2d(const double U[8], const double *gcov, const double *gcon,
                    double gdet, double prim[8], double gam, int *n_iter, int *k_conv);
int  orc_utoprim_1dvsq2fix1(const double U[8], const double *gcov, const double *gcon,
                            double gdet, double prim[8], double K, double gam);
int  orc_utoprim_1dfix1(const double U[8], const double *gcov, const double *gcon,
                        double gdet, double prim[8], double K, double gam);
void orc_primtoU_g(const double prim[8], const double *gcov, const double *gcon,
                   double gdet, double gam, double U[8]);
/* HARM's conserved vector (phys.c:100-105) from primitives. */
void orc_primtoU(const double prim[8], const double *gcov, const double *gcon,
                 double gdet, double gam, double U[8]);
/* Learned Newton seed (no reference counterpart).  Raw network inputs of one
 * zone: log D, asinh(E/D), asinh(Qt^2/D^2), asinh(B^2/D), asinh((Q.B)^2/D^3),
 * lapse, asinh((E-D-B^2/2)/D), asinh(B^2/E) with E = -Q.n; -100 if U[0]<=0. */
int  orc_u2p_features(const double U[8], const double *gcov, const double *gcon,
                      double gdet, double feat[8]);
/* Utoprim_2d started from (W0, vsq0) instead of the guess built from prim. */
int  orc_utoprim_2d_seeded(const double U[8], const double *gcov, const double *gcon,
                           double gdet, double prim[8], double gam, double W0, double vsq0,
                           int vsq_from_W, int *n_iter, int *k_conv);
/* Seeded solve with fallback to Utoprim_2d from prim; counts add up. */
int  orc_utoprim_2d_learned(const double U[8], const double *gcov, const double *gcon,
                            double gdet, double prim[8], double gam, double W0, double vsq0,
                            int vsq_from_W, int *n_iter, int *k_conv, int *fell_back);
/* Gated learned solve: the seed only where the first Newton step from prim
 * exceeds tau or prim gives no start; otherwise Utoprim_2d bit for bit. */
int  orc_utoprim_2d_gated(const double U[8], const double *gcov, const double *gcon, double gdet, double prim[8],
                          double gam, double W0, double vsq0, int vsq_from_W, double tau, int *n_iter,
                          int *k_conv, int *seeded);
/* Tabulated EOS: rho[nr], T[nt] increasing; e, p [nr][nt] (code units);
 * gam is the Gamma-law fallback outside the table. */
typedef struct orc_eos {
    int nr, nt;
    const double *rho, *T, *e, *p;
    double gam;
} orc_eos;
/* p(rho, u) by piecewise-bilinear interpolation inverted along T, its
 * derivatives at fixed u / fixed rho, and T; returns 1 on the Gamma-law branch. */
int  orc_eos_table_p(const orc_eos *t, double rho, double u, double *p, double *dpdrho,
                     double *dpdu, double *T);
/* Step with the tabulated EOS (the reference's COOL_EOS build); NULL: Gamma-law. */
void orc_set_eos(orc_ctx *o, const orc_eos *eos, int ref_derivs);
/* The step's COOL_EOS pressure on the cooling-table cache (cooleos.c:314-444,
 * 667-741, 770-807, 826-930): a temperature Newton solve per call on the cache
 * of the zone's radial index, from the zone's last temperature.  The caches
 * share the (rho, T) grid vx[X] x vy[Y] (cgs, z = 1 plane); values holds n of
 * them, [n][X][Y][D] (component 0 e, 1 p, cgs), n = 1 (every radius) or N1+4
 * (one per radial index, ghosts included); rho_min..T_max are rhoMin, rhoMax,
 * TMin, TMax.  The arrays must outlive the context's use of them. */
typedef struct orc_ceos {
    int X, Y, D, n;
    const double *vx, *vy, *values;
    double rho_unit, u_unit, p_unit, gam;
    double rho_min, rho_max, T_min, T_max;
} orc_ceos;
/* Step with the cache EOS (NULL: back to orc_set_eos's EOS); the zones' last
 * temperatures start at 0 (never set). */
int  orc_set_eos_cache(orc_ctx *o, const orc_ceos *c);
/* The zones' last temperatures, canonical zone layout (N1+4) x (N2+4) x N3. */
int  orc_get_lastT(const orc_ctx *o, double *T);
int  orc_set_lastT(orc_ctx *o, const double *T);
/* pressure_rho0_u_CoolEOS for one call on zone (h, i, j) (array indices),
 * with InitCache_CoolEOS's selection; the zone's last T is updated. */
double orc_ceos_pressure(orc_ctx *o, int h, int i, int j, double rho0, double u);
/* The cache's lookup and temperature solve (cache_oracle.c). */
int  orc_cache_get(int X, int Y, int Z, int D, const int *types, int whole, const double *vx, const double *vy,
                   const double *vz, const double *values, const uint8_t *counted, double x, double y, double z,
                   int zindex, int spline, const int *dimmask, double *out, double *der, size_t ostride);
int  orc_cache_solve_T(int X, int Y, int Z, int D, const int *types, int whole, const double *vx, const double *vy,
                       const double *vz, const double *values, const uint8_t *counted, double rho, double e,
                       int zindex, double T0, double lo, double hi, int ntrial, double tolf, double *T,
                       double *berr);
/* Utoprim_5d (utoprim_5d.c:179-239); eos NULL = Gamma-law. */
int  orc_utoprim_5d(const double U[8], const double *gcov, const double *gcon, double gdet,
                    double prim[8], double gam, const orc_eos *eos, int *n_iter);
void orc_primtoU_eos(const double prim[8], const double *gcov, const double *gcon, double gdet,
                     double gam, const orc_eos *eos, double U[8]);
void orc_dudp_5d(const double prim[8], const double *gcov, const double *gcon, double gdet,
                 double gam, const orc_eos *eos, double *Am /* 5x5 */);
/* Geometry of one point in code coordinates X (X[0] ignored). */
void orc_gcov_func(const orc_params *p, const double X[4], double *gcov);
void orc_dxdxp_func(const orc_params *p, const double X[4], double *dxdxp);
void orc_bl_coord_vec(const orc_params *p, const double X[4], double *V);
void orc_conn_func(const orc_params *p, const double X[4], const double *gcon, double *conn);
double orc_gdet_func(const double *gcov);
int  orc_invert_matrix(const double *A, double *Ainv);
double orc_slope_lim(int lim, double y1, double y2, double y3);
/* PPM face values of one zone from its value, its neighbours and their MC
 * slopes (Colella & Woodward 1984; lim 3, not in the reference); returns r - l. */
double orc_para(double ym, double y0, double yp, double dqm, double dq0, double dqp, double *l, double *r);
double orc_kiss_uniform(uint32_t z, uint32_t w, uint32_t jsr, uint32_t jcong);
/* One face of fluxcalc (step_ch.c:407-487): left/right states -> flux, ctop. */
void orc_face_flux(const double *p_l, const double *p_r, const double *gcov,
                   const double *gcon, double gdet, int dir, double gam,
                   double hllf, double laxf, double *F, double *ctop);
/* vchar (phys.c:337-476) for one state; returns the fail() code or 0. */
int  orc_vchar(const double *pr, const double *gcov, const double *gcon, double gdet,
               int dir, double gam, double *cmax, double *cmin);

/* Restart file (restart.c:73-383).  Same field layout as harm_restart_hdr in
 * include/harm_cool.h; prim in the canonical layout, temp [N1+4][N2+4][N3]. */
typedef struct orc_restart_hdr {
    int32_t N1, N2, N3, nstep;
    double t, tf, a, gam, cour, DTd, DTl, DTi;
    int32_t DTr, dump_cnt, image_cnt, rdump_cnt;
    double dt;
    int32_t lim, failed;
    double Rin, Rout, hslope, R0, mass_start, mass_in, mass_out, tdump, timage, tlog, defcon;
    int32_t nfailed, ener_size;
    double Rbr, cyl_R0, cyl_Th0;
} orc_restart_hdr;
/* restart_write(): the header (restart.c:105-138), the body (restart.c:144-147),
 * the ener.out size when ener_size >= 0 (restart.c:150-157), the COOL_EOS
 * temperatures when temp != NULL (cooleos.c:996-1014). */
int orc_restart_write(const char *path, const orc_restart_hdr *h, const double *prim, const double *temp);
/* The same with, after the ener.out size, the tracer block of o's tracers
 * (restart_write_tracers, tracers.c:565-606) when o has them. */
int orc_restart_write_tr(const char *path, const orc_restart_hdr *h, const double *prim, const double *temp,
                         const orc_ctx *o);
/* dump() (dump.c:1710-1990, HDFDUMP = 0): hdr supplies the driver's counters;
 * the context's parameters, t and dt replace the rest. */
int orc_dump_write(orc_ctx *o, const char *path, const orc_restart_hdr *hdr);
/* restart_read() (restart.c:317-376) with fscanf; read_ener as rank 0 does,
 * temp (may be NULL) as restart_read_CoolEOS (cooleos.c:1018-1030). */
int orc_restart_read(const char *path, orc_restart_hdr *h, double *prim, double *temp, int read_ener);

#ifdef __cplusplus
}
#endif
#endif
