/*
 * harm_cool.h -- C ABI of libharmcool.so, the MI355X (gfx950) implementation of
 * HARM-COOL's per-timestep GRMHD update.
 *
 * The reference (agnieszkajaniuk/HARM_COOL) has no plugin or FFI layer: its hot
 * path sits behind plain C functions over process-global arrays.  Each entry point
 * below names the reference interface it replaces.  All functions are extern "C",
 * take plain pointers and sizes, return 0 on success and a nonzero code on error
 * (never exit()), and record a message readable with harm_last_error().
 *
 * Layouts
 *   Zone index: i in [0,N1+4) (r, active 2..N1+1), j in [0,N2+4) (theta, active
 *   2..N2+1), h in [0,N3) (phi, periodic, no ghost planes).
 *   Host exchange layout ("canonical"): prim[v][i][j][h], v = RHO,UU,U1,U2,U3,B1,B2,B3
 *   (decs.h:190-197), phi fastest; flags int8 [i][j][h].
 */
#ifndef HARM_COOL_H
#define HARM_COOL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run parameters.  The reference fixes these at compile time (decs.h:59-258,
 * init.c:90-161, step_ch.c:66-67, u2p_defs.h:79-96). */
typedef struct harm_params {
    int32_t n1, n2, n3;      /* active zones in X1 (r), X2 (theta), X3 (phi); n3 = 1 or even >= 4 */
    int32_t lim;             /* slope limiter: 0 MC, 1 van Leer, 2 minmod (interp.c:49-93);
                                3 PPM (Colella & Woodward 1984 on MC slopes; not in the reference) */
    int32_t font_fix;        /* 1: Font-fix fallbacks on inversion failure (step_ch.c:277-300) */
    int32_t pad0;
    double a, gam;           /* BH spin, adiabatic index */
    double rin, rmax;        /* FM torus inner edge, pressure maximum (init.c:111-112) */
    double kappa, beta;      /* polytropic constant, plasma-beta normalization */
    double Rin, Rout;        /* radial extent; Rin <= 0 selects 0.98*(1+sqrt(1-a^2)) */
    double hslope, R0, Rbr;  /* MKS theta concentration, r offset, hyperexp break (<0: off) */
    double cour, dt0, tf, defcon;
    double hllf, laxf;       /* weights of the HLL and LLF fluxes (step_ch.c:66-67) */
    double pert_amp;         /* u *= 1 + pert_amp*(U[0,1]-0.5) (init.c:273) */
    uint32_t seed[4];        /* KISS base seeds (init.c:272) */
} harm_params;

/* Per-context event counters, cumulative over steps (reference: failimage[0..4],
 * step_ch.c:274-295, fixup.c:96,354,525; Newton iteration counts are new). */
enum {
    HARM_C2P_CALLS = 0,      /* Utoprim_2d calls */
    HARM_NEWTON_ITERS,       /* Newton iterations incl. the 2 extra (utoprim_2d.c:467-520) */
    HARM_NEWTON_KCONV,       /* iterations until |dx/x| <= 1e-10 first held */
    HARM_FAIL_2D,            /* Utoprim_2d returned nonzero */
    HARM_FAIL_FONT1,         /* Utoprim_1dvsq2fix1 also failed */
    HARM_FAIL_FONT2,         /* Utoprim_1dfix1 also failed */
    HARM_FIXUP_GAMMA,        /* fixup1zone gamma failure, pflag = -33 */
    HARM_FIXUP_REPAIRS,      /* zones repaired by fixup_utoprim */
    HARM_FIXUP_NOSTENCIL,    /* ... of which used the v=0 fallback branch */
    HARM_FLOOR_ZONES,        /* fixup1zone calls that applied a rho/u floor */
    HARM_GAMMA_CAP,          /* Lorentz factor capped at GAMMAMAX */
    HARM_FATAL,              /* events where the reference calls fail() and exits */
    HARM_NCOUNTERS
};

/* Kernel classes timed with HIP events when timing is enabled. */
enum {
    HARM_K_FLUX1 = 0, HARM_K_FLUX2, HARM_K_FLUX3, HARM_K_EMF, HARM_K_UPDATE,
    HARM_K_BOUND, HARM_K_FIXUTOPRIM, HARM_K_STEPEND, HARM_K_C2P_BATCH, HARM_K_MLP, HARM_K_FONTFIX,
    HARM_K_DIV, HARM_K_TRACERS, HARM_NKERNELS
};

typedef struct harm_ctx harm_ctx;

/* Last error message of the calling thread ("" if none). */
const char *harm_last_error(void);

/* Context = the reference's globals (decs.h:310-378) made explicit: device-resident
 * SoA state, metric tables, one HIP stream.  Replaces set_arrays()+set_grid()
 * (main.c:288-508).  device < 0 keeps the current device. */
harm_ctx *harm_create(const harm_params *p, int device);
void      harm_destroy(harm_ctx *c);

/* Fishbone-Moncrief torus initial data; replaces init() (init.c:70-472). */
int harm_init_torus(harm_ctx *c);

/* nsteps predictor/corrector steps with floors, bounds, fixups and dt control;
 * replaces step_ch() (step_ch.c:90-140, decs.h:600), called nsteps times.
 * Where the reference exits, the batch stops after that step and returns:
 *   11  dt < 1e-9 (step_ch.c:126-129): the step's state and mass budget are
 *       kept, t is not advanced;
 *   14  a fail() event in the step (diag.c:381-397: gamma, vchar or metric
 *       failure; the reference exits inside the step, here the step completes),
 *       t is not advanced.
 * The steps after it launch nothing that changes the context.  A stopped
 * context refuses further steps with the same code until harm_set_time,
 * harm_init_torus or harm_restart_read resets it. */
int harm_step(harm_ctx *c, int nsteps);

/* Host <-> device state in the canonical layout (8*(N1+4)*(N2+4)*N3 doubles). */
int harm_get_prims(harm_ctx *c, double *prim, int8_t *pflag);
int harm_set_prims(harm_ctx *c, const double *prim, const int8_t *pflag);
int harm_get_time(harm_ctx *c, double *t, double *dt);
int harm_set_time(harm_ctx *c, double t, double dt);
int harm_get_counters(harm_ctx *c, int64_t *counters /* HARM_NCOUNTERS */);
int harm_reset_counters(harm_ctx *c);

/* Diagnostics (diag.c:156-209 divB; diag.c:477-580 mass budget):
 * out[0] mass, out[1] sum U[UU] dV, out[2] mdot_in, out[3] mdot_out,
 * out[4] accumulated mass_in, out[5] accumulated mass_out, out[6] max |divB|. */
int harm_diag(harm_ctx *c, double *out /* 7 */);

/* The record diag() writes to ener.out (diag.c:59-326), replacing diag(LOG_OUT)
 * with the diag_flux (diag.c:437-465) of the last full step and the diag_mass
 * (diag.c:477-580) of the current state.  mdot/edot/ldot and mdot_in/mdot_out
 * are summed in the reference's (h, j) order; the volume sums in a fixed tree
 * order.  The neutrino luminosities need the cooling table's Lambda, Qnu and
 * H_d, which the tabulated EOS here does not carry: lnu, lnu_rho and lnu2 are
 * 0; lnu_rho_2 = sum gdet rho with a tabulated EOS (the COOL_EOS build's weight),
 * else 0. */
typedef struct harm_diag_rec {
    double t, rmed, pp, e;                 /* t, sum U[RHO] dV, sum U[U3] dV, sum U[UU] dV */
    double mdot, edot, ldot;               /* inner face, last full step (0 before one) */
    double lnu, lnu_rho, lnu_rho_2, lnu2;
    double mdot_in, mdot_out;              /* first ghost zones (MDOTTYPE NEU_AGA_GHOST1) */
    double mass_start, mass_curr, mass_in, mass_out;
    double divbmax;
    int32_t imax[3];    /* zero-based (i, j, h) of the first zone with divbmax; -1 if 0 */
    int32_t flux_valid; /* mdot/edot/ldot come from a completed full step */
} harm_diag_rec;
int harm_diag_log(harm_ctx *c, harm_diag_rec *r);
/* One ener.out line in the format of diag.c:317-326, newline included;
 * returns its length, -1 if it does not fit in n bytes. */
int harm_ener_line(const harm_diag_rec *r, char *buf, size_t n);
/* mass_start (main.c:173-176; set by harm_init_torus and harm_restart_read) and
 * the mass_in / mass_out that diag_mass accumulates (diag.c:577-578). */
int harm_set_mass_budget(harm_ctx *c, double mass_start, double mass_in, double mass_out);

/* Lagrangian tracers (tracers.c:109-654).  harm_tracers_init seeds one tracer
 * at every active zone whose density exceeds RHOMIN, at X = (zone centre in
 * X1, X2; phi face in X3), in (h, i, j) order (init_tracers, tracers.c:150-170);
 * afterwards every harm_step advances them with the interpolated u^i/u^t of
 * the new state and the step's dt (update_tracers, tracers.c:298-430) and
 * records the ones that leave (r < Rin_tr or theta outside [0, pi]: IN;
 * r >= Rout_tr: OUT, or JET outside [0.02 pi, 0.98 pi]).  Slot k of the
 * per-tracer arrays is the seed zone k = (h N1 + i) N2 + j (zero-based). */
enum {
    HARM_TRACER_NONE = 0, /* no tracer seeded at this zone */
    HARM_TRACER_LIVE = 1,
    HARM_TRACER_IN = 2,
    HARM_TRACER_OUT = 3,
    HARM_TRACER_JET = 4,
    HARM_TRACER_LOST = 5 /* position outside the searchable grid (tracers.c:322-328) */
};
/* One line of a tracer file (dump_one_tracer, tracers.c:470-510): t T_UNIT,
 * rho RHO_UNIT, T, Ye, r, theta, phi; (h, i, j) name the tracer's file. */
typedef struct harm_tracer_rec {
    double t, rho, T, ye, r, th, ph;
    int32_t h, i, j, kind;
} harm_tracer_rec;
/* rho_unit, t_unit: the run's RHO_UNIT and T_UNIT (cooleos.c:120-140).  T is
 * the Gamma-law temperature of rhoTYe_fun (tracers.c:446-459) with Ye = -1, or
 * with a tabulated EOS the table's temperature (Ye = -1: the table carries no
 * electron fraction).  Not supported for N3 == 2.  *n_live may be NULL. */
int harm_tracers_init(harm_ctx *c, double rho_unit, double t_unit, int64_t *n_live);
int harm_tracers_stop(harm_ctx *c);
/* dump_tracers (tracers.c:512-560): one record per live tracer, in slot order;
 * live tracers whose position no longer lies in the grid become LOST.
 * Returns -1 with *n = the count if cap is too small. */
int harm_tracers_dump(harm_ctx *c, harm_tracer_rec *recs, int64_t cap, int64_t *n);
/* The records of tracers that left since the previous call, in (step, slot)
 * order, as update_tracers writes them before it renames the files into
 * tracers/in, out or jet.  Returns -1 with *n = the count if cap is too small. */
int harm_tracers_events(harm_ctx *c, harm_tracer_rec *recs, int64_t cap, int64_t *n);
/* update_tracers(dt) on the current state and time without a step (harm_step
 * already calls it after every step). */
int harm_tracers_update(harm_ctx *c, double dt);
/* Positions (X[3][N1 N2 N3], X1..X3) and states (HARM_TRACER_*) of every slot. */
int harm_tracers_get(harm_ctx *c, double *X, int8_t *state);
/* Replace positions and states (restart_read_tracers, tracers.c:608-654). */
int harm_tracers_set(harm_ctx *c, const double *X, const int8_t *state);

/* Device pointers of the state (canonical layout, var stride (N1+4)*(N2+4)*N3). */
int harm_device_state(harm_ctx *c, double **prim, int8_t **pflag, void **stream);

/* Kernel timing with HIP events on the context's stream.  ms[k] = summed elapsed
 * time of kernel class k since the last reset, launches[k] = number of launches. */
int harm_set_timing(harm_ctx *c, int enable);
int harm_get_timing(harm_ctx *c, double *ms, int64_t *launches);

/* Per-stage entry points for parity tests (one half-step pieces of advance(),
 * step_ch.c:153-318).  Fluxes after flux_ct are returned in the canonical layout
 * (8 x (N1+4) x (N2+4) x N3 doubles each). */
int harm_debug_fluxes(harm_ctx *c, int fullstep, double *F1, double *F2, double *F3, double *ndt3);
int harm_debug_bound(harm_ctx *c);
/* fixup_utoprim (fixup.c:432-538) then bound_prim on the current state and flags. */
int harm_debug_fixup_utoprim(harm_ctx *c);
/* advance() then fixup() of one half of step_ch (step_ch.c:101-103, 112-114;
 * advance step_ch.c:153-318, fixup fixup.c:51-60), without the bound_prim and
 * fixup_utoprim that follow them.  fullstep 0: the predictor, state -> ph with
 * dt/2, the state unchanged and ph copied out (its active zones; the ghosts hold
 * what the half-step buffer held before, as in the reference until the
 * bound_prim that follows).  fullstep 1: the corrector, from
 * the state and the given ph -> the state with dt.  ph is 8 x (N1+4) x (N2+4) x N3
 * doubles in the canonical layout, ghosts included; flags are those advance()
 * leaves for the active zones (read them with harm_get_state). */
int harm_debug_advance(harm_ctx *c, int fullstep, double *ph);
/* fixup() alone on the current state (fixup.c:51-60, fixup1zone fixup.c:142-337):
 * floors, the drift-frame velocity blend, the gamma cap; flag -33 where gamma fails. */
int harm_debug_fixup(harm_ctx *c);
/* The device pow used by the Font-fix solvers (correctly rounded), on host arrays;
 * for tests against the host libm pow the reference calls. */
int harm_debug_pow(const double *x, const double *y, double *out, size_t n);
/* A plain device-to-device copy of nbytes (1 + repeats launches): the reference
 * reading for the HBM traffic counters and the achievable copy bandwidth. */
int harm_debug_copy(size_t nbytes, int repeats, double *ms_per_copy);

/* Geometry tables built on the host (main.c:413-508), canonical per-(i,j) layout:
 * gcov/gcon [5][16][N1+4][N2+4], gdet [5][N1+4][N2+4], conn [64][N1+4][N2+4]. */
int harm_get_geometry(harm_ctx *c, double *gcov, double *gcon, double *gdet, double *conn);

/* Con2prim options for the batched inversion. */
typedef struct harm_u2p_opts {
    double gam;              /* adiabatic index */
    int32_t font_fix;        /* run the Font-fix fallbacks on failure */
    int32_t guess_mode;      /* 0: guess from prim_in, 1: MLP guess with prim_in fallback,
                                2: gated -- the MLP guess only for zones whose first Newton step
                                from prim_in exceeds gate_tau, the plain solve for the others */
    const void *mlp;         /* harm_mlp handle for guess_mode 1 and 2 (may be NULL otherwise) */
    double K;                /* isentrope constant for the Font fix */
    double gate_tau;         /* guess_mode 2: threshold on |dW/W| + |dv2/v2| of the first step (> 0) */
} harm_u2p_opts;

/* Batched Noble 2D inversion on the device; replaces n calls of Utoprim_2d
 * (utoprim_2d.c:126, decs.h:608-609).  All pointers are device pointers in SoA
 * form: U[8][n], gcov[16][n], gcon[16][n], gdet[n], prim[8][n] (guess in, result
 * out; B always overwritten, the rest only where ret == 0), ret[n] (int8, the
 * reference's char-truncated code), iters[n] (Newton iterations, may be NULL),
 * kconv[n] (may be NULL).  stream may be NULL (default stream). */
int harm_utoprim2d_batch(const double *U, const double *gcov, const double *gcon, const double *gdet,
                         double *prim, int8_t *ret, uint8_t *iters, uint8_t *kconv, size_t n,
                         const harm_u2p_opts *opts, void *stream);

/* ---- Learned Newton seed (no reference counterpart) -----------------------
 * A small network predicts the Newton start (W, v^2) of the Noble 2D inversion
 * in place of the guess block utoprim_2d.c:268-307; the FP64 Newton solve
 * (utoprim_2d.c:442-541) refines it, and the previous-step guess is the
 * fallback where the seeded solve fails.  Network 8 -> 64 -> 64 -> 64 -> 2
 * (tanh), FP32 master parameters (HARM_MLP_NPARAMS floats, row-major W[out][in]:
 * W1 64x8, b1, W2 64x64, b2, W3 64x64, b3, W4 2x64, b4), trained on the GPU in
 * FP32; inference in BF16 on MFMA with FP32 accumulation.
 * Inputs (raw, standardized by harm_mlp_fit_norm): log D, asinh(E/D),
 * asinh(Qt^2/D^2), asinh(B^2/D), asinh((Q.B)^2/D^3), lapse, asinh((E-D-B^2/2)/D),
 * asinh(B^2/E) with D, Q, B of utoprim_2d.c:240-266 and E = -Q.n.
 * Outputs: log(W/Wref) (Wref = E - B^2/2 from the features), v^2.  Device buffers are on the handle's device. */
#define HARM_MLP_NPARAMS 9026
#define HARM_MLP_NORM 20      /* feature mean[8], 1/std[8], output mean[2], std[2] */
typedef struct harm_mlp harm_mlp;
harm_mlp *harm_mlp_create(int device);
void harm_mlp_destroy(harm_mlp *m);
/* Random initialization (uniform +-sqrt(3/fan_in), zero biases) from seed. */
int harm_mlp_init(harm_mlp *m, uint64_t seed);
int harm_mlp_set_params(harm_mlp *m, const float *params);  /* host array */
int harm_mlp_get_params(harm_mlp *m, float *params);
int harm_mlp_set_norm(harm_mlp *m, const double *norm);     /* host array */
int harm_mlp_get_norm(harm_mlp *m, double *norm);
/* 1: seed v^2 from the predicted W as utoprim_2d.c:306-307 does (default);
 * 0: use the network's second output. */
int harm_mlp_set_vsq_from_w(harm_mlp *m, int on);
/* Weights of the two outputs in the training loss (default 1, 1); with v^2 taken
 * from W the second output is unused and (1, 0) spends the network on W. */
int harm_mlp_set_loss_weights(harm_mlp *m, float w_logw, float w_vsq);
/* Training set (device): feat[8][n] raw FP64 features, targ[2][n] raw FP64
 * targets (log(W/Wref), v^2), valid[n] (0 rows carry no loss). */
int harm_mlp_fit_norm(harm_mlp *m, const double *feat, const double *targ, const uint8_t *valid, size_t n);
/* steps Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) on batches of `batch`
 * rows drawn with replacement from the set by a seeded hash; learning rate
 * decays geometrically from lr0 to lr1.  loss_trace (host, may be NULL) gets
 * the batch loss of every step. */
int harm_mlp_train(harm_mlp *m, const double *feat, const double *targ, const uint8_t *valid, size_t n,
                   int batch, int steps, float lr0, float lr1, uint64_t seed, float *loss_trace);
/* One Adam step on a standardized batch (device X[n][8], T[n][2], mask[n] or NULL). */
int harm_mlp_train_batch(harm_mlp *m, const float *X, const float *T, const uint8_t *mask, int n, float lr);
/* Loss 1/(2n) sum mask*|O-T|^2 and its gradient (host array of HARM_MLP_NPARAMS)
 * on a standardized batch, with the FP32 training kernels. */
int harm_mlp_grad(harm_mlp *m, const float *X, const float *T, const uint8_t *mask, int n, float *grad,
                  double *loss);
/* Forward pass on standardized device inputs X[n][8] -> Y[n][2]: bf16 = 0 runs
 * the FP32 training kernels, bf16 = 1 the BF16 MFMA inference path. */
int harm_mlp_forward(harm_mlp *m, const float *X, float *Y, int n, int bf16);
/* Raw features of n zones (device SoA as harm_utoprim2d_batch: U[8][n],
 * gcov[16][n], gcon[16][n], gdet[n]) -> feat[8][n]; with W0/vsq0 non-NULL also
 * the network's seed per zone (BF16 inference). */
int harm_mlp_seed(harm_mlp *m, const double *U, const double *gcov, const double *gcon, const double *gdet,
                  size_t n, double *feat, double *W0, double *vsq0, void *stream);

/* Training pairs from solved zones (device, SoA as harm_utoprim2d_batch):
 * feat[8][n] raw features of U, targ[2][n] = (log(W/Wref), v^2) of the converged
 * primitives prim, valid[n] = (ret[n] == 0). */
int harm_mlp_pairs(const double *U, const double *gcov, const double *gcon, const double *gdet, const double *prim,
                   const int8_t *ret, size_t n, double gam, double *feat, double *targ, uint8_t *valid, void *stream);

/* Step integration of the learned seed.  mode 0: previous-step guess (the
 * reference's Utoprim_2d); mode 1: learned seed with previous-step fallback;
 * mode 2: gated -- each zone starts from the previous-step guess, and only a
 * zone whose first Newton step |dW/W| + |dv2/v2| exceeds the gate threshold
 * (harm_set_guess_gate, default 0.1), or that has no valid start, solves from
 * the network's seed (a wave runs the network only when one of its zones asks);
 * the other zones give Utoprim_2d's result bit for bit.
 * The handle must outlive its use by the context. */
int harm_set_guess(harm_ctx *c, int mode, const harm_mlp *m);
int harm_set_guess_gate(harm_ctx *c, double tau);
/* Capture (U, P) training pairs from the plain inversion of later half-steps:
 * every stride-th active zone (phase rotating per half-step) appends its raw
 * features, targets and a valid flag to the device buffers feat[8][cap],
 * targ[2][cap], valid[cap] until cap rows are filled.  feat == NULL stops. */
int harm_set_capture(harm_ctx *c, double *feat, double *targ, uint8_t *valid, size_t cap, int stride);
int harm_get_capture_count(harm_ctx *c, size_t *n);

/* Synthetic stress batch for the inversion alone (no reference counterpart):
 * n zones with the Kerr-Schild metric of spin a in MKS coordinates (R0 = 0,
 * hslope) at r log-uniform in [r_lo, r_hi] and theta uniform in [th_lo, th_hi];
 * rho log-uniform, u/rho log-uniform, Lorentz factor uniform, isotropic velocity
 * and field directions, |B| set so that b^2/rho is log-uniform in
 * [sigma_lo, sigma_hi]; U = primtoU(P) (phys.c:100-105).  Zone z draws from a
 * KISS stream (random.c:30-43) seeded seed[k] + z.  guess_kind 0: the true P
 * with rho, u, U1..U3 each scaled by 1 + guess_pert*(2x-1); 1: cold (rho = D,
 * u = rho, zero velocity).  Device outputs in the harm_utoprim2d_batch layout:
 * U[8][n], gcov[16][n], gcon[16][n], gdet[n], prim_true[8][n], guess[8][n]. */
typedef struct harm_stress_spec {
    double a, hslope, gam;
    double r_lo, r_hi, th_lo, th_hi;
    double rho_lo, rho_hi, urho_lo, urho_hi;
    double lorentz_lo, lorentz_hi, sigma_lo, sigma_hi;
    double guess_pert;
    uint32_t seed[4];
    int32_t guess_kind, pad0;
} harm_stress_spec;
int harm_stress_batch(const harm_stress_spec *spec, size_t n, double *U, double *gcov, double *gcon, double *gdet,
                      double *prim_true, double *guess, void *stream);

/* ---- Tabulated EOS and the 5D inversion (utoprim_5d.c) --------------------
 * Table: rho[nr], T[nt] increasing; e[nr][nt], p[nr][nt] in code units (e is
 * the internal energy density u).  p(rho, u) = the piecewise-bilinear
 * interpolant in (rho, T), inverted along T at fixed rho; Gamma-law pressure
 * outside the table (as cooleos.c:342-366).  Replaces
 * pressure_rho0_u_CoolEOS / dpdrho_calc_CoolEOS / dpdu_calc_CoolEOS
 * (cooleos.c:314-444, 667-741) and the lazily filled Cache4d tables
 * (cache4d.c:863-1134). */
typedef struct harm_eos_spec {
    double rho_unit;               /* g/cm^3 per code density unit */
    double lrho_lo, lrho_hi;       /* log10 rho range, g/cm^3 */
    double lT_lo, lT_hi;           /* log10 T range, K */
    int32_t nr, nt;
    double ye;                     /* electron fraction */
} harm_eos_spec;
/* Fill a table (host arrays rho[nr], T[nt], e[nr*nt], p[nr*nt]) from the
 * build's closed-form EOS (degenerate electrons + ideal gas + radiation). */
int harm_eos_table_build(const harm_eos_spec *spec, double *rho, double *T, double *e, double *p);
typedef struct harm_eos_table {
    int32_t nr, nt;
    const double *rho, *T, *e, *p; /* device pointers */
    double gam;                    /* Gamma-law fallback outside the table */
} harm_eos_table;
/* p, dp/drho (fixed u), dp/du (fixed rho) for n (rho, u) pairs (device arrays). */
int harm_eos_eval(const harm_eos_table *tab, const double *rho, const double *u, double *p, double *dpdrho,
                  double *dpdu, size_t n, void *stream);
/* Run the step with the tabulated EOS (the reference's COOL_EOS build,
 * decs.h:110): pressures in fluxes, sources and U from the table, the
 * tabulated branch of vchar (phys.c:359-378), Utoprim_5d in place of
 * Utoprim_2d and no Font fix (step_ch.c:265-300).  ref_derivs = 1 zeroes
 * dp/drho and dp/du in vchar as the reference's vanishing finite-difference
 * step does (cooleos.c:667-741).  tab NULL returns to the Gamma-law EOS.  The
 * table's device arrays must outlive their use by the context. */
int harm_set_eos(harm_ctx *c, const harm_eos_table *tab, int ref_derivs);
/* Batched Utoprim_5d (utoprim_5d.c:179-239; decs.h:610-611): SoA device arrays
 * as harm_utoprim2d_batch; tab NULL selects the Gamma-law EOS with gam. */
int harm_utoprim5d_batch(const double *U, const double *gcov, const double *gcon, const double *gdet, double *prim,
                         int8_t *ret, uint8_t *iters, size_t n, double gam, const harm_eos_table *tab, void *stream);

/* Restart files (restart.c:73-383): the reference's ASCII layout, one header
 * line ("%10d" / "%32.22e", restart.c:105-138), one line of the 8 primitives
 * per zone in (h, i, j) order over all zones including ghosts, the ener.out
 * size when ener_size >= 0 (rank 0 in the reference, restart.c:150-157), and
 * with a tabulated EOS one temperature per zone (cooleos.c:996-1030).
 * Header fields in file order; the context owns N1..N3, t, tf, a, gam, cour,
 * dt, lim, Rin, Rout, hslope, R0, defcon, Rbr, mass_start, mass_in and
 * mass_out, the caller (the run driver) the dump/image counters and clocks. */
typedef struct harm_restart_hdr {
    int32_t N1, N2, N3, nstep;
    double t, tf, a, gam, cour, DTd, DTl, DTi;
    int32_t DTr, dump_cnt, image_cnt, rdump_cnt;
    double dt;
    int32_t lim, failed;
    double Rin, Rout, hslope, R0, mass_start, mass_in, mass_out, tdump, timage, tlog, defcon;
    int32_t nfailed, ener_size; /* ener_size < 0: no ener.out size in the file */
    double Rbr, cyl_R0, cyl_Th0;
} harm_restart_hdr;
/* restart_write(), restart.c:73-184: context-owned fields of hdr are replaced by
 * the context's values; hdr may be NULL (driver fields zero, no ener.out size).
 * With a tabulated EOS set, the zone temperatures of the table inversion follow
 * (0 where a zone lies outside the table). */
int harm_restart_write(harm_ctx *c, const char *path, const harm_restart_hdr *hdr);
/* restart_init()/restart_read(), restart.c:195-376: grid size, a, gam, Rin,
 * Rout, hslope, R0 and Rbr must equal the context's (create the context from
 * harm_restart_peek's header); cour, lim, tf and defcon are taken from the
 * file as the reference does.  Loads p, pflag = 0, t, dt, mass_start, mass_in,
 * mass_out, then bound_prim (restart.c:278-279).  hdr (may be NULL) receives the header. */
int harm_restart_read(harm_ctx *c, const char *path, harm_restart_hdr *hdr);
/* Gas physics of the cooling-table generator (the reference's Fortran chain
 * behind cooleos.c's cache, abundance.f90:6-243).  Device pointers; stream
 * may be NULL.
 * harm_fermi_batch: dfermi (gauss_integral.f90:241-349), the generalized
 * Fermi-Dirac integral F_k(eta, theta) = int_0^inf x^k sqrt(1 + x theta / 2)
 * / (exp(x - eta) + 1) dx and its eta and theta derivatives (fdeta, fdtheta
 * may be NULL), by Aparicio's four-interval 20-point quadrature. */
int harm_fermi_batch(const double *dk, const double *eta, const double *theta, double *fd, double *fdeta,
                     double *fdtheta, size_t n, void *stream);
/* The rules it uses: out[0..9] 20-point Gauss-Legendre positive nodes,
 * [10..19] their weights, [20..39] 20-point Gauss-Laguerre nodes, [40..59] their
 * weights times exp(node), [60..79] / [80..99] the 40-point Gauss-Legendre
 * positive nodes / weights of the neutrino-rate integrals (host call). */
int harm_fermi_rules(double *out /* 100 */);
/* One cooling-table point per element (abundance.f90:6-243): rho [g cm^-3], T
 * [K] and the disk-height factor HdFac -> out[6][n]: e, p [erg cm^-3], Qnu
 * [erg cm^-2 s^-1], optical depth tau, n_e-, n_e+ [fm^-3].  The chemical
 * equilibrium solve, the neutrino trapping iteration and all rates of a point
 * run in one group of lanes, which shares each Fermi-Dirac integral's
 * quadrature nodes (the sums in the serial order). */
int harm_eos_point_batch(const double *rho, const double *T, const double *hdfac, double *out, size_t n,
                         void *stream);
/* The same with one point per thread (every integral serial); the same bits. */
int harm_eos_point_batch_thread(const double *rho, const double *T, const double *hdfac, double *out, size_t n,
                                void *stream);
/* The step's tabulated EOS (harm_eos_table) from this generator: e and p of
 * abundance() at every point of the spec's log-spaced (rho, T) grid at one
 * disk-height factor hdfac [s] (the reference's per-radius cooling tables,
 * cooleos.c:908-973, with HdFac = 1/Omega(r)), in code units: rho / rho_unit,
 * e and p / (rho_unit c^2) (cooleos.c:114-135).  Host arrays rho[nr], T[nt],
 * e[nr*nt], p[nr*nt]; raw (may be NULL) receives the generator's own output
 * [6][nr*nt] in harm_eos_point_batch's units and order.  spec->ye is not used
 * (the generator's electron fraction follows from chemical equilibrium).
 * Returns 5 when a point has no positive finite e, p or e does not grow with T
 * (the inversion along T needs it). */
int harm_eos_table_fermi(const harm_eos_spec *spec, double hdfac, double *rho, double *T, double *e, double *p,
                         double *raw);
/* Per-species EOS of eos.f90 (energy density e and pressure p in MeV fm^-3,
 * entropy density s in k_B fm^-3) for n points; in is [nin][n], out [3][n]:
 *   HARM_EOS_NPE (eos_npe, eos.f90:31-113) nin 8: kT [MeV], eta_n, eta_p,
 *     eta_e (chemical potentials over kT, rest mass excluded), n_n, n_p, n_e-,
 *     n_e+ [fm^-3]; free nucleons and e-/e+ pairs;
 *   HARM_EOS_RAD (eos_rad, eos.f90:122-132) nin 1: kT;
 *   HARM_EOS_NEU (eos_neu, eos.f90:143-153) nin 3: kT, blocking factors b1, b2;
 *   HARM_EOS_HE  (eos_he, eos.f90:164-181) nin 2: kT, n_He. */
enum { HARM_EOS_NPE = 0, HARM_EOS_RAD = 1, HARM_EOS_NEU = 2, HARM_EOS_HE = 3 };
int harm_eos_species_batch(int kind, const double *in, double *out, size_t n, void *stream);

/* Cooling-table cache (cache4d.h:17-53, cache4d.c): D values at every point of
 * an (x, y, z) = (rho, T, HdFac) grid.  The header is the reference's packed
 * record, byte for byte; the file holds it, the three axes, the values
 * [z][x][y][d] and one computed flag per grid point (cache4d.c:263-316). */
#pragma pack(push, 1)
typedef struct harm_cache_desc {
    int32_t mode; /* 0: constant grid (the only kind looked up by grid), 1: prepared points */
    char ver[13]; /* "AGACACHE_2.0" ("AGACACHE_1.0" files store int flags) */
    int32_t x;
    double xmin, xmax, xrat;
    int32_t xtype; /* HARM_AXIS_* */
    int32_t y;
    double ymin, ymax, yrat;
    int32_t ytype;
    int32_t z;
    double zmin, zmax, zrat;
    int32_t ztype;
    int32_t d; /* values per grid point */
    int32_t unused, autosave, wholecachecalculated;
    uint64_t xcust, ycust, zcust; /* the writer's pointers to custom axes, kept as stored, never read */
    char reserved[76];
} harm_cache_desc;
#pragma pack(pop)
enum { HARM_AXIS_LOG = 0, HARM_AXIS_LINEAR = 1, HARM_AXIS_CUSTOM = 2, HARM_AXIS_NONUNIFORM = 3, HARM_AXIS_ROLLING = 4 };
enum { HARM_INTERP_LIN = 0, HARM_INTERP_SPLINE = 1 };
/* Lookup status per point: 1 done; 0 outside the grid (getvalues() returns 0);
 * 2 outside a rolling axis' inner range (the reference re-centres the grid and
 * recomputes it); 3 the window holds points not yet computed (the reference
 * computes them first); 4 z index outside [0, z). */
enum { HARM_CACHE_OUT_OF_RANGE = 0, HARM_CACHE_OK = 1, HARM_CACHE_NEEDS_RECENTRE = 2, HARM_CACHE_MISSING = 3,
       HARM_CACHE_BAD_ZINDEX = 4 };
typedef struct harm_cache harm_cache;
/* Axis values for an axis type (setindices, cache4d.c:329-388); host call. */
int harm_cache_axis(int type, int n, double vmin, double vmax, double rat, const double *cust, double *out);
/* Header of a cache file; then harm_cache_read fills host arrays sized from it
 * (vx[x], vy[y], vz[z], values[z*x*y*d], counted[z*x*y]) (load, cache4d.c:152-260). */
int harm_cache_peek(const char *path, harm_cache_desc *desc);
int harm_cache_read(const char *path, harm_cache_desc *desc, double *vx, double *vy, double *vz, double *values,
                    uint8_t *counted);
/* save(), cache4d.c:263-316: the same layout from host arrays. */
int harm_cache_write(const char *path, const harm_cache_desc *desc, const double *vx, const double *vy,
                     const double *vz, const double *values, const uint8_t *counted);
/* A device copy of a constant-grid cache on the current device (x, y >= 3
 * points, d <= 64, axes increasing by >= 1e-12), from host arrays or a file;
 * NULL on error.  It counts as whole when its header says so or every flag is 1. */
harm_cache *harm_cache_create(const harm_cache_desc *desc, const double *vx, const double *vy, const double *vz,
                              const double *values, const uint8_t *counted);
harm_cache *harm_cache_open(const char *path);
int harm_cache_info(const harm_cache *c, harm_cache_desc *desc, int *whole);
void harm_cache_free(harm_cache *c);
/* n lookups, one getvalues() each (cache4d.c:863-1134): x, y (device, n); z
 * (device, n; may be NULL when zindex is given) and zindex (device, n, may be
 * NULL: every point interpolates in z) where zindex[i] >= 0 takes that z plane
 * and -1 interpolates in z; out[d][n]; dery[d][n] (may be NULL) the derivative
 * along y (spline only); status[n] (may be NULL) HARM_CACHE_*.  dimmask (host,
 * d ints, may be NULL = all) selects the components the spline computes; the
 * others, like every output of a point whose status is not 1, are not written. */
int harm_cache_lookup(const harm_cache *c, int mode, const int *dimmask, const double *x, const double *y,
                      const double *z, const int *zindex, double *out, double *dery, int *status, size_t n,
                      void *stream);
/* Temperature solve of the COOL_EOS pressure and cooling routines
 * (cooleos.c:365-404, 545-577; mnewt.h:104-205 mnewtc<1>): for n zones (device
 * arrays, cgs) find T with cache component 0 (e) at (rho, T, z plane zindex
 * (NULL: plane 0)) equal to e, by up to ntrial bounded Newton steps from the
 * T given (the zone's last T), each kept inside [Tlo, Thi] (the reference uses
 * 300, TMin (1 + 1e-6), TMax (1 - 1e-6), tolf 1e-6).  T is overwritten with the
 * result, berr (may be NULL) gets |F| at it and ret (may be NULL) the
 * reference's code: 1 converged, 2 or 0 not (T = the iterate of least error),
 * -1 a lookup outside the cache.  Tlo = Thi = 0 selects the reference's bounds
 * from the ends of the cache's y axis; otherwise Tlo < Thi is required.  The
 * density-range and negative-e branches of those routines, and the unit
 * conversions, stay with the caller. */
int harm_cache_solve_T(const harm_cache *c, const double *rho, const double *e, const int *zindex, double *T,
                       double Tlo, double Thi, int ntrial, double tolf, double *berr, int *ret, size_t n, void *stream);

/* The step's COOL_EOS pressure on the cooling-table cache, as the reference's
 * COOL_EOS build computes it (replaces pressure_rho0_u_CoolEOS,
 * dpdrho_calc_CoolEOS and dpdu_calc_CoolEOS, cooleos.c:314-444 and 667-741,
 * with the LastT array and per-radius caches of InitCache_CoolEOS and
 * SetupCache_CoolEOS, cooleos.c:770-807, 826-930).  Every pressure the step
 * takes -- the face states of fluxcalc, the source and primtoU of the update,
 * the 5D inversion's residual and Jacobian, vchar's sound speed -- is a
 * temperature Newton solve (mnewtc<1>, 300 trials, 1e-6) on the cache of the
 * zone's radial index for e = u U_UNIT at rho = rho0 RHO_UNIT, started from the
 * zone's last temperature, which every converged solve replaces; then
 * p = P(rho, T) / P_UNIT from the same cache (component 0 e, 1 p, cgs, z plane
 * 0).  Outside [rho_min, rho_max] and where the solve does not converge the
 * pressure is (gam - 1) u; for e < 0 it is 0; a zone whose last temperature is
 * unset starts from (gam - 1) u / rho0 / k * m_av c^2 (at least
 * T_min (1 + 1e-6)), or from its j-1 neighbour's in the reference's loop order.
 * The step inverts with Utoprim_5d and has no Font fix, as with harm_set_eos.
 * caches: n = 1 (every radius) or N1+4 (one per radial index, ghosts included),
 * whole constant-grid caches sharing one (rho, T) grid (increasing axes, as
 * harm_cache_create requires), on the context's device; they must outlive
 * their use.  Zero ranges take the cache's axis ends
 * (xmin, xmax, ymin, ymax).  Every zone's last temperature starts unset.
 * Reconstruction MC, van Leer or minmod (lim 0-2).  caches NULL returns to the
 * Gamma-law EOS.  The ASCII dump, the tracers and the ener.out record are not
 * available with this EOS (they would change the zones' last temperatures). */
typedef struct harm_ceos_units {
    double rho_unit, u_unit, p_unit;         /* RHO_UNIT, U_UNIT, P_UNIT (cooleos.c:114-135) */
    double rho_min, rho_max, T_min, T_max;   /* rhoMin, rhoMax, TMin, TMax (cgs) */
} harm_ceos_units;
int harm_set_eos_cache(harm_ctx *c, const harm_cache *const *caches, int n, const harm_ceos_units *u);
/* The zones' last temperatures (restart_write_CoolEOS / restart_read_CoolEOS,
 * cooleos.c:1000-1030), (N1+4) x (N2+4) x N3 doubles, 0 = unset. */
int harm_get_last_T(harm_ctx *c, double *T);
int harm_set_last_T(harm_ctx *c, const double *T);
/* Temperature solves the cache EOS ran since it was set (repeated calls with
 * the same arguments excluded: they return the same bits without solving). */
int harm_get_eos_solves(harm_ctx *c, int64_t *n);

/* dump() ASCII file (dump.c:1710-1990; the HDFDUMP=0 build): a header line
 * (t, N1, N2 and the theta origin/spacing in 2D or N3 and the phi ones in 3D,
 * then tf, nstep, a, gam, cour, DTd, DTl, DTi, DTr, dump/image/rdump counters,
 * dt, lim, failed, Rin, Rout, hslope, R0), then one line per active zone in
 * (i, j, h) order of HARM_DUMP_NCOL "%32.22e" values (HARM_DUMP_NCOL_FAILED when
 * hdr->failed: the derived block is skipped, dump.c:1917).  The context owns
 * N1..N3, t, dt, tf, a, gam, cour, lim, Rin, Rout, hslope and R0; hdr (may be
 * NULL) gives the driver's counters and clocks. */
#define HARM_DUMP_NCOL 46
#define HARM_DUMP_NCOL_FAILED 20
int harm_dump_write(harm_ctx *c, const char *path, const harm_restart_hdr *hdr);
/* The same columns for active zones z0 .. z0+n-1 in (i, j, h) order, row-major
 * [n][HARM_DUMP_NCOL]: X1, X2 (3D: X3), r, theta (3D: phi), the 8 primitives,
 * Lambda, P, T, H_d, Qnu, tau, Ye (zero with the Gamma law; the table's P and
 * T with the tabulated EOS, the others zero), |div B| (the dump's corner
 * stencil), u^mu, u_mu, b^mu, b_mu, vmin/vmax along X1 and X2, gdet, |v|,
 * p = (gam-1) u, c_s, u^r (sqrt(g_rr) u^r/u^t), Mach. */
int harm_dump_columns(harm_ctx *c, size_t z0, size_t n, double *cols);
/* harm_dump_columns without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The columns
 * are staged in a fixed device array of HARM_DUMP_NOALLOC_MAX_ZONES zones
 * (2.9 MiB per GPU, held until the process ends); n above that is rejected
 * (error 3).  All calls on one GPU share that array: calls that overlap on one
 * GPU (two threads, two contexts on the same device) race. */
#define HARM_DUMP_NOALLOC_MAX_ZONES 8192
int harm_dump_columns_noalloc(harm_ctx *c, size_t z0, size_t n, double *cols);
/* The header alone (ener_size, which follows the body, reads as -1). */
int harm_restart_peek(const char *path, harm_restart_hdr *hdr);
/* Host-array forms (no device): prim in the canonical layout of
 * harm_get_prims (8 x (N1+4) x (N2+4) x N3), temp (may be NULL) (N1+4) x (N2+4)
 * x N3.  Reading fills prim (NULL: header only) and, when temp is not NULL and
 * the file holds the block, temp; *has_temp (may be NULL) says which. */
int harm_restart_write_host(const char *path, const harm_restart_hdr *hdr, const double *prim, const double *temp);
int harm_restart_read_host(const char *path, harm_restart_hdr *hdr, double *prim, double *temp, int *has_temp);

/* Exact reference signatures, evaluated on the GPU (current device) for one zone.
 * The reference takes the EOS from its globals (gam, init.c:91; the COOL_EOS
 * tables); here each calling thread sets its own: harm_set_gamma (default 4/3)
 * and harm_set_u2p_eos (a tabulated EOS for Utoprim_5d, device arrays that must
 * stay alive; NULL: the Gamma law).  Calls from several threads are serialized.
 * Utoprim_2d: utoprim_2d.c:126, decs.h:608-609.
 * Utoprim_5d: utoprim_5d.c:179-239, decs.h:610-611 (return code i*100 + j as
 * utoprim_5d.c:162 defines it, not truncated). */
int Utoprim_2d(double U[8], double gcov[4][4], double gcon[4][4], double gdet, double prim[8]);
int Utoprim_5d(double U[8], double gcov[4][4], double gcon[4][4], double gdet, double prim[8]);
void harm_set_gamma(double gam);
int harm_set_u2p_eos(const harm_eos_table *tab);

#ifdef __cplusplus
}
#endif
#endif
