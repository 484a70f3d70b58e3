/*
 * ref_driver.cpp -- TEST INFRASTRUCTURE.  Thin extern "C" shim that links the
 * reference's own per-zone sources, compiled where they lie in the reference
 * tree (src/utoprim_2d.c, u2p_util.c, utoprim_1dvsq2fix1.c, utoprim_1dfix1.c,
 * coord.c, metric.c, lu.c, interp.c, random.c), into oracle/_ref/libharmref.so.
 *
 * It exists only to pin oracle/harm_oracle.c: tests call the same inputs through
 * both and require bitwise-equal outputs.  primtoU_g is not exported: as compiled
 * from the reference tree (COOL_EOS=1 in decs.h) it calls the tabulated-EOS
 * pressure, which lives in files that need the Fortran EOS chain.  The reference's globals are defined
 * by including its own defs.h; fail() is the only symbol the chosen files use
 * from a file that is not compiled (diag.c), and here it just records the code.
 */
#include "decs.h"
#include "defs.h"
#include "u2p_defs.h"
#include "random.h"

extern "C" {

static int g_last_fail = 0;
}
void fail(int fail_type) { g_last_fail = fail_type; }

void bl_coord_vec(double *X, double *V);

extern "C" {

int ref_last_fail(void) { return g_last_fail; }

void ref_set_params(double a_, double gam_, double hslope_, double R0_, double Rbr_)
{
    a = a_;
    gam = gam_;
    hslope = hslope_;
    R0 = R0_;
    Rbr = Rbr_;
}

int ref_utoprim_2d(const double *U, const double *gcov, const double *gcon, double gdet, double *prim)
{
    double Uc[NPR], gc[NDIM][NDIM], gn[NDIM][NDIM];
    memcpy(Uc, U, sizeof(Uc));
    memcpy(gc, gcov, sizeof(gc));
    memcpy(gn, gcon, sizeof(gn));
    return Utoprim_2d(Uc, gc, gn, gdet, prim);
}

int ref_utoprim_1dvsq2fix1(const double *U, const double *gcov, const double *gcon, double gdet, double *prim,
                           double K)
{
    double Uc[NPR], gc[NDIM][NDIM], gn[NDIM][NDIM];
    memcpy(Uc, U, sizeof(Uc));
    memcpy(gc, gcov, sizeof(gc));
    memcpy(gn, gcon, sizeof(gn));
    return Utoprim_1dvsq2fix1(Uc, gc, gn, gdet, prim, K);
}

int ref_utoprim_1dfix1(const double *U, const double *gcov, const double *gcon, double gdet, double *prim, double K)
{
    double Uc[NPR], gc[NDIM][NDIM], gn[NDIM][NDIM];
    memcpy(Uc, U, sizeof(Uc));
    memcpy(gc, gcov, sizeof(gc));
    memcpy(gn, gcon, sizeof(gn));
    return Utoprim_1dfix1(Uc, gc, gn, gdet, prim, K);
}

void ref_gcov_func(const double *X, double *gcov)
{
    double Xc[NDIM];
    memcpy(Xc, X, sizeof(Xc));
    gcov_func(Xc, (double(*)[NDIM])gcov);
}

void ref_dxdxp_func(const double *X, double *dxdxp)
{
    double Xc[NDIM];
    memcpy(Xc, X, sizeof(Xc));
    dxdxp_func(Xc, (double(*)[NDIM])dxdxp);
}

void ref_bl_coord_vec(const double *X, double *V)
{
    double Xc[NDIM];
    memcpy(Xc, X, sizeof(Xc));
    bl_coord_vec(Xc, V);
}

void ref_conn_func(const double *X, const double *gcon, double *conn)
{
    double Xc[NDIM];
    struct of_geom geom;
    memcpy(Xc, X, sizeof(Xc));
    memcpy(geom.gcon, gcon, sizeof(geom.gcon));
    conn_func(Xc, &geom, (double(*)[NDIM][NDIM])conn);
}

double ref_gdet_func(const double *gcov)
{
    double gc[NDIM][NDIM];
    memcpy(gc, gcov, sizeof(gc));
    return gdet_func(gc);
}

int ref_invert_matrix(const double *A, double *Ainv)
{
    double a_[NDIM][NDIM];
    memcpy(a_, A, sizeof(a_));
    return invert_matrix(a_, (double(*)[NDIM])Ainv);
}

double ref_slope_lim(int lim_, double y1, double y2, double y3)
{
    lim = lim_;
    return slope_lim(y1, y2, y3);
}

double ref_kiss_uniform(unsigned int z, unsigned int w, unsigned int jsr, unsigned int jcong)
{
    RANDOMDEF rd = {z, w, jsr, jcong};
    return UniformRandomRange(&rd, 0.0, 1.0);
}

/* metric.c:62-115 for a single rank: startx/dx set by the caller */
void ref_coord(int h, int i, int j, int loc, const double *startx_, const double *dx_, double *X)
{
    for (int k = 0; k < NDIM; k++) {
        startx[k] = startx_[k];
        dx[k] = dx_[k];
    }
    coord(h, i, j, loc, X);
}
}
