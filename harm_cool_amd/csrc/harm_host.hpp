// harm_host.hpp -- host-side grid description and setup tables of libharmcool.so.
#pragma once
#include <cstddef>
#include <functional>
#include <string>
#include <vector>

#include "../../include/harm_cool.h"

namespace harm_host {

struct Grid {
    int N1, N2, N3, SN1, SN2;
    int is, ie, js, je;
    double startx[4], dx[4];
};

// Tables uploaded once per context.  ncol = SN1*SN2, column col = i*SN2 + j.
struct GeomTables {
    std::vector<double> geom[5];  // per location (FACE1,FACE2,CORN,CENT,FACE3): [33][ncol]
    std::vector<double> conn;     // [64][ncol], index (k*4+mu)*4+j = Gamma^k_{mu j}
    std::vector<double> coords;   // [4][ncol] Boyer-Lindquist (t, r, theta, phi) at CENT
    std::vector<double> floors;   // [2][ncol] radius-only rho and u floors
    std::vector<double> katm;     // [SN1] Font-fix isentrope constant per radius
};

int build_geometry(const harm_params &P, Grid &g, GeomTables &t);
// Sets the message harm_last_error() returns; returns code.
int set_error(const std::string &msg, int code);

// Restart file text (harm_restart.cpp); prim/temp in the canonical layout.
// The tracer block of a restart file (restart_write_tracers, tracers.c:565-606):
// Rin, Rout, ThetaMin, ThetaMax, Rank_Rin, Rank_Rout, Rank_Phin, Rank_Phout,
// the count, then per tracer i, j, h (zero-based), X[0..3] and its file size.
struct TracerBlock {
    double par[8];
    std::vector<int32_t> i, j, h, size;
    std::vector<double> X;  // 4 per tracer
};
int restart_write_file(const char *path, const harm_restart_hdr &h, const double *prim, const double *temp,
                       const TracerBlock *tb = nullptr);
// tb (may be NULL) receives the tracer block when the file holds one; *has_tb
// says whether it did.
int restart_read_file(const char *path, harm_restart_hdr &h, double *prim, double *temp, int *has_temp,
                      TracerBlock *tb = nullptr, int *has_tb = nullptr);
// dump() text (dump.c:1710-1990): the header line, then one line of ncol
// values per zone; fill(z0, n, cols) puts zones z0..z0+n-1, stride values each.
std::string dump_header_text(const harm_restart_hdr &h, const double *startx, const double *dx);
int dump_write_file(const char *path, const std::string &header, size_t nzones, int ncol, int stride,
                    const std::function<int(size_t, size_t, double *)> &fill);

int build_torus2d(const harm_params &P, const Grid &g, const GeomTables &t, std::vector<double> &tor, double &rhomax,
                  double &umax);

}  // namespace harm_host
