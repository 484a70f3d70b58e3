// harm_cool.hip -- gfx950 kernels and the C ABI of libharmcool.so.
//
// One step of the reference (step_ch.c:90-140) maps to this launch sequence on
// the context's stream, with the whole domain resident in HBM:
//   per half-step:  flux X1, flux X2, [flux X3]   (fluxcalc + fix_flux fused)
//                   emf                           (flux_ct EMF averages)
//                   update                        (source + flux divergence +
//                                                  con2prim + Font fix + fixup)
//                   bound, fixutoprim scan/serial, bound
//   end of step:    mdot partials, stepend        (diag_mass, dt control)
// dt and t live in device memory, so a step needs no host round trip.
//
// Layout: SoA, phi fastest.  Zone z = (i*SN2 + j)*N3 + h, variable stride nz.
#include <hip/hip_runtime.h>

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "harm_host.hpp"
#include "harm_phys.hpp"
#include "harm_mlp.hpp"
#include "harm_eos5d.hpp"
#include "harm_ceos.hpp"

int harm_cache_view_of(const harm_cache *c, harm::CacheView &v, harm_cache_desc &d, int &device);

using namespace harm;

namespace {

thread_local std::string g_err;

int set_err(const std::string &m, int code = 1)
{
    g_err = m;
    return code;
}

#define HIPCHK(x)                                                                               \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess)                                                                   \
            return set_err(std::string(#x) + ": " + hipGetErrorString(e_) + " @" + std::to_string(__LINE__), 2); \
    } while (0)

constexpr int TPB = 256;

struct DGrid {
    int N1, N2, N3, SN1, SN2;
    int is, ie, js, je;
    int lim, font_fix, polefix;
    size_t ncol, nz;
    double dx1, dx2, dx3, gam, cour, hllf, laxf, defcon, tf;
    double rdx1, rdx2, rdx3;  // 1/dx, correctly rounded (div_r in the flux divergence)
    const double *geom[5];
    const double *conn;
    const double *floors;
    const double *katm;
    const double *cblk;     // column-major CENT block, CENT_STRIDE per column
    const double *fblk[3];  // column-major FACE1/FACE2/FACE3 blocks, FACE_STRIDE per column
    EosTab tab;             // tabulated EOS (used by the TAB kernel variants)
    int ref_derivs;         // tabulated EOS: zero dp/drho, dp/du as the reference's
    // Inside harm_step: the step-state error word (tstate[4]).  Once a step has
    // ended in an error every later kernel of the batch returns at once, so the
    // state stays as the reference leaves it when it exits (step_ch.c:126-129).
    // nullptr outside harm_step.
    const double *halt;
};

// Wave-uniform: true when the batch has stopped (see DGrid::halt).
__device__ __forceinline__ bool halted(const double *halt) { return halt && *halt != 0.; }
__device__ __forceinline__ bool halted(const DGrid &g) { return halted(g.halt); }

__device__ __forceinline__ size_t zidx(const DGrid &g, int i, int j, int h)
{
    return ((size_t)i * g.SN2 + j) * g.N3 + h;
}

// Element zz of an array whose base is wave-uniform, addressed as base plus a
// 32-bit byte offset (the SGPR-base form of global memory instructions: one
// offset register serves the same zone of every variable).  Valid while
// 8 * nz < 2^32 (OFF32_MAX_ZONES), which the host checks before it selects
// the kernels that use it.
constexpr size_t OFF32_MAX_ZONES = (size_t)1 << 29;
__device__ __forceinline__ const double &at32(const double *base, unsigned zz)
{
    return *(const double *)((const char *)base + (size_t)(zz * 8u));
}
__device__ __forceinline__ double &at32(double *base, unsigned zz)
{
    return *(double *)((char *)base + (size_t)(zz * 8u));
}

// Logical block of this workgroup.  Consecutive hardware workgroups are
// dispatched to the 8 XCDs in turn, and each XCD has its own L2; with this
// mapping each XCD walks one contiguous range of logical blocks, so the stencil
// neighbours that adjacent logical blocks read come from the same L2.  Grids
// using it are launched with a multiple of 8 workgroups.
__device__ __forceinline__ unsigned xcd_block()
{
    return (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
}

// Zone (i, j, h) of logical index idx over an ni x nj x N3 box at (i0, j0).
// I_FAST orders the columns with i fastest (for stencils along X1), else j.
template <bool I_FAST>
__device__ __forceinline__ void box_zone(unsigned idx, unsigned n3, int i0, int j0, unsigned ni, unsigned nj, int &i,
                                         int &j, int &h)
{
    const unsigned rest = idx / n3;
    h = (int)(idx - rest * n3);
    if (I_FAST) {
        const unsigned rj = rest / ni;
        i = i0 + (int)(rest - rj * ni);
        j = j0 + (int)rj;
    } else {
        const unsigned ri = rest / nj;
        j = j0 + (int)(rest - ri * nj);
        i = i0 + (int)ri;
    }
}

// Same box, columns walked in tiles of TI rows of i with i fastest inside a
// tile: the i+-1 and j+-1 neighbours of a column are visited at most TI columns
// away, so 2D stencil kernels find them in L2 (the last tile may be shorter).
template <unsigned TI>
__device__ __forceinline__ void box_zone_tiled(unsigned idx, unsigned n3, int i0, int j0, unsigned ni, unsigned nj,
                                               int &i, int &j, int &h)
{
    const unsigned rest = idx / n3;
    h = (int)(idx - rest * n3);
    const unsigned tile = rest / (TI * nj);
    const unsigned r = rest - tile * (TI * nj);
    const unsigned ti0 = tile * TI;
    const unsigned th = min(TI, ni - ti0);
    const unsigned jj = r / th;
    i = i0 + (int)(ti0 + (r - jj * th));
    j = j0 + (int)jj;
}
constexpr unsigned STENCIL_TILE = 8;

// ---- wave-level reductions (64 lanes) ----
__device__ __forceinline__ int wave_sum_int(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_min_dbl(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, 64);
        v = (w < v) ? w : v;
    }
    return v;
}
__device__ __forceinline__ void add_counter(unsigned long long *cnt, int idx, int v)
{
    int s = wave_sum_int(v);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&cnt[idx], (unsigned long long)s);
}

// Per-block totals of N per-thread counts, stored (no atomics) at
// out[blockIdx.x*N + k]; every thread of the block must call it.  Device-wide
// atomics from every wave to one address serialize across the XCDs and stall
// the memory pipeline of the whole kernel; k_sum_partials folds these instead.
template <int N>
__device__ __forceinline__ void block_counts(const int (&v)[N], unsigned long long *__restrict__ out)
{
    __shared__ int red[TPB / 64][N];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; k++) {
        const int s = wave_sum_int(v[k]);
        if ((threadIdx.x & 63) == 0) red[w][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        unsigned long long t = 0;
#pragma unroll
        for (int q = 0; q < TPB / 64; q++) t += (unsigned long long)red[q][threadIdx.x];
        out[(size_t)blockIdx.x * N + threadIdx.x] = t;
    }
}

// Per-block minimum of v stored at out[blockIdx.x]; every thread of the NT-thread
// block must call it.
template <int NT = TPB>
__device__ __forceinline__ void block_min(double v, double *__restrict__ out)
{
    __shared__ double red[NT / 64];
    const double m = wave_min_dbl(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
#pragma unroll
        for (int q = 1; q < NT / 64; q++) t = (red[q] < t) ? red[q] : t;
        out[blockIdx.x] = t;
    }
}

struct CntMap {
    int n;
    int idx[8];
};

// cnt[map.idx[k]] += sum over blocks b of part[b*map.n + k].  Grid
// (RED_BLOCKS, map.n): block (x, k) folds every RED_BLOCKS-th partial of
// counter k and adds its total with one atomic.
constexpr int RED_BLOCKS = 64;
__global__ void __launch_bounds__(256) k_sum_partials(const unsigned long long *__restrict__ part, size_t nb,
                                                      CntMap map, unsigned long long *__restrict__ cnt,
                                                      const double *__restrict__ halt)
{
    if (halted(halt)) return;
    __shared__ unsigned long long red[256];
    const int k = blockIdx.y;
    unsigned long long s = 0;
    for (size_t b = (size_t)blockIdx.x * 256 + threadIdx.x; b < nb; b += (size_t)RED_BLOCKS * 256)
        s += part[b * map.n + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicAdd(&cnt[map.idx[k]], red[0]);
}

// slot = min(slot, min over blocks of part[b]) where that minimum is < 1e9
// (the CFL limit of a direction, step_ch.c:477-484); grid RED_BLOCKS.
__global__ void __launch_bounds__(256) k_min_partials(const double *__restrict__ part, size_t nb,
                                                      unsigned long long *__restrict__ slot,
                                                      const double *__restrict__ halt)
{
    if (halted(halt)) return;
    __shared__ double red[256];
    double m = 1.e300;
    for (size_t b = (size_t)blockIdx.x * 256 + threadIdx.x; b < nb; b += (size_t)RED_BLOCKS * 256)
        m = (part[b] < m) ? part[b] : m;
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o && red[threadIdx.x + o] < red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] < 1.e9) atomicMin(slot, (unsigned long long)__double_as_longlong(red[0]));
}

// ---------------------------------------------------------------------------
// fluxcalc (step_ch.c:331-504) + fix_flux (coord.c:264-324), one face per thread.
// DIR 1,2,3 = X1,X2,X3.  Slopes of the two zones adjacent to the face are formed
// on the fly from four zones along DIR, as the reference's dq array holds them.
// ---------------------------------------------------------------------------
// Columns are walked with the stencil direction fastest (X1: i, else j) and
// XCD-contiguous logical blocks, so the four zones each face reads are mostly
// L2 hits; 4 waves per SIMD (128 VGPRs) measured faster than 3 with a few
// spilled registers.
// LIM >= 0 fixes the slope limiter at compile time (the hot MC path); -1 reads
// it from the grid.  ROWS (DIR 3 with N3 dividing TPB): the block holds whole
// phi rows, so each thread loads only its own zone and forms only its own
// zone's slopes; the neighbour's values come through LDS.
template <int DIR, bool UNI, bool TAB, int LIM, bool ROWS = false>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_flux(DGrid g, const double *__restrict__ pb, double *__restrict__ F, double *__restrict__ dtpart, int fullstep,
       unsigned long long *__restrict__ cnt)
{
    static_assert(!ROWS || DIR == 3, "row-local reconstruction is along phi only");
    if (halted(g)) return;
    const int i0 = (DIR == 1) ? g.is : g.is - 1;
    const int j0 = (DIR == 2) ? g.js : g.js - 1;
    const int ni = g.ie + 1 - i0 + 1;
    const int nj = g.je + 1 - j0 + 1;
    const unsigned idx = xcd_block() * TPB + threadIdx.x;
    const unsigned total = (unsigned)ni * nj * g.N3;
    const bool act = idx < total;
    const int lim = (LIM >= 0) ? LIM : g.lim;
    const size_t nz = g.nz;
    double dtij = 1.e300;
    int fatal = 0;
    int i = 0, j = 0, h = 0;
    size_t z = 0;
    if (act) {
        box_zone<DIR == 1>(idx, g.N3, i0, j0, ni, nj, i, j, h);
        z = zidx(g, i, j, h);
    }
    const bool zero_face = act && DIR == 2 && (j == g.js || j == g.je + 1);
    double p_l[NPR], p_r[NPR];
    if constexpr (ROWS) {
        // whole rows per block (a row is all active or all inactive); every
        // thread reaches both barriers
        __shared__ double sp[NPR][TPB], sdq[NPR][TPB];
        const unsigned t = threadIdx.x;
        double pc[NPR], pm[NPR], dq[NPR];
        if (act) {
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                pc[k] = pb[k * nz + z];
                sp[k][t] = pc[k];
            }
        }
        __syncthreads();
        const unsigned rs = t - (unsigned)h;
        const unsigned tm1 = rs + (unsigned)((h + g.N3 - 1) % g.N3), tp1 = rs + (unsigned)((h + 1) % g.N3);
        if constexpr (LIM == LIM_PPM) {
            // each zone's MC slope, then its parabola from its two neighbours'
            // values and slopes; the face takes zone h-1's right value and zone
            // h's left value (the right values go through sp once every thread
            // has read the primitives)
            if (act) {
#pragma unroll
                for (int k = 0; k < NPR; k++) {
                    dq[k] = slope_lim(LIM_MC, sp[k][tm1], pc[k], sp[k][tp1]);
                    sdq[k][t] = dq[k];
                }
            }
            __syncthreads();
            double qr[NPR];
            if (act) {
#pragma unroll
                for (int k = 0; k < NPR; k++)
                    ppm_para(sp[k][tm1], pc[k], sp[k][tp1], sdq[k][tm1], dq[k], sdq[k][tp1], p_r[k], qr[k]);
            }
            __syncthreads();
            if (act) {
#pragma unroll
                for (int k = 0; k < NPR; k++) sp[k][t] = qr[k];
            }
            __syncthreads();
            if (act) {
#pragma unroll
                for (int k = 0; k < NPR; k++) p_l[k] = sp[k][tm1];
            }
        } else {
            if (act) {
#pragma unroll
                for (int k = 0; k < NPR; k++) {
                    pm[k] = sp[k][tm1];
                    dq[k] = slope_lim(lim, pm[k], pc[k], sp[k][tp1]);
                    sdq[k][t] = dq[k];
                }
            }
            __syncthreads();
            if (act) {
#pragma unroll
                for (int k = 0; k < NPR; k++) {
                    p_l[k] = pm[k] + 0.5 * sdq[k][tm1];
                    p_r[k] = pc[k] - 0.5 * dq[k];
                }
            }
        }
    } else if (LIM == LIM_PPM && act && !zero_face) {
        // six zones m-3 .. m+2 along DIR; a zone whose five-zone stencil leaves
        // the two ghost layers (the first and last radial zones' neighbours
        // outside) keeps the linear MC state, so m-3 / m+2 are read only where
        // the stencil is inside the grid
        size_t zz[6];
        bool ppm_l = true, ppm_r = true;
        if (DIR == 3) {
            const size_t base = z - h;
#pragma unroll
            for (int q = 0; q < 6; q++) zz[q] = base + (h + 3 * g.N3 + q - 3) % g.N3;
        } else {
            const size_t s = (DIR == 1) ? (size_t)g.SN2 * g.N3 : (size_t)g.N3;
            const int m = (DIR == 1) ? i : j, lo = (DIR == 1) ? g.is : g.js, hi = (DIR == 1) ? g.ie : g.je;
            ppm_l = m - 1 >= lo && m - 1 <= hi;
            ppm_r = m >= lo && m <= hi;
            zz[0] = ppm_l ? z - 3 * s : z - 2 * s;
            zz[1] = z - 2 * s;
            zz[2] = z - s;
            zz[3] = z;
            zz[4] = z + s;
            zz[5] = ppm_r ? z + 2 * s : z + s;
        }
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            double y[6], dq[4];
#pragma unroll
            for (int q = 0; q < 6; q++) y[q] = pb[k * nz + zz[q]];
#pragma unroll
            for (int q = 0; q < 4; q++) dq[q] = slope_lim(LIM_MC, y[q], y[q + 1], y[q + 2]);
            // dq[q] is the slope of zone q+1 (zones m-2 .. m+1)
            double l, r;
            if (ppm_l) {
                ppm_para(y[1], y[2], y[3], dq[0], dq[1], dq[2], l, r);
                p_l[k] = r;
            } else {
                p_l[k] = y[2] + 0.5 * dq[1];
            }
            if (ppm_r) {
                ppm_para(y[2], y[3], y[4], dq[1], dq[2], dq[3], l, r);
                p_r[k] = l;
            } else {
                p_r[k] = y[3] - 0.5 * dq[2];
            }
        }
    } else if (act && !zero_face) {
        size_t zm2, zm1, zp1;
        if (DIR == 1) {
            const size_t s = (size_t)g.SN2 * g.N3;
            zm2 = z - 2 * s;
            zm1 = z - s;
            zp1 = z + s;
        } else if (DIR == 2) {
            const size_t s = (size_t)g.N3;
            zm2 = z - 2 * s;
            zm1 = z - s;
            zp1 = z + s;
        } else {
            const size_t base = z - h;
            zm2 = base + (h + 2 * g.N3 - 2) % g.N3;
            zm1 = base + (h + g.N3 - 1) % g.N3;
            zp1 = base + (h + 1) % g.N3;
        }
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            const double a = pb[k * nz + zm2], b = pb[k * nz + zm1], c = pb[k * nz + z], d = pb[k * nz + zp1];
            const double dqm = slope_lim(lim, a, b, c);
            const double dq0 = slope_lim(lim, b, c, d);
            p_l[k] = b + 0.5 * dqm;
            p_r[k] = c - 0.5 * dq0;
        }
    }
    if (zero_face) {
#pragma unroll
        for (int k = 0; k < NPR; k++) F[k * nz + z] = 0.;
    } else if (act) {
        Geom geom;
        const int loc = (DIR == 1) ? FACE1 : (DIR == 2) ? FACE2 : FACE3;
        if (UNI)
            load_geom_block<FACE_DERIVED>(g.fblk[DIR - 1] + uni_col<UNI>((size_t)i * g.SN2 + j) * FACE_STRIDE, geom);
        else
            load_geom(g.geom[loc], g.ncol, (size_t)i * g.SN2 + j, geom);
        double Fz[NPR], ctop;
        if constexpr (TAB)
            face_flux(p_l, p_r, geom, DIR, EosTable{g.tab, g.ref_derivs}, g.hllf, g.laxf, Fz, ctop, fatal);
        else
            face_flux(p_l, p_r, geom, DIR, g.gam, g.hllf, g.laxf, Fz, ctop, fatal);
        if (DIR == 1) {
            // coord.c:310-320: no inflow of mass through the radial boundaries
            if (i == g.is && Fz[RHO] > 0) Fz[RHO] = 0;
            if (i == g.ie + 1 && Fz[RHO] < 0) Fz[RHO] = 0;
        }
        // coord.c:290-306: antisymmetric B2 flux in the first polar ghost face
        const bool skip_b2 = (DIR != 2) && (j == g.js - 1 || j == g.je + 1);
#pragma unroll
        for (int k = 0; k < NPR; k++)
            if (!(k == B2 && skip_b2)) F[k * nz + z] = Fz[k];
        if (DIR != 2) {
            if (j == g.js) F[B2 * nz + zidx(g, i, g.js - 1, h)] = -Fz[B2];
            if (j == g.je) F[B2 * nz + zidx(g, i, g.je + 1, h)] = -Fz[B2];
        }
        if (fullstep && i >= g.is && i <= g.ie && j >= g.js && j <= g.je) {
            const double dxd = (DIR == 1) ? g.dx1 : (DIR == 2) ? g.dx2 : g.dx3;
            const double d = g.cour * dxd / ctop;
            if (d < 1.e9) dtij = d;
        }
    }
    if (fullstep) block_min(dtij, dtpart);
    add_counter(cnt, HARM_FATAL, fatal);
}

// The X1 / X2 fluxes of the wave-uniform path with the reconstruction shared
// along the stencil direction (TAB, LIM as k_flux).  A block is STRIP waves:
// wave w computes face f0 + w (f0 = first face of the block's strip) for 64 phi
// zones of one (i, j) column.  The block loads the STRIP+3 zone rows
// f0-2 .. f0+STRIP once and forms the slopes of rows f0-1 .. f0+STRIP-1 once
// (per face (STRIP+3)/STRIP loads instead of 4 and (STRIP+1)/STRIP slopes
// instead of 2); the reconstructed states p[r] +- dq[r]/2 are exchanged through
// LDS.  Same arithmetic as k_flux.
constexpr int STRIP = 4;                // faces per block (8 and 16 measured 3% and 11% slower)
constexpr int STRIP_TPB = 64 * STRIP;

// The face of a strip kernel (DIR 1 or 2, face index f along DIR, other index
// o, phi zone h): step_ch.c:429-475 from the reconstructed states, fix_flux
// (coord.c:264-324) and the CFL limit.
template <int DIR, bool TAB>
__device__ __forceinline__ void strip_face(const DGrid &g, double *__restrict__ F, const double *p_l,
                                           const double *p_r, int f, int o, int h, int fullstep, double &dtij,
                                           int &fatal)
{
    const size_t nz = g.nz;
    const int i = (DIR == 1) ? f : o, j = (DIR == 1) ? o : f;
    const size_t z = zidx(g, i, j, h);
    if (DIR == 2 && (j == g.js || j == g.je + 1)) {
#pragma unroll
        for (int k = 0; k < NPR; k++) F[k * nz + z] = 0.;
        return;
    }
    Geom geom;
    load_geom_block<FACE_DERIVED>(g.fblk[DIR - 1] + uni_col<true>((size_t)i * g.SN2 + j) * FACE_STRIDE, geom);
    double Fz[NPR], ctop;
    if constexpr (TAB)
        face_flux(p_l, p_r, geom, DIR, EosTable{g.tab, g.ref_derivs}, g.hllf, g.laxf, Fz, ctop, fatal);
    else
        face_flux(p_l, p_r, geom, DIR, g.gam, g.hllf, g.laxf, Fz, ctop, fatal);
    if (DIR == 1) {
        // coord.c:310-320: no inflow of mass through the radial boundaries
        if (i == g.is && Fz[RHO] > 0) Fz[RHO] = 0;
        if (i == g.ie + 1 && Fz[RHO] < 0) Fz[RHO] = 0;
    }
    // coord.c:290-306: antisymmetric B2 flux in the first polar ghost face
    const bool skip_b2 = (DIR != 2) && (j == g.js - 1 || j == g.je + 1);
#pragma unroll
    for (int k = 0; k < NPR; k++)
        if (!(k == B2 && skip_b2)) F[k * nz + z] = Fz[k];
    if (DIR != 2) {
        if (j == g.js) F[B2 * nz + zidx(g, i, g.js - 1, h)] = -Fz[B2];
        if (j == g.je) F[B2 * nz + zidx(g, i, g.je + 1, h)] = -Fz[B2];
    }
    if (fullstep && i >= g.is && i <= g.ie && j >= g.js && j <= g.je) {
        const double dxd = (DIR == 1) ? g.dx1 : g.dx2;
        const double d = g.cour * dxd / ctop;
        if (d < 1.e9) dtij = d;
    }
}

template <int DIR, bool TAB, int LIM>
__global__ void __launch_bounds__(STRIP_TPB) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_flux_strip(DGrid g, const double *__restrict__ pb, double *__restrict__ F, double *__restrict__ dtpart,
             int fullstep, unsigned long long *__restrict__ cnt)
{
    static_assert(DIR == 1 || DIR == 2, "strips run along X1 or X2");
    if (halted(g)) return;
    // faces along DIR: [fs, fe]; the other index o over [os, oe]
    const int fs = (DIR == 1) ? g.is : g.js, fe = (DIR == 1) ? g.ie + 1 : g.je + 1;
    const int os = (DIR == 1) ? g.js - 1 : g.is - 1, oe = (DIR == 1) ? g.je + 1 : g.ie + 1;
    const unsigned nstrip = (unsigned)(fe - fs + STRIP) / STRIP, no = (unsigned)(oe - os + 1);
    const unsigned nseg = (unsigned)g.N3 / 64;
    const unsigned b = xcd_block();
    const bool blk_act = b < nstrip * no * nseg;
    const unsigned s = b % nstrip, rest = b / nstrip;
    const int o = os + (int)(rest % no);
    const unsigned seg = rest / no;
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int h = (int)seg * 64 + lane;
    const int f0 = fs + (int)s * STRIP;
    const int sn = (DIR == 1) ? g.SN1 : g.SN2;
    const int lim = (LIM >= 0) ? LIM : g.lim;
    const size_t nz = g.nz;
    // zone of row r (r = 0..STRIP+2 is zone index f0 - 2 + r along DIR) in this lane's column
    auto zone = [&](int r) -> size_t {
        const int zd = f0 - 2 + r;
        return (DIR == 1) ? zidx(g, zd, o, h) : zidx(g, o, zd, h);
    };
    // LDS: B[STRIP+4 rows][8 vars][64 lanes]; first the primitives of rows
    // 0..STRIP+2, then the right states p-dq/2 of rows 2..STRIP+1 in
    // B[4..STRIP+3] (left states stay in registers)
    __shared__ double B[STRIP + 4][NPR][64];
    // phase 1: row w+1 by every wave, rows 0, STRIP+1, STRIP+2 by waves 0, 1, 2
    const int r1 = w + 1, r2 = (w == 0) ? 0 : (w == 1) ? STRIP + 1 : (w == 2) ? STRIP + 2 : -1;
    double pc[NPR];
    if (blk_act && f0 - 2 + r1 < sn) {
        const size_t zz = zone(r1);
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            pc[k] = pb[k * nz + zz];
            B[r1][k][lane] = pc[k];
        }
    }
    if (blk_act && r2 >= 0 && f0 - 2 + r2 < sn) {
        const size_t zz = zone(r2);
#pragma unroll
        for (int k = 0; k < NPR; k++) B[r2][k][lane] = pb[k * nz + zz];
    }
    __syncthreads();
    // phase 2: slopes of row w+1, and of row STRIP+1 by the last wave
    double dq[NPR], dq5[NPR], p5[NPR];
    const bool row1_ok = blk_act && f0 - 1 + r1 < sn;  // rows r1-1 .. r1+1 loaded
    const bool row5_ok = blk_act && w == STRIP - 1 && f0 + STRIP + 1 <= sn;
    if (row1_ok) {
#pragma unroll
        for (int k = 0; k < NPR; k++) dq[k] = slope_lim(lim, B[r1 - 1][k][lane], pc[k], B[r1 + 1][k][lane]);
    }
    if (row5_ok) {
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            p5[k] = B[STRIP + 1][k][lane];
            dq5[k] = slope_lim(lim, B[STRIP][k][lane], p5[k], B[STRIP + 2][k][lane]);
        }
    }
    __syncthreads();
    // phase 3: left state of row w+1 (kept), right states of rows 2..STRIP+1 to LDS
    double p_l[NPR];
    if (row1_ok) {
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            p_l[k] = pc[k] + 0.5 * dq[k];
            if (r1 >= 2) B[r1 + 2][k][lane] = pc[k] - 0.5 * dq[k];
        }
    }
    if (row5_ok) {
#pragma unroll
        for (int k = 0; k < NPR; k++) B[STRIP + 3][k][lane] = p5[k] - 0.5 * dq5[k];
    }
    __syncthreads();
    // phase 4: face f0 + w between rows w+1 and w+2
    double dtij = 1.e300;
    int fatal = 0;
    if (blk_act && f0 + w <= fe) {
        double p_r[NPR];
#pragma unroll
        for (int k = 0; k < NPR; k++) p_r[k] = B[w + 4][k][lane];
        strip_face<DIR, TAB>(g, F, p_l, p_r, f0 + w, o, h, fullstep, dtij, fatal);
    }
    if (fullstep) block_min<STRIP_TPB>(dtij, dtpart);
    add_counter(cnt, HARM_FATAL, fatal);
}

// The X1 / X2 fluxes of the wave-uniform path with PPM (lim 3), in the strip
// layout of k_flux_strip (wave w computes face f0 + w for 64 phi zones of one
// column): the zone rows, their MC slopes and their parabolas are formed once
// per block and shared through LDS.  Row r = 0..STRIP+4 holds zone f0-3+r along
// DIR; face f0+w takes the right value of row w+2 and the left value of row
// w+3, and a zone whose five-zone stencil leaves the two ghost layers keeps its
// MC state, as in k_flux<..., LIM_PPM>.  The eight variables pass through LDS
// in two halves of four (37 KB per block, four blocks per CU).
constexpr int PPM_ROWS = STRIP + 5, PPM_HALF = NPR / 2;

template <int DIR, bool TAB>
__global__ void __launch_bounds__(STRIP_TPB) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_flux_strip_ppm(DGrid g, const double *__restrict__ pb, double *__restrict__ F, double *__restrict__ dtpart,
                 int fullstep, unsigned long long *__restrict__ cnt)
{
    static_assert(DIR == 1 || DIR == 2, "strips run along X1 or X2");
    if (halted(g)) return;
    const int fs = (DIR == 1) ? g.is : g.js, fe = (DIR == 1) ? g.ie + 1 : g.je + 1;
    const int os = (DIR == 1) ? g.js - 1 : g.is - 1, oe = (DIR == 1) ? g.je + 1 : g.ie + 1;
    const int lo = (DIR == 1) ? g.is : g.js, hi = (DIR == 1) ? g.ie : g.je;
    const unsigned nstrip = (unsigned)(fe - fs + STRIP) / STRIP, no = (unsigned)(oe - os + 1);
    const unsigned nseg = (unsigned)g.N3 / 64;
    const unsigned b = xcd_block();
    const bool blk_act = b < nstrip * no * nseg;
    const unsigned s = b % nstrip, rest = b / nstrip;
    const int o = os + (int)(rest % no);
    const unsigned seg = rest / no;
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int h = (int)seg * 64 + lane;
    const int f0 = fs + (int)s * STRIP;
    const int sn = (DIR == 1) ? g.SN1 : g.SN2;
    const size_t nz = g.nz;
    auto loaded = [&](int r) -> bool {
        const int zd = f0 - 3 + r;
        return blk_act && zd >= 0 && zd < sn;
    };
    auto zone = [&](int r) -> size_t {
        const int zd = f0 - 3 + r;
        return (DIR == 1) ? zidx(g, zd, o, h) : zidx(g, o, zd, h);
    };
    __shared__ double V[PPM_ROWS][PPM_HALF][64];
    __shared__ double S[PPM_ROWS][PPM_HALF][64];  // MC slopes, then the left values
    // rows of wave w: values w, w+4 (and 8: wave 0); slopes w+1, w+5 (w <= 2);
    // parabolas w+2 (and 6: wave 0)
    const int rv2 = (w == 0) ? PPM_ROWS - 1 : -1, rs2 = (w <= 2) ? w + 5 : -1, rp2 = (w == 0) ? 6 : -1;
    double p_l[NPR], p_r[NPR];
#pragma unroll
    for (int half = 0; half < 2; half++) {
        const int k0 = half * PPM_HALF;
        // phase 1: the zone rows
        for (int rr : {w, w + 4, rv2}) {
            if (rr < 0 || !loaded(rr)) continue;
            const size_t zz = zone(rr);
#pragma unroll
            for (int q = 0; q < PPM_HALF; q++) V[rr][q][lane] = pb[(k0 + q) * nz + zz];
        }
        __syncthreads();
        // phase 2: MC slopes of rows 1..STRIP+3
        for (int rr : {w + 1, rs2}) {
            if (rr < 0 || !loaded(rr - 1) || !loaded(rr + 1)) continue;
#pragma unroll
            for (int q = 0; q < PPM_HALF; q++)
                S[rr][q][lane] = slope_lim(LIM_MC, V[rr - 1][q][lane], V[rr][q][lane], V[rr + 1][q][lane]);
        }
        __syncthreads();
        // phase 3: the parabolas of rows 2..6 (the MC state for a zone next to
        // the radial edges); the right value of row w+2 is face f0+w's left state
        double lv[2][PPM_HALF];
#pragma unroll
        for (int a = 0; a < 2; a++) {
#pragma unroll
            for (int q = 0; q < PPM_HALF; q++) lv[a][q] = 0.;
            const int rr = (a == 0) ? w + 2 : rp2;
            if (rr < 0 || !loaded(rr - 1) || !loaded(rr + 1)) continue;
            const int m = f0 - 3 + rr;
            const bool ppm = m >= lo && m <= hi;
#pragma unroll
            for (int q = 0; q < PPM_HALF; q++) {
                const double y = V[rr][q][lane], dq = S[rr][q][lane];
                double l, r;
                if (ppm) {
                    ppm_para(V[rr - 1][q][lane], y, V[rr + 1][q][lane], S[rr - 1][q][lane], dq, S[rr + 1][q][lane],
                             l, r);
                } else {
                    l = y - 0.5 * dq;
                    r = y + 0.5 * dq;
                }
                lv[a][q] = l;
                if (a == 0) p_l[k0 + q] = r;
            }
        }
        __syncthreads();
        // phase 4: the left values through LDS; face f0+w takes row w+3's
#pragma unroll
        for (int q = 0; q < PPM_HALF; q++) {
            S[w + 2][q][lane] = lv[0][q];
            if (rp2 >= 0) S[rp2][q][lane] = lv[1][q];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PPM_HALF; q++) p_r[k0 + q] = S[w + 3][q][lane];
        __syncthreads();
    }
    double dtij = 1.e300;
    int fatal = 0;
    if (blk_act && f0 + w <= fe) strip_face<DIR, TAB>(g, F, p_l, p_r, f0 + w, o, h, fullstep, dtij, fatal);
    if (fullstep) block_min<STRIP_TPB>(dtij, dtpart);
    add_counter(cnt, HARM_FATAL, fatal);
}

// flux_ct EMFs (step_ch.c:514-629 for N3 <= 2, step_ch.c:726-798 for N3 > 2).
__global__ void __launch_bounds__(TPB) k_emf(DGrid g, const double *__restrict__ F1, const double *__restrict__ F2,
                                             const double *__restrict__ F3, double *__restrict__ e1,
                                             double *__restrict__ e2, double *__restrict__ e3)
{
    if (halted(g)) return;
    const int ni = g.ie + 1 - g.is + 1, nj = g.je + 1 - g.js + 1;
    const unsigned idx = xcd_block() * TPB + threadIdx.x;
    if (idx >= (unsigned)ni * nj * g.N3) return;
    int i, j, h;
    box_zone_tiled<STENCIL_TILE>(idx, g.N3, g.is, g.js, ni, nj, i, j, h);
    const size_t nz = g.nz;
    const size_t z = zidx(g, i, j, h);
    const size_t zjm = zidx(g, i, j - 1, h), zim = zidx(g, i - 1, j, h);
    if (g.N3 <= 2) {
        e3[z] = 0.25 * (F1[B2 * nz + z] + F1[B2 * nz + zjm] - F2[B1 * nz + z] - F2[B1 * nz + zim]);
        return;
    }
    const size_t zhm = z - h + (h + g.N3 - 1) % g.N3;
    e1[z] = 0.25 * ((F2[B3 * nz + z] + F2[B3 * nz + zhm]) - (F3[B2 * nz + z] + F3[B2 * nz + zjm]));
    e2[z] = 0.25 * ((F3[B1 * nz + z] + F3[B1 * nz + zim]) - (F1[B3 * nz + z] + F1[B3 * nz + zhm]));
    e3[z] = 0.25 * ((F1[B2 * nz + z] + F1[B2 * nz + zjm]) - (F2[B1 * nz + z] + F2[B1 * nz + zim]));
}

// The flux_ct-rewritten B-components of F1/F2/F3 at a face (step_ch.c:541-554,
// step_ch.c:783-797), evaluated from the EMFs where they are consumed.
__device__ __forceinline__ void ct_f1(const DGrid &g, const double *e1, const double *e2, const double *e3, int i,
                                      int j, int h, double *fb)
{
    const size_t z = zidx(g, i, j, h);
    fb[0] = 0.;
    fb[1] = 0.5 * (e3[z] + e3[zidx(g, i, j + 1, h)]);
    if (g.N3 > 2) fb[2] = -0.5 * (e2[z] + e2[z - h + (h + 1) % g.N3]);
}
__device__ __forceinline__ void ct_f2(const DGrid &g, const double *e1, const double *e2, const double *e3, int i,
                                      int j, int h, double *fb)
{
    const size_t z = zidx(g, i, j, h);
    fb[0] = -0.5 * (e3[z] + e3[zidx(g, i + 1, j, h)]);
    fb[1] = 0.;
    if (g.N3 > 2) fb[2] = 0.5 * (e1[z] + e1[z - h + (h + 1) % g.N3]);
}
__device__ __forceinline__ void ct_f3(const DGrid &g, const double *e1, const double *e2, int i, int j, int h,
                                      double *fb)
{
    const size_t z = zidx(g, i, j, h);
    fb[0] = 0.5 * (e2[z] + e2[zidx(g, i + 1, j, h)]);
    fb[1] = -0.5 * (e1[z] + e1[zidx(g, i, j + 1, h)]);
    fb[2] = 0.;
}

// Zones whose Noble 2D inversion failed, handed from k_update to k_fontfix with
// their conserved vector: the rarely taken Font-fix solvers then do not inflate
// the register allocation of the per-zone kernel.
struct FontList {
    int *count;
    int *zone;
    double *U;  // [8][cap]
    int cap;
    int *overflow;
};

// Flux divergence of the zone update (step_ch.c:250-263) at active zone
// (i, j, h) with the flux_ct B-fluxes formed from the EMFs:
//   d = -(F1[i+1] - F1[i])/dx1 - (F2[j+1] - F2[j])/dx2 [- (F3[h+1] - F3[h])/dx3]
// so that U += Dt*(d + dU) repeats the reference's operation order exactly.
struct DivArgs {
    const double *F1, *F2, *F3, *e1, *e2, *e3;
};

__device__ __forceinline__ void flux_div(const DGrid &g, const DivArgs &a, int i, int j, int h, double *d)
{
    const size_t z = zidx(g, i, j, h);
    const size_t nz = g.nz;
    const size_t zi1 = zidx(g, i + 1, j, h), zj1 = zidx(g, i, j + 1, h);
    const size_t zh1 = z - h + (h + 1) % g.N3;
    double f1a[3], f1b[3], f2a[3], f2b[3], f3a[3], f3b[3];
    ct_f1(g, a.e1, a.e2, a.e3, i, j, h, f1a);
    ct_f1(g, a.e1, a.e2, a.e3, i + 1, j, h, f1b);
    ct_f2(g, a.e1, a.e2, a.e3, i, j, h, f2a);
    ct_f2(g, a.e1, a.e2, a.e3, i, j + 1, h, f2b);
    if (g.N3 > 2) {
        ct_f3(g, a.e1, a.e2, i, j, h, f3a);
        ct_f3(g, a.e1, a.e2, i, j, (h + 1) % g.N3, f3b);
    }
#pragma unroll
    for (int k = 0; k < NPR; k++) {
        double a1, b1, a2, b2;
        if (k < B1 || (k == B3 && g.N3 <= 2)) {
            a1 = a.F1[k * nz + z];
            b1 = a.F1[k * nz + zi1];
            a2 = a.F2[k * nz + z];
            b2 = a.F2[k * nz + zj1];
        } else {
            a1 = f1a[k - B1];
            b1 = f1b[k - B1];
            a2 = f2a[k - B1];
            b2 = f2b[k - B1];
        }
        if (g.N3 > 2) {
            double a3, b3;
            if (k < B1) {
                a3 = a.F3[k * nz + z];
                b3 = a.F3[k * nz + zh1];
            } else {
                a3 = f3a[k - B1];
                b3 = f3b[k - B1];
            }
            d[k] = -div_r(b1 - a1, g.dx1, g.rdx1) - div_r(b2 - a2, g.dx2, g.rdx2) - div_r(b3 - a3, g.dx3, g.rdx3);
        } else {
            d[k] = -div_r(b1 - a1, g.dx1, g.rdx1) - div_r(b2 - a2, g.dx2, g.rdx2);
        }
    }
}

// flux_div with 32-bit zone offsets (nz < OFF32_MAX_ZONES, N3 > 2); the same
// operations in the same order.  The twelve corner EMFs it needs are formed
// here from the face B-fluxes with k_emf's expressions (step_ch.c:773-781), so
// the fused update needs no EMF pass.
__device__ __forceinline__ void flux_div32(const DGrid &g, const DivArgs &a, int i, int j, int h, double *d)
{
    const unsigned sI = (unsigned)g.SN2 * (unsigned)g.N3, sJ = (unsigned)g.N3;
    const unsigned z = (unsigned)zidx(g, i, j, h);
    const unsigned zh1 = z - (unsigned)h + (unsigned)((h + 1) % g.N3);
    const unsigned zhm = z - (unsigned)h + (unsigned)((h + g.N3 - 1) % g.N3);
    const unsigned zi1 = z + sI, zj1 = z + sJ, zi1h1 = zh1 + sI, zj1h1 = zh1 + sJ, zi1j1 = zi1 + sJ;
    const size_t nz = g.nz;
    const double *F1B2 = a.F1 + B2 * nz, *F1B3 = a.F1 + B3 * nz, *F2B1 = a.F2 + B1 * nz, *F2B3 = a.F2 + B3 * nz,
                 *F3B1 = a.F3 + B1 * nz, *F3B2 = a.F3 + B2 * nz;
    // e1 = ((F2[B3] + F2[B3](h-1)) - (F3[B2] + F3[B2](j-1))) / 4 at (j, h), (j, h+1), (j+1, h), (j+1, h+1)
    const double e1z = 0.25 * ((at32(F2B3, z) + at32(F2B3, zhm)) - (at32(F3B2, z) + at32(F3B2, z - sJ)));
    const double e1h = 0.25 * ((at32(F2B3, zh1) + at32(F2B3, z)) - (at32(F3B2, zh1) + at32(F3B2, zh1 - sJ)));
    const double e1j = 0.25 * ((at32(F2B3, zj1) + at32(F2B3, zhm + sJ)) - (at32(F3B2, zj1) + at32(F3B2, z)));
    const double e1jh = 0.25 * ((at32(F2B3, zj1h1) + at32(F2B3, zj1)) - (at32(F3B2, zj1h1) + at32(F3B2, zh1)));
    // e2 = ((F3[B1] + F3[B1](i-1)) - (F1[B3] + F1[B3](h-1))) / 4 at (i, h), (i, h+1), (i+1, h), (i+1, h+1)
    const double e2z = 0.25 * ((at32(F3B1, z) + at32(F3B1, z - sI)) - (at32(F1B3, z) + at32(F1B3, zhm)));
    const double e2h = 0.25 * ((at32(F3B1, zh1) + at32(F3B1, zh1 - sI)) - (at32(F1B3, zh1) + at32(F1B3, z)));
    const double e2i = 0.25 * ((at32(F3B1, zi1) + at32(F3B1, z)) - (at32(F1B3, zi1) + at32(F1B3, zhm + sI)));
    const double e2ih = 0.25 * ((at32(F3B1, zi1h1) + at32(F3B1, zh1)) - (at32(F1B3, zi1h1) + at32(F1B3, zi1)));
    // e3 = ((F1[B2] + F1[B2](j-1)) - (F2[B1] + F2[B1](i-1))) / 4 at (i, j), (i, j+1), (i+1, j), (i+1, j+1)
    const double e3z = 0.25 * ((at32(F1B2, z) + at32(F1B2, z - sJ)) - (at32(F2B1, z) + at32(F2B1, z - sI)));
    const double e3j = 0.25 * ((at32(F1B2, zj1) + at32(F1B2, z)) - (at32(F2B1, zj1) + at32(F2B1, zj1 - sI)));
    const double e3i = 0.25 * ((at32(F1B2, zi1) + at32(F1B2, zi1 - sJ)) - (at32(F2B1, zi1) + at32(F2B1, z)));
    const double e3ij = 0.25 * ((at32(F1B2, zi1j1) + at32(F1B2, zi1)) - (at32(F2B1, zi1j1) + at32(F2B1, zj1)));
    // ct_f1 at (i,j,h) and (i+1,j,h), ct_f2 at (i,j,h) and (i,j+1,h), ct_f3 at (i,j,h) and (i,j,h+1)
    const double f1a1 = 0.5 * (e3z + e3j), f1a2 = -0.5 * (e2z + e2h);
    const double f1b1 = 0.5 * (e3i + e3ij), f1b2 = -0.5 * (e2i + e2ih);
    const double f2a0 = -0.5 * (e3z + e3i), f2a2 = 0.5 * (e1z + e1h);
    const double f2b0 = -0.5 * (e3j + e3ij), f2b2 = 0.5 * (e1j + e1jh);
    const double f3a0 = 0.5 * (e2z + e2i), f3a1 = -0.5 * (e1z + e1j);
    const double f3b0 = 0.5 * (e2h + e2ih), f3b1 = -0.5 * (e1h + e1jh);
    const double f1a[3] = {0., f1a1, f1a2}, f1b[3] = {0., f1b1, f1b2};
    const double f2a[3] = {f2a0, 0., f2a2}, f2b[3] = {f2b0, 0., f2b2};
    const double f3a[3] = {f3a0, f3a1, 0.}, f3b[3] = {f3b0, f3b1, 0.};
#pragma unroll
    for (int k = 0; k < NPR; k++) {
        double a1, b1, a2, b2, a3, b3;
        if (k < B1) {
            const double *F1 = a.F1 + k * nz, *F2 = a.F2 + k * nz, *F3 = a.F3 + k * nz;
            a1 = at32(F1, z);
            b1 = at32(F1, zi1);
            a2 = at32(F2, z);
            b2 = at32(F2, zj1);
            a3 = at32(F3, z);
            b3 = at32(F3, zh1);
        } else {
            a1 = f1a[k - B1];
            b1 = f1b[k - B1];
            a2 = f2a[k - B1];
            b2 = f2b[k - B1];
            a3 = f3a[k - B1];
            b3 = f3b[k - B1];
        }
        d[k] = -div_r(b1 - a1, g.dx1, g.rdx1) - div_r(b2 - a2, g.dx2, g.rdx2) - div_r(b3 - a3, g.dx3, g.rdx3);
    }
}

// The divergence as its own light, high-occupancy stencil pass into the compact
// array S[k][idx] (idx = active-zone index in the update kernel's order), for
// the update variants that read it instead of forming it.
__global__ void __launch_bounds__(TPB) k_div(DGrid g, DivArgs a, double *__restrict__ S)
{
    if (halted(g)) return;
    const int ni = g.ie - g.is + 1, nj = g.je - g.js + 1;
    const size_t na = (size_t)ni * nj * g.N3;
    const unsigned idx = xcd_block() * TPB + threadIdx.x;
    if (idx >= na) return;
    int i, j, h;
    box_zone_tiled<STENCIL_TILE>(idx, g.N3, g.is, g.js, ni, nj, i, j, h);
    const size_t sidx = ((size_t)(i - g.is) * nj + (j - g.js)) * g.N3 + h;
    double d[NPR];
    flux_div(g, a, i, j, h, d);
#pragma unroll
    for (int k = 0; k < NPR; k++) S[k * na + sidx] = d[k];
}

// Zones that the gated update hands to k_seedfix: the zone and its conserved
// vector (k_seedfix repeats the plain start and its first iteration, which
// give the same bits as in the update).
struct SeedList {
    int *count;
    int *zone;
    double *U;  // [8][cap]
    int cap;
};

// Puts a zone on the seed list; false when the list is full.
__device__ __forceinline__ bool seed_defer(const SeedList &sl, int z, const double *U)
{
    const int slot = atomicAdd(sl.count, 1);
    if (slot >= sl.cap) return false;
    const size_t cap = (size_t)sl.cap;
    sl.zone[slot] = z;
#pragma unroll
    for (int k = 0; k < NPR; k++) sl.U[k * cap + slot] = U[k];
    return true;
}

// utoprim_2d (utoprim_2d.c:126-362) with the gate of the learned seed after
// its first Newton iteration: a zone with no start from the guess (code 2), or
// whose first step |dW/W| + |dv2/v2| exceeds tau, goes to the seed list and
// returns SEED_DEFERRED (k_seedfix finishes it); every other zone, and a zone
// that finds the list full, runs utoprim_2d's operations to the end.
constexpr int SEED_DEFERRED = 1000;
__device__ int utoprim_2d_gate(const double *U, const Geom &geom, double gam, double *prim, double tau,
                               const SeedList &sl, int z, int &n_iter, int &k_conv)
{
    n_iter = 0;
    k_conv = 0;
#pragma unroll
    for (int i = B1; i <= B3; i++) prim[i] = div_r(U[i], geom.g, geom.rg);
    if (U[0] <= 0.) return -100;
    double U_tmp[NPR];
    u2p_transform(U, geom, U_tmp);
    U2P s;
    s.gam = gam;
    s.rgam = 1. / gam;
    s.K_atm = 0.;
    double Bcon[4], Qcon[4], ncon[4], QdotB;
    u2p_scalars(s, U_tmp, geom, Bcon, Qcon, ncon, QdotB);
    double W_last;
    if (!guess_W(s, geom, prim, W_last)) return seed_defer(sl, z, U) ? SEED_DEFERRED : 2;
    W_last = raise_W(s, W_last);
    Gnr2 st;
    gnr_start(st, fabs(W_last), x1_of_x0(s, W_last));
    while (st.keep) {
        gnr_step(s, st);
        // the gate, after the first iteration (inside the loop, so that the loop
        // is the plain solve's and is not peeled)
        if (st.n_iter == 1 && st.keep && fabs(st.errx) > tau && seed_defer(sl, z, U)) return SEED_DEFERRED;
    }
    const int retval = gnr_end(st, n_iter, k_conv);
    double pt[5];
    const int ret = finish_2d(s, Bcon, Qcon, ncon, QdotB, retval, st.x0, st.x1, pt);
    if (ret == 0) {
#pragma unroll
        for (int i = 0; i < 5; i++) prim[i] = pt[i];
    }
    return ret;
}

// Optional parts of k_update: the learned Newton seed and the capture of
// training pairs (no counterpart in the reference).
struct UpdateExtra {
    MlpDev mlp;
    double *cap_feat;   // [8][cap] raw features
    double *cap_targ;   // [2][cap] log(W/Wref), v^2
    uint8_t *cap_valid; // [cap]
    size_t cap, cap_base;
    int stride, phase;
    double gate_tau;  // UPD_GATED: zones whose first Newton step exceeds it go to the seed list
    SeedList seeds;
};

enum { UPD_PLAIN = 0, UPD_CAPTURE = 1, UPD_LEARNED = 2, UPD_GATED = 3 };

// Row (i - is) * nj + (j - js) of launch block b over an ni x nj box of phi rows:
// XCD-contiguous logical blocks (xcd_block's order, G a multiple of 8) walked
// in UPD_TI x UPD_TJ tiles of (i, j), j fastest inside a tile; the last tile
// row and column may be partial.  (Measured at 288x256x256 against 4x8, 4x32,
// 16x8, 2x64 and XCD order alone: 8x16 took the least time, by <1%.)
constexpr unsigned UPD_TI = 8, UPD_TJ = 16;
__device__ __forceinline__ unsigned upd_tiled_row(unsigned b, unsigned G, unsigned ni, unsigned nj)
{
    const unsigned lb = (b & 7u) * (G >> 3) + (b >> 3);
    const unsigned ti = lb / (UPD_TI * nj), r = lb - ti * UPD_TI * nj;
    const unsigned th = min(UPD_TI, ni - ti * UPD_TI);
    const unsigned tj = r / (th * UPD_TJ), r2 = r - tj * th * UPD_TJ;
    const unsigned tw = min(UPD_TJ, nj - tj * UPD_TJ);
    const unsigned ii = ti * UPD_TI + r2 / tw, jj = tj * UPD_TJ + r2 % tw;
    return ii * nj + jj;
}

// advance() zone loop (step_ch.c:237-306) + fixup() (fixup.c:51-60), one active
// zone per thread: conserved update from pi with sources at pb, Noble 2D
// inversion seeded with pi, then floors / gamma limit.  Zones whose inversion
// fails go to k_fontfix (step_ch.c:277-300) with the guess left in pf.
// MODE UPD_CAPTURE also stores training pairs of sampled zones; UPD_LEARNED
// seeds the Newton solve from the network (wave-level BF16 MFMA forward) and
// falls back to the pi guess where the seeded solve fails.
// FUSED forms the flux divergence in place (flux_div) instead of reading S.
// The wave-uniform UPD_GATED update is held to the plain one's 4 waves per SIMD
// (its seed-list store would otherwise take it past 128 VGPRs); every other
// instance keeps the compiler's choice.
template <bool UNI, int MODE, bool TAB, bool FUSED>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu((UNI && MODE == UPD_GATED) ? 4 : 1, 8)))
k_update(DGrid g, const double *__restrict__ pi, const double *__restrict__ pb, double *__restrict__ pf,
         int8_t *__restrict__ pflag, const double *__restrict__ S, DivArgs dv, const double *__restrict__ tstate,
         double dtfac, int pi_is_pb, FontList fl, unsigned long long *__restrict__ cpart, UpdateExtra ex)
{
    if (halted(g)) return;
    const int ni = g.ie - g.is + 1, nj = g.je - g.js + 1;
    const size_t na = (size_t)ni * nj * g.N3;
    size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x;
    if constexpr (FUSED) {
        // One block is one phi row (N3 == TPB): the rows go to the XCDs as
        // contiguous ranges of UPD_TI x UPD_TJ (i, j) tiles, j fastest, so the
        // rows that share face fluxes run on one XCD at about the same time
        // and find them in its L2 (288x256x256: 357 instead of 485 HBM bytes
        // per zone; the kernel's time is set by its instruction count).
        const unsigned G = gridDim.x;
        if (g.N3 == TPB && (G & 7u) == 0 && G == (unsigned)ni * (unsigned)nj)
            idx = (size_t)upd_tiled_row(blockIdx.x, G, (unsigned)ni, (unsigned)nj) * TPB + threadIdx.x;
    }
    const bool active = idx < na;
    int c_call = 0, c_iter = 0, c_kconv = 0, c_f2d = 0, c_gam = 0, c_flr = 0, c_cap = 0, fatal = 0;
    int i = 0, j = 0, h = 0;
    size_t z = 0, col = 0;
    Geom geom;
    const double *blk = nullptr;
    double piz[NPR], U[NPR];
    if (active) {
        const unsigned rest = (unsigned)idx / (unsigned)g.N3;
        h = (int)((unsigned)idx - rest * (unsigned)g.N3);
        const unsigned ri = rest / (unsigned)nj;
        j = g.js + (int)(rest - ri * (unsigned)nj);
        i = g.is + (int)ri;
        z = zidx(g, i, j, h);
        const size_t nz = g.nz;
        col = uni_col<UNI>((size_t)i * g.SN2 + j);
        const double Dt = dtfac * tstate[1];
        blk = g.cblk + col * CENT_STRIDE;
        if (UNI)
            load_geom_block<CENT_DERIVED>(blk, geom);
        else
            load_geom(g.geom[CENT], g.ncol, col, geom);
        double dU[NPR];
        {
            double pbz[NPR];
#pragma unroll
            for (int k = 0; k < NPR; k++) pbz[k] = FUSED ? at32(pb + k * nz, (unsigned)z) : pb[k * nz + z];
            State q;
            if constexpr (TAB) {
                const EosTable eos{g.tab, g.ref_derivs};
                if (UNI)
                    source(pbz, geom, blk + CENT_CONN, 1, 0, eos, dU, q, fatal);
                else
                    source(pbz, geom, g.conn, g.ncol, col, eos, dU, q, fatal);
            } else {
                if (UNI)
                    source(pbz, geom, blk + CENT_CONN, 1, 0, g.gam, dU, q, fatal);
                else
                    source(pbz, geom, g.conn, g.ncol, col, g.gam, dU, q, fatal);
            }
#pragma unroll
            for (int k = 0; k < NPR; k++)
                piz[k] = pi_is_pb ? pbz[k] : (FUSED ? at32(pi + k * nz, (unsigned)z) : pi[k * nz + z]);
            if (!pi_is_pb) get_state(piz, geom, q, fatal);
            if constexpr (TAB)
                primtoflux(piz, q, 0, geom, EosTable{g.tab, g.ref_derivs}, U);
            else
                primtoflux(piz, q, 0, geom, g.gam, U);
        }
        if constexpr (FUSED) {
            double d[NPR];
            flux_div32(g, dv, i, j, h, d);
#pragma unroll
            for (int k = 0; k < NPR; k++) U[k] += Dt * (d[k] + dU[k]);
        } else {
#pragma unroll
            for (int k = 0; k < NPR; k++) U[k] += Dt * (S[k * na + idx] + dU[k]);
        }
    }
    double pr[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) pr[k] = piz[k];
    int niter = 0, kconv = 0, ret = 0;
    bool seed_deferred = false;  // UPD_GATED: k_seedfix finishes this zone
    if constexpr (TAB) {
        // advance() with COOL_EOS (step_ch.c:265-268): Utoprim_5d, no Font fix
        if (active) {
            Eos5 e5;
            e5.tab = &g.tab;
            e5.gam = g.gam;
            ret = utoprim_5d(U, geom, e5, pr, niter);
            kconv = niter;
        }
    } else if constexpr (MODE == UPD_GATED) {
        // the plain start and its first Newton iteration; a zone that asks for
        // the network's seed goes to the seed list (k_seedfix), the others
        // resume the plain solve (the network is not in this kernel, which keeps
        // the plain update's register budget)
        if (active) {
            ret = utoprim_2d_gate(U, geom, g.gam, pr, ex.gate_tau, ex.seeds, (int)z, niter, kconv);
            seed_deferred = (ret == SEED_DEFERRED);
        }
    } else if constexpr (MODE == UPD_LEARNED) {
      {
        double f[MLP_NIN];
        const bool ok = active && mlp_features_U(U, geom, f);
        float x[MLP_NIN];
        if (ok)
            mlp_standardize(ex.mlp, f, x);
        else
#pragma unroll
            for (int k = 0; k < MLP_NIN; k++) x[k] = 0.f;
        float y0, y1;
        mlp_forward_wave(ex.mlp, x, y0, y1);
        if (active) {
            double W0 = 0., vsq0 = 0.;
            if (ok) mlp_unstandardize(ex.mlp, y0, y1, mlp_wref(f), W0, vsq0);
            int fb;
            ret = utoprim_2d_learned(U, geom, g.gam, pr, W0, vsq0, ex.mlp.vsq_from_W, niter, kconv, fb);
        }
      }
    } else {
        if (active) ret = utoprim_2d(U, geom, g.gam, pr, niter, kconv);
        if constexpr (MODE == UPD_CAPTURE) {
            if (active && (int)(idx % (size_t)ex.stride) == ex.phase) {
                const size_t slot = ex.cap_base + idx / (size_t)ex.stride;
                if (slot < ex.cap) {
                    double f[MLP_NIN];
                    const bool ok = mlp_features_U(U, geom, f) && ret == 0 && mlp_wref(f) > 0.;
                    double t0 = 0., t1 = 0.;
                    if (ok) {
                        const double utsq = utsq_of(geom, pr);
                        t0 = log((pr[RHO] + g.gam * pr[UU]) * (1. + utsq)) - log(mlp_wref(f));
                        t1 = utsq / (1. + utsq);
                    }
#pragma unroll
                    for (int k = 0; k < MLP_NIN; k++) ex.cap_feat[(size_t)k * ex.cap + slot] = f[k];
                    ex.cap_targ[slot] = t0;
                    ex.cap_targ[ex.cap + slot] = t1;
                    ex.cap_valid[slot] = ok ? 1 : 0;
                }
            }
        }
    }
    if (active && seed_deferred) {
        c_call = 1;  // its iterations are counted by k_seedfix
    } else if (active) {
        c_call = 1;
        c_iter = niter;
        c_kconv = kconv;
        int8_t flag = (int8_t)ret;
        bool deferred = false;
        if (flag) {
            c_f2d = 1;
            if (!TAB && g.font_fix) {
                const int slot = atomicAdd(fl.count, 1);
                if (slot < fl.cap) {
                    fl.zone[slot] = (int)z;
#pragma unroll
                    for (int k = 0; k < NPR; k++) fl.U[(size_t)k * fl.cap + slot] = U[k];
                    deferred = true;
                } else {
                    atomicOr(fl.overflow, 1);
                }
            }
        }
        if (!deferred) {
            const double rf = UNI ? ldc(blk, 33) : ldc(g.floors, col);
            const double uf = UNI ? ldc(blk, 34) : ldc(g.floors, g.ncol + col);
            const int fx = fixup1zone(pr, geom, rf, uf, g.gam, fatal);
            if (fx & FIX_FLOOR) c_flr = 1;
            if (fx & FIX_GAMMA_CAP) c_cap = 1;
            if (fx & FIX_GAMMA_FAIL) {
                flag = -33;
                c_gam = 1;
            }
        }
        const size_t nz = g.nz;
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            if (FUSED)
                at32(pf + k * nz, (unsigned)z) = pr[k];
            else
                pf[k * nz + z] = pr[k];
        }
        pflag[z] = flag;
    }
    const int v[8] = {c_call, c_iter, c_kconv, c_f2d, c_gam, c_flr, c_cap, fatal};
    block_counts<8>(v, cpart);
}


// ---------------------------------------------------------------------------
// The step with the cache EOS (the reference's COOL_EOS build): every face and
// every zone update threads its zone's last temperature through its pressure
// calls (harm_ceos.hpp).  lastT is zone-indexed like the primitives.
// ---------------------------------------------------------------------------

#ifndef CEOS_WPE
#define CEOS_WPE 2
#endif
#ifndef CEOS_EOS_WPE
#define CEOS_EOS_WPE 2
#endif

// The coefficient tables of the splines along T of e and p of cache blockIdx.y
// (cache_ycoef_one): one (component, x row, bracket) per thread.
__global__ void __launch_bounds__(TPB) k_ceos_ycoef(CacheView v, const double *const *values,
                                                   double *const *yco)
{
    const size_t per = (size_t)v.X * v.Y;
    const size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (idx >= 2 * per) return;
    const int comp = (int)(idx / per);
    const size_t r = idx - (size_t)comp * per;
    const int ix = (int)(r / (size_t)v.Y), left = (int)(r % (size_t)v.Y);
    if (left < 1) return;
    CacheView c = v;
    c.values = values[blockIdx.y];
    cache_ycoef_one(c, 0, comp, ix, left, yco[blockIdx.y] + idx * CACHE_YCO);
}

// The stencil of a face along DIR: zones m-2, m-1 (left of the face), m, m+1.
template <int DIR>
__device__ __forceinline__ void ceos_stencil(const DGrid &g, size_t z, int h, size_t &zm2, size_t &zm1, size_t &zp1)
{
    if (DIR == 1) {
        const size_t s = (size_t)g.SN2 * g.N3;
        zm2 = z - 2 * s;
        zm1 = z - s;
        zp1 = z + s;
    } else if (DIR == 2) {
        const size_t s = (size_t)g.N3;
        zm2 = z - 2 * s;
        zm1 = z - s;
        zp1 = z + s;
    } else {
        const size_t base = z - h;
        zm2 = base + (h + 2 * g.N3 - 2) % g.N3;
        zm1 = base + (h + g.N3 - 1) % g.N3;
        zp1 = base + (h + 1) % g.N3;
    }
}

// The face range of a sweep along DIR (k_flux's) and, with chain, one (h, i)
// column per thread walked in j order; calls f(i, j, h) per face.
template <int DIR, class Fn>
__device__ __forceinline__ void ceos_faces(const DGrid &g, int chain, Fn &&f)
{
    const int i0 = (DIR == 1) ? g.is : g.is - 1;
    const int j0 = (DIR == 2) ? g.js : g.js - 1;
    const int ni = g.ie + 1 - i0 + 1, nj = g.je + 1 - j0 + 1;
    const size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x;
    const size_t n = chain ? (size_t)ni * g.N3 : (size_t)ni * nj * g.N3;
    if (idx >= n) return;
    const int h = (int)(idx % (size_t)g.N3);
    const size_t rest = idx / (size_t)g.N3;
    const int i = chain ? i0 + (int)rest : i0 + (int)(rest / (size_t)nj);
    const int ja = chain ? j0 : j0 + (int)(rest % (size_t)nj), jb = chain ? g.je + 1 : ja;
    for (int j = ja; j <= jb; j++) f(i, j, h);
}

// The pressures of fluxcalc's face (h, i, j) with InitCache_CoolEOS before it
// (step_ch.c:439-441): its left state's pressure and vchar sound speed (with
// dp/drho, dp/du by their pressure calls), then its right state's, in the
// order primtoflux and vchar take them (step_ch.c:450-456), on the zone's last
// T; out: p_l, cs2_l, p_r, cs2_r at the face's zone.  Only density and energy
// of the two states are reconstructed (the pressure sees nothing else).  chain:
// the zones' last temperatures may be unset, and the loop over j runs in
// order, so a zone's j-1 neighbour has its value when it is read
// (cooleos.c:784-799).
template <int DIR, bool V>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(CEOS_EOS_WPE, 8)))
k_face_eos(DGrid g, CeosParams ce, double *__restrict__ lastT, const double *__restrict__ pb,
           double *__restrict__ fe, int chain, unsigned long long *__restrict__ solves)
{
    if (halted(g)) return;
    int nsolve = 0;
    ceos_faces<DIR>(g, chain, [&](int i, int j, int h) {
        if (DIR == 2 && (j == g.js || j == g.je + 1)) return;  // pole faces: no InitCache, no pressure
        const size_t nz = g.nz, z = zidx(g, i, j, h);
        size_t zm2, zm1, zp1;
        ceos_stencil<DIR>(g, z, h, zm2, zm1, zp1);
        double rl, ul, rr, ur;
        {
            const double a = pb[RHO * nz + zm2], b = pb[RHO * nz + zm1], c = pb[RHO * nz + z],
                         d = pb[RHO * nz + zp1];
            rl = b + 0.5 * slope_lim(g.lim, a, b, c);
            rr = c - 0.5 * slope_lim(g.lim, b, c, d);
        }
        {
            const double a = pb[UU * nz + zm2], b = pb[UU * nz + zm1], c = pb[UU * nz + z],
                         d = pb[UU * nz + zp1];
            ul = b + 0.5 * slope_lim(g.lim, a, b, c);
            ur = c - 0.5 * slope_lim(g.lim, b, c, d);
        }
        CeosZone zs;
        zs.T = lastT[z];
        if (chain && zs.T == 0.0 && j > g.js) zs.T = lastT[z - g.N3];
        const EosCacheT<V> eos{&ce, ceos_radius(ce, i), &zs};
        const double pl = eos.pres(rl, ul);
        const double cl = eos.cs2(rl, ul, pl);
        const double pr = eos.pres(rr, ur);
        const double cr = eos.cs2(rr, ur, pr);
        lastT[z] = zs.T;
        nsolve += zs.nsolve;
        fe[0 * nz + z] = pl;
        fe[1 * nz + z] = cl;
        fe[2 * nz + z] = pr;
        fe[3 * nz + z] = cr;
    });
    add_counter(solves, 0, nsolve);
}

// The EOS policy of a face whose two states' pressures and sound speeds
// k_face_eos formed: a state is told apart by its density and energy bits.
struct EosGiven {
    static constexpr bool TAB = true;
    double rl, ul, pl, cl, pr, cr;
    __device__ bool left(double rho, double u) const
    {
        return __double_as_longlong(rho) == __double_as_longlong(rl) && __double_as_longlong(u) == __double_as_longlong(ul);
    }
    __device__ double pres(double rho, double u) const { return left(rho, u) ? pl : pr; }
    __device__ double cs2(double rho, double u, double) const { return left(rho, u) ? cl : cr; }
};

// fluxcalc (step_ch.c:331-504) with the face pressures k_face_eos formed:
// slopes from four zones along DIR (MC / van Leer / minmod), fix_flux
// (coord.c:264-324) and the CFL limit, as k_flux.  One face per thread.
template <int DIR>
__global__ void __launch_bounds__(TPB) k_flux_ceos(DGrid g, const double *__restrict__ fe,
                                                   const double *__restrict__ pb, double *__restrict__ F,
                                                   double *__restrict__ dtpart, int fullstep,
                                                   unsigned long long *__restrict__ cnt)
{
    if (halted(g)) return;
    double dtij = 1.e300;
    int fatal = 0;
    ceos_faces<DIR>(g, 0, [&](int i, int j, int h) {
        const size_t nz = g.nz, z = zidx(g, i, j, h);
        if (DIR == 2 && (j == g.js || j == g.je + 1)) {
#pragma unroll
            for (int k = 0; k < NPR; k++) F[k * nz + z] = 0.;
            return;
        }
        size_t zm2, zm1, zp1;
        ceos_stencil<DIR>(g, z, h, zm2, zm1, zp1);
        double p_l[NPR], p_r[NPR];
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            const double a = pb[k * nz + zm2], b = pb[k * nz + zm1], c = pb[k * nz + z], d = pb[k * nz + zp1];
            p_l[k] = b + 0.5 * slope_lim(g.lim, a, b, c);
            p_r[k] = c - 0.5 * slope_lim(g.lim, b, c, d);
        }
        Geom geom;
        const int loc = (DIR == 1) ? FACE1 : (DIR == 2) ? FACE2 : FACE3;
        load_geom(g.geom[loc], g.ncol, (size_t)i * g.SN2 + j, geom);
        const EosGiven eos{p_l[RHO], p_l[UU], fe[0 * nz + z], fe[1 * nz + z], fe[2 * nz + z], fe[3 * nz + z]};
        double Fz[NPR], ctop;
        face_flux(p_l, p_r, geom, DIR, eos, g.hllf, g.laxf, Fz, ctop, fatal);
        if (DIR == 1) {
            if (i == g.is && Fz[RHO] > 0) Fz[RHO] = 0;
            if (i == g.ie + 1 && Fz[RHO] < 0) Fz[RHO] = 0;
        }
        const bool skip_b2 = (DIR != 2) && (j == g.js - 1 || j == g.je + 1);
#pragma unroll
        for (int k = 0; k < NPR; k++)
            if (!(k == B2 && skip_b2)) F[k * nz + z] = Fz[k];
        if (DIR != 2) {
            if (j == g.js) F[B2 * nz + zidx(g, i, g.js - 1, h)] = -Fz[B2];
            if (j == g.je) F[B2 * nz + zidx(g, i, g.je + 1, h)] = -Fz[B2];
        }
        if (fullstep && i >= g.is && i <= g.ie && j >= g.js && j <= g.je) {
            const double dxd = (DIR == 1) ? g.dx1 : (DIR == 2) ? g.dx2 : g.dx3;
            const double d = g.cour * dxd / ctop;
            if (d < 1.e9 && d < dtij) dtij = d;
        }
    });
    if (fullstep) block_min(dtij, dtpart);
    add_counter(cnt, HARM_FATAL, fatal);
}

// advance()'s zone loop with COOL_EOS (step_ch.c:237-306): InitCache_CoolEOS,
// the source at pb, primtoU at pi, the update with the flux divergence S,
// Utoprim_5d seeded with pi (no Font fix), then fixup1zone (fixup.c:51-60).
template <bool V>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(CEOS_WPE, 8))) k_update_ceos(DGrid g, CeosParams ce, double *__restrict__ lastT,
                                                     const double *__restrict__ pi, const double *__restrict__ pb,
                                                     double *__restrict__ pf, int8_t *__restrict__ pflag,
                                                     const double *__restrict__ S,
                                                     const double *__restrict__ tstate, double dtfac, int pi_is_pb,
                                                     unsigned long long *__restrict__ cpart,
                                                     unsigned long long *__restrict__ solves)
{
    if (halted(g)) return;
    const int ni = g.ie - g.is + 1, nj = g.je - g.js + 1;
    const size_t na = (size_t)ni * nj * g.N3;
    const size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x;
    int c_call = 0, c_iter = 0, c_kconv = 0, c_f2d = 0, c_gam = 0, c_flr = 0, c_cap = 0, fatal = 0, nsolve = 0;
    if (idx < na) {
        const unsigned rest = (unsigned)idx / (unsigned)g.N3;
        const int h = (int)((unsigned)idx - rest * (unsigned)g.N3);
        const unsigned ri = rest / (unsigned)nj;
        const int j = g.js + (int)(rest - ri * (unsigned)nj);
        const int i = g.is + (int)ri;
        const size_t z = zidx(g, i, j, h), nz = g.nz, col = (size_t)i * g.SN2 + j;
        const double Dt = dtfac * tstate[1];
        Geom geom;
        load_geom(g.geom[CENT], g.ncol, col, geom);
        CeosZone zs;
        zs.T = lastT[z];
        const EosCacheT<V> eos{&ce, ceos_radius(ce, i), &zs};
        double pbz[NPR], piz[NPR], dU[NPR], U[NPR];
#pragma unroll
        for (int k = 0; k < NPR; k++) pbz[k] = pb[k * nz + z];
        State q;
        source(pbz, geom, g.conn, g.ncol, col, eos, dU, q, fatal);
#pragma unroll
        for (int k = 0; k < NPR; k++) piz[k] = pi_is_pb ? pbz[k] : pi[k * nz + z];
        if (!pi_is_pb) get_state(piz, geom, q, fatal);
        primtoflux(piz, q, 0, geom, eos, U);
#pragma unroll
        for (int k = 0; k < NPR; k++) U[k] += Dt * (S[k * na + idx] + dU[k]);
        double pr[NPR];
#pragma unroll
        for (int k = 0; k < NPR; k++) pr[k] = piz[k];
        int niter = 0;
        const Eos5CacheT<V> e5{eos, g.gam};
        const int ret = utoprim_5d(U, geom, e5, pr, niter);
        lastT[z] = zs.T;
        nsolve = zs.nsolve;
        c_call = 1;
        c_iter = niter;
        c_kconv = niter;
        int8_t flag = (int8_t)ret;
        if (flag) c_f2d = 1;
        const int fx = fixup1zone(pr, geom, ldc(g.floors, col), ldc(g.floors, g.ncol + col), g.gam, fatal);
        if (fx & FIX_FLOOR) c_flr = 1;
        if (fx & FIX_GAMMA_CAP) c_cap = 1;
        if (fx & FIX_GAMMA_FAIL) {
            flag = -33;
            c_gam = 1;
        }
#pragma unroll
        for (int k = 0; k < NPR; k++) pf[k * nz + z] = pr[k];
        pflag[z] = flag;
    }
    const int v[8] = {c_call, c_iter, c_kconv, c_f2d, c_gam, c_flr, c_cap, fatal};
    block_counts<8>(v, cpart);
    add_counter(solves, 0, nsolve);
}

// The zones the gated update put on the seed list: the network's seed for
// them (64 listed zones per wavefront, so every lane of the MFMA pass is a
// zone that asked), the seeded solve, where it fails the plain solve resumed
// from its first iteration; then, as in k_update, the Font-fix list or
// fixup1zone.  pi holds the zones' guess primitives (unchanged by the update
// for a listed zone).  Grid-stride over the list with a grid-uniform number of
// rounds, so whole wavefronts reach the network pass; the count is reset by
// k_bound.
constexpr int SEEDFIX_BLOCKS = 256;
__global__ void __launch_bounds__(TPB) k_seedfix(DGrid g, const double *__restrict__ pi, double *__restrict__ pf,
                                                 int8_t *__restrict__ pflag, SeedList sl, FontList fl, MlpDev mlp,
                                                 unsigned long long *__restrict__ cnt)
{
    if (halted(g)) return;
    const int n = min(*sl.count, sl.cap);
    int c_iter = 0, c_kconv = 0, c_f2d = 0, c_gam = 0, c_flr = 0, c_cap = 0, fatal = 0;
    const int stride = gridDim.x * TPB;
    const int nround = (n + stride - 1) / stride;
    const size_t cap = (size_t)sl.cap, nz = g.nz;
    for (int r = 0; r < nround; r++) {
        const int e = r * stride + blockIdx.x * TPB + threadIdx.x;
        const bool act = e < n;
        size_t z = 0, col = 0;
        Geom geom;
        double U[NPR], pr[NPR];
        if (act) {
            z = (size_t)sl.zone[e];
            col = z / g.N3;
            load_geom(g.geom[CENT], g.ncol, col, geom);
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                U[k] = sl.U[k * cap + e];
                pr[k] = pi[k * nz + z];
            }
        }
        double f[MLP_NIN];
        const bool ok = act && mlp_features_U(U, geom, f);
        float x[MLP_NIN];
        if (ok)
            mlp_standardize(mlp, f, x);
        else
#pragma unroll
            for (int k = 0; k < MLP_NIN; k++) x[k] = 0.f;
        float y0, y1;
        mlp_forward_wave(mlp, x, y0, y1);
        if (!act) continue;
        double W0 = 0., vsq0 = 0.;
        if (ok) mlp_unstandardize(mlp, y0, y1, mlp_wref(f), W0, vsq0);
        // the plain start and its first iteration again (as in the update: same bits)
        Gnr2 st;
        bool need;
        const int ph1 = u2p_gate_first(U, geom, g.gam, pr, 0., st, need);
        int niter, kconv, seeded;
        const int ret = u2p_gate_finish(U, geom, g.gam, pr, ph1, st, true, W0, vsq0, mlp.vsq_from_W, niter, kconv,
                                        seeded);
        c_iter += niter;
        c_kconv += kconv;
        int8_t flag = (int8_t)ret;
        bool deferred = false;
        if (flag) {
            c_f2d++;
            if (g.font_fix) {
                const int slot = atomicAdd(fl.count, 1);
                if (slot < fl.cap) {
                    fl.zone[slot] = (int)z;
#pragma unroll
                    for (int k = 0; k < NPR; k++) fl.U[(size_t)k * fl.cap + slot] = U[k];
                    deferred = true;
                } else {
                    atomicOr(fl.overflow, 1);
                }
            }
        }
        if (!deferred) {
            const int fx = fixup1zone(pr, geom, g.floors[col], g.floors[g.ncol + col], g.gam, fatal);
            if (fx & FIX_FLOOR) c_flr++;
            if (fx & FIX_GAMMA_CAP) c_cap++;
            if (fx & FIX_GAMMA_FAIL) {
                flag = -33;
                c_gam++;
            }
        }
#pragma unroll
        for (int k = 0; k < NPR; k++) pf[k * nz + z] = pr[k];
        pflag[z] = flag;
    }
    add_counter(cnt, HARM_NEWTON_ITERS, c_iter);
    add_counter(cnt, HARM_NEWTON_KCONV, c_kconv);
    add_counter(cnt, HARM_FAIL_2D, c_f2d);
    add_counter(cnt, HARM_FIXUP_GAMMA, c_gam);
    add_counter(cnt, HARM_FLOOR_ZONES, c_flr);
    add_counter(cnt, HARM_GAMMA_CAP, c_cap);
    add_counter(cnt, HARM_FATAL, fatal);
}

// Font fix (step_ch.c:277-300) + fixup1zone for the zones k_update deferred.
// Grid-stride over the list; the list count is reset by the next kernel on the
// stream (k_bound), after every thread here has read it.
// A fixed grid-stride grid: a launch sized for the list's capacity costs tens of
// microseconds even when the list is empty, and every listed zone is a serial
// 1D Newton solve, so too few threads serialize the long lists of later times.
// Over 20 steps of the 288x256x256 torus: 16 blocks 121.5 us per call on
// average (5.1 ms worst), 256 blocks 15.0 us (0.42 ms), 1024 blocks 18.8 us;
// an empty list costs 3.4 us with 16 or 256 blocks, 7.2 us with 1024.
constexpr int FONTFIX_BLOCKS = 256;
__global__ void __launch_bounds__(TPB) k_fontfix(DGrid g, double *__restrict__ pf, int8_t *__restrict__ pflag,
                                                 FontList fl, unsigned long long *__restrict__ cnt)
{
    if (halted(g)) return;
    const int n = min(*fl.count, fl.cap);
    int c_ff1 = 0, c_ff2 = 0, c_gam = 0, c_flr = 0, c_cap = 0, fatal = 0;
    const int stride = gridDim.x * TPB;
    const int nround = (n + stride - 1) / stride;
    for (int r = 0; r < nround; r++) {
        const int e = r * stride + blockIdx.x * TPB + threadIdx.x;
        if (e < n) {
            const size_t z = (size_t)fl.zone[e];
            const size_t nz = g.nz;
            const size_t col = z / g.N3;
            const int i = (int)(col / g.SN2);
            Geom geom;
            load_geom(g.geom[CENT], g.ncol, col, geom);
            double U[NPR], pr[NPR];
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                U[k] = fl.U[(size_t)k * fl.cap + e];
                pr[k] = pf[k * nz + z];
            }
            const double K = g.katm[i];
            int8_t flag = (int8_t)utoprim_1dvsq2fix1(U, geom, g.gam, K, pr);
            if (flag) {
                c_ff1++;
                flag = (int8_t)utoprim_1dfix1(U, geom, g.gam, K, pr);
                if (flag) c_ff2++;
            }
            const int fx = fixup1zone(pr, geom, g.floors[col], g.floors[g.ncol + col], g.gam, fatal);
            if (fx & FIX_FLOOR) c_flr++;
            if (fx & FIX_GAMMA_CAP) c_cap++;
            if (fx & FIX_GAMMA_FAIL) {
                flag = -33;
                c_gam++;
            }
#pragma unroll
            for (int k = 0; k < NPR; k++) pf[k * nz + z] = pr[k];
            pflag[z] = flag;
        }
    }
    add_counter(cnt, HARM_FAIL_FONT1, c_ff1);
    add_counter(cnt, HARM_FAIL_FONT2, c_ff2);
    add_counter(cnt, HARM_FIXUP_GAMMA, c_gam);
    add_counter(cnt, HARM_FLOOR_ZONES, c_flr);
    add_counter(cnt, HARM_GAMMA_CAP, c_cap);
    add_counter(cnt, HARM_FATAL, fatal);
}

// fixup() alone over all active zones (fixup.c:51-60): the initial-data path
// (init.c:311,420) with cnt null, the per-stage test entry with counters.
__global__ void __launch_bounds__(TPB) k_fixup(DGrid g, double *__restrict__ p, int8_t *__restrict__ pflag,
                                               unsigned long long *__restrict__ cnt)
{
    const int ni = g.ie - g.is + 1, nj = g.je - g.js + 1;
    const size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x;
    int c_gam = 0, c_flr = 0, c_cap = 0, fatal = 0;
    if (idx < (size_t)ni * nj * g.N3) {
        const unsigned rest = (unsigned)idx / (unsigned)g.N3;
        const int h = (int)((unsigned)idx - rest * (unsigned)g.N3);
        const unsigned ri = rest / (unsigned)nj;
        const int j = g.js + (int)(rest - ri * (unsigned)nj);
        const int i = g.is + (int)ri;
        const size_t z = zidx(g, i, j, h), nz = g.nz, col = (size_t)i * g.SN2 + j;
        Geom geom;
        load_geom(g.geom[CENT], g.ncol, col, geom);
        double pr[NPR];
        for (int k = 0; k < NPR; k++) pr[k] = p[k * nz + z];
        const int fx = fixup1zone(pr, geom, g.floors[col], g.floors[g.ncol + col], g.gam, fatal);
        for (int k = 0; k < NPR; k++) p[k * nz + z] = pr[k];
        if (fx & FIX_GAMMA_FAIL) {
            pflag[z] = -33;
            c_gam++;
        }
        if (fx & FIX_FLOOR) c_flr++;
        if (fx & FIX_GAMMA_CAP) c_cap++;
    }
    if (cnt) {
        add_counter(cnt, HARM_FIXUP_GAMMA, c_gam);
        add_counter(cnt, HARM_FLOOR_ZONES, c_flr);
        add_counter(cnt, HARM_GAMMA_CAP, c_cap);
        add_counter(cnt, HARM_FATAL, fatal);
    }
}

// ---------------------------------------------------------------------------
// bound_prim (bounds.c:54-331, polefix bounds.c:381-436): one thread per zone it
// writes -- the four radial ghost columns (all rows), the four polar ghost rows
// of the active columns, and the polefix rows.  Every value a thread reads comes
// from active rows no thread of this launch writes, so the reference's
// copy -> polefix -> pole ghosts -> sign flip -> inflow sequence is reproduced
// without inter-thread ordering.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double bval(const DGrid &g, const double *p, int hh, int ii, int j, int k)
{
    const int c = ii < g.is ? g.is : (ii > g.ie ? g.ie : ii);
    const size_t nz = g.nz;
    if (g.polefix && k < B1) {
        const int jref = g.polefix;
        if (j < g.js + jref) {
            const double src = p[k * nz + zidx(g, c, g.js + jref, hh)];
            return (k == U2) ? ((j - g.js) + 0.5) / (jref + 0.5) * src : src;
        }
        if (j > g.je - jref) {
            const double src = p[k * nz + zidx(g, c, g.je - jref, hh)];
            return (k == U2) ? ((g.je - j) + 0.5) / (jref + 0.5) * src : src;
        }
    }
    return p[k * nz + zidx(g, c, j, hh)];
}

__device__ __forceinline__ size_t bound_nthreads(const DGrid &g)
{
    return ((size_t)4 * g.SN2 + (size_t)g.N1 * (4 + 2 * g.polefix)) * g.N3;
}

__global__ void __launch_bounds__(TPB) k_bound(DGrid g, double *__restrict__ p, int8_t *__restrict__ pflag,
                                               int *__restrict__ font_count, int *__restrict__ seed_count,
                                               unsigned long long *__restrict__ cnt)
{
    if (halted(g)) return;
    const size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (idx == 0 && font_count) *font_count = 0;
    if (idx == 0 && seed_count) *seed_count = 0;
    int fatal = 0;
    if (idx < bound_nthreads(g)) {
        const int h = (int)(idx % g.N3);
        size_t t = idx / g.N3;
        int i, j;
        bool poleonly = false;
        if (t < (size_t)4 * g.SN2) {  // radial ghost columns, every row
            const int q = (int)(t / g.SN2);
            i = (q < 2) ? g.is - 2 + q : g.ie + 1 + (q - 2);
            j = (int)(t % g.SN2);
        } else {
            t -= (size_t)4 * g.SN2;
            const int nr = 4 + 2 * g.polefix;
            i = g.is + (int)(t / nr);
            const int r = (int)(t % nr);
            const int rows[8] = {g.js - 2, g.js - 1, g.je + 1, g.je + 2, g.js, g.js + 1, g.je, g.je - 1};
            j = rows[r];
            poleonly = (r >= 4);
        }
        const size_t nz = g.nz;
        const bool ghost = (i < g.is || i > g.ie);
        const int c = i < g.is ? g.is : (i > g.ie ? g.ie : i);
        const size_t z = zidx(g, i, j, h);
        double pr[NPR];
        int8_t fl;
        if (j >= g.js && j <= g.je) {
#pragma unroll
            for (int k = 0; k < NPR; k++) pr[k] = bval(g, p, h, i, j, k);
            fl = pflag[zidx(g, c, j, h)];
        } else {
            const int hs = (g.N3 > 1) ? (h + g.N3 / 2) % g.N3 : h;
            const int src = (j == g.js - 1) ? g.js : (j == g.js - 2) ? g.js + 1 : (j == g.je + 1) ? g.je : g.je - 1;
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                double v = bval(g, p, hs, i, src, k);
                if (k == U2 || k == B2) v *= -1.;
                pr[k] = v;
            }
            fl = pflag[zidx(g, c, src, h)];
        }
        if (ghost) {
            Geom geom;
            load_geom(g.geom[CENT], g.ncol, (size_t)i * g.SN2 + j, geom);
            inflow_check(pr, geom, (i < g.is) ? 0 : 1, fatal);
        }
        if (poleonly) {
#pragma unroll
            for (int k = 0; k < B1; k++) p[k * nz + z] = pr[k];
        } else {
#pragma unroll
            for (int k = 0; k < NPR; k++) p[k * nz + z] = pr[k];
            pflag[z] = fl;
        }
    }
    add_counter(cnt, HARM_FATAL, fatal);
}

// ---------------------------------------------------------------------------
// fixup_utoprim (fixup.c:432-538).  The reference walks zones in (h,i,j) order
// and its fallback branch reads neighbours repaired earlier in that walk.  A bad
// zone whose 8 neighbours are all good reads only zones nothing modifies, so it
// is repaired in parallel; bad zones with a bad neighbour are collected and
// repaired by one thread in the reference's order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void repair_zone(const DGrid &g, double *p, int8_t *pf_live, int i, int j, int h,
                                            bool live_flags, const int8_t *pf_in, int &c_rep, int &c_nost,
                                            int &c_flr, int &c_cap, int &c_gam, int &fatal)
{
    const size_t nz = g.nz;
    const size_t z = zidx(g, i, j, h);
    auto good = [&](int ii, int jj) -> bool {
        const size_t zz = zidx(g, ii, jj, h);
        return live_flags ? (pf_live[zz] != 0) : (pf_in[zz] == 0);
    };
    const bool p1 = good(i - 1, j + 1), p2 = good(i, j + 1), p3 = good(i + 1, j + 1), p8 = good(i - 1, j);
    const bool p4 = good(i + 1, j), p7 = good(i - 1, j - 1), p6 = good(i, j - 1), p5 = good(i + 1, j - 1);
    const size_t zN = zidx(g, i, j + 1, h), zS = zidx(g, i, j - 1, h), zW = zidx(g, i - 1, j, h), zE = zidx(g, i + 1, j, h);
    const size_t zNE = zidx(g, i + 1, j + 1, h), zSE = zidx(g, i + 1, j - 1, h), zNW = zidx(g, i - 1, j + 1, h),
                 zSW = zidx(g, i - 1, j - 1, h);
    double pr[NPR];
    for (int k = 0; k < NPR; k++) pr[k] = p[k * nz + z];
    if (p2 && p4 && p6 && p8) {
        for (int k = 0; k < B1; k++)
            pr[k] = 0.25 * (p[k * nz + zN] + p[k * nz + zS] + p[k * nz + zW] + p[k * nz + zE]);
    } else if (p1 && p3 && p5 && p7) {
        for (int k = 0; k < B1; k++)
            pr[k] = 0.25 * (p[k * nz + zNE] + p[k * nz + zSE] + p[k * nz + zNW] + p[k * nz + zSW]);
    } else {
        for (int k = RHO; k <= UU; k++) {
            const double a4 = 0.25 * (p[k * nz + zN] + p[k * nz + zS] + p[k * nz + zW] + p[k * nz + zE]);
            const double a2 = 0.25 * (p[k * nz + zNE] + p[k * nz + zSE] + p[k * nz + zNW] + p[k * nz + zSW]);
            pr[k] = 0.5 * (a4 + a2);
        }
        pr[U1] = pr[U2] = pr[U3] = 0.;
        c_nost++;
    }
    c_rep++;
    const size_t col = (size_t)i * g.SN2 + j;
    Geom geom;
    load_geom(g.geom[CENT], g.ncol, col, geom);
    const int fx = fixup1zone(pr, geom, g.floors[col], g.floors[g.ncol + col], g.gam, fatal);
    int8_t fl = 0;
    if (fx & FIX_FLOOR) c_flr++;
    if (fx & FIX_GAMMA_CAP) c_cap++;
    if (fx & FIX_GAMMA_FAIL) {
        fl = -33;
        c_gam++;
    }
    for (int k = 0; k < NPR; k++) p[k * nz + z] = pr[k];
    pf_live[z] = fl;
}

// Clustered bad zones are marked in a per-plane bitmap, bit (i*SN2 + j) of
// plane h's row of nwords 64-bit words, and counted per plane: the repair walk
// of a plane reads and writes that plane only (fixup_utoprim's stencil is
// i+-1, j+-1 at fixed h), so the planes are repaired independently and each in
// the reference's (i, j) order with no list, sort or capacity.
__global__ void __launch_bounds__(TPB) k_fixutoprim_scan(DGrid g, double *__restrict__ p,
                                                         const int8_t *__restrict__ fin, int8_t *__restrict__ fout,
                                                         unsigned long long *__restrict__ bits,
                                                         int *__restrict__ pcount, unsigned nwords,
                                                         unsigned long long *__restrict__ cnt)
{
    if (halted(g)) return;
    // 8 zones per thread: the common all-good octet is one 8-byte load and store.
    const size_t z0 = ((size_t)blockIdx.x * TPB + threadIdx.x) * 8;
    int c_rep = 0, c_nost = 0, c_flr = 0, c_cap = 0, c_gam = 0, fatal = 0;
    if (z0 < g.nz) {
        const bool full = z0 + 8 <= g.nz;
        const unsigned long long f8 = full ? *reinterpret_cast<const unsigned long long *>(fin + z0) : 1ull;
        if (full && f8 == 0ull) {
            *reinterpret_cast<unsigned long long *>(fout + z0) = 0x0101010101010101ull;
        } else {
            for (int q = 0; q < 8; q++) {
                const size_t idx = z0 + q;
                if (idx >= g.nz) break;
                const unsigned col = (unsigned)idx / (unsigned)g.N3;
                const int h = (int)((unsigned)idx - col * (unsigned)g.N3);
                const unsigned ci = col / (unsigned)g.SN2;
                const int j = (int)(col - ci * (unsigned)g.SN2), i = (int)ci;
                const int8_t f = fin[idx];
                fout[idx] = (int8_t)(!f);
                if (f != 0 && i >= g.is && i <= g.ie && j >= g.js && j <= g.je) {
                    bool isolated = true;
                    for (int di = -1; di <= 1; di++)
                        for (int dj = -1; dj <= 1; dj++)
                            if ((di || dj) && fin[zidx(g, i + di, j + dj, h)] != 0) isolated = false;
                    if (isolated) {
                        repair_zone(g, p, fout, i, j, h, false, fin, c_rep, c_nost, c_flr, c_cap, c_gam, fatal);
                    } else {
                        atomicOr(&bits[(size_t)h * nwords + (col >> 6)], 1ull << (col & 63u));
                        atomicAdd(&pcount[h], 1);
                    }
                }
            }
        }
    }
    add_counter(cnt, HARM_FIXUP_REPAIRS, c_rep);
    add_counter(cnt, HARM_FIXUP_NOSTENCIL, c_nost);
    add_counter(cnt, HARM_FLOOR_ZONES, c_flr);
    add_counter(cnt, HARM_GAMMA_CAP, c_cap);
    add_counter(cnt, HARM_FIXUP_GAMMA, c_gam);
    add_counter(cnt, HARM_FATAL, fatal);
}

// Serial repair of the clustered bad zones of plane h = blockIdx.x (one wave
// per plane) in the reference's (i, j) order.  The wave reads 64 bitmap words
// at a time and lane 0 repairs the marked zones of the nonzero words in word
// and bit order; each word is cleared after its zones are done, and the plane
// count is reset, ready for the next call.
__global__ void __launch_bounds__(64) k_fixutoprim_serial(DGrid g, double *__restrict__ p, int8_t *__restrict__ f,
                                                          unsigned long long *__restrict__ bits,
                                                          int *__restrict__ pcount, unsigned nwords,
                                                          unsigned long long *__restrict__ cnt)
{
    if (halted(g)) return;
    const int h = blockIdx.x;
    if (pcount[h] == 0) return;
    __shared__ unsigned long long sw[64];
    unsigned long long *row = bits + (size_t)h * nwords;
    int c_rep = 0, c_nost = 0, c_flr = 0, c_cap = 0, c_gam = 0, fatal = 0;
    for (unsigned w0 = 0; w0 < nwords; w0 += 64) {
        const unsigned w = w0 + threadIdx.x;
        const unsigned long long word = (w < nwords) ? row[w] : 0ull;
        sw[threadIdx.x] = word;
        unsigned long long lanes = __ballot(word != 0ull);
        __syncthreads();
        if (threadIdx.x == 0) {
            while (lanes) {
                const int l = __ffsll((long long)lanes) - 1;
                lanes &= lanes - 1;
                unsigned long long b = sw[l];
                while (b) {
                    const int bit = __ffsll((long long)b) - 1;
                    b &= b - 1;
                    const unsigned col = (w0 + (unsigned)l) * 64u + (unsigned)bit;
                    repair_zone(g, p, f, (int)(col / (unsigned)g.SN2), (int)(col % (unsigned)g.SN2), h, true, nullptr,
                                c_rep, c_nost, c_flr, c_cap, c_gam, fatal);
                }
            }
        }
        __syncthreads();
        if (word) row[w] = 0ull;
    }
    if (threadIdx.x == 0) {
        atomicAdd(&cnt[HARM_FIXUP_REPAIRS], (unsigned long long)c_rep);
        atomicAdd(&cnt[HARM_FIXUP_NOSTENCIL], (unsigned long long)c_nost);
        atomicAdd(&cnt[HARM_FLOOR_ZONES], (unsigned long long)c_flr);
        atomicAdd(&cnt[HARM_GAMMA_CAP], (unsigned long long)c_cap);
        atomicAdd(&cnt[HARM_FIXUP_GAMMA], (unsigned long long)c_gam);
        atomicAdd(&cnt[HARM_FATAL], (unsigned long long)fatal);
        pcount[h] = 0;
    }
}

// ---------------------------------------------------------------------------
// diag_mass (diag.c:477-580, MDOTTYPE=NEU_AGA_GHOST1) partial sums of the mass
// flux through the first ghost columns, one block per phi plane.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(TPB) k_mdot_partials(DGrid g, const double *__restrict__ p, double *__restrict__ part)
{
    if (halted(g)) return;
    __shared__ double s_in[TPB], s_out[TPB];
    const int h = blockIdx.x;
    double a_in = 0., a_out = 0.;
    int fatal = 0;
    for (int j = g.js + threadIdx.x; j <= g.je; j += TPB) {
        for (int side = 0; side < 2; side++) {
            const int i = side ? g.ie + 1 : g.is - 1;
            const size_t z = zidx(g, i, j, h);
            double pr[NPR], ucon[4];
            for (int k = 0; k < NPR; k++) pr[k] = p[k * g.nz + z];
            Geom geom;
            load_geom(g.geom[CENT], g.ncol, (size_t)i * g.SN2 + j, geom);
            ucon_calc(pr, geom, ucon, fatal);
            const double v = pr[RHO] * ucon[1] * geom.g * g.dx2 * g.dx3;
            if (side) a_out -= v;
            else a_in += v;
        }
    }
    s_in[threadIdx.x] = a_in;
    s_out[threadIdx.x] = a_out;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s_in[threadIdx.x] += s_in[threadIdx.x + s];
            s_out[threadIdx.x] += s_out[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * h] = s_in[0];
        part[2 * h + 1] = s_out[0];
    }
}

// Start of a harm_step batch: tstate[5] = the fail() event count so far, so
// that k_stepend sees the events of its own step.  One thread.
__global__ void k_step_begin(double *tstate, const unsigned long long *cnt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (tstate[4] == 0.) tstate[5] = (double)cnt[HARM_FATAL];
}

// End of step_ch (step_ch.c:119-137): mass budget, dt control.  One thread.
// tstate: [0] t, [1] dt, [2] mass_in, [3] mass_out, [4] error code (11: dt
// below 1e-9, 14: a fail() event in this step), [5] fail() count at the start
// of the step.  The reference exits inside fail() (diag.c:381-397); here the
// step in which such an event happened completes, t is not advanced, and the
// batch stops.  A dt below 1e-9 exits after diag_mass and before t += dt
// (step_ch.c:119-129), which is where this kernel stops too.
__global__ void k_stepend(DGrid g, double *tstate, unsigned long long *ndt_slot, const double *part, int nparts,
                          const unsigned long long *cnt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (halted(g)) return;
    const double nfatal = (double)cnt[HARM_FATAL];
    if (nfatal != tstate[5]) {
        tstate[5] = nfatal;
        tstate[4] = 14.;
        return;
    }
    double mdot_in = 0., mdot_out = 0.;
    for (int q = 0; q < nparts; q++) {
        mdot_in += part[2 * q];
        mdot_out += part[2 * q + 1];
    }
    const double dt = tstate[1];
    tstate[2] += mdot_in * dt;
    tstate[3] += mdot_out * dt;
    const double ndt1 = __longlong_as_double((long long)ndt_slot[0]);
    const double ndt2 = __longlong_as_double((long long)ndt_slot[1]);
    const double ndt3 = __longlong_as_double((long long)ndt_slot[2]);
    double ndt;
    if (g.N3 > 2)
        ndt = g.defcon * 1. / (1. / ndt1 + 1. / ndt2 + 1. / ndt3);
    else
        ndt = g.defcon * 1. / (1. / ndt1 + 1. / ndt2);
    if (dt < 1.e-9) {
        tstate[4] = 11.;
        return;
    }
    double t = tstate[0] + dt;
    if (ndt > 1.3 * dt) ndt = 1.3 * dt;
    double ndt_out = ndt;
    if (t + ndt_out > g.tf) ndt_out = g.tf - t;
    tstate[0] = t;
    tstate[1] = ndt_out;
    const unsigned long long big = (unsigned long long)__double_as_longlong(1.e9);
    ndt_slot[0] = big;
    ndt_slot[1] = big;
    ndt_slot[2] = big;
}

// ---------------------------------------------------------------------------
// Initial data (init.c:180-472) on the device; the axisymmetric torus values come
// from the host, the per-zone KISS perturbation (random.c:44-55) is drawn here.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double kiss_uniform(uint32_t z, uint32_t w, uint32_t jsr, uint32_t jcong)
{
    z = 36969u * (z & 65535u) + (z >> 16);
    w = 18000u * (w & 65535u) + (w >> 16);
    const uint32_t mwc = (z << 16) + w;
    jcong = 69069u * jcong + 1234567u;
    jsr ^= (jsr << 17);
    jsr ^= (jsr >> 13);
    jsr ^= (jsr << 5);
    const uint32_t kiss = (mwc ^ jcong) + jsr;
    const double u = ((double)kiss) / 4294967295.0;
    return 0.0 + (1.0 - 0.0) * u;
}

__global__ void __launch_bounds__(TPB) k_init_expand(DGrid g, double *__restrict__ p, const double *__restrict__ tor,
                                                     double rhomax, double pert_amp, uint4 seed)
{
    const int ni = g.ie - g.is + 1, nj = g.je - g.js + 1;
    const size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (idx >= (size_t)ni * nj * g.N3) return;
    const unsigned rest = (unsigned)idx / (unsigned)g.N3;
    const int h = (int)((unsigned)idx - rest * (unsigned)g.N3);
    const unsigned ri = rest / (unsigned)nj;
    const int j = g.js + (int)(rest - ri * (unsigned)nj);
    const int i = g.is + (int)ri;
    const size_t z = zidx(g, i, j, h), nz = g.nz, col = (size_t)i * g.SN2 + j, nc = g.ncol;
    double rho = tor[col], u = tor[nc + col];
    double u1 = 0., u2 = 0., u3 = 0.;
    if (tor[5 * nc + col] != 0.) {
        const unsigned int ii = (unsigned int)(i - 2), jj = (unsigned int)(j - 2);
        const unsigned int hh = (g.N3 == 1) ? 0u : (unsigned int)h;
        const unsigned int off = hh * ii * jj + ii * jj + jj;
        const double rnd = kiss_uniform(seed.x + off, seed.y + off, seed.z + off, seed.w + off);
        u = u * (1. + pert_amp * (rnd - 0.5));
        u1 = tor[2 * nc + col];
        u2 = tor[3 * nc + col];
        u3 = tor[4 * nc + col];
    }
    p[RHO * nz + z] = rho / rhomax;
    p[UU * nz + z] = u / rhomax;
    p[U1 * nz + z] = u1;
    p[U2 * nz + z] = u2;
    p[U3 * nz + z] = u3;
    p[B1 * nz + z] = 0.;
    p[B2 * nz + z] = 0.;
    p[B3 * nz + z] = 0.;
}

// Corner vector potential A_phi = max(<rho>/rho_max - 0.2, 0) (init.c:330-365).
__global__ void __launch_bounds__(TPB) k_init_A(DGrid g, const double *__restrict__ p, double *__restrict__ A)
{
    const int ni = g.ie + 1 - g.is + 1, nj = g.je + 1 - g.js + 1;
    const size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (idx >= (size_t)ni * nj * g.N3) return;
    const unsigned rest = (unsigned)idx / (unsigned)g.N3;
    const int h = (int)((unsigned)idx - rest * (unsigned)g.N3);
    const unsigned ri = rest / (unsigned)nj;
    const int j = g.js + (int)(rest - ri * (unsigned)nj);
    const int i = g.is + (int)ri;
    const double *r = p + RHO * g.nz;
    const double rho_av = 0.25 * (r[zidx(g, i, j, h)] + r[zidx(g, i - 1, j, h)] + r[zidx(g, i, j - 1, h)] +
                                  r[zidx(g, i - 1, j - 1, h)]);
    const double q = rho_av / 1. - 0.2;
    A[zidx(g, i, j, h)] = (q > 0.) ? q : 0.;
}

// B = curl A at zone centres and b^2 (init.c:374-386); per-block max of b^2.
__global__ void __launch_bounds__(TPB) k_init_B(DGrid g, double *__restrict__ p, const double *__restrict__ A,
                                                double norm, int apply_norm, double *__restrict__ bsq_part)
{
    __shared__ double s[TPB];
    const int ni = g.ie - g.is + 1, nj = g.je - g.js + 1;
    const size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x;
    double bsq = 0.;
    if (idx < (size_t)ni * nj * g.N3) {
        const unsigned rest = (unsigned)idx / (unsigned)g.N3;
        const int h = (int)((unsigned)idx - rest * (unsigned)g.N3);
        const unsigned ri = rest / (unsigned)nj;
        const int j = g.js + (int)(rest - ri * (unsigned)nj);
        const int i = g.is + (int)ri;
        const size_t z = zidx(g, i, j, h), nz = g.nz, col = (size_t)i * g.SN2 + j;
        Geom geom;
        load_geom(g.geom[CENT], g.ncol, col, geom);
        double pr[NPR];
        for (int k = 0; k < NPR; k++) pr[k] = p[k * nz + z];
        if (apply_norm) {
            pr[B1] *= norm;
            pr[B2] *= norm;
        } else {
            const double a00 = A[zidx(g, i, j, h)], a01 = A[zidx(g, i, j + 1, h)];
            const double a10 = A[zidx(g, i + 1, j, h)], a11 = A[zidx(g, i + 1, j + 1, h)];
            pr[B1] = -(a00 - a01 + a10 - a11) / (2. * g.dx2 * geom.g);
            pr[B2] = (a00 + a01 - a10 - a11) / (2. * g.dx1 * geom.g);
            pr[B3] = 0.;
        }
        int fatal = 0;
        State q;
        get_state(pr, geom, q, fatal);
        bsq = dot4(q.bcon, q.bcov);
        for (int k = B1; k <= B3; k++) p[k * nz + z] = pr[k];
    }
    s[threadIdx.x] = bsq;
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) s[threadIdx.x] = (s[threadIdx.x + st] > s[threadIdx.x]) ? s[threadIdx.x + st] : s[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsq_part[blockIdx.x] = s[0];
}

// Diagnostics of diag() (diag.c:115-209) and diag_mass (diag.c:485-488): per
// block, the sums of U[RHO] dV, U[U3] dV, U[UU] dV, rho gdet dx1 dx2 dx3 and
// gdet rho (the COOL_EOS luminosity weight, diag.c:203), the largest |div B|
// and the ZLOOP rank ((h ni + i - is) nj + j - js) of the first zone that
// attains it (diag.c:190-195 keeps the first strict maximum).
constexpr int DIAG_NPART = 7;
constexpr double DIAG_NO_KEY = 9007199254740992.;  // 2^53, above every rank

template <bool TAB>
__global__ void __launch_bounds__(TPB) k_diag(DGrid g, const double *__restrict__ p, double *__restrict__ part)
{
    __shared__ double s[DIAG_NPART][TPB];
    const int ni = g.ie - g.is + 1, nj = g.je - g.js + 1;
    const size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x;
    const double dV = g.dx1 * g.dx2 * g.dx3;  // main.c:427
    double acc[5] = {0., 0., 0., 0., 0.}, divb = 0., key = DIAG_NO_KEY;
    if (idx < (size_t)ni * nj * g.N3) {
        const unsigned rest = (unsigned)idx / (unsigned)g.N3;
        const int h = (int)((unsigned)idx - rest * (unsigned)g.N3);
        const unsigned ri = rest / (unsigned)nj;
        const int j = g.js + (int)(rest - ri * (unsigned)nj);
        const int i = g.is + (int)ri;
        const size_t z = zidx(g, i, j, h), nz = g.nz, col = (size_t)i * g.SN2 + j;
        Geom geom;
        load_geom(g.geom[CENT], g.ncol, col, geom);
        double pr[NPR], U[NPR];
        for (int k = 0; k < NPR; k++) pr[k] = p[k * nz + z];
        int fatal = 0;
        State q;
        get_state(pr, geom, q, fatal);
        if (TAB)
            primtoflux(pr, q, 0, geom, EosTable{g.tab, g.ref_derivs}, U);
        else
            primtoflux(pr, q, 0, geom, g.gam, U);
        acc[0] = U[RHO] * dV;
        acc[1] = U[U3] * dV;
        acc[2] = U[UU] * dV;
        acc[3] = pr[RHO] * geom.g * g.dx1 * g.dx2 * g.dx3;
        acc[4] = geom.g * pr[RHO];
        const double *GD = g.geom[CENT] + 32 * g.ncol;
        const double gd00 = GD[col], gdjm = GD[col - 1], gdim = GD[col - g.SN2], gdimjm = GD[col - g.SN2 - 1];
        auto P = [&](int k, int ii, int jj, int hh) { return p[k * nz + zidx(g, ii, jj, hh)]; };
        if (g.N3 > 2) {
            const int hm = (h + g.N3 - 1) % g.N3;
            divb = fabs(0.25 *
                            (P(B1, i, j, h) * gd00 + P(B1, i, j, hm) * gd00 + P(B1, i, j - 1, h) * gdjm +
                             P(B1, i, j - 1, hm) * gdjm - P(B1, i - 1, j, h) * gdim - P(B1, i - 1, j, hm) * gdim -
                             P(B1, i - 1, j - 1, h) * gdimjm - P(B1, i - 1, j - 1, hm) * gdimjm) /
                            g.dx1 +
                        0.25 *
                            (P(B2, i, j, h) * gd00 + P(B2, i, j, hm) * gd00 + P(B2, i - 1, j, h) * gdim +
                             P(B2, i - 1, j, hm) * gdim - P(B2, i, j - 1, h) * gdjm - P(B2, i, j - 1, hm) * gdjm -
                             P(B2, i - 1, j - 1, h) * gdimjm - P(B2, i - 1, j - 1, hm) * gdimjm) /
                            g.dx2 +
                        0.25 *
                            (P(B3, i, j, h) * gd00 + P(B3, i - 1, j, h) * gdim + P(B3, i, j - 1, h) * gdjm +
                             P(B3, i - 1, j - 1, h) * gdimjm - P(B3, i, j, hm) * gd00 - P(B3, i - 1, j, hm) * gdim -
                             P(B3, i, j - 1, hm) * gdjm - P(B3, i - 1, j - 1, hm) * gdimjm) /
                            g.dx3);
        } else {
            divb = fabs(0.5 *
                            (P(B1, i, j, h) * gd00 + P(B1, i, j - 1, h) * gdjm - P(B1, i - 1, j, h) * gdim -
                             P(B1, i - 1, j - 1, h) * gdimjm) /
                            g.dx1 +
                        0.5 *
                            (P(B2, i, j, h) * gd00 + P(B2, i - 1, j, h) * gdim - P(B2, i, j - 1, h) * gdjm -
                             P(B2, i - 1, j - 1, h) * gdimjm) /
                            g.dx2);
        }
        if (divb > 0.) key = (double)(((size_t)h * ni + ri) * nj + (size_t)(j - g.js));
        else divb = 0.;
    }
    for (int k = 0; k < 5; k++) s[k][threadIdx.x] = acc[k];
    s[5][threadIdx.x] = divb;
    s[6][threadIdx.x] = key;
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const int o = threadIdx.x + st;
            for (int k = 0; k < 5; k++) s[k][threadIdx.x] += s[k][o];
            const double d0 = s[5][threadIdx.x], d1 = s[5][o];
            if (d1 > d0 || (d1 == d0 && s[6][o] < s[6][threadIdx.x])) {
                s[5][threadIdx.x] = d1;
                s[6][threadIdx.x] = s[6][o];
            }
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < DIAG_NPART) part[DIAG_NPART * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// diag_flux (diag.c:437-465) from the inner-face F1 of the last full step, and
// the ghost-face mass fluxes of diag_mass (diag.c:498-512), each summed in the
// reference's (h, j) loop order: the block forms the terms of a chunk, thread 0
// adds them up.  out: mdot, edot, ldot, mdot_in, mdot_out.
constexpr int DIAG_SEQ_CHUNK = 1024;
__global__ void __launch_bounds__(TPB) k_diag_seq(DGrid g, const double *__restrict__ p,
                                                  const double *__restrict__ F1, int flux_valid, double *out)
{
    __shared__ double t[5][DIAG_SEQ_CHUNK];
    const int nj = g.je - g.js + 1, n = g.N3 * nj;
    const size_t nz = g.nz;
    double acc[5] = {0., 0., 0., 0., 0.};
    for (int base = 0; base < n; base += DIAG_SEQ_CHUNK) {
        const int m = min(DIAG_SEQ_CHUNK, n - base);
        for (int q = threadIdx.x; q < m; q += TPB) {
            const int r = base + q, h = r / nj, j = g.js + (r - h * nj);
            double fm = 0., fe = 0., fl = 0.;
            if (flux_valid) {
                const size_t z = zidx(g, g.is, j, h);
                const double FR = F1[RHO * nz + z], FU = F1[UU * nz + z], FL = F1[U3 * nz + z];
                fm = FR * g.dx2 * g.dx3;
                fe = (FU - FR) * g.dx2 * g.dx3;
                fl = FL * g.dx2 * g.dx3;
            }
            double v[2];
            int fatal = 0;
            for (int side = 0; side < 2; side++) {
                const int i = side ? g.ie + 1 : g.is - 1;
                const size_t z = zidx(g, i, j, h);
                double pr[NPR], ucon[4];
                for (int k = 0; k < NPR; k++) pr[k] = p[k * nz + z];
                Geom geom;
                load_geom(g.geom[CENT], g.ncol, (size_t)i * g.SN2 + j, geom);
                ucon_calc(pr, geom, ucon, fatal);
                v[side] = pr[RHO] * ucon[1] * geom.g * g.dx2 * g.dx3;
            }
            t[0][q] = fm;
            t[1][q] = fe;
            t[2][q] = fl;
            t[3][q] = v[0];
            t[4][q] = v[1];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int q = 0; q < m; q++) {
                acc[0] += t[0][q];
                acc[1] -= t[1][q];
                acc[2] += t[2][q];
                acc[3] += t[3][q];
                acc[4] -= t[4][q];
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; k++) out[k] = acc[k];
}

// ---------------------------------------------------------------------------
// Lagrangian tracers (tracers.c:37-510).  Slot k of the tracer arrays is the
// seed zone k = (h N1 + i) N2 + j of init_tracers' loop (tracers.c:150-170);
// a tracer that leaves keeps its slot with a non-live state, so slot order is
// the reference's vector order without compaction.  Positions are code
// coordinates, SoA X[3][n].
// ---------------------------------------------------------------------------
struct TrArgs {
    double *X;
    int8_t *state;
    harm_tracer_rec *ev;  // removal records, appended
    unsigned long long *nev;
    unsigned long long cap_ev;
    long long n;
    double sx1, sx2, sx3;  // startx
    double R0, hslope, Rbr;
    double Rin, Rout, thmin, thmax;  // tracers.c:126-131
    double rho_unit, t_unit;
};

// metric.c:62-115 centre coordinate of active-relative index q
__device__ __forceinline__ double tr_cent(double sx, int q, double dx) { return sx + ((double)q + 0.5) * dx; }
__device__ __forceinline__ double tr_phcent(const DGrid &g, const TrArgs &a, int h)
{
    return g.N3 > 2 ? a.sx3 + ((double)h + 0.5) * g.dx3 : a.sx3 + (double)h * g.dx3;
}

// The k in [lo, hi] with edge(k) <= x < edge(k+1) for increasing edges, started
// from a guess: the same k as the reference's linear scans (tracers.c:50-95).
template <class F>
__device__ bool tr_interval(double x, int lo, int hi, double guess, F edge, int &k)
{
    const double gk = fmin(fmax(floor(guess), (double)lo), (double)hi);
    int q = (int)gk;
    while (q > lo && x < edge(q)) q--;
    while (q < hi && x >= edge(q + 1)) q++;
    if (x >= edge(q) && x < edge(q + 1)) {
        k = q;
        return true;
    }
    return false;
}

// search_tracer (tracers.c:37-98): hh = phi plane (0-based), ii, jj = array
// indices of the cell corner below the tracer.
__device__ bool tr_search(const DGrid &g, const TrArgs &a, const double *X, int &hh, int &ii, int &jj)
{
    if (X[0] < tr_cent(a.sx1, g.is - 3, g.dx1) || X[1] < tr_cent(a.sx2, g.js - 3, g.dx2)) return false;
    if (!isfinite(X[0]) || !isfinite(X[1])) return false;
    hh = 0;
    if (!tr_interval(X[0], g.is - 1, g.ie, (X[0] - a.sx1) / g.dx1 + 1.5,
                     [&](int q) { return tr_cent(a.sx1, q - 2, g.dx1); }, ii))
        return false;
    if (!tr_interval(X[1], g.js - 1, g.je, (X[1] - a.sx2) / g.dx2 + 1.5,
                     [&](int q) { return tr_cent(a.sx2, q - 2, g.dx2); }, jj))
        return false;
    if (g.N3 > 2) {
        if (!isfinite(X[2])) return false;
        if (!tr_interval(X[2], 0, g.N3 - 1, (X[2] - a.sx3) / g.dx3, [&](int q) { return a.sx3 + (double)q * g.dx3; },
                         hh))
            return false;
    }
    return true;
}

// bl_coord (coord.c:61-125) for r and theta
__device__ void tr_bl(const TrArgs &a, const double *X, double &r, double &th)
{
    double theexp = X[0];
    if (a.Rbr >= 0.) {
        const double x1br = log(a.Rbr - a.R0);
        if (X[0] > x1br) theexp += 1.0 * pow(X[0] - x1br, 4.0);
    }
    r = exp(theexp) + a.R0;
    th = M_PI_2 * (1. + X[1]) + ((1. - a.hslope) / 2.) * sin(M_PI * (1. + X[1]));
    if (fabs(th) < 1.E-20) th = (th >= 0) ? 1.E-20 : -1.E-20;
    if (fabs(M_PI - th) < 1.E-20) th = (th >= M_PI) ? M_PI + 1.E-20 : M_PI - 1.E-20;
}

// linear_interpolation (tracers.c:174-245) of NV values of the zones around
// (hh, ii, jj) to X; the plane past the last is the periodic ghost plane.
template <int NV, class VF>
__device__ void tr_interp(const DGrid &g, const TrArgs &a, const double *X, int hh, int ii, int jj, VF vf,
                          double *val)
{
    const int wph = g.N3 > 2 ? 2 : 1;
    double tmp2[2][NV];
    const double x0 = tr_cent(a.sx2, jj - 2, g.dx2), x1 = tr_cent(a.sx2, jj - 1, g.dx2);
    const double r0 = tr_cent(a.sx1, ii - 2, g.dx1), r1 = tr_cent(a.sx1, ii - 1, g.dx1);
    for (int l3 = 0; l3 < wph; l3++) {
        const int hz = (hh + l3 < g.N3) ? hh + l3 : hh + l3 - g.N3;
        double tmp1[2][NV];
        for (int l1 = 0; l1 < 2; l1++) {
            double v0[NV], v1[NV];
            vf(hz, ii + l1, jj, v0);
            vf(hz, ii + l1, jj + 1, v1);
            const double T1 = 1.0 / (x1 - x0), T2 = X[1] - x0;
            for (int k = 0; k < NV; k++) tmp1[l1][k] = v0[k] + T2 * ((v1[k] - v0[k]) * T1);
        }
        const double T1 = 1.0 / (r1 - r0), T2 = X[0] - r0;
        for (int k = 0; k < NV; k++) tmp2[l3][k] = tmp1[0][k] + T2 * ((tmp1[1][k] - tmp1[0][k]) * T1);
    }
    if (wph == 1) {
        for (int k = 0; k < NV; k++) val[k] = tmp2[0][k];
        return;
    }
    const double p0 = tr_phcent(g, a, hh), p1 = tr_phcent(g, a, hh + 1);
    const double T1 = 1.0 / (p1 - p0), T2 = X[2] - p0;
    for (int k = 0; k < NV; k++) val[k] = tmp2[0][k] + T2 * ((tmp2[1][k] - tmp2[0][k]) * T1);
}

// ucon_fun (tracers.c:248-260): u^t, u^i / u^t of one zone
__device__ void tr_vel(const DGrid &g, const double *__restrict__ p, int h, int i, int j, double *v)
{
    const size_t z = zidx(g, i, j, h), nz = g.nz;
    double pr[NPR];
    for (int k = 0; k < NPR; k++) pr[k] = p[k * nz + z];
    Geom geom;
    load_geom(g.geom[CENT], g.ncol, (size_t)i * g.SN2 + j, geom);
    int fatal = 0;
    double ucon[4];
    ucon_calc(pr, geom, ucon, fatal);
    v[0] = ucon[0];
    v[1] = ucon[1] / ucon[0];
    v[2] = ucon[2] / ucon[0];
    v[3] = ucon[3] / ucon[0];
}

// rhoTYe_fun (tracers.c:432-466): rho RHO_UNIT, T, Ye of one zone
template <bool TAB>
__device__ void tr_rhoT(const DGrid &g, const TrArgs &a, const double *__restrict__ p, int h, int i, int j, double *v)
{
    const size_t z = zidx(g, i, j, h), nz = g.nz;
    const double rho = p[RHO * nz + z], u = p[UU * nz + z];
    v[0] = rho * a.rho_unit;
    if (TAB) {
        double pp, d1, d2, T;
        eos_table_p(g.tab, rho, u, pp, d1, d2, &T);
        v[1] = T;
    } else {
        constexpr double CL = 2.99792458e10, KBOL = 1.38e-16, mn = 1.67492729e-24, me = 9.109565e-28,
                         mp = 1.673e-24, mhe = 6.6465e-24;
        constexpr double mav = (mp + mn + 2. * me + mhe) / 5.;
        v[1] = (g.gam - 1.) * u / rho / KBOL * mav * CL * CL;
    }
    v[2] = -1.;
}

// dump_one_tracer (tracers.c:470-510) at X with the cell found before the move
template <bool TAB>
__device__ void tr_record(const DGrid &g, const TrArgs &a, const double *__restrict__ p, const double *X, int hh,
                          int ii, int jj, double t, long long k, int kind, harm_tracer_rec &rec)
{
    double val[3];
    tr_interp<3>(g, a, X, hh, ii, jj, [&](int h, int i, int j, double *v) { tr_rhoT<TAB>(g, a, p, h, i, j, v); },
                 val);
    double r, th, ph;
    tr_bl(a, X, r, th);
    if (g.N3 > 2) {
        ph = X[2];
        if (ph > 2. * M_PI)
            ph -= 2. * M_PI;
        else if (ph < 0.)
            ph += 2. * M_PI;
    } else {
        ph = 0.;
    }
    const long long n12 = (long long)g.N1 * g.N2;
    rec.t = t * a.t_unit;
    rec.rho = val[0];
    rec.T = val[1];
    rec.ye = val[2];
    rec.r = r;
    rec.th = th;
    rec.ph = ph;
    rec.h = (int32_t)(k / n12);
    rec.i = (int32_t)((k % n12) / g.N2);
    rec.j = (int32_t)(k % g.N2);
    rec.kind = kind;
}

// init_tracers (tracers.c:109-172)
__global__ void __launch_bounds__(TPB) k_tracer_init(DGrid g, TrArgs a, const double *__restrict__ p,
                                                     unsigned long long *count)
{
    const long long k = (long long)blockIdx.x * TPB + threadIdx.x;
    int live = 0;
    if (k < a.n) {
        const long long n12 = (long long)g.N1 * g.N2;
        const int h = (int)(k / n12), i = (int)((k % n12) / g.N2), j = (int)(k % g.N2);
        double X[3];
        X[0] = a.sx1 + (i + 0.5) * g.dx1;
        X[1] = a.sx2 + (j + 0.5) * g.dx2;
        X[2] = a.sx3 + h * g.dx3;
        int hh, ii, jj;
        if (tr_search(g, a, X, hh, ii, jj)) live = p[RHO * g.nz + zidx(g, ii, jj, hh)] > RHOMIN;
        a.X[k] = X[0];
        a.X[a.n + k] = X[1];
        a.X[2 * a.n + k] = X[2];
        a.state[k] = live ? HARM_TRACER_LIVE : HARM_TRACER_NONE;
    }
    const unsigned long long b = __ballot(live);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, (unsigned long long)__popcll(b));
}

// update_tracers (tracers.c:298-430) after a step: t and dt are the step's
// (tstate before k_stepend advances t).
template <bool TAB>
__global__ void __launch_bounds__(TPB) k_tracer_update(DGrid g, TrArgs a, const double *__restrict__ p,
                                                       const double *__restrict__ tstate,
                                                       const unsigned long long *__restrict__ cnt)
{
    if (halted(g)) return;
    // a step with a fail() event stops before update_tracers (tstate[5]: the
    // event count at the start of the step, k_step_begin / k_stepend)
    if (cnt && (double)cnt[HARM_FATAL] != tstate[5]) return;
    const long long k = (long long)blockIdx.x * TPB + threadIdx.x;
    if (k >= a.n || a.state[k] != HARM_TRACER_LIVE) return;
    const double t = tstate[0], Dt = tstate[1];
    double X[3] = {a.X[k], a.X[a.n + k], a.X[2 * a.n + k]};
    int hh, ii, jj;
    if (!tr_search(g, a, X, hh, ii, jj)) {
        a.state[k] = HARM_TRACER_LOST;
        return;
    }
    double ucon[4];
    tr_interp<4>(g, a, X, hh, ii, jj, [&](int h, int i, int j, double *v) { tr_vel(g, p, h, i, j, v); }, ucon);
    for (int d = 0; d < 3; d++) X[d] += ucon[d + 1] * Dt;
    double r, th;
    tr_bl(a, X, r, th);
    int kind = 0;
    if (r < a.Rin || th < 0 || th > M_PI)
        kind = HARM_TRACER_IN;
    else if (r >= a.Rout)
        kind = (th >= a.thmin && th <= a.thmax) ? HARM_TRACER_OUT : HARM_TRACER_JET;
    if (kind) {
        harm_tracer_rec rec;
        tr_record<TAB>(g, a, p, X, hh, ii, jj, t, k, kind, rec);
        const unsigned long long slot = atomicAdd(a.nev, 1ull);
        if (slot < a.cap_ev) a.ev[slot] = rec;
        a.state[k] = (int8_t)kind;
    } else if (g.N3 > 2) {  // tracers.c:386-392 (phi-periodic on one rank)
        if (X[2] < 0.)
            X[2] += 2. * M_PI;
        else if (X[2] > 2. * M_PI)
            X[2] -= 2. * M_PI;
    }
    a.X[k] = X[0];
    a.X[a.n + k] = X[1];
    a.X[2 * a.n + k] = X[2];
}

// dump_tracers (tracers.c:512-560), pass 1: per-block count of the live
// tracers that still lie in the grid (the others become LOST).
__global__ void __launch_bounds__(TPB) k_tracer_count(DGrid g, TrArgs a, unsigned *bcount)
{
    __shared__ unsigned s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const long long k = (long long)blockIdx.x * TPB + threadIdx.x;
    if (k < a.n && a.state[k] == HARM_TRACER_LIVE) {
        const double X[3] = {a.X[k], a.X[a.n + k], a.X[2 * a.n + k]};
        int hh, ii, jj;
        if (tr_search(g, a, X, hh, ii, jj))
            atomicAdd(&s, 1u);
        else
            a.state[k] = HARM_TRACER_LOST;
    }
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = s;
}

// pass 2: records at base[block] + rank of the tracer in its block
template <bool TAB>
__global__ void __launch_bounds__(TPB) k_tracer_dump(DGrid g, TrArgs a, const double *__restrict__ p,
                                                     const double *__restrict__ tstate,
                                                     const unsigned long long *__restrict__ base,
                                                     harm_tracer_rec *__restrict__ out)
{
    __shared__ unsigned wsum[TPB / 64];
    const long long k = (long long)blockIdx.x * TPB + threadIdx.x;
    const bool live = k < a.n && a.state[k] == HARM_TRACER_LIVE;
    const unsigned long long b = __ballot(live);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = (unsigned)__popcll(b);
    __syncthreads();
    if (!live) return;
    unsigned off = (unsigned)__popcll(b & ((1ull << lane) - 1ull));
    for (int q = 0; q < w; q++) off += wsum[q];
    const double X[3] = {a.X[k], a.X[a.n + k], a.X[2 * a.n + k]};
    int hh, ii, jj;
    tr_search(g, a, X, hh, ii, jj);  // found in pass 1
    tr_record<TAB>(g, a, p, X, hh, ii, jj, tstate[0], k, HARM_TRACER_LIVE, out[base[blockIdx.x] + off]);
}

// ---------------------------------------------------------------------------
// dump() (dump.c:1710-1990): the per-zone columns of the ASCII dump, zones in
// (i, j, h) order.  Columns 0-3 (code and BL coordinates) are filled on the
// host from its coordinate tables; this kernel writes 4-45: the 8 primitives,
// the 7 cooling values (Lambda, P, T, H_d, Qnu, tau, Ye: zeros with the Gamma
// law; the table's P and T with the tabulated EOS), the dump's own corner
// |div B| and, unless the run failed, u^mu, u_mu, b^mu, b_mu, the dir-1 and
// dir-2 characteristic speeds, gdet, |v|, and the (p, c_s, u^r, Mach) block.
// ---------------------------------------------------------------------------
constexpr int DUMP_NCOL = 46, DUMP_NCOL_FAILED = 20;

template <bool TAB>
__global__ void __launch_bounds__(TPB) k_dump_cols(DGrid g, const double *__restrict__ p, size_t z0, size_t n,
                                                   double *__restrict__ out)
{
    const size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= n) return;
    const size_t zl = z0 + t, nj = (size_t)(g.je - g.js + 1);
    const int h = (int)(zl % g.N3);
    const size_t rest = zl / g.N3;
    const int j = g.js + (int)(rest % nj), i = g.is + (int)(rest / nj);
    const size_t nz = g.nz, z = zidx(g, i, j, h);
    double *o = out + t * DUMP_NCOL;
    double pr[NPR];
    for (int k = 0; k < NPR; k++) {
        pr[k] = p[k * nz + z];
        o[4 + k] = pr[k];
    }
    for (int k = 0; k < 7; k++) o[12 + k] = 0.0;
    if (TAB) {
        double pp, d1, d2, T;
        eos_table_p(g.tab, pr[RHO], pr[UU], pp, d1, d2, &T);
        o[13] = pp;
        o[14] = T;
    }
    const double *GD = g.geom[CENT] + 32 * g.ncol;
    const size_t col = (size_t)i * g.SN2 + j;
    const double g00 = GD[col], gjm = GD[col - 1], gim = GD[col - g.SN2], gimjm = GD[col - g.SN2 - 1];
    auto P = [&](int k, int ii, int jj, int hh) { return p[k * nz + zidx(g, ii, jj, hh)]; };
    double divb;
    if (g.N3 == 1) {
        // dump.c:1864-1879: interior corners only (array indices)
        if (i > 0 && j > 0 && i < g.N1 && j < g.N2)
            divb = fabs(0.5 *
                            (P(B1, i, j, h) * g00 + P(B1, i, j - 1, h) * gjm - P(B1, i - 1, j, h) * gim -
                             P(B1, i - 1, j - 1, h) * gimjm) /
                            g.dx1 +
                        0.5 *
                            (P(B2, i, j, h) * g00 + P(B2, i - 1, j, h) * gim - P(B2, i, j - 1, h) * gjm -
                             P(B2, i - 1, j - 1, h) * gimjm) /
                            g.dx2);
        else
            divb = 0.;
    } else {
        // dump.c:1881-1915, term for term (the (i, j-1) B3 term takes gdet of column (i, j))
        const int m = (h + g.N3 - 1) % g.N3;
        divb = fabs(0.25 *
                        (+P(B1, i, j, m) * g00 + P(B1, i, j - 1, m) * gjm - P(B1, i - 1, j, m) * gim -
                         P(B1, i - 1, j - 1, m) * gimjm

                         + P(B1, i, j, h) * g00 + P(B1, i, j - 1, h) * gjm - P(B1, i - 1, j, h) * gim -
                         P(B1, i - 1, j - 1, h) * gimjm) /
                        g.dx1 +
                    0.25 *
                        (+P(B2, i, j, m) * g00 + P(B2, i - 1, j, m) * gim - P(B2, i, j - 1, m) * gjm -
                         P(B2, i - 1, j - 1, m) * gimjm

                         + P(B2, i, j, h) * g00 + P(B2, i - 1, j, h) * gim - P(B2, i, j - 1, h) * gjm -
                         P(B2, i - 1, j - 1, h) * gimjm) /
                        g.dx2 +
                    0.25 *
                        (+P(B3, i, j, h) * g00 + P(B3, i - 1, j, h) * gim - P(B3, i, j, m) * g00 -
                         P(B3, i - 1, j, m) * gim

                         + P(B3, i, j - 1, h) * g00 + P(B3, i - 1, j - 1, h) * gimjm - P(B3, i, j - 1, m) * gjm -
                         P(B3, i - 1, j - 1, m) * gimjm) /
                        g.dx3);
    }
    o[19] = divb;
    Geom geom;
    load_geom(g.geom[CENT], g.ncol, col, geom);
    State q;
    int fatal = 0;
    get_state(pr, geom, q, fatal);
    for (int k = 0; k < 4; k++) {
        o[20 + k] = q.ucon[k];
        o[24 + k] = q.ucov[k];
        o[28 + k] = q.bcon[k];
        o[32 + k] = q.bcov[k];
    }
    double vmax, vmin;
    for (int dir = 1; dir <= 2; dir++) {
        if (TAB)
            vchar(pr, q, geom, dir, EosTable{g.tab, g.ref_derivs}, &vmax, &vmin);
        else
            vchar(pr, q, geom, dir, g.gam, &vmax, &vmin);
        o[36 + 2 * (dir - 1)] = vmin;
        o[37 + 2 * (dir - 1)] = vmax;
    }
    o[40] = geom.g;
    double vsq = 0.;
    for (int a = 1; a < 4; a++)
        for (int b = 1; b < 4; b++) vsq += geom.gcov[a][b] * q.ucon[a] * q.ucon[b];
    const double gsq = 1. + vsq;
    vsq /= gsq;
    o[41] = sqrt(vsq);
    const double ps_p = pr[UU] * (g.gam - 1.0);
    const double ps_cs = sqrt(4.0 / 3.0 * ps_p / (pr[RHO] + 4.0 * ps_p));
    const double uur = sqrt(geom.gcov[1][1]) * q.ucon[1] / q.ucon[0];
    o[42] = ps_p;
    o[43] = ps_cs;
    o[44] = uur;
    o[45] = -uur / ps_cs;
}

// Batched Noble 2D inversion over SoA arrays (utoprim_2d.c:126, plus the Font fix).
// Batched inversion; LEARNED seeds Newton from the network (all lanes take
// part in the wave-level forward pass, so there is no early exit).
template <bool LEARNED>
__global__ void __launch_bounds__(TPB) k_c2p_batch(const double *__restrict__ U, const double *__restrict__ gcov,
                                                   const double *__restrict__ gcon, const double *__restrict__ gdet,
                                                   double *__restrict__ prim, int8_t *__restrict__ ret,
                                                   uint8_t *__restrict__ iters, uint8_t *__restrict__ kconv, size_t n,
                                                   double gam, int font_fix, double K, int *__restrict__ ret_full,
                                                   MlpDev mlp, double gate_tau)
{
    const size_t z = (size_t)blockIdx.x * TPB + threadIdx.x;
    const bool act = z < n;
    if (!LEARNED && !act) return;
    Geom geom;
    double Uz[NPR], pr[NPR];
    if (act) {
#pragma unroll
        for (int a = 0; a < 16; a++) {
            geom.gcov[a / 4][a % 4] = gcov[(size_t)a * n + z];
            geom.gcon[a / 4][a % 4] = gcon[(size_t)a * n + z];
        }
        geom.g = gdet[z];
        geom_derive(geom);
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            Uz[k] = U[(size_t)k * n + z];
            pr[k] = prim[(size_t)k * n + z];
        }
    }
    int niter = 0, kc = 0, r = 0;
    if (LEARNED && gate_tau > 0.) {
        // gated: the network only for the waves with a zone whose plain start is poor
        Gnr2 st;
        bool need = false;
        int ph1 = 0;
        if (act) ph1 = u2p_gate_first(Uz, geom, gam, pr, gate_tau, st, need);
        double W0 = 0., vsq0 = 0.;
        if (__any(need)) {
            double f[MLP_NIN];
            const bool ok = need && mlp_features_U(Uz, geom, f);
            float x[MLP_NIN];
            if (ok)
                mlp_standardize(mlp, f, x);
            else
#pragma unroll
                for (int k = 0; k < MLP_NIN; k++) x[k] = 0.f;
            float y0, y1;
            mlp_forward_wave(mlp, x, y0, y1);
            if (ok) mlp_unstandardize(mlp, y0, y1, mlp_wref(f), W0, vsq0);
        }
        if (!act) return;
        int seeded;
        r = u2p_gate_finish(Uz, geom, gam, pr, ph1, st, need, W0, vsq0, mlp.vsq_from_W, niter, kc, seeded);
    } else if constexpr (LEARNED) {
        double f[MLP_NIN];
        const bool ok = act && mlp_features_U(Uz, geom, f);
        float x[MLP_NIN];
        if (ok)
            mlp_standardize(mlp, f, x);
        else
#pragma unroll
            for (int k = 0; k < MLP_NIN; k++) x[k] = 0.f;
        float y0, y1;
        mlp_forward_wave(mlp, x, y0, y1);
        if (!act) return;
        double W0 = 0., vsq0 = 0.;
        if (ok) mlp_unstandardize(mlp, y0, y1, mlp_wref(f), W0, vsq0);
        int fb;
        r = utoprim_2d_learned(Uz, geom, gam, pr, W0, vsq0, mlp.vsq_from_W, niter, kc, fb);
    } else {
        r = utoprim_2d(Uz, geom, gam, pr, niter, kc);
    }
    int8_t flag = (int8_t)r;
    if (flag && font_fix) {
        flag = (int8_t)utoprim_1dvsq2fix1(Uz, geom, gam, K, pr);
        if (flag) flag = (int8_t)utoprim_1dfix1(Uz, geom, gam, K, pr);
    }
#pragma unroll
    for (int k = 0; k < NPR; k++) prim[(size_t)k * n + z] = pr[k];
    ret[z] = flag;
    if (ret_full) ret_full[z] = r;
    if (iters) iters[z] = (uint8_t)niter;
    if (kconv) kconv[z] = (uint8_t)kc;
}

// ---------------------------------------------------------------------------
// Synthetic stress batch for the inversion alone (no reference counterpart):
// zones with a Kerr-Schild metric in MKS code coordinates at random (r, theta)
// and random primitive states, and the conserved vector U = primtoU(P)
// (phys.c:100-105).  Each zone draws from its own KISS stream (random.c:30-43)
// seeded {s0+z, s1+z, s2+z, s3+z}.
// ---------------------------------------------------------------------------
struct KissStream {
    uint32_t z, w, jsr, jcong;
    __device__ double next()
    {
        z = 36969u * (z & 65535u) + (z >> 16);
        w = 18000u * (w & 65535u) + (w >> 16);
        const uint32_t mwc = (z << 16) + w;
        jcong = 69069u * jcong + 1234567u;
        jsr ^= (jsr << 17);
        jsr ^= (jsr >> 13);
        jsr ^= (jsr << 5);
        return (double)((mwc ^ jcong) + jsr) / 4294967295.0;
    }
};

struct StressSpec {
    double a, hslope, gam;
    double r_lo, r_hi, th_lo, th_hi;      // r log-uniform, theta uniform
    double rho_lo, rho_hi;                // log-uniform
    double urho_lo, urho_hi;              // u/rho log-uniform
    double lor_lo, lor_hi;                // Lorentz factor uniform
    double sig_lo, sig_hi;                // sigma = b^2/rho log-uniform
    double pert;                          // relative perturbation of the guess
    uint32_t seed[4];
    int guess_kind;                       // 0: perturbed truth, 1: cold (rho = D, u = rho, v = 0)
};

// Kerr-Schild metric (coord.c:128-199 before the Jacobian) at (r, th) in MKS
// coordinates with R0 = 0: Jacobian diag(1, r, dth/dX2, 1) taken analytically,
// X2 found from th by Newton on th(X2) = pi/2 (1+X2) + (1-h)/2 sin(pi (1+X2)).
__device__ void ks_mks_geom(double a, double h, double r, double th, Geom &g)
{
    double X2 = 2. * th / M_PI - 1.;
    for (int it = 0; it < 40; it++) {
        const double f = 0.5 * M_PI * (1. + X2) + 0.5 * (1. - h) * sin(M_PI * (1. + X2)) - th;
        const double df = 0.5 * M_PI * (1. + (1. - h) * cos(M_PI * (1. + X2)));
        X2 -= f / df;
    }
    const double dth = 0.5 * M_PI * (1. + (1. - h) * cos(M_PI * (1. + X2)));
    const double sth = sin(th), cth = cos(th), s2 = sth * sth, rho2 = r * r + a * a * cth * cth;
    double gc[4][4] = {}, gn[4][4] = {};
    gc[0][0] = -1. + 2. * r / rho2;
    gc[0][1] = gc[1][0] = 2. * r / rho2;
    gc[0][3] = gc[3][0] = -2. * a * r * s2 / rho2;
    gc[1][1] = 1. + 2. * r / rho2;
    gc[1][3] = gc[3][1] = -a * s2 * (1. + 2. * r / rho2);
    gc[2][2] = rho2;
    gc[3][3] = s2 * (rho2 + a * a * s2 * (1. + 2. * r / rho2));
    gn[0][0] = -(1. + 2. * r / rho2);
    gn[0][1] = gn[1][0] = 2. * r / rho2;
    gn[1][1] = (r * r - 2. * r + a * a) / rho2;
    gn[1][3] = gn[3][1] = a / rho2;
    gn[2][2] = 1. / rho2;
    gn[3][3] = 1. / (rho2 * s2);
    const double J[4] = {1., r, dth, 1.};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            g.gcov[i][j] = gc[i][j] * J[i] * J[j];
            g.gcon[i][j] = gn[i][j] / (J[i] * J[j]);
        }
    g.g = rho2 * fabs(sth) * r * dth;
    geom_derive(g);
}

// x = s L^{-T} n for the Cholesky factor L of the spatial metric gamma_ij =
// gcov[1..3][1..3]: gamma(x, x) = s^2 for a Euclidean unit vector n.
__device__ void spatial_scale(const Geom &g, const double *n, double s, double *x)
{
    const double l00 = sqrt(g.gcov[1][1]);
    const double l10 = g.gcov[2][1] / l00, l20 = g.gcov[3][1] / l00;
    const double l11 = sqrt(g.gcov[2][2] - l10 * l10);
    const double l21 = (g.gcov[3][2] - l20 * l10) / l11;
    const double l22 = sqrt(g.gcov[3][3] - l20 * l20 - l21 * l21);
    x[2] = n[2] / l22;
    x[1] = (n[1] - l21 * x[2]) / l11;
    x[0] = (n[0] - l10 * x[1] - l20 * x[2]) / l00;
    for (int k = 0; k < 3; k++) x[k] *= s;
}

__device__ void unit3(KissStream &ks, double *n)
{
    const double c = 2. * ks.next() - 1., ph = 2. * M_PI * ks.next();
    const double sn = sqrt(fmax(0., 1. - c * c));
    n[0] = sn * cos(ph);
    n[1] = sn * sin(ph);
    n[2] = c;
}

__global__ void __launch_bounds__(TPB) k_stress_gen(StressSpec sp, size_t n, double *__restrict__ U,
                                                    double *__restrict__ gcov, double *__restrict__ gcon,
                                                    double *__restrict__ gdet, double *__restrict__ ptrue,
                                                    double *__restrict__ guess)
{
    const size_t z = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (z >= n) return;
    const uint32_t off = (uint32_t)z;
    KissStream ks = {sp.seed[0] + off, sp.seed[1] + off, sp.seed[2] + off, sp.seed[3] + off};
    auto logu = [&](double lo, double hi) { return exp(log(lo) + (log(hi) - log(lo)) * ks.next()); };
    const double r = logu(sp.r_lo, sp.r_hi);
    const double th = sp.th_lo + (sp.th_hi - sp.th_lo) * ks.next();
    Geom geom;
    ks_mks_geom(sp.a, sp.hslope, r, th, geom);
    double pr[NPR];
    pr[RHO] = logu(sp.rho_lo, sp.rho_hi);
    pr[UU] = pr[RHO] * logu(sp.urho_lo, sp.urho_hi);
    const double lor = sp.lor_lo + (sp.lor_hi - sp.lor_lo) * ks.next();
    double nv[3], ut[3], nb[3], bh[3];
    unit3(ks, nv);
    spatial_scale(geom, nv, sqrt(lor * lor - 1.), ut);
    for (int k = 0; k < 3; k++) pr[U1 + k] = ut[k];
    unit3(ks, nb);
    spatial_scale(geom, nb, 1., bh);
    for (int k = 0; k < 3; k++) pr[B1 + k] = bh[k];
    {
        // scale B so that b^2 = sigma rho (b^mu is linear in B)
        State q;
        int fatal = 0;
        get_state(pr, geom, q, fatal);
        const double bsq1 = dot4(q.bcon, q.bcov);
        const double sig = logu(sp.sig_lo, sp.sig_hi);
        const double sc = sqrt(sig * pr[RHO] / bsq1);
        for (int k = 0; k < 3; k++) pr[B1 + k] *= sc;
    }
    double Uz[NPR];
    {
        State q;
        int fatal = 0;
        get_state(pr, geom, q, fatal);
        primtoflux(pr, q, 0, geom, sp.gam, Uz);
    }
    double gs[NPR];
    if (sp.guess_kind == 0) {
        for (int k = 0; k < NPR; k++) gs[k] = pr[k];
        for (int k = RHO; k <= U3; k++) gs[k] = pr[k] * (1. + sp.pert * (2. * ks.next() - 1.));
    } else {
        const double alpha = 1. / sqrt(-geom.gcon[0][0]);
        gs[RHO] = alpha * Uz[RHO] / geom.g;
        gs[UU] = gs[RHO];
        gs[U1] = gs[U2] = gs[U3] = 0.;
        for (int k = B1; k <= B3; k++) gs[k] = pr[k];
    }
    for (int k = 0; k < NPR; k++) {
        U[(size_t)k * n + z] = Uz[k];
        ptrue[(size_t)k * n + z] = pr[k];
        guess[(size_t)k * n + z] = gs[k];
    }
    for (int a = 0; a < 16; a++) {
        gcov[(size_t)a * n + z] = geom.gcov[a / 4][a % 4];
        gcon[(size_t)a * n + z] = geom.gcon[a / 4][a % 4];
    }
    gdet[z] = geom.g;
}

// Batched Utoprim_5d with the Gamma-law (tab.nr == 0) or the tabulated EOS.
__global__ void __launch_bounds__(TPB) k_c2p5_batch(const double *__restrict__ U, const double *__restrict__ gcov,
                                                    const double *__restrict__ gcon, const double *__restrict__ gdet,
                                                    double *__restrict__ prim, int8_t *__restrict__ ret,
                                                    uint8_t *__restrict__ iters, size_t n, double gam, EosTab tab,
                                                    int *__restrict__ ret_full)
{
    const size_t z = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (z >= n) return;
    Geom geom;
#pragma unroll
    for (int a = 0; a < 16; a++) {
        geom.gcov[a / 4][a % 4] = gcov[(size_t)a * n + z];
        geom.gcon[a / 4][a % 4] = gcon[(size_t)a * n + z];
    }
    geom.g = gdet[z];
    geom_derive(geom);
    double Uz[NPR], pr[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) {
        Uz[k] = U[(size_t)k * n + z];
        pr[k] = prim[(size_t)k * n + z];
    }
    Eos5 eos;
    eos.tab = tab.nr > 0 ? &tab : nullptr;
    eos.gam = gam;
    int niter = 0;
    const int r = utoprim_5d(Uz, geom, eos, pr, niter);
#pragma unroll
    for (int k = 0; k < NPR; k++) prim[(size_t)k * n + z] = pr[k];
    ret[z] = (int8_t)r;
    if (ret_full) ret_full[z] = r;
    if (iters) iters[z] = (uint8_t)niter;
}

__global__ void __launch_bounds__(TPB) k_eos_eval(EosTab tab, const double *__restrict__ rho,
                                                  const double *__restrict__ u, double *__restrict__ p,
                                                  double *__restrict__ dpdrho, double *__restrict__ dpdu, size_t n)
{
    const size_t z = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (z >= n) return;
    double pp, a, b;
    eos_table_p(tab, rho[z], u[z], pp, a, b);
    p[z] = pp;
    if (dpdrho) dpdrho[z] = a;
    if (dpdu) dpdu[z] = b;
}

__global__ void __launch_bounds__(TPB) k_pow(const double *__restrict__ x, const double *__restrict__ y,
                                             double *__restrict__ out, size_t n)
{
    const size_t z = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (z < n) out[z] = cr_pow(x[z], y[z]);
}

// Zone temperatures of the tabulated EOS for the restart file (all zones).
__global__ void __launch_bounds__(TPB) k_zone_temp(EosTab tab, const double *__restrict__ p, double *__restrict__ T,
                                                   size_t nz)
{
    const size_t z = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (z >= nz) return;
    double pp, a, b, tz;
    eos_table_p(tab, p[z], p[nz + z], pp, a, b, &tz);
    T[z] = tz;
}

// Plain device copy of known size: calibrates the FETCH_SIZE / WRITE_SIZE counters.
__global__ void __launch_bounds__(TPB) k_copy(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n)
{
    for (size_t z = (size_t)blockIdx.x * TPB + threadIdx.x; z < n; z += (size_t)gridDim.x * TPB) b[z] = a[z];
}

inline unsigned nblocks(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }
// Grid for kernels that use xcd_block(): a multiple of 8 workgroups.
inline unsigned nblocks8(size_t n) { return (nblocks(n) + 7u) & ~7u; }

}  // namespace

// ===========================================================================
// Context
// ===========================================================================
// Staging of harm_dump_columns_noalloc: the columns of at most
// HARM_DUMP_NOALLOC_MAX_ZONES zones (8192 x 46 doubles, 2.9 MiB), one copy per
// device for the life of the process.  In host code `scratch` names a host-side
// stand-in: kernels and copies take the device address harm_create looks up for
// its device (scratch_addr).
static __device__ __align__(256) double scratch[HARM_DUMP_NOALLOC_MAX_ZONES * DUMP_NCOL];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[SCRATCH_MAX_DEVICES];

// Once per device, with that device current (called from harm_create).
int scratch_resolve(int dev)
{
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return set_err("device index out of range", 4);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[dev]) return 0;
    double *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_err(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e), 2);
    g_scratch_dev[dev] = ptr;
    return 0;
}

double *scratch_addr(int dev)
{
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_dev[dev];
}

}  // namespace

struct harm_ctx {
    harm_params P;
    harm_host::Grid hg;
    harm_host::GeomTables tab;
    DGrid g;
    int device;
    hipStream_t stream;
    size_t nz;
    double *p, *ph, *F1, *F2, *F3, *e1, *e2, *e3;
    double *S;  // [8][active zones] flux divergence
    int8_t *pflag, *pflag2;
    double *d_geom[5], *d_conn, *d_floors, *d_katm, *d_cblk, *d_fblk;
    double *tstate;                   // t, dt, mass_in, mass_out, err
    unsigned long long *ndt_slot;     // 3
    double *dtpart;                   // per-block CFL minima of one flux launch
    unsigned long long *cpart;        // per-block counter totals of one update launch
    unsigned long long *cnt;          // HARM_NCOUNTERS
    double *part;                     // reduction partials
    size_t npart;
    unsigned long long *fix_bits;  // fixup_utoprim: clustered bad zones, [N3][fix_nwords] bitmap
    int *fix_pcount;               // [N3] marked zones per plane
    unsigned fix_nwords;
    int *fix_err;
    FontList font;
    SeedList seeds;  // guess mode 2; allocated by harm_set_guess
    bool uniform;  // N3 % 64 == 0: a wave's lanes share one (i,j) column
    int halted;    // error code of the step that stopped the last batch (11, 14), else 0
    int eos_kind;  // 0 Gamma-law, 1 tabulated (the table's device arrays belong to the caller), 2 cache
    // the cache EOS (eos_kind 2): parameters, the zones' last temperatures, the
    // caches' value pointers, the solve count and, per direction, whether the
    // next flux sweep may meet unset temperatures (then it runs in j order)
    CeosParams ceos;
    double *lastT;
    const double **ceos_vals;
    std::vector<double *> ceos_ycoef;  // per cache: the splines along T of e and p
    double **ceos_yco;                 // device array of ceos_ycoef
    double *ceos_fe;                   // [4][nz]: each face's p_l, cs2_l, p_r, cs2_r (k_face_eos)
    bool ceos_vanish;                  // the units make dp/drho, dp/du repeat calls (ceos_vanishing_steps)
    unsigned long long *ceos_solves;
    bool ceos_chain[3];
    int *ceos_hint;                    // the rho and T axes' bracket-start tables (CeosHint)
    // learned Newton seed and training-pair capture
    int guess_mode;
    double gate_tau = 0.1;  // guess mode 2
    const harm_mlp *mlp;
    double *cap_feat, *cap_targ;
    uint8_t *cap_valid;
    size_t cap, cap_count;
    int cap_stride, cap_phase;
    bool timing;
    struct Ev {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Ev> ev_live, ev_pool;
    double ms[HARM_NKERNELS];
    long long launches[HARM_NKERNELS];
    double mass_start;  // main.c:173-176
    bool flux_valid;    // F1 holds the fluxes of a completed full step (diag_flux)
    struct Tracers {
        bool on;
        TrArgs a;                   // device arrays and parameters
        harm_tracer_rec *dump;      // dump_tracers output, a.cap_ev records
        unsigned *bcount;           // per-block live counts
        unsigned long long *base;   // per-block output offsets
    } tr;
};

namespace {

int ev_begin(harm_ctx *c, int cls, harm_ctx::Ev &e)
{
    if (!c->timing) return 0;
    if (c->ev_pool.empty()) {
        harm_ctx::Ev n;
        HIPCHK(hipEventCreate(&n.a));
        HIPCHK(hipEventCreate(&n.b));
        c->ev_pool.push_back(n);
    }
    e = c->ev_pool.back();
    c->ev_pool.pop_back();
    e.cls = cls;
    HIPCHK(hipEventRecord(e.a, c->stream));
    return 0;
}
int ev_end(harm_ctx *c, harm_ctx::Ev &e)
{
    if (!c->timing) return 0;
    HIPCHK(hipEventRecord(e.b, c->stream));
    c->ev_live.push_back(e);
    return 0;
}
int ev_collect(harm_ctx *c)
{
    if (c->ev_live.empty()) return 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto &e : c->ev_live) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
        c->ms[e.cls] += ms;
        c->launches[e.cls]++;
        c->ev_pool.push_back(e);
    }
    c->ev_live.clear();
    return 0;
}

#define LAUNCH(cls, ...)                                       \
    do {                                                       \
        harm_ctx::Ev ev_;                                      \
        int rc_ = ev_begin(c, cls, ev_);                       \
        if (rc_) return rc_;                                   \
        __VA_ARGS__;                                           \
        HIPCHK(hipGetLastError());                             \
        rc_ = ev_end(c, ev_);                                  \
        if (rc_) return rc_;                                   \
    } while (0)

template <int DIR, bool UNI, bool TAB, int LIM = -1, bool ROWS = false>
void flux_kernel(harm_ctx *c, dim3 grid, const double *pb, double *F, int fullstep)
{
    hipLaunchKernelGGL((k_flux<DIR, UNI, TAB, LIM, ROWS>), grid, dim3(TPB), 0, c->stream, c->g, pb, F, c->dtpart,
                       fullstep, c->cnt);
}

template <int DIR>
void flux_kernel_any(harm_ctx *c, dim3 &grid, const double *pb, double *F, int fullstep)
{
    const bool tab = c->eos_kind == 1, mc = c->g.lim == LIM_MC;
    if (c->g.lim == LIM_PPM) {
        // wave-uniform grids share the zones' slopes and parabolas through LDS
        // (strips along X1 / X2, whole rows along phi where they divide the
        // block); otherwise six zones per face, one face per thread
        if (c->uniform && DIR != 3) {
            const DGrid &g = c->g;
            const size_t nf = (DIR == 1) ? (size_t)(g.ie + 2 - g.is) : (size_t)(g.je + 2 - g.js);
            const size_t no = (DIR == 1) ? (size_t)(g.je + 3 - g.js) : (size_t)(g.ie + 3 - g.is);
            const size_t nb = (nf + STRIP - 1) / STRIP * no * (size_t)(g.N3 / 64);
            grid = dim3((unsigned)((nb + 7) / 8 * 8));
            constexpr int D = (DIR == 3) ? 1 : DIR;
            if (tab)
                hipLaunchKernelGGL((k_flux_strip_ppm<D, true>), grid, dim3(STRIP_TPB), 0, c->stream, c->g, pb, F,
                                   c->dtpart, fullstep, c->cnt);
            else
                hipLaunchKernelGGL((k_flux_strip_ppm<D, false>), grid, dim3(STRIP_TPB), 0, c->stream, c->g, pb, F,
                                   c->dtpart, fullstep, c->cnt);
        } else if (c->uniform && DIR == 3 && TPB % c->g.N3 == 0) {
            if (tab)
                flux_kernel<DIR, true, true, LIM_PPM, DIR == 3>(c, grid, pb, F, fullstep);
            else
                flux_kernel<DIR, true, false, LIM_PPM, DIR == 3>(c, grid, pb, F, fullstep);
        } else if (c->uniform) {
            if (tab)
                flux_kernel<DIR, true, true, LIM_PPM>(c, grid, pb, F, fullstep);
            else
                flux_kernel<DIR, true, false, LIM_PPM>(c, grid, pb, F, fullstep);
        } else {
            if (tab)
                flux_kernel<DIR, false, true, LIM_PPM>(c, grid, pb, F, fullstep);
            else
                flux_kernel<DIR, false, false, LIM_PPM>(c, grid, pb, F, fullstep);
        }
    } else if (c->uniform && DIR != 3) {
        // strips of STRIP faces along DIR x one other index x 64 phi zones
        const DGrid &g = c->g;
        const size_t nf = (DIR == 1) ? (size_t)(g.ie + 2 - g.is) : (size_t)(g.je + 2 - g.js);
        const size_t no = (DIR == 1) ? (size_t)(g.je + 3 - g.js) : (size_t)(g.ie + 3 - g.is);
        const size_t nb = (nf + STRIP - 1) / STRIP * no * (size_t)(g.N3 / 64);
        grid = dim3((unsigned)((nb + 7) / 8 * 8));
        constexpr int D = (DIR == 3) ? 1 : DIR;
        if (tab)
            hipLaunchKernelGGL((k_flux_strip<D, true, -1>), grid, dim3(STRIP_TPB), 0, c->stream, c->g, pb, F,
                               c->dtpart, fullstep, c->cnt);
        else if (mc)
            hipLaunchKernelGGL((k_flux_strip<D, false, 0>), grid, dim3(STRIP_TPB), 0, c->stream, c->g, pb, F,
                               c->dtpart, fullstep, c->cnt);
        else
            hipLaunchKernelGGL((k_flux_strip<D, false, -1>), grid, dim3(STRIP_TPB), 0, c->stream, c->g, pb, F,
                               c->dtpart, fullstep, c->cnt);
    } else if (c->uniform && TPB % c->g.N3 == 0) {
        // DIR 3, whole phi rows per block
        if (tab)
            flux_kernel<DIR, true, true, -1, DIR == 3>(c, grid, pb, F, fullstep);
        else if (mc)
            flux_kernel<DIR, true, false, 0, DIR == 3>(c, grid, pb, F, fullstep);
        else
            flux_kernel<DIR, true, false, -1, DIR == 3>(c, grid, pb, F, fullstep);
    } else if (c->uniform) {
        if (tab)
            flux_kernel<DIR, true, true>(c, grid, pb, F, fullstep);
        else if (mc)
            flux_kernel<DIR, true, false, 0>(c, grid, pb, F, fullstep);
        else
            flux_kernel<DIR, true, false>(c, grid, pb, F, fullstep);
    } else {
        if (tab)
            flux_kernel<DIR, false, true>(c, grid, pb, F, fullstep);
        else
            flux_kernel<DIR, false, false>(c, grid, pb, F, fullstep);
    }
}

int flux_launch(harm_ctx *c, int dir, const double *pb, double *F, int fullstep)
{
    const DGrid &g = c->g;
    const int i0 = (dir == 1) ? g.is : g.is - 1;
    const int j0 = (dir == 2) ? g.js : g.js - 1;
    const size_t total = (size_t)(g.ie + 2 - i0) * (g.je + 2 - j0) * g.N3;
    dim3 grid(nblocks8(total));  // a variant may launch another grid and report it here
    if (c->eos_kind == 2) {
        if (g.lim < 0 || g.lim > 2)
            return set_err("the cache EOS step reconstructs with MC, van Leer or minmod (lim 0-2)", 4);
        const int chain = c->ceos_chain[dir - 1] ? 1 : 0;
        const size_t nc = (size_t)(g.ie + 2 - i0) * g.N3;
        const dim3 ge(nblocks(chain ? nc : total));
        grid = dim3(nblocks(total));
#define CEOS_FLUX(D_, K_)                                                                                          \
    LAUNCH(K_, if (c->ceos_vanish) hipLaunchKernelGGL((k_face_eos<D_, true>), ge, dim3(TPB), 0, c->stream, g, c->ceos, \
                                                      c->lastT, pb, c->ceos_fe, chain, c->ceos_solves);            \
               else hipLaunchKernelGGL((k_face_eos<D_, false>), ge, dim3(TPB), 0, c->stream, g, c->ceos, c->lastT, pb, \
                                       c->ceos_fe, chain, c->ceos_solves);                                         \
           hipLaunchKernelGGL(k_flux_ceos<D_>, grid, dim3(TPB), 0, c->stream, g, c->ceos_fe, pb, F, c->dtpart,     \
                              fullstep, c->cnt))
        if (dir == 1)
            CEOS_FLUX(1, HARM_K_FLUX1);
        else if (dir == 2)
            CEOS_FLUX(2, HARM_K_FLUX2);
        else
            CEOS_FLUX(3, HARM_K_FLUX3);
#undef CEOS_FLUX
        c->ceos_chain[dir - 1] = false;
    } else if (dir == 1)
        LAUNCH(HARM_K_FLUX1, flux_kernel_any<1>(c, grid, pb, F, fullstep));
    else if (dir == 2)
        LAUNCH(HARM_K_FLUX2, flux_kernel_any<2>(c, grid, pb, F, fullstep));
    else
        LAUNCH(HARM_K_FLUX3, flux_kernel_any<3>(c, grid, pb, F, fullstep));
    if (fullstep) {
        hipLaunchKernelGGL(k_min_partials, dim3(RED_BLOCKS), dim3(256), 0, c->stream, c->dtpart, (size_t)grid.x,
                           c->ndt_slot + (dir - 1), c->g.halt);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

// emf = false: the update that follows forms the corner EMFs itself (flux_div32)
int fluxes_and_emf(harm_ctx *c, const double *pb, int fullstep, bool emf = true)
{
    const DGrid &g = c->g;
    int rc;
    if ((rc = flux_launch(c, 1, pb, c->F1, fullstep))) return rc;
    if ((rc = flux_launch(c, 2, pb, c->F2, fullstep))) return rc;
    if (g.N3 > 2 && (rc = flux_launch(c, 3, pb, c->F3, fullstep))) return rc;
    if (!emf) return 0;
    const size_t ne = (size_t)(g.ie + 2 - g.is) * (g.je + 2 - g.js) * g.N3;
    LAUNCH(HARM_K_EMF, hipLaunchKernelGGL(k_emf, dim3(nblocks8(ne)), dim3(TPB), 0, c->stream, g, c->F1, c->F2, c->F3,
                                          c->e1, c->e2, c->e3));
    return 0;
}

int bound(harm_ctx *c, double *p)
{
    const DGrid &g = c->g;
    const size_t nb = ((size_t)4 * g.SN2 + (size_t)g.N1 * (4 + 2 * g.polefix)) * g.N3;
    LAUNCH(HARM_K_BOUND, hipLaunchKernelGGL(k_bound, dim3(nblocks(nb)), dim3(TPB), 0, c->stream, g, p, c->pflag,
                                            c->font.count, c->seeds.count, c->cnt));
    return 0;
}

int fixup_utoprim(harm_ctx *c, double *p)
{
    const DGrid &g = c->g;
    LAUNCH(HARM_K_FIXUTOPRIM,
           hipLaunchKernelGGL(k_fixutoprim_scan, dim3(nblocks((g.nz + 7) / 8)), dim3(TPB), 0, c->stream, g, p, c->pflag,
                              c->pflag2, c->fix_bits, c->fix_pcount, c->fix_nwords, c->cnt);
           hipLaunchKernelGGL(k_fixutoprim_serial, dim3(g.N3), dim3(64), 0, c->stream, g, p, c->pflag2, c->fix_bits,
                              c->fix_pcount, c->fix_nwords, c->cnt));
    std::swap(c->pflag, c->pflag2);
    return 0;
}

// advance() (step_ch.c:153-318) with fixup() folded into the update and the
// Font-fix pass (step_ch.c:101-103, 112-114).  The Font-fix and seed lists it
// fills are emptied by the next bound().
// advance() with the cache EOS: the fluxes, their divergence, then the zone
// update (k_update_ceos) with its fixup.
int advance_ceos(harm_ctx *c, double *pi, double *pb, double *pf, double dtfac, int fullstep)
{
    const DGrid &g = c->g;
    int rc;
    const size_t na = (size_t)(g.ie - g.is + 1) * (g.je - g.js + 1) * g.N3;
    const DivArgs dv = {c->F1, c->F2, c->F3, c->e1, c->e2, c->e3};
    if ((rc = fluxes_and_emf(c, pb, fullstep, true))) return rc;
    LAUNCH(HARM_K_DIV, hipLaunchKernelGGL(k_div, dim3(nblocks8(na)), dim3(TPB), 0, c->stream, g, dv, c->S));
#define CEOS_UPD(V_)                                                                                             \
    LAUNCH(HARM_K_UPDATE, hipLaunchKernelGGL(k_update_ceos<V_>, dim3(nblocks(na)), dim3(TPB), 0, c->stream, g,     \
                                             c->ceos, c->lastT, pi, pb, pf, c->pflag, c->S, c->tstate, dtfac,      \
                                             (pi == pb) ? 1 : 0, c->cpart, c->ceos_solves))
    if (c->ceos_vanish)
        CEOS_UPD(true);
    else
        CEOS_UPD(false);
#undef CEOS_UPD
    CntMap map = {8, {HARM_C2P_CALLS, HARM_NEWTON_ITERS, HARM_NEWTON_KCONV, HARM_FAIL_2D, HARM_FIXUP_GAMMA,
                      HARM_FLOOR_ZONES, HARM_GAMMA_CAP, HARM_FATAL}};
    hipLaunchKernelGGL(k_sum_partials, dim3(RED_BLOCKS, 8), dim3(256), 0, c->stream, c->cpart, (size_t)nblocks(na), map,
                       c->cnt, g.halt);
    HIPCHK(hipGetLastError());
    return 0;
}

int advance(harm_ctx *c, double *pi, double *pb, double *pf, double dtfac, int fullstep)
{
    if (c->eos_kind == 2) return advance_ceos(c, pi, pb, pf, dtfac, fullstep);
    const DGrid &g = c->g;
    int rc;
    const size_t na = (size_t)(g.ie - g.is + 1) * (g.je - g.js + 1) * g.N3;
    const int same = (pi == pb) ? 1 : 0;
    UpdateExtra ex;
    std::memset(&ex, 0, sizeof(ex));
    int mode = UPD_PLAIN;
    if (c->eos_kind == 1) {
        // the tabulated-EOS path has its own (5D) solver: no learned seed, no capture
    } else if (c->guess_mode == 1 && c->mlp) {
        mode = UPD_LEARNED;
        ex.mlp = harm_mlp_view(c->mlp);
    } else if (c->guess_mode == 2 && c->mlp && c->seeds.count) {
        mode = UPD_GATED;
        ex.mlp = harm_mlp_view(c->mlp);
        ex.gate_tau = c->gate_tau;
        ex.seeds = c->seeds;
    } else if (c->cap_feat && c->cap_count < c->cap) {
        mode = UPD_CAPTURE;
        ex.cap_feat = c->cap_feat;
        ex.cap_targ = c->cap_targ;
        ex.cap_valid = c->cap_valid;
        ex.cap = c->cap;
        ex.cap_base = c->cap_count;
        ex.stride = c->cap_stride;
        ex.phase = c->cap_phase;
        const size_t ph = (size_t)c->cap_phase;
        const size_t taken = (na > ph) ? (na - ph + c->cap_stride - 1) / c->cap_stride : 0;
        c->cap_count = std::min(c->cap, c->cap_count + taken);
        c->cap_phase = (int)((ph + 7919) % (size_t)c->cap_stride);
    }
    const DivArgs dv = {c->F1, c->F2, c->F3, c->e1, c->e2, c->e3};
    // the plain wave-uniform update forms the EMFs and the divergence itself; the
    // others read S, which k_div forms from the EMF arrays
    const bool fused = c->uniform && (mode != UPD_CAPTURE || c->eos_kind == 1) && c->nz < OFF32_MAX_ZONES;
    if ((rc = fluxes_and_emf(c, pb, fullstep, !fused))) return rc;
    if (!fused)
        LAUNCH(HARM_K_DIV, hipLaunchKernelGGL(k_div, dim3(nblocks8(na)), dim3(TPB), 0, c->stream, g, dv, c->S));
#define UPD_LAUNCH(U_, M_, T_, F_)                                                                                  \
    LAUNCH(HARM_K_UPDATE, hipLaunchKernelGGL((k_update<U_, M_, T_, F_>), dim3(nblocks(na)), dim3(TPB), 0, c->stream, \
                                             g, pi, pb, pf, c->pflag, c->S, dv, c->tstate, dtfac, same, c->font,   \
                                             c->cpart, ex))
#define GATED_LAUNCH(U_, F_) UPD_LAUNCH(U_, UPD_GATED, false, F_)
    if (c->eos_kind == 1) {
        if (fused)
            UPD_LAUNCH(true, UPD_PLAIN, true, true);
        else if (c->uniform)
            UPD_LAUNCH(true, UPD_PLAIN, true, false);
        else
            UPD_LAUNCH(false, UPD_PLAIN, true, false);
    } else if (c->uniform) {
        if (mode == UPD_PLAIN && fused)
            UPD_LAUNCH(true, UPD_PLAIN, false, true);
        else if (mode == UPD_PLAIN)
            UPD_LAUNCH(true, UPD_PLAIN, false, false);
        else if (mode == UPD_CAPTURE)
            UPD_LAUNCH(true, UPD_CAPTURE, false, false);
        else if (mode == UPD_GATED && fused)
            GATED_LAUNCH(true, true);
        else if (mode == UPD_GATED)
            GATED_LAUNCH(true, false);
        else if (fused)
            UPD_LAUNCH(true, UPD_LEARNED, false, true);
        else
            UPD_LAUNCH(true, UPD_LEARNED, false, false);
    } else {
        if (mode == UPD_PLAIN)
            UPD_LAUNCH(false, UPD_PLAIN, false, false);
        else if (mode == UPD_CAPTURE)
            UPD_LAUNCH(false, UPD_CAPTURE, false, false);
        else if (mode == UPD_GATED)
            GATED_LAUNCH(false, false);
        else
            UPD_LAUNCH(false, UPD_LEARNED, false, false);
    }
#undef UPD_LAUNCH
#undef GATED_LAUNCH
    {
        CntMap map = {8, {HARM_C2P_CALLS, HARM_NEWTON_ITERS, HARM_NEWTON_KCONV, HARM_FAIL_2D, HARM_FIXUP_GAMMA,
                          HARM_FLOOR_ZONES, HARM_GAMMA_CAP, HARM_FATAL}};
        hipLaunchKernelGGL(k_sum_partials, dim3(RED_BLOCKS, 8), dim3(256), 0, c->stream, c->cpart, (size_t)nblocks(na), map,
                           c->cnt, g.halt);
        HIPCHK(hipGetLastError());
    }
    if (mode == UPD_GATED)
        LAUNCH(HARM_K_MLP, hipLaunchKernelGGL(k_seedfix, dim3(SEEDFIX_BLOCKS), dim3(TPB), 0, c->stream, g, pi, pf,
                                              c->pflag, c->seeds, c->font, ex.mlp, c->cnt));
    if (g.font_fix)
        LAUNCH(HARM_K_FONTFIX, hipLaunchKernelGGL(k_fontfix, dim3(FONTFIX_BLOCKS), dim3(TPB), 0, c->stream, g, pf, c->pflag,
                                                  c->font, c->cnt));
    return 0;
}

// advance() and fixup(), then bound_prim(), fixup_utoprim(), bound_prim()
// (step_ch.c:103-117).
int half_step(harm_ctx *c, double *pi, double *pb, double *pf, double dtfac, int fullstep)
{
    int rc;
    if ((rc = advance(c, pi, pb, pf, dtfac, fullstep))) return rc;
    if ((rc = bound(c, pf))) return rc;
    if ((rc = fixup_utoprim(c, pf))) return rc;
    if ((rc = bound(c, pf))) return rc;
    return 0;
}

int one_step(harm_ctx *c)
{
    int rc;
    const DGrid &g = c->g;
    if ((rc = half_step(c, c->p, c->p, c->ph, 0.5, 0))) return rc;
    if ((rc = half_step(c, c->p, c->ph, c->p, 1.0, 1))) return rc;
    c->flux_valid = true;
    if (c->tr.on) {
        const unsigned nb = nblocks((size_t)c->tr.a.n);
        if (c->eos_kind == 1)
            LAUNCH(HARM_K_TRACERS, hipLaunchKernelGGL(k_tracer_update<true>, dim3(nb), dim3(TPB), 0, c->stream, g,
                                                      c->tr.a, c->p, c->tstate, c->cnt));
        else
            LAUNCH(HARM_K_TRACERS, hipLaunchKernelGGL(k_tracer_update<false>, dim3(nb), dim3(TPB), 0, c->stream, g,
                                                      c->tr.a, c->p, c->tstate, c->cnt));
    }
    LAUNCH(HARM_K_STEPEND,
           hipLaunchKernelGGL(k_mdot_partials, dim3(g.N3), dim3(TPB), 0, c->stream, g, c->p, c->part);
           hipLaunchKernelGGL(k_stepend, dim3(1), dim3(1), 0, c->stream, g, c->tstate, c->ndt_slot, c->part, g.N3,
                              c->cnt));
    return 0;
}

const char *halt_message(int code)
{
    return code == 11 ? "timestep too small (dt < 1e-9, step_ch.c:126-129); the batch stopped after that step"
                      : "a fail() event (diag.c:381-397: gamma, vchar or metric failure) in the last step; the batch "
                        "stopped after that step";
}

int check_flags(harm_ctx *c)
{
    double ts[5];
    int fe = 0;
    HIPCHK(hipMemcpyAsync(ts, c->tstate, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&fe, c->fix_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (ts[4] == 11. || ts[4] == 14.) {
        c->halted = (int)ts[4];
        return set_err(halt_message(c->halted), c->halted);
    }
    if (fe & 1) return set_err("more failed inversions in one half-step than the Font-fix list holds", 13);
    return 0;
}

// Clears a stopped batch's error (the caller resets the clock or the state).
int clear_halt(harm_ctx *c)
{
    const double z = 0.;
    HIPCHK(hipMemcpyAsync(c->tstate + 4, &z, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->halted = 0;
    return 0;
}

template <class T>
int dalloc(T **p, size_t n)
{
    HIPCHK(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    HIPCHK(hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    // hipMemset runs on the null stream, which the library's non-blocking stream
    // does not wait for: finish it before any kernel on that stream can touch
    // the buffer (allocation happens at set-up and in occasional calls only)
    HIPCHK(hipDeviceSynchronize());
    return 0;
}

int ctx_alloc(harm_ctx *c)
{
    const harm_params &P = c->P;
    auto &hg = c->hg;
    hg.N1 = P.n1;
    hg.N2 = P.n2;
    hg.N3 = P.n3;
    hg.SN1 = P.n1 + 4;
    hg.SN2 = P.n2 + 4;
    hg.is = 2;
    hg.ie = P.n1 + 1;
    hg.js = 2;
    hg.je = P.n2 + 1;
    if (harm_host::build_geometry(P, hg, c->tab)) return set_err("metric setup failed (singular metric)", 3);
    const size_t ncol = (size_t)hg.SN1 * hg.SN2;
    c->nz = ncol * hg.N3;
    int rc;
    for (double **q : {&c->p, &c->ph, &c->F1, &c->F2, &c->F3})
        if ((rc = dalloc(q, 8 * c->nz))) return rc;
    if ((rc = dalloc(&c->e3, c->nz))) return rc;
    if ((rc = dalloc(&c->S, 8 * (size_t)hg.N1 * hg.N2 * hg.N3))) return rc;
    if (hg.N3 > 2) {
        if ((rc = dalloc(&c->e1, c->nz))) return rc;
        if ((rc = dalloc(&c->e2, c->nz))) return rc;
    }
    if ((rc = dalloc(&c->pflag, c->nz))) return rc;
    if ((rc = dalloc(&c->pflag2, c->nz))) return rc;
    // device tables: the host's GEO_BASE rows per location plus the derived rows
    std::vector<double> gx[5];
    for (int l = 0; l < 5; l++) {
        const std::vector<double> &T = c->tab.geom[l];
        gx[l].assign((size_t)GEO_ROWS * ncol, 0.);
        std::copy(T.begin(), T.begin() + (size_t)GEO_BASE * ncol, gx[l].begin());
        for (size_t col = 0; col < ncol; col++) {
            double d[GEO_DERIVED];
            geom_derived(T[16 * ncol + col], T[17 * ncol + col], T[18 * ncol + col], T[19 * ncol + col],
                         T[32 * ncol + col], d);
            for (int k = 0; k < GEO_DERIVED; k++) gx[l][(size_t)(GEO_BASE + k) * ncol + col] = d[k];
        }
        if ((rc = dalloc(&c->d_geom[l], (size_t)GEO_ROWS * ncol))) return rc;
        HIPCHK(hipMemcpy(c->d_geom[l], gx[l].data(), (size_t)GEO_ROWS * ncol * sizeof(double), hipMemcpyHostToDevice));
    }
    if ((rc = dalloc(&c->d_conn, 64 * ncol))) return rc;
    HIPCHK(hipMemcpy(c->d_conn, c->tab.conn.data(), 64 * ncol * sizeof(double), hipMemcpyHostToDevice));
    if ((rc = dalloc(&c->d_floors, 2 * ncol))) return rc;
    HIPCHK(hipMemcpy(c->d_floors, c->tab.floors.data(), 2 * ncol * sizeof(double), hipMemcpyHostToDevice));
    if ((rc = dalloc(&c->d_katm, hg.SN1))) return rc;
    HIPCHK(hipMemcpy(c->d_katm, c->tab.katm.data(), hg.SN1 * sizeof(double), hipMemcpyHostToDevice));
    {
        // column-major copies for the wave-uniform kernels
        std::vector<double> cb((size_t)CENT_STRIDE * ncol, 0.), fb((size_t)3 * FACE_STRIDE * ncol, 0.);
        const int floc[3] = {FACE1, FACE2, FACE3};
        for (size_t col = 0; col < ncol; col++) {
            for (int v = 0; v < GEO_BASE; v++) cb[col * CENT_STRIDE + v] = gx[CENT][(size_t)v * ncol + col];
            for (int v = 0; v < GEO_DERIVED; v++)
                cb[col * CENT_STRIDE + CENT_DERIVED + v] = gx[CENT][(size_t)(GEO_BASE + v) * ncol + col];
            cb[col * CENT_STRIDE + 33] = c->tab.floors[col];
            cb[col * CENT_STRIDE + 34] = c->tab.floors[ncol + col];
            for (int v = 0; v < 64; v++) cb[col * CENT_STRIDE + CENT_CONN + v] = c->tab.conn[(size_t)v * ncol + col];
            for (int d = 0; d < 3; d++)
                for (int v = 0; v < GEO_ROWS; v++)
                    fb[((size_t)d * ncol + col) * FACE_STRIDE + v] = gx[floc[d]][(size_t)v * ncol + col];
        }
        if ((rc = dalloc(&c->d_cblk, cb.size()))) return rc;
        HIPCHK(hipMemcpy(c->d_cblk, cb.data(), cb.size() * sizeof(double), hipMemcpyHostToDevice));
        if ((rc = dalloc(&c->d_fblk, fb.size()))) return rc;
        HIPCHK(hipMemcpy(c->d_fblk, fb.data(), fb.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if ((rc = dalloc(&c->tstate, 8))) return rc;
    if ((rc = dalloc(&c->ndt_slot, 3))) return rc;
    if ((rc = dalloc(&c->dtpart, (size_t)nblocks8(c->nz)))) return rc;
    if ((rc = dalloc(&c->cpart, 8 * (size_t)nblocks(c->nz)))) return rc;
    if ((rc = dalloc(&c->cnt, HARM_NCOUNTERS))) return rc;
    c->npart = std::max<size_t>(DIAG_NPART * nblocks(c->nz) + 16, 2 * (size_t)hg.N3 + 16);
    if ((rc = dalloc(&c->part, c->npart))) return rc;
    // one slot per active zone: every inversion of a half-step may fail
    c->font.cap = (int)((size_t)hg.N1 * hg.N2 * hg.N3);
    if ((rc = dalloc(&c->font.count, 1))) return rc;
    if ((rc = dalloc(&c->font.zone, (size_t)c->font.cap))) return rc;
    if ((rc = dalloc(&c->font.U, 8 * (size_t)c->font.cap))) return rc;
    c->uniform = (hg.N3 % 64) == 0;
    c->fix_nwords = (unsigned)((ncol + 63) / 64);
    if ((rc = dalloc(&c->fix_bits, (size_t)hg.N3 * c->fix_nwords))) return rc;
    if ((rc = dalloc(&c->fix_pcount, (size_t)hg.N3))) return rc;
    if ((rc = dalloc(&c->fix_err, 1))) return rc;
    c->font.overflow = c->fix_err;  // bit 0: Font-fix list (cannot fill: one slot per active zone)
    unsigned long long big[3];
    const double b = 1.e9;
    for (int k = 0; k < 3; k++) std::memcpy(&big[k], &b, 8);
    HIPCHK(hipMemcpy(c->ndt_slot, big, sizeof(big), hipMemcpyHostToDevice));
    double ts[8] = {0., P.dt0, 0., 0., 0., 0., 0., 0.};
    HIPCHK(hipMemcpy(c->tstate, ts, sizeof(ts), hipMemcpyHostToDevice));

    DGrid &g = c->g;
    g.N1 = hg.N1;
    g.N2 = hg.N2;
    g.N3 = hg.N3;
    g.SN1 = hg.SN1;
    g.SN2 = hg.SN2;
    g.is = hg.is;
    g.ie = hg.ie;
    g.js = hg.js;
    g.je = hg.je;
    g.lim = P.lim;
    g.font_fix = P.font_fix;
    g.polefix = (P.n3 > 1 && 2 < P.n2 / 2) ? 2 : 0;  // decs.h:91-95, bounds.c:149
    g.ncol = ncol;
    g.nz = c->nz;
    g.dx1 = hg.dx[1];
    g.dx2 = hg.dx[2];
    g.dx3 = hg.dx[3];
    g.rdx1 = 1. / g.dx1;
    g.rdx2 = 1. / g.dx2;
    g.rdx3 = 1. / g.dx3;
    g.gam = P.gam;
    g.cour = P.cour;
    g.hllf = P.hllf;
    g.laxf = P.laxf;
    g.defcon = P.defcon;
    g.tf = P.tf;
    for (int l = 0; l < 5; l++) g.geom[l] = c->d_geom[l];
    g.conn = c->d_conn;
    g.floors = c->d_floors;
    g.katm = c->d_katm;
    g.cblk = c->d_cblk;
    for (int d = 0; d < 3; d++) g.fblk[d] = c->d_fblk + (size_t)d * FACE_STRIDE * ncol;
    return 0;
}

void ctx_free(harm_ctx *c)
{
    for (void *q : {(void *)c->p, (void *)c->ph, (void *)c->F1, (void *)c->F2, (void *)c->F3, (void *)c->e1,
                    (void *)c->e2, (void *)c->e3, (void *)c->S, (void *)c->pflag, (void *)c->pflag2, (void *)c->d_conn,
                    (void *)c->d_floors, (void *)c->d_katm, (void *)c->tstate, (void *)c->ndt_slot, (void *)c->dtpart, (void *)c->cpart, (void *)c->cnt,
                    (void *)c->part, (void *)c->fix_bits, (void *)c->fix_pcount, (void *)c->fix_err, (void *)c->font.count, (void *)c->font.zone, (void *)c->font.U,
                    (void *)c->d_cblk, (void *)c->d_fblk, (void *)c->seeds.count, (void *)c->seeds.zone,
                    (void *)c->seeds.U, (void *)c->lastT, (void *)c->ceos_vals, (void *)c->ceos_solves,
                    (void *)c->ceos_yco, (void *)c->ceos_fe, (void *)c->ceos_hint})
        if (q) (void)hipFree(q);
    for (double *q : c->ceos_ycoef) (void)hipFree(q);
    for (int l = 0; l < 5; l++)
        if (c->d_geom[l]) (void)hipFree(c->d_geom[l]);
    for (auto &e : c->ev_pool) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto &e : c->ev_live) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

// Canonical layout <-> device layout: identical (SoA [v][i][j][h]), so the copies are flat.
int copy_state(harm_ctx *c, double *host_prim, int8_t *host_flag, bool to_host)
{
    const size_t n = 8 * c->nz;
    if (to_host) {
        if (host_prim) HIPCHK(hipMemcpyAsync(host_prim, c->p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (host_flag) HIPCHK(hipMemcpyAsync(host_flag, c->pflag, c->nz, hipMemcpyDeviceToHost, c->stream));
    } else {
        if (host_prim) HIPCHK(hipMemcpyAsync(c->p, host_prim, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        if (host_flag) HIPCHK(hipMemcpyAsync(c->pflag, host_flag, c->nz, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// The bracket-start table of an increasing axis (CeosHint): for each key
// between those of the axis's ends, the bracket of the smallest double with
// that key, with the shift that keeps the table within 8192 entries (at the
// cache grids' spacing a key spans well under one grid interval).  An axis
// with a value that is not positive and finite gets none (nk = 0).
CeosHint ceos_hint_table(const std::vector<double> &ax, std::vector<int> &tab)
{
    CeosHint h{};
    tab.clear();
    for (double a : ax)
        if (!(a > 0.) || !std::isfinite(a)) return h;
    auto key = [](double x, int s) {
        unsigned long long b;
        std::memcpy(&b, &x, sizeof b);
        return b >> s;
    };
    int s = 40;
    while (key(ax.back(), s) - key(ax.front(), s) + 1 > 8192) s++;
    const unsigned long long k0 = key(ax.front(), s), k1 = key(ax.back(), s);
    tab.resize((size_t)(k1 - k0 + 1));
    for (size_t k = 0; k < tab.size(); k++) {
        const unsigned long long b = (k0 + k) << s;
        double x;
        std::memcpy(&x, &b, sizeof x);
        tab[k] = cache_bracket((int)ax.size(), ax.data(), x);
    }
    h.k0 = k0;
    h.nk = (int)tab.size();
    h.shift = s;
    return h;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
int harm_host::set_error(const std::string &msg, int code) { return set_err(msg, code); }

extern "C" {

const char *harm_last_error(void) { return g_err.c_str(); }

harm_ctx *harm_create(const harm_params *p, int device)
{
    g_err.clear();
    if (!p) {
        set_err("null params");
        return nullptr;
    }
    if (p->n1 < 4 || p->n2 < 8 || p->n3 < 1 || p->n3 == 2 || (p->n3 > 1 && (p->n3 % 2) != 0)) {
        set_err("unsupported grid: need n1 >= 4, n2 >= 8, n3 == 1 or even n3 >= 4 (mpi.c:376)");
        return nullptr;
    }
    if ((double)(p->n1 + 4) * (p->n2 + 4) * p->n3 >= 2147483648.0) {
        set_err("grid too large: (n1+4)*(n2+4)*n3 must stay below 2^31 zones per device (32-bit zone indices)");
        return nullptr;
    }
    if (p->lim < 0 || p->lim > 3) {
        set_err("unknown slope limiter");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_err("no HIP device available: libharmcool.so runs only on a gfx950 GPU");
        return nullptr;
    }
    if (device >= 0 && hipSetDevice(device) != hipSuccess) {
        set_err("hipSetDevice failed");
        return nullptr;
    }
    harm_ctx *c = new harm_ctx();
    std::memset(c->ms, 0, sizeof(c->ms));
    std::memset(c->launches, 0, sizeof(c->launches));
    c->P = *p;
    if (c->P.Rin <= 0.) c->P.Rin = 0.98 * (1. + sqrt(1. - c->P.a * c->P.a));  // init.c:125
    (void)hipGetDevice(&c->device);
    if (scratch_resolve(c->device)) {
        delete c;
        return nullptr;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_err("hipStreamCreate failed");
        delete c;
        return nullptr;
    }
    if (ctx_alloc(c)) {
        std::string m = g_err;
        ctx_free(c);
        delete c;
        g_err = m;
        return nullptr;
    }
    return c;
}

namespace {
void tracers_free(harm_ctx *c)
{
    auto &t = c->tr;
    for (void *q : {(void *)t.a.X, (void *)t.a.state, (void *)t.a.ev, (void *)t.a.nev, (void *)t.dump,
                    (void *)t.bcount, (void *)t.base})
        if (q) (void)hipFree(q);
    std::memset(&t, 0, sizeof(t));
}
}  // namespace

void harm_destroy(harm_ctx *c)
{
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    tracers_free(c);
    ctx_free(c);
    delete c;
}

namespace {
int diag_volume(harm_ctx *c, double *sums, double &divbmax, double &key);
}

int harm_init_torus(harm_ctx *c)
{
    if (!c) return set_err("null context");
    const DGrid &g = c->g;
    std::vector<double> tor;
    double rhomax, umax;
    harm_host::build_torus2d(c->P, c->hg, c->tab, tor, rhomax, umax);
    umax /= rhomax;
    double *d_tor = nullptr;
    int rc;
    if ((rc = dalloc(&d_tor, tor.size()))) return rc;
    HIPCHK(hipMemcpy(d_tor, tor.data(), tor.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(c->p, 0, 8 * c->nz * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(c->pflag, 0, c->nz, c->stream));
    const size_t na = (size_t)(g.ie - g.is + 1) * (g.je - g.js + 1) * g.N3;
    uint4 seed = make_uint4(c->P.seed[0], c->P.seed[1], c->P.seed[2], c->P.seed[3]);
    hipLaunchKernelGGL(k_init_expand, dim3(nblocks(na)), dim3(TPB), 0, c->stream, g, c->p, d_tor, rhomax,
                       c->P.pert_amp, seed);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_fixup, dim3(nblocks(na)), dim3(TPB), 0, c->stream, g, c->p, c->pflag, nullptr);
    if ((rc = bound(c, c->p))) return rc;
    const size_t nA = (size_t)(g.ie + 2 - g.is) * (g.je + 2 - g.js) * g.N3;
    HIPCHK(hipMemsetAsync(c->e3, 0, c->nz * sizeof(double), c->stream));
    hipLaunchKernelGGL(k_init_A, dim3(nblocks(nA)), dim3(TPB), 0, c->stream, g, c->p, c->e3);
    const unsigned nb = nblocks(na);
    hipLaunchKernelGGL(k_init_B, dim3(nb), dim3(TPB), 0, c->stream, g, c->p, c->e3, 1.0, 0, c->part);
    HIPCHK(hipGetLastError());
    std::vector<double> part(nb);
    HIPCHK(hipMemcpyAsync(part.data(), c->part, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double bsq_max = 0.;
    for (double v : part)
        if (v > bsq_max) bsq_max = v;
    const double beta_act = (c->P.gam - 1.) * umax / (0.5 * bsq_max);  // init.c:391
    const double norm = sqrt(beta_act / c->P.beta);
    hipLaunchKernelGGL(k_init_B, dim3(nb), dim3(TPB), 0, c->stream, g, c->p, c->e3, norm, 1, c->part);
    hipLaunchKernelGGL(k_fixup, dim3(nblocks(na)), dim3(TPB), 0, c->stream, g, c->p, c->pflag, nullptr);
    HIPCHK(hipGetLastError());
    if ((rc = bound(c, c->p))) return rc;
    double ts[8] = {0., c->P.dt0, 0., 0., 0., 0., 0., 0.};
    HIPCHK(hipMemcpyAsync(c->tstate, ts, sizeof(ts), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->halted = 0;
    HIPCHK(hipFree(d_tor));
    // main.c:173-176
    double sums[5], dv, key;
    if ((rc = diag_volume(c, sums, dv, key))) return rc;
    c->mass_start = sums[3];
    c->flux_valid = false;
    return 0;
}

int harm_step(harm_ctx *c, int nsteps)
{
    if (!c) return set_err("null context");
    if (c->halted) return set_err(halt_message(c->halted), c->halted);
    if (nsteps <= 0) return 0;
    // the step kernels read the error word and return once a step has set it
    struct Gate {
        harm_ctx *c;
        ~Gate() { c->g.halt = nullptr; }
    } gate{c};
    c->g.halt = c->tstate + 4;
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, c->stream, c->tstate, c->cnt);
    HIPCHK(hipGetLastError());
    for (int s = 0; s < nsteps; s++) {
        int rc = one_step(c);
        if (rc) return rc;
    }
    int rc = ev_collect(c);
    if (rc) return rc;
    return check_flags(c);
}

int harm_get_prims(harm_ctx *c, double *prim, int8_t *pflag) { return c ? copy_state(c, prim, pflag, true) : 1; }
int harm_set_prims(harm_ctx *c, const double *prim, const int8_t *pflag)
{
    if (!c) return 1;
    c->flux_valid = false;
    return copy_state(c, const_cast<double *>(prim), const_cast<int8_t *>(pflag), false);
}

int harm_get_time(harm_ctx *c, double *t, double *dt)
{
    double ts[2];
    HIPCHK(hipMemcpyAsync(ts, c->tstate, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *t = ts[0];
    *dt = ts[1];
    return 0;
}

int harm_set_time(harm_ctx *c, double t, double dt)
{
    double ts[2] = {t, dt};
    HIPCHK(hipMemcpyAsync(c->tstate, ts, sizeof(ts), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return clear_halt(c);
}

int harm_get_counters(harm_ctx *c, int64_t *out)
{
    unsigned long long v[HARM_NCOUNTERS];
    HIPCHK(hipMemcpyAsync(v, c->cnt, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < HARM_NCOUNTERS; k++) out[k] = (int64_t)v[k];
    return 0;
}

int harm_reset_counters(harm_ctx *c)
{
    HIPCHK(hipMemsetAsync(c->cnt, 0, HARM_NCOUNTERS * sizeof(unsigned long long), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

namespace {

// k_diag over the active zones, its block partials finished on the host in
// block order: sums[5] as k_diag documents, the largest |div B| and its rank.
int diag_volume(harm_ctx *c, double *sums, double &divbmax, double &key)
{
    const DGrid &g = c->g;
    const size_t na = (size_t)(g.ie - g.is + 1) * (g.je - g.js + 1) * g.N3;
    const unsigned nb = nblocks(na);
    if (c->eos_kind == 1)
        hipLaunchKernelGGL(k_diag<true>, dim3(nb), dim3(TPB), 0, c->stream, g, c->p, c->part);
    else
        hipLaunchKernelGGL(k_diag<false>, dim3(nb), dim3(TPB), 0, c->stream, g, c->p, c->part);
    HIPCHK(hipGetLastError());
    std::vector<double> part(DIAG_NPART * (size_t)nb);
    HIPCHK(hipMemcpyAsync(part.data(), c->part, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 5; k++) sums[k] = 0.;
    divbmax = 0.;
    key = DIAG_NO_KEY;
    for (unsigned b = 0; b < nb; b++) {
        const double *q = &part[DIAG_NPART * (size_t)b];
        for (int k = 0; k < 5; k++) sums[k] += q[k];
        if (q[5] > divbmax || (q[5] == divbmax && q[6] < key)) {
            divbmax = q[5];
            key = q[6];
        }
    }
    // the energy needs the EOS; the cache EOS's pressure would move the zones'
    // last temperatures (the reference's diag() does), so it is not formed here
    if (c->eos_kind == 2) sums[2] = NAN;
    return 0;
}

}  // namespace

int harm_diag(harm_ctx *c, double *out)
{
    const DGrid &g = c->g;
    double sums[5], dv, key;
    int rc;
    if ((rc = diag_volume(c, sums, dv, key))) return rc;
    std::vector<double> md(2 * (size_t)g.N3);
    hipLaunchKernelGGL(k_mdot_partials, dim3(g.N3), dim3(TPB), 0, c->stream, g, c->p, c->part);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(md.data(), c->part, md.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    double ts[4];
    HIPCHK(hipMemcpyAsync(ts, c->tstate, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double mi = 0., mo = 0.;
    for (int h = 0; h < g.N3; h++) {
        mi += md[2 * h];
        mo += md[2 * h + 1];
    }
    out[0] = sums[3];
    out[1] = sums[2];
    out[2] = mi;
    out[3] = mo;
    out[4] = ts[2];
    out[5] = ts[3];
    out[6] = dv;
    return 0;
}

int harm_diag_log(harm_ctx *c, harm_diag_rec *r)
{
    if (!c || !r) return set_err("null argument");
    if (c->eos_kind == 2) return set_err("not available with the cache EOS (it would move the zones' last temperatures)", 4);
    const DGrid &g = c->g;
    double sums[5], dv, key;
    int rc;
    if ((rc = diag_volume(c, sums, dv, key))) return rc;
    hipLaunchKernelGGL(k_diag_seq, dim3(1), dim3(TPB), 0, c->stream, g, c->p, c->F1, c->flux_valid ? 1 : 0, c->part);
    HIPCHK(hipGetLastError());
    double seq[5], ts[4];
    HIPCHK(hipMemcpyAsync(seq, c->part, sizeof(seq), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(ts, c->tstate, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::memset(r, 0, sizeof(*r));
    r->t = ts[0];
    r->rmed = sums[0];
    r->pp = sums[1];
    r->e = sums[2];
    r->mdot = seq[0];
    r->edot = seq[1];
    r->ldot = seq[2];
    // diag.c:198-205, 263-265: without the cooling table's Lambda, Qnu and H_d
    // only the density weight of the COOL_EOS build is known.
    r->lnu_rho_2 = (c->eos_kind == 1) ? sums[4] : 0.;
    r->mdot_in = seq[3];
    r->mdot_out = seq[4];
    r->mass_start = c->mass_start;
    r->mass_curr = sums[3];
    r->mass_in = ts[2];
    r->mass_out = ts[3];
    r->divbmax = dv;
    r->imax[0] = r->imax[1] = r->imax[2] = -1;
    if (key < DIAG_NO_KEY) {
        const size_t ni = (size_t)(g.ie - g.is + 1), nj = (size_t)(g.je - g.js + 1), k = (size_t)key;
        r->imax[1] = (int32_t)(k % nj);
        r->imax[0] = (int32_t)((k / nj) % ni);
        r->imax[2] = (int32_t)(k / nj / ni);
    }
    r->flux_valid = c->flux_valid ? 1 : 0;
    return 0;
}

int harm_ener_line(const harm_diag_rec *r, char *buf, size_t n)
{
    if (!r || !buf) return -1;
    // diag.c:317-326
    const int len =
        std::snprintf(buf, n,
                      "%10.5g %10.5g %10.5g %10.5g %15.8g %15.8g "
                      "%15.8g %15.8g %15.8g %15.8g %15.8g %15.8g %15.8g"
                      " -0 %15.8g %15.8g %15.8g %15.8g %15.8g %15.8g %15.8g\n",
                      r->t, r->rmed, r->pp, r->e, 0.0, 0.0, r->mdot, r->edot, r->ldot, r->lnu, r->lnu_rho,
                      r->lnu_rho_2, r->lnu2, r->mdot_in, r->mdot_out, r->mass_start, r->mass_curr, r->mass_in,
                      r->mass_out, (r->mass_curr - r->mass_in - r->mass_out) / r->mass_start * 100.0);
    return (len < 0 || (size_t)len >= n) ? -1 : len;
}

int harm_set_mass_budget(harm_ctx *c, double mass_start, double mass_in, double mass_out)
{
    if (!c) return set_err("null argument");
    c->mass_start = mass_start;
    double ts[4];
    HIPCHK(hipMemcpyAsync(ts, c->tstate, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    ts[2] = mass_in;
    ts[3] = mass_out;
    HIPCHK(hipMemcpyAsync(c->tstate, ts, sizeof(ts), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int harm_device_state(harm_ctx *c, double **prim, int8_t **pflag, void **stream)
{
    if (prim) *prim = c->p;
    if (pflag) *pflag = c->pflag;
    if (stream) *stream = (void *)c->stream;
    return 0;
}

int harm_set_timing(harm_ctx *c, int enable)
{
    int rc = ev_collect(c);
    if (rc) return rc;
    c->timing = enable != 0;
    std::memset(c->ms, 0, sizeof(c->ms));
    std::memset(c->launches, 0, sizeof(c->launches));
    return 0;
}

int harm_get_timing(harm_ctx *c, double *ms, int64_t *launches)
{
    int rc = ev_collect(c);
    if (rc) return rc;
    for (int k = 0; k < HARM_NKERNELS; k++) {
        if (ms) ms[k] = c->ms[k];
        if (launches) launches[k] = c->launches[k];
    }
    return 0;
}

int harm_debug_fluxes(harm_ctx *c, int fullstep, double *F1, double *F2, double *F3, double *ndt3)
{
    int rc;
    const DGrid &g = c->g;
    c->flux_valid = false;
    if ((rc = fluxes_and_emf(c, c->p, fullstep))) return rc;
    // materialize the flux_ct rewrite on the host from the EMFs (step_ch.c:541-554, 783-797)
    const size_t n = 8 * c->nz;
    std::vector<double> e1, e2, e3(c->nz);
    HIPCHK(hipMemcpyAsync(F1, c->F1, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(F2, c->F2, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (g.N3 > 2) HIPCHK(hipMemcpyAsync(F3, c->F3, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(e3.data(), c->e3, c->nz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (g.N3 > 2) {
        e1.resize(c->nz);
        e2.resize(c->nz);
        HIPCHK(hipMemcpyAsync(e1.data(), c->e1, c->nz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(e2.data(), c->e2, c->nz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    unsigned long long nd[3];
    HIPCHK(hipMemcpyAsync(nd, c->ndt_slot, sizeof(nd), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; k++) std::memcpy(&ndt3[k], &nd[k], 8);
    unsigned long long big;
    const double b = 1.e9;
    std::memcpy(&big, &b, 8);
    unsigned long long bigs[3] = {big, big, big};
    HIPCHK(hipMemcpy(c->ndt_slot, bigs, sizeof(bigs), hipMemcpyHostToDevice));
    auto Z = [&](int i, int j, int h) { return ((size_t)i * g.SN2 + j) * g.N3 + h; };
    auto hp = [&](int h) { return (h + 1) % g.N3; };
    const size_t nz = c->nz;
    for (int h = 0; h < g.N3; h++) {
        for (int i = g.is; i <= g.ie + 1; i++)
            for (int j = g.js; j <= g.je; j++) {
                F1[B1 * nz + Z(i, j, h)] = 0.;
                F1[B2 * nz + Z(i, j, h)] = 0.5 * (e3[Z(i, j, h)] + e3[Z(i, j + 1, h)]);
                if (g.N3 > 2) F1[B3 * nz + Z(i, j, h)] = -0.5 * (e2[Z(i, j, h)] + e2[Z(i, j, hp(h))]);
            }
        for (int i = g.is; i <= g.ie; i++)
            for (int j = g.js; j <= g.je + 1; j++) {
                F2[B1 * nz + Z(i, j, h)] = -0.5 * (e3[Z(i, j, h)] + e3[Z(i + 1, j, h)]);
                F2[B2 * nz + Z(i, j, h)] = 0.;
                if (g.N3 > 2) F2[B3 * nz + Z(i, j, h)] = 0.5 * (e1[Z(i, j, h)] + e1[Z(i, j, hp(h))]);
            }
        if (g.N3 > 2)
            for (int i = g.is; i <= g.ie; i++)
                for (int j = g.js; j <= g.je; j++) {
                    F3[B1 * nz + Z(i, j, h)] = 0.5 * (e2[Z(i, j, h)] + e2[Z(i + 1, j, h)]);
                    F3[B2 * nz + Z(i, j, h)] = -0.5 * (e1[Z(i, j, h)] + e1[Z(i, j + 1, h)]);
                    F3[B3 * nz + Z(i, j, h)] = 0.;
                }
    }
    return 0;
}

int harm_debug_bound(harm_ctx *c)
{
    int rc = bound(c, c->p);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int harm_debug_fixup_utoprim(harm_ctx *c)
{
    int rc = fixup_utoprim(c, c->p);
    if (rc) return rc;
    if ((rc = bound(c, c->p))) return rc;
    return check_flags(c);
}

int harm_debug_advance(harm_ctx *c, int fullstep, double *ph)
{
    if (!ph) return set_err("harm_debug_advance: ph is null", 2);
    int rc;
    const size_t bytes = 8 * c->nz * sizeof(double);
    c->flux_valid = false;
    if (fullstep) {
        HIPCHK(hipMemcpyAsync(c->ph, ph, bytes, hipMemcpyHostToDevice, c->stream));
        if ((rc = advance(c, c->p, c->ph, c->p, 1.0, 1))) return rc;
    } else {
        if ((rc = advance(c, c->p, c->p, c->ph, 0.5, 0))) return rc;
        HIPCHK(hipMemcpyAsync(ph, c->ph, bytes, hipMemcpyDeviceToHost, c->stream));
    }
    // the step's next bound() would empty the Font-fix and seed lists
    HIPCHK(hipMemsetAsync(c->font.count, 0, sizeof(int), c->stream));
    if (c->seeds.count) HIPCHK(hipMemsetAsync(c->seeds.count, 0, sizeof(int), c->stream));
    return check_flags(c);
}

int harm_debug_fixup(harm_ctx *c)
{
    const DGrid &g = c->g;
    const size_t na = (size_t)(g.ie - g.is + 1) * (g.je - g.js + 1) * g.N3;
    hipLaunchKernelGGL(k_fixup, dim3(nblocks(na)), dim3(TPB), 0, c->stream, g, c->p, c->pflag, c->cnt);
    HIPCHK(hipGetLastError());
    return check_flags(c);
}

int harm_get_geometry(harm_ctx *c, double *gcov, double *gcon, double *gdet, double *conn)
{
    const size_t ncol = c->g.ncol;
    for (int l = 0; l < 5; l++) {
        const std::vector<double> &T = c->tab.geom[l];
        if (gcov) std::memcpy(gcov + (size_t)l * 16 * ncol, T.data(), 16 * ncol * sizeof(double));
        if (gcon) std::memcpy(gcon + (size_t)l * 16 * ncol, T.data() + 16 * ncol, 16 * ncol * sizeof(double));
        if (gdet) std::memcpy(gdet + (size_t)l * ncol, T.data() + 32 * ncol, ncol * sizeof(double));
    }
    if (conn) std::memcpy(conn, c->tab.conn.data(), 64 * ncol * sizeof(double));
    return 0;
}

static int eos_tab_of(const harm_eos_table *t, EosTab &e);

int harm_set_eos(harm_ctx *c, const harm_eos_table *tab, int ref_derivs)
{
    if (!c) return set_err("null context");
    if (!tab) {
        c->eos_kind = 0;
        std::memset(&c->g.tab, 0, sizeof(c->g.tab));
        return 0;
    }
    EosTab e;
    int rc;
    if ((rc = eos_tab_of(tab, e))) return rc;
    if (!(e.gam > 1.)) return set_err("EOS table: Gamma-law fallback index must be > 1", 4);
    c->g.tab = e;
    c->g.ref_derivs = ref_derivs ? 1 : 0;
    c->eos_kind = 1;
    return 0;
}

int harm_set_eos_cache(harm_ctx *c, const harm_cache *const *caches, int n, const harm_ceos_units *u)
{
    if (!c) return set_err("null context");
    if (!caches) {
        if (c->eos_kind == 2) c->eos_kind = 0;
        return 0;
    }
    if (!u) return set_err("harm_set_eos_cache: null units", 1);
    if (n != 1 && n != c->g.SN1) return set_err("harm_set_eos_cache: one cache, or one per radial index (N1+4)", 4);
    if (c->g.lim < 0 || c->g.lim > 2)
        return set_err("harm_set_eos_cache: the cache EOS step reconstructs with MC, van Leer or minmod (lim 0-2)", 4);
    if (c->tr.on) return set_err("harm_set_eos_cache: not with tracers (their records take the EOS temperature)", 4);
    if (!(u->rho_unit > 0.) || !(u->u_unit > 0.) || !(u->p_unit > 0.))
        return set_err("harm_set_eos_cache: units must be positive", 4);
    CacheView v0{};
    harm_cache_desc d0;
    std::vector<const double *> vals((size_t)n);
    std::vector<double> ax0, ay0;
    for (int k = 0; k < n; k++) {
        CacheView v;
        harm_cache_desc d;
        int dev, rc;
        if ((rc = harm_cache_view_of(caches[k], v, d, dev))) return rc;
        if (dev != c->device) return set_err("harm_set_eos_cache: a cache lives on another device", 4);
        if (!v.whole) return set_err("harm_set_eos_cache: every cache must be whole (all points computed)", 4);
        if (v.D < 2) return set_err("harm_set_eos_cache: a cache needs components e and p", 4);
        if (v.xtype == AXIS_ROLLING || v.ytype == AXIS_ROLLING)
            return set_err("harm_set_eos_cache: rolling caches are not supported", 4);
        std::vector<double> ax((size_t)v.X), ay((size_t)v.Y);
        HIPCHK(hipMemcpy(ax.data(), v.vx, ax.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ay.data(), v.vy, ay.size() * sizeof(double), hipMemcpyDeviceToHost));
        if (k == 0) {
            v0 = v;
            d0 = d;
            ax0 = ax;
            ay0 = ay;
        } else if (v.X != v0.X || v.Y != v0.Y || v.D != v0.D || ax != ax0 || ay != ay0) {
            return set_err("harm_set_eos_cache: the caches must share one (rho, T) grid", 4);
        }
        // z plane 0 of each cache: [X][Y][D]
        vals[(size_t)k] = v.values;
    }
    for (const std::vector<double> *a : {&ax0, &ay0})
        for (size_t k = 1; k < a->size(); k++)
            if (!((*a)[k - 1] < (*a)[k])) return set_err("harm_set_eos_cache: the cache axes must increase", 4);
    CeosParams P{};
    P.view = v0;
    P.view.whole = 1;
    P.n = n;
    P.rho_unit = u->rho_unit;
    P.u_unit = u->u_unit;
    P.p_unit = u->p_unit;
    P.gam = c->P.gam;
    P.rho_min = u->rho_min != 0. ? u->rho_min : d0.xmin;
    P.rho_max = u->rho_max != 0. ? u->rho_max : d0.xmax;
    const double Tmin = u->T_min != 0. ? u->T_min : d0.ymin;
    const double Tmax = u->T_max != 0. ? u->T_max : d0.ymax;
    P.T_lo = Tmin * (1.0 + CEOS_HH);  // cooleos.c:384-385
    P.T_hi = Tmax * (1.0 - CEOS_HH);
    if (!(P.rho_min < P.rho_max) || !(P.T_lo < P.T_hi))
        return set_err("harm_set_eos_cache: empty density or temperature range", 4);
    int rc;
    if (c->ceos_vals) (void)hipFree((void *)c->ceos_vals);
    c->ceos_vals = nullptr;
    if ((rc = dalloc(&c->ceos_vals, (size_t)n))) return rc;
    HIPCHK(hipMemcpy((void *)c->ceos_vals, vals.data(), (size_t)n * sizeof(double *), hipMemcpyHostToDevice));
    P.values = c->ceos_vals;
    // the splines along T of e and p, formed once per cache
    const size_t nco = 2 * (size_t)v0.X * v0.Y * CACHE_YCO;
    for (double *q : c->ceos_ycoef) (void)hipFree(q);
    c->ceos_ycoef.assign((size_t)n, nullptr);
    for (int k = 0; k < n; k++)
        if ((rc = dalloc(&c->ceos_ycoef[(size_t)k], nco))) return rc;
    if (c->ceos_yco) (void)hipFree((void *)c->ceos_yco);
    c->ceos_yco = nullptr;
    if ((rc = dalloc(&c->ceos_yco, (size_t)n))) return rc;
    HIPCHK(hipMemcpy((void *)c->ceos_yco, c->ceos_ycoef.data(), (size_t)n * sizeof(double *), hipMemcpyHostToDevice));
    P.yco = c->ceos_yco;
    std::vector<int> hx, hy;
    P.hx = ceos_hint_table(ax0, hx);
    P.hy = ceos_hint_table(ay0, hy);
    if (c->ceos_hint) (void)hipFree((void *)c->ceos_hint);
    c->ceos_hint = nullptr;
    if ((rc = dalloc(&c->ceos_hint, hx.size() + hy.size()))) return rc;
    if (!hx.empty())
        HIPCHK(hipMemcpy(c->ceos_hint, hx.data(), hx.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!hy.empty())
        HIPCHK(hipMemcpy(c->ceos_hint + hx.size(), hy.data(), hy.size() * sizeof(int), hipMemcpyHostToDevice));
    P.hx.idx = c->ceos_hint;
    P.hy.idx = c->ceos_hint + hx.size();
    hipLaunchKernelGGL(k_ceos_ycoef, dim3(nblocks(2 * (size_t)v0.X * v0.Y), (unsigned)n), dim3(TPB), 0, c->stream, v0,
                       P.values, c->ceos_yco);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->lastT && (rc = dalloc(&c->lastT, c->nz))) return rc;
    if (!c->ceos_fe && (rc = dalloc(&c->ceos_fe, 4 * c->nz))) return rc;
    if (!c->ceos_solves && (rc = dalloc(&c->ceos_solves, 1))) return rc;
    HIPCHK(hipMemset(c->lastT, 0, c->nz * sizeof(double)));  // calloc'ed by SetupCache_CoolEOS, cooleos.c:834
    HIPCHK(hipMemset(c->ceos_solves, 0, sizeof(unsigned long long)));
    HIPCHK(hipDeviceSynchronize());
    c->ceos = P;
    c->ceos_vanish = ceos_vanishing_steps(P.rho_unit, P.u_unit);
    for (bool &b : c->ceos_chain) b = true;
    c->eos_kind = 2;
    return 0;
}

int harm_get_last_T(harm_ctx *c, double *T)
{
    if (!c || !T) return set_err("null argument");
    if (!c->lastT) return set_err("harm_get_last_T: no cache EOS was set", 4);
    HIPCHK(hipMemcpyAsync(T, c->lastT, c->nz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int harm_set_last_T(harm_ctx *c, const double *T)
{
    if (!c || !T) return set_err("null argument");
    if (!c->lastT) return set_err("harm_set_last_T: no cache EOS was set", 4);
    HIPCHK(hipMemcpyAsync(c->lastT, T, c->nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // unset temperatures may now sit anywhere: the next sweep of each direction runs in j order
    for (bool &b : c->ceos_chain) b = true;
    return 0;
}

int harm_get_eos_solves(harm_ctx *c, int64_t *n)
{
    if (!c || !n) return set_err("null argument");
    unsigned long long v = 0;
    if (c->ceos_solves) {
        HIPCHK(hipMemcpyAsync(&v, c->ceos_solves, sizeof(v), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    *n = (int64_t)v;
    return 0;
}

namespace {

// tracer_name (tracers.c:100-107) under the run directory's tracers/
std::string tracer_file(const harm_ctx *c, int h, int i, int j)
{
    char b[96];
    if (c->hg.N3 > 2)
        snprintf(b, sizeof(b), "tracers/tracer_%03d_%03d_%03d.dat", i, j, h);
    else
        snprintf(b, sizeof(b), "tracers/tracer_%03d_%03d.dat", i, j);
    return b;
}

// restart_write_tracers (tracers.c:565-606): the live tracers in slot order,
// with the current size of their files (0 when absent)
int tracer_block(harm_ctx *c, harm_host::TracerBlock &tb)
{
    const auto &hg = c->hg;
    const TrArgs &a = c->tr.a;
    std::vector<double> X(3 * (size_t)a.n);
    std::vector<int8_t> st((size_t)a.n);
    HIPCHK(hipMemcpyAsync(X.data(), a.X, X.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(st.data(), a.state, st.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const bool d3 = hg.N3 > 2;
    // tracers.c:126-145 (the phi limits are set only in 3D)
    const double par[8] = {a.Rin, a.Rout, a.thmin, a.thmax, hg.startx[1] + (0 + 0.5) * hg.dx[1],
                           hg.startx[1] + (hg.N1 + 0.5) * hg.dx[1], d3 ? hg.startx[3] + (0 + 0.5) * hg.dx[3] : 0.0,
                           d3 ? hg.startx[3] + (hg.N3 + 0.5) * hg.dx[3] : 0.0};
    std::copy(par, par + 8, tb.par);
    const long long n12 = (long long)hg.N1 * hg.N2;
    for (long long k = 0; k < a.n; k++) {
        if (st[k] != HARM_TRACER_LIVE) continue;
        const int h = (int)(k / n12), i = (int)((k % n12) / hg.N2), j = (int)(k % hg.N2);
        tb.i.push_back(i);
        tb.j.push_back(j);
        tb.h.push_back(h);
        tb.X.push_back(0.0);  // X[TT] (tracers.c:153)
        for (int d = 0; d < 3; d++) tb.X.push_back(X[(size_t)d * a.n + k]);
        struct stat sb;
        const std::string f = tracer_file(c, h, i, j);
        tb.size.push_back(stat(f.c_str(), &sb) == 0 ? (int32_t)sb.st_size : 0);
    }
    return 0;
}

}  // namespace

int harm_restart_write(harm_ctx *c, const char *path, const harm_restart_hdr *hdr)
{
    if (!c || !path) return set_err("null argument");
    harm_restart_hdr h;
    if (hdr) {
        h = *hdr;
    } else {
        std::memset(&h, 0, sizeof(h));
        h.ener_size = -1;
    }
    const harm_params &P = c->P;
    h.N1 = P.n1;
    h.N2 = P.n2;
    h.N3 = P.n3;
    h.tf = P.tf;
    h.a = P.a;
    h.gam = P.gam;
    h.cour = P.cour;
    h.lim = P.lim;
    h.Rin = P.Rin;
    h.Rout = P.Rout;
    h.hslope = P.hslope;
    h.R0 = P.R0;
    h.defcon = P.defcon;
    h.Rbr = P.Rbr;
    double ts[4];
    HIPCHK(hipMemcpyAsync(ts, c->tstate, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    std::vector<double> prim(8 * c->nz), temp;
    HIPCHK(hipMemcpyAsync(prim.data(), c->p, prim.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (c->eos_kind == 1) {
        // the temperatures go through the divergence buffer (8*N1*N2*N3 >= nz doubles)
        temp.resize(c->nz);
        hipLaunchKernelGGL(k_zone_temp, dim3(nblocks(c->nz)), dim3(TPB), 0, c->stream, c->g.tab, c->p, c->S, c->nz);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(temp.data(), c->S, c->nz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    } else if (c->eos_kind == 2) {
        // restart_write_CoolEOS (cooleos.c:1000-1014): the zones' last temperatures
        temp.resize(c->nz);
        HIPCHK(hipMemcpyAsync(temp.data(), c->lastT, c->nz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    h.t = ts[0];
    h.dt = ts[1];
    h.mass_in = ts[2];
    h.mass_out = ts[3];
    h.mass_start = c->mass_start;
    harm_host::TracerBlock tb;
    if (c->tr.on) {
        int rc = tracer_block(c, tb);
        if (rc) return rc;
    }
    return harm_host::restart_write_file(path, h, prim.data(), temp.empty() ? nullptr : temp.data(),
                                         c->tr.on ? &tb : nullptr);
}

int harm_restart_read(harm_ctx *c, const char *path, harm_restart_hdr *hdr)
{
    if (!c || !path) return set_err("null argument");
    harm_restart_hdr h;
    int rc;
    if ((rc = harm_host::restart_read_file(path, h, nullptr, nullptr, nullptr))) return rc;
    const harm_params &P = c->P;
    if (h.N1 != P.n1 || h.N2 != P.n2 || h.N3 != P.n3) return set_err("restart file: grid size differs from the context's", 3);
    const struct {
        const char *name;
        double file, ctx;
    } same[] = {{"a", h.a, P.a},         {"gam", h.gam, P.gam},          {"Rin", h.Rin, P.Rin},
                {"Rout", h.Rout, P.Rout}, {"hslope", h.hslope, P.hslope}, {"R0", h.R0, P.R0},
                {"Rbr", h.Rbr, P.Rbr}};
    for (const auto &f : same)
        if (!(f.file == f.ctx))
            return set_err(std::string("restart file: ") + f.name +
                               " differs from the context's (create the context from harm_restart_peek's header)",
                           3);
    if (h.lim < 0 || h.lim > 3) return set_err("restart file: unknown slope limiter", 3);
    std::vector<double> prim(8 * c->nz);
    harm_host::TracerBlock tb;
    int has_tb = 0;
    std::vector<double> temp(c->eos_kind == 2 ? c->nz : 0);
    int has_temp = 0;
    if ((rc = harm_host::restart_read_file(path, h, prim.data(), temp.empty() ? nullptr : temp.data(),
                                           temp.empty() ? nullptr : &has_temp, &tb, &has_tb)))
        return rc;
    if (has_temp) {
        // restart_read_CoolEOS (cooleos.c:1016-1030)
        if ((rc = harm_set_last_T(c, temp.data()))) return rc;
    }
    // cour, lim, tf, defcon come from the file (restart.c:335-367)
    c->P.cour = c->g.cour = h.cour;
    c->P.lim = c->g.lim = h.lim;
    c->P.tf = c->g.tf = h.tf;
    c->P.defcon = c->g.defcon = h.defcon;
    HIPCHK(hipMemcpyAsync(c->p, prim.data(), prim.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->pflag, 0, c->nz, c->stream));
    const double ts[5] = {h.t, h.dt, h.mass_in, h.mass_out, 0.};
    HIPCHK(hipMemcpyAsync(c->tstate, ts, sizeof(ts), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->halted = 0;
    c->mass_start = h.mass_start;
    c->flux_valid = false;
    if (has_tb && c->tr.on) {
        // restart_read_tracers (tracers.c:608-654): a listed tracer whose file is
        // gone is dropped; an existing file is cut back to its recorded size
        const auto &hg = c->hg;
        const long long n = c->tr.a.n;
        std::vector<double> X(3 * (size_t)n, 0.0);
        std::vector<int8_t> st((size_t)n, (int8_t)HARM_TRACER_NONE);
        for (size_t q = 0; q < tb.i.size(); q++) {
            if (tb.i[q] < 0 || tb.i[q] >= hg.N1 || tb.j[q] < 0 || tb.j[q] >= hg.N2 || tb.h[q] < 0 || tb.h[q] >= hg.N3)
                return set_err("restart file: tracer outside the grid", 3);
            const std::string f = tracer_file(c, tb.h[q], tb.i[q], tb.j[q]);
            if (access(f.c_str(), F_OK) != 0) continue;
            if (truncate(f.c_str(), (off_t)tb.size[q]) != 0) return set_err("restart: cannot truncate " + f, 2);
            const long long k = ((long long)tb.h[q] * hg.N1 + tb.i[q]) * hg.N2 + tb.j[q];
            st[k] = HARM_TRACER_LIVE;
            for (int d = 0; d < 3; d++) X[(size_t)d * n + k] = tb.X[4 * q + 1 + d];
        }
        if ((rc = harm_tracers_set(c, X.data(), st.data()))) return rc;
    }
    if ((rc = bound(c, c->p))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (hdr) *hdr = h;
    return 0;
}

int harm_set_guess(harm_ctx *c, int mode, const harm_mlp *m)
{
    if (!c) return set_err("null context");
    if (mode < 0 || mode > 2) return set_err("unknown guess mode", 4);
    if (mode != 0) {
        if (!m) return set_err("guess modes 1 and 2 need a harm_mlp handle", 4);
        if (harm_mlp_device(m) != c->device) return set_err("network and context are on different devices", 4);
    }
    if (mode == 2 && !c->seeds.count) {
        // the seed list of the gated update: one slot per 16 active zones (at
        // least 65536); a zone that finds it full resumes the plain solve
        HIPCHK(hipStreamSynchronize(c->stream));
        const size_t na = (size_t)c->hg.N1 * c->hg.N2 * c->hg.N3;
        const size_t cap = std::max<size_t>(65536, na / 16);
        SeedList &sl = c->seeds;
        sl.cap = (int)cap;
        int rc;
        if ((rc = dalloc(&sl.count, 1)) || (rc = dalloc(&sl.zone, cap)) || (rc = dalloc(&sl.U, 8 * cap))) return rc;
        HIPCHK(hipMemsetAsync(sl.count, 0, sizeof(int), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->guess_mode = mode;
    c->mlp = (mode != 0) ? m : nullptr;
    return 0;
}

int harm_set_guess_gate(harm_ctx *c, double tau)
{
    if (!c) return set_err("null context");
    if (!(tau > 0.)) return set_err("the gate threshold must be > 0", 4);
    c->gate_tau = tau;
    return 0;
}

int harm_set_capture(harm_ctx *c, double *feat, double *targ, uint8_t *valid, size_t cap, int stride)
{
    if (!c) return set_err("null context");
    if (!feat) {
        c->cap_feat = nullptr;
        c->cap_targ = nullptr;
        c->cap_valid = nullptr;
        c->cap = 0;
        return 0;
    }
    if (!targ || !valid || cap == 0 || stride < 1) return set_err("capture needs feat, targ, valid, cap > 0, stride >= 1", 4);
    c->cap_feat = feat;
    c->cap_targ = targ;
    c->cap_valid = valid;
    c->cap = cap;
    c->cap_count = 0;
    c->cap_stride = stride;
    c->cap_phase = 0;
    return 0;
}

int harm_get_capture_count(harm_ctx *c, size_t *n)
{
    if (!c || !n) return set_err("null argument");
    *n = c->cap_count;
    return 0;
}

int harm_utoprim2d_batch(const double *U, const double *gcov, const double *gcon, const double *gdet, double *prim,
                         int8_t *ret, uint8_t *iters, uint8_t *kconv, size_t n, const harm_u2p_opts *opts,
                         void *stream)
{
    if (!opts) return set_err("null options");
    if (n == 0) return 0;
    MlpDev mv;
    std::memset(&mv, 0, sizeof(mv));
    double gate = 0.;
    if (opts->guess_mode == 1 || opts->guess_mode == 2) {
        if (!opts->mlp) return set_err("guess_mode 1 and 2 need a harm_mlp handle", 4);
        if (opts->guess_mode == 2 && !(opts->gate_tau > 0.)) return set_err("guess_mode 2 needs gate_tau > 0", 4);
        if (opts->guess_mode == 2) gate = opts->gate_tau;
        mv = harm_mlp_view((const harm_mlp *)opts->mlp);
        hipLaunchKernelGGL((k_c2p_batch<true>), dim3(nblocks(n)), dim3(TPB), 0, (hipStream_t)stream, U, gcov, gcon,
                           gdet, prim, ret, iters, kconv, n, opts->gam, opts->font_fix, opts->K, (int *)nullptr, mv, gate);
    } else if (opts->guess_mode == 0) {
        hipLaunchKernelGGL((k_c2p_batch<false>), dim3(nblocks(n)), dim3(TPB), 0, (hipStream_t)stream, U, gcov, gcon,
                           gdet, prim, ret, iters, kconv, n, opts->gam, opts->font_fix, opts->K, (int *)nullptr, mv, gate);
    } else {
        return set_err("unknown guess_mode", 4);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int harm_stress_batch(const harm_stress_spec *spec, size_t n, double *U, double *gcov, double *gcon, double *gdet,
                      double *prim_true, double *guess, void *stream)
{
    if (!spec || !U || !gcov || !gcon || !gdet || !prim_true || !guess) return set_err("null argument");
    if (n == 0) return 0;
    if (n > ((size_t)1 << 32)) return set_err("stress batch too large (n <= 2^32)", 4);
    StressSpec sp;
    sp.a = spec->a;
    sp.hslope = spec->hslope;
    sp.gam = spec->gam;
    sp.r_lo = spec->r_lo;
    sp.r_hi = spec->r_hi;
    sp.th_lo = spec->th_lo;
    sp.th_hi = spec->th_hi;
    sp.rho_lo = spec->rho_lo;
    sp.rho_hi = spec->rho_hi;
    sp.urho_lo = spec->urho_lo;
    sp.urho_hi = spec->urho_hi;
    sp.lor_lo = spec->lorentz_lo;
    sp.lor_hi = spec->lorentz_hi;
    sp.sig_lo = spec->sigma_lo;
    sp.sig_hi = spec->sigma_hi;
    sp.pert = spec->guess_pert;
    for (int k = 0; k < 4; k++) sp.seed[k] = spec->seed[k];
    sp.guess_kind = spec->guess_kind;
    if (!(sp.r_lo > 0.) || !(sp.r_hi > sp.r_lo) || !(sp.th_lo > 0.) ||
        !(sp.th_hi < M_PI) || !(sp.rho_lo > 0.) || !(sp.urho_lo > 0.) || !(sp.lor_lo >= 1.) ||
        !(sp.sig_lo > 0.) || sp.guess_kind < 0 || sp.guess_kind > 1)
        return set_err("bad stress-batch ranges (r > 0, theta in (0, pi), positive scales)", 4);
    hipLaunchKernelGGL(k_stress_gen, dim3(nblocks(n)), dim3(TPB), 0, (hipStream_t)stream, sp, n, U, gcov, gcon, gdet,
                       prim_true, guess);
    HIPCHK(hipGetLastError());
    return 0;
}

static int eos_tab_of(const harm_eos_table *t, EosTab &e)
{
    std::memset(&e, 0, sizeof(e));
    if (!t) return 0;
    if (t->nr < 2 || t->nt < 2 || !t->rho || !t->T || !t->e || !t->p) return set_err("bad EOS table", 4);
    e.nr = t->nr;
    e.nt = t->nt;
    e.rho = t->rho;
    e.T = t->T;
    e.e = t->e;
    e.p = t->p;
    e.gam = t->gam;
    return 0;
}

int harm_eos_eval(const harm_eos_table *tab, const double *rho, const double *u, double *p, double *dpdrho,
                  double *dpdu, size_t n, void *stream)
{
    if (!tab || !rho || !u || !p) return set_err("null argument");
    if (n == 0) return 0;
    EosTab e;
    int rc;
    if ((rc = eos_tab_of(tab, e))) return rc;
    hipLaunchKernelGGL(k_eos_eval, dim3(nblocks(n)), dim3(TPB), 0, (hipStream_t)stream, e, rho, u, p, dpdrho, dpdu,
                       n);
    HIPCHK(hipGetLastError());
    return 0;
}

int harm_utoprim5d_batch(const double *U, const double *gcov, const double *gcon, const double *gdet, double *prim,
                         int8_t *ret, uint8_t *iters, size_t n, double gam, const harm_eos_table *tab, void *stream)
{
    if (!U || !gcov || !gcon || !gdet || !prim || !ret) return set_err("null argument");
    if (n == 0) return 0;
    EosTab e;
    int rc;
    if ((rc = eos_tab_of(tab, e))) return rc;
    hipLaunchKernelGGL(k_c2p5_batch, dim3(nblocks(n)), dim3(TPB), 0, (hipStream_t)stream, U, gcov, gcon, gdet, prim,
                       ret, iters, n, gam, e, (int *)nullptr);
    HIPCHK(hipGetLastError());
    return 0;
}

int harm_debug_pow(const double *x, const double *y, double *out, size_t n)
{
    double *d = nullptr;
    HIPCHK(hipMalloc((void **)&d, 3 * n * sizeof(double)));
    HIPCHK(hipMemcpy(d, x, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d + n, y, n * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_pow, dim3(nblocks(n)), dim3(TPB), 0, 0, d, d + n, d + 2 * n, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipFree(d));
    return 0;
}

int harm_debug_copy(size_t nbytes, int repeats, double *ms_per_copy)
{
    double2 *a = nullptr, *b = nullptr;
    const size_t n = nbytes / sizeof(double2);
    HIPCHK(hipMalloc((void **)&a, n * sizeof(double2)));
    HIPCHK(hipMalloc((void **)&b, n * sizeof(double2)));
    HIPCHK(hipMemset(a, 0, n * sizeof(double2)));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(TPB), 0, 0, a, b, n);
    HIPCHK(hipEventRecord(e0, 0));
    for (int r = 0; r < repeats; r++) hipLaunchKernelGGL(k_copy, dim3(4096), dim3(TPB), 0, 0, a, b, n);
    HIPCHK(hipEventRecord(e1, 0));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (ms_per_copy) *ms_per_copy = ms / std::max(repeats, 1);
    HIPCHK(hipEventDestroy(e0));
    HIPCHK(hipEventDestroy(e1));
    HIPCHK(hipFree(a));
    HIPCHK(hipFree(b));
    return 0;
}

// The reference-signature inversions (Utoprim_2d, Utoprim_5d) take their EOS
// from the reference's globals (gam, the COOL_EOS tables); here each calling
// thread sets its own with harm_set_gamma / harm_set_u2p_eos.
static thread_local double g_gamma = 4. / 3.;
static thread_local EosTab g_u2p_tab;  // nr == 0: the Gamma law
static std::mutex g_u2p_mu;

void harm_set_gamma(double gam) { g_gamma = gam; }

int harm_set_u2p_eos(const harm_eos_table *tab) { return eos_tab_of(tab, g_u2p_tab); }

namespace {

// Staging for one zone on the current device: 49 doubles (U, gcov, gcon, gdet,
// prim), an int8 and an int code; allocated on a device's first call and kept
// (the caller holds g_u2p_mu).
struct U2pStage {
    double *buf;
    int8_t *ret8;
    int *ret;
};

int u2p_stage(U2pStage &st, const double *U, const double (*gcov)[4], const double (*gcon)[4], double gdet,
              const double *prim)
{
    static U2pStage per_dev[64];
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_err("device index out of range", 4);
    U2pStage &s = per_dev[dev];
    if (!s.buf) {
        HIPCHK(hipMalloc((void **)&s.buf, 49 * sizeof(double)));
        HIPCHK(hipMalloc((void **)&s.ret8, 8));
        HIPCHK(hipMalloc((void **)&s.ret, sizeof(int)));
    }
    double h[49];
    std::memcpy(h, U, 8 * sizeof(double));
    std::memcpy(h + 8, gcov, 16 * sizeof(double));
    std::memcpy(h + 24, gcon, 16 * sizeof(double));
    h[40] = gdet;
    std::memcpy(h + 41, prim, 8 * sizeof(double));
    HIPCHK(hipMemcpy(s.buf, h, sizeof(h), hipMemcpyHostToDevice));
    st = s;
    return 0;
}

int u2p_unstage(const U2pStage &s, double *prim, int &r)
{
    HIPCHK(hipMemcpy(prim, s.buf + 41, 8 * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&r, s.ret, sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

}  // namespace

int Utoprim_2d(double U[8], double gcov[4][4], double gcon[4][4], double gdet, double prim[8])
{
    std::lock_guard<std::mutex> lk(g_u2p_mu);
    U2pStage s;
    int rc = u2p_stage(s, U, gcov, gcon, gdet, prim);
    if (rc) return rc;
    MlpDev mv;
    std::memset(&mv, 0, sizeof(mv));
    hipLaunchKernelGGL((k_c2p_batch<false>), dim3(1), dim3(TPB), 0, 0, s.buf, s.buf + 8, s.buf + 24, s.buf + 40,
                       s.buf + 41, s.ret8, (uint8_t *)nullptr, (uint8_t *)nullptr, (size_t)1, g_gamma, 0, 0., s.ret,
                       mv, 0.);
    HIPCHK(hipGetLastError());
    int r = 0;
    if ((rc = u2p_unstage(s, prim, r))) return rc;
    return r;
}

int Utoprim_5d(double U[8], double gcov[4][4], double gcon[4][4], double gdet, double prim[8])
{
    std::lock_guard<std::mutex> lk(g_u2p_mu);
    U2pStage s;
    int rc = u2p_stage(s, U, gcov, gcon, gdet, prim);
    if (rc) return rc;
    hipLaunchKernelGGL(k_c2p5_batch, dim3(1), dim3(TPB), 0, 0, s.buf, s.buf + 8, s.buf + 24, s.buf + 40, s.buf + 41,
                       s.ret8, (uint8_t *)nullptr, (size_t)1, g_gamma, g_u2p_tab, s.ret);
    HIPCHK(hipGetLastError());
    int r = 0;
    if ((rc = u2p_unstage(s, prim, r))) return rc;
    return r;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Tracers (tracers.c:109-654)
// ---------------------------------------------------------------------------
int harm_tracers_init(harm_ctx *c, double rho_unit, double t_unit, int64_t *n_live)
{
    if (!c) return set_err("null context");
    if (c->eos_kind == 2) return set_err("not available with the cache EOS (it would move the zones' last temperatures)", 4);
    if (c->P.n3 == 2) return set_err("tracers: N3 == 2 is not supported", 3);
    HIPCHK(hipStreamSynchronize(c->stream));
    tracers_free(c);
    auto &t = c->tr;
    TrArgs &a = t.a;
    const auto &hg = c->hg;
    const harm_params &P = c->P;
    int rc;
    a.n = (long long)hg.N1 * hg.N2 * hg.N3;
    const unsigned nb = nblocks((size_t)a.n);
    if ((rc = dalloc(&a.X, 3 * (size_t)a.n))) return rc;
    if ((rc = dalloc(&a.state, (size_t)a.n))) return rc;
    if ((rc = dalloc(&a.nev, 1))) return rc;
    if ((rc = dalloc(&t.bcount, nb))) return rc;
    if ((rc = dalloc(&t.base, nb))) return rc;
    a.sx1 = hg.startx[1];
    a.sx2 = hg.startx[2];
    a.sx3 = hg.startx[3];
    a.R0 = P.R0;
    a.hslope = P.hslope;
    a.Rbr = P.Rbr;
    a.Rin = exp(hg.startx[1] - 0.5 * hg.dx[1]) + P.R0;        // tracers.c:126
    a.Rout = 0.8 * (exp(hg.startx[1] + hg.N1 * hg.dx[1]) + P.R0);  // tracers.c:127
    a.thmin = 0.02 * M_PI;
    a.thmax = 0.98 * M_PI;
    a.rho_unit = rho_unit;
    a.t_unit = t_unit;
    hipLaunchKernelGGL(k_tracer_init, dim3(nb), dim3(TPB), 0, c->stream, c->g, a, c->p, a.nev);
    HIPCHK(hipGetLastError());
    unsigned long long live = 0;
    HIPCHK(hipMemcpyAsync(&live, a.nev, sizeof(live), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemsetAsync(a.nev, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // each tracer leaves at most once, and the live count never grows
    a.cap_ev = live;
    if ((rc = dalloc(&a.ev, (size_t)live))) return rc;
    if ((rc = dalloc(&t.dump, (size_t)live))) return rc;
    t.on = true;
    if (n_live) *n_live = (int64_t)live;
    return 0;
}

int harm_tracers_stop(harm_ctx *c)
{
    if (!c) return set_err("null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    tracers_free(c);
    return 0;
}

int harm_tracers_dump(harm_ctx *c, harm_tracer_rec *recs, int64_t cap, int64_t *n)
{
    if (!c || !n) return set_err("null argument");
    auto &t = c->tr;
    if (!t.on) return set_err("tracers: harm_tracers_init has not been called");
    const unsigned nb = nblocks((size_t)t.a.n);
    hipLaunchKernelGGL(k_tracer_count, dim3(nb), dim3(TPB), 0, c->stream, c->g, t.a, t.bcount);
    HIPCHK(hipGetLastError());
    std::vector<unsigned> cnt(nb);
    std::vector<unsigned long long> base(nb);
    HIPCHK(hipMemcpyAsync(cnt.data(), t.bcount, nb * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    unsigned long long total = 0;
    for (unsigned b = 0; b < nb; b++) {
        base[b] = total;
        total += cnt[b];
    }
    *n = (int64_t)total;
    if ((int64_t)total > cap || (total && !recs)) return set_err("harm_tracers_dump: record buffer too small", -1);
    if (total > t.a.cap_ev) return set_err("harm_tracers_dump: more live tracers than were seeded");
    HIPCHK(hipMemcpyAsync(t.base, base.data(), nb * sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
    if (c->eos_kind == 1)
        hipLaunchKernelGGL(k_tracer_dump<true>, dim3(nb), dim3(TPB), 0, c->stream, c->g, t.a, c->p, c->tstate, t.base,
                           t.dump);
    else
        hipLaunchKernelGGL(k_tracer_dump<false>, dim3(nb), dim3(TPB), 0, c->stream, c->g, t.a, c->p, c->tstate,
                           t.base, t.dump);
    HIPCHK(hipGetLastError());
    if (total)
        HIPCHK(hipMemcpyAsync(recs, t.dump, total * sizeof(harm_tracer_rec), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int harm_tracers_events(harm_ctx *c, harm_tracer_rec *recs, int64_t cap, int64_t *n)
{
    if (!c || !n) return set_err("null argument");
    auto &t = c->tr;
    if (!t.on) return set_err("tracers: harm_tracers_init has not been called");
    unsigned long long ne = 0;
    HIPCHK(hipMemcpyAsync(&ne, t.a.nev, sizeof(ne), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *n = (int64_t)ne;
    if ((int64_t)ne > cap || (ne && !recs)) return set_err("harm_tracers_events: record buffer too small", -1);
    if (ne > t.a.cap_ev) return set_err("harm_tracers_events: more removals than tracers");
    if (ne) HIPCHK(hipMemcpyAsync(recs, t.a.ev, ne * sizeof(harm_tracer_rec), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemsetAsync(t.a.nev, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // the atomic append order is arbitrary within a step: restore (step, slot)
    const long long n12 = (long long)c->hg.N1 * c->hg.N2;
    auto slot = [&](const harm_tracer_rec &r) { return (long long)r.h * n12 + (long long)r.i * c->hg.N2 + r.j; };
    std::sort(recs, recs + ne, [&](const harm_tracer_rec &x, const harm_tracer_rec &y) {
        return x.t < y.t || (x.t == y.t && slot(x) < slot(y));
    });
    return 0;
}

int harm_tracers_update(harm_ctx *c, double dt)
{
    if (!c) return set_err("null context");
    auto &t = c->tr;
    if (!t.on) return set_err("tracers: harm_tracers_init has not been called");
    // k_tracer_update reads (t, dt) from a two-double state: the divergence
    // buffer's first two doubles hold them for this launch
    double ts[2];
    HIPCHK(hipMemcpyAsync(ts, c->tstate, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    ts[1] = dt;
    HIPCHK(hipMemcpyAsync(c->S, ts, sizeof(ts), hipMemcpyHostToDevice, c->stream));
    const unsigned nb = nblocks((size_t)t.a.n);
    if (c->eos_kind == 1)
        hipLaunchKernelGGL(k_tracer_update<true>, dim3(nb), dim3(TPB), 0, c->stream, c->g, t.a, c->p, c->S,
                           (const unsigned long long *)nullptr);
    else
        hipLaunchKernelGGL(k_tracer_update<false>, dim3(nb), dim3(TPB), 0, c->stream, c->g, t.a, c->p, c->S,
                           (const unsigned long long *)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int harm_tracers_get(harm_ctx *c, double *X, int8_t *state)
{
    if (!c) return set_err("null context");
    auto &t = c->tr;
    if (!t.on) return set_err("tracers: harm_tracers_init has not been called");
    if (X) HIPCHK(hipMemcpyAsync(X, t.a.X, 3 * (size_t)t.a.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (state) HIPCHK(hipMemcpyAsync(state, t.a.state, (size_t)t.a.n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int harm_tracers_set(harm_ctx *c, const double *X, const int8_t *state)
{
    if (!c) return set_err("null context");
    auto &t = c->tr;
    if (!t.on) return set_err("tracers: harm_tracers_init has not been called");
    if (state) {
        long long live = 0;
        for (long long k = 0; k < t.a.n; k++) {
            if (state[k] < HARM_TRACER_NONE || state[k] > HARM_TRACER_LOST)
                return set_err("harm_tracers_set: unknown tracer state", 3);
            live += state[k] == HARM_TRACER_LIVE;
        }
        if ((unsigned long long)live > t.a.cap_ev) {
            // room for every live tracer's removal record and dump line
            for (void *q : {(void *)t.a.ev, (void *)t.dump})
                if (q) HIPCHK(hipFree(q));
            t.a.ev = nullptr;
            t.dump = nullptr;
            int rc;
            if ((rc = dalloc(&t.a.ev, (size_t)live))) return rc;
            if ((rc = dalloc(&t.dump, (size_t)live))) return rc;
            t.a.cap_ev = (unsigned long long)live;
        }
    }
    if (X) HIPCHK(hipMemcpyAsync(t.a.X, X, 3 * (size_t)t.a.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (state) HIPCHK(hipMemcpyAsync(t.a.state, state, (size_t)t.a.n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(t.a.nev, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// ASCII dump (dump.c:1710-1990)
// ---------------------------------------------------------------------------

namespace {

// The 46 columns of zones z0..z0+n-1 of the active grid in (i, j, h) order.
int dump_fill(harm_ctx *c, double *dcols, size_t cap, size_t z0, size_t n, double *cols)
{
    const DGrid &g = c->g;
    const auto &hg = c->hg;
    for (size_t a = 0; a < n; a += cap) {
        const size_t m = std::min(cap, n - a);
        if (c->eos_kind == 1)
            hipLaunchKernelGGL(k_dump_cols<true>, dim3(nblocks(m)), dim3(TPB), 0, c->stream, g, c->p, z0 + a, m, dcols);
        else
            hipLaunchKernelGGL(k_dump_cols<false>, dim3(nblocks(m)), dim3(TPB), 0, c->stream, g, c->p, z0 + a, m,
                               dcols);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(cols + a * DUMP_NCOL, dcols, m * DUMP_NCOL * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    // coord() and bl_coord() (metric.c:62-115, coord.c:115-125) from the host tables
    const size_t nj = (size_t)(g.je - g.js + 1), ncol = (size_t)hg.SN1 * hg.SN2;
    const double *bl = c->tab.coords.data();
    for (size_t t = 0; t < n; t++) {
        const size_t zl = z0 + t;
        const int h = (int)(zl % g.N3);
        const size_t rest = zl / g.N3;
        const int j = g.js + (int)(rest % nj), i = g.is + (int)(rest / nj);
        const size_t col = (size_t)i * hg.SN2 + j;
        double *o = cols + t * DUMP_NCOL;
        const double X1 = hg.startx[1] + ((i - 2) + 0.5) * hg.dx[1];
        const double X2 = hg.startx[2] + ((j - 2) + 0.5) * hg.dx[2];
        const double X3 = hg.N3 > 2 ? hg.startx[3] + (h + 0.5) * hg.dx[3] : hg.startx[3] + h * hg.dx[3];
        o[0] = X1;
        o[1] = hg.N3 > 1 ? X3 : X2;
        o[2] = bl[1 * ncol + col];
        o[3] = hg.N3 > 1 ? X3 : bl[2 * ncol + col];
    }
    return 0;
}

}  // namespace

int harm_dump_columns(harm_ctx *c, size_t z0, size_t n, double *cols)
{
    if (!c || (n && !cols)) return set_err("null argument");
    if (c->eos_kind == 2) return set_err("not available with the cache EOS (it would move the zones' last temperatures)", 4);
    const DGrid &g = c->g;
    const size_t na = (size_t)(g.ie - g.is + 1) * (g.je - g.js + 1) * g.N3;
    if (z0 > na || n > na - z0) return set_err("harm_dump_columns: zone range outside the grid", 3);
    if (!n) return 0;
    const size_t cap = std::min<size_t>(n, (size_t)1 << 18);
    double *d = nullptr;
    int rc;
    if ((rc = dalloc(&d, cap * DUMP_NCOL))) return rc;
    rc = dump_fill(c, d, cap, z0, n, cols);
    HIPCHK(hipFree(d));
    return rc;
}

int harm_dump_columns_noalloc(harm_ctx *c, size_t z0, size_t n, double *cols)
{
    if (!c || (n && !cols)) return set_err("null argument");
    if (c->eos_kind == 2) return set_err("not available with the cache EOS (it would move the zones' last temperatures)", 4);
    const DGrid &g = c->g;
    const size_t na = (size_t)(g.ie - g.is + 1) * (g.je - g.js + 1) * g.N3;
    if (z0 > na || n > na - z0) return set_err("harm_dump_columns_noalloc: zone range outside the grid", 3);
    if (n > (size_t)HARM_DUMP_NOALLOC_MAX_ZONES)
        return set_err("harm_dump_columns_noalloc: n = " + std::to_string(n) +
                           " zones exceeds HARM_DUMP_NOALLOC_MAX_ZONES (" +
                           std::to_string(HARM_DUMP_NOALLOC_MAX_ZONES) + "); use harm_dump_columns",
                       3);
    if (!n) return 0;
    double *d = scratch_addr(c->device);
    if (!d) return set_err("harm_dump_columns_noalloc: no staging address for this device", 2);
    // every column of the n rows copied back is written by k_dump_cols (4..45)
    // or on the host (0..3), so nothing left by an earlier call shows through
    return dump_fill(c, d, n, z0, n, cols);
}

int harm_dump_write(harm_ctx *c, const char *path, const harm_restart_hdr *hdr)
{
    if (!c || !path) return set_err("null argument");
    if (c->eos_kind == 2) return set_err("not available with the cache EOS (it would move the zones' last temperatures)", 4);
    harm_restart_hdr h;
    if (hdr) {
        h = *hdr;
    } else {
        std::memset(&h, 0, sizeof(h));
    }
    const harm_params &P = c->P;
    h.N1 = P.n1;
    h.N2 = P.n2;
    h.N3 = P.n3;
    h.tf = P.tf;
    h.a = P.a;
    h.gam = P.gam;
    h.cour = P.cour;
    h.lim = P.lim;
    h.Rin = P.Rin;
    h.Rout = P.Rout;
    h.hslope = P.hslope;
    h.R0 = P.R0;
    double ts[2];
    HIPCHK(hipMemcpyAsync(ts, c->tstate, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    h.t = ts[0];
    h.dt = ts[1];
    const DGrid &g = c->g;
    const size_t na = (size_t)(g.ie - g.is + 1) * (g.je - g.js + 1) * g.N3;
    const size_t cap = std::min<size_t>(na, (size_t)1 << 18);
    double *d = nullptr;
    int rc;
    if ((rc = dalloc(&d, std::max<size_t>(cap, 1) * DUMP_NCOL))) return rc;
    // dump.c:1917: the derived block only while the run has not failed
    const int ncol = h.failed ? DUMP_NCOL_FAILED : DUMP_NCOL;
    rc = harm_host::dump_write_file(path, harm_host::dump_header_text(h, c->hg.startx, c->hg.dx), na, ncol, DUMP_NCOL,
                                    [&](size_t z0, size_t n, double *cols) { return dump_fill(c, d, cap, z0, n, cols); });
    HIPCHK(hipFree(d));
    return rc;
}
