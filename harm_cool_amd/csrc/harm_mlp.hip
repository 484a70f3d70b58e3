// harm_mlp.hip -- the learned Newton seed: GPU training (FP32 MFMA GEMMs,
// Adam), packing of the BF16 inference operands, and its C ABI.
//
// No counterpart in the reference (the seed replaces utoprim_2d.c:268-307).
// Training works on batches of B standardized rows:
//   forward   H1 = tanh(X W1^T + b1), H2, H3 likewise, O = H3 W4^T + b4
//   loss      L = 1/(2B) sum_b mask_b |O_b - T_b|^2
//   backward  dO = mask (O - T) / B;  dW_l = dZ_l^T H_{l-1};  db_l = dZ_l^T 1
//             dZ_{l-1} = (dZ_l W_l) * (1 - H_{l-1}^2)
// Every product is one launch of k_gemm (v_mfma_f32_16x16x4_f32, one 16x16
// tile per wavefront, split-K over the batch for the weight gradients with a
// fixed-order reduction, so a training step is deterministic).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "harm_host.hpp"
#include "harm_phys.hpp"
#include "harm_mlp.hpp"

using namespace harm;

struct harm_mlp {
    int device;
    hipStream_t stream;
    float *params, *adam_m, *adam_v, *grad;
    uint4 *wpack;
    float *bias;
    double *norm;
    int vsq_from_W;
    float loss_w[2];  // weights of the two outputs in the training loss
    long long adam_t;
    // training workspace, sized for wsB rows
    int wsB;
    float *X, *T, *H1, *H2, *H3, *O, *dO, *dZa, *dZb, *ones, *part, *lossp;
    uint8_t *M;
    double *normp;
};

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

#define MCHK(x)                                                                                            \
    do {                                                                                                   \
        hipError_t e_ = (x);                                                                               \
        if (e_ != hipSuccess)                                                                              \
            return harm_host::set_error(std::string(#x) + ": " + hipGetErrorString(e_) + " @harm_mlp.hip:" + \
                                            std::to_string(__LINE__),                                      \
                                        2);                                                                \
    } while (0)

constexpr int KCHUNK = 256;  // split-K length of the weight-gradient products

// C(m,n) = sum_k A(m,k) B(k,n), strided operands; epilogue:
//   0: C = acc   1: C = acc + bias[n]   2: C = tanh(acc + bias[n])
//   3: C = acc * (1 - H(m,n)^2)
// grid = (ceil(N/16), ceil(M/16), splits), 64 threads; split z covers
// k in [z*kchunk, min(K, (z+1)*kchunk)) and writes C + z*cz.
struct GemmArgs {
    int M, N, K, kchunk, epi;
    const float *A;
    long sam, sak;
    const float *B;
    long sbk, sbn;
    float *C;
    long scm, scn, cz;
    const float *bias;
    const float *H;
    long shm, shn;
};

__global__ void __launch_bounds__(64) k_gemm(GemmArgs a)
{
    const int lane = threadIdx.x, r = lane & 15, q = lane >> 4;
    const int m0 = blockIdx.y * 16, n0 = blockIdx.x * 16;
    const int kb = blockIdx.z * a.kchunk;
    const int ke = min(a.K, kb + a.kchunk);
    const int mr = m0 + r, nr = n0 + r;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k = kb; k < ke; k += 4) {
        const int kk = k + q;
        const float av = (mr < a.M && kk < ke) ? a.A[mr * a.sam + kk * a.sak] : 0.f;
        const float bv = (nr < a.N && kk < ke) ? a.B[kk * a.sbk + nr * a.sbn] : 0.f;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
    }
    float *C = a.C + (size_t)blockIdx.z * a.cz;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int m = m0 + 4 * q + i, n = nr;
        if (m < a.M && n < a.N) {
            float v = acc[i];
            if (a.epi == 1 || a.epi == 2) v += a.bias[n];
            if (a.epi == 2) v = tanhf(v);
            if (a.epi == 3) {
                const float h = a.H[m * a.shm + n * a.shn];
                v *= (1.f - h * h);
            }
            C[m * a.scm + n * a.scn] = v;
        }
    }
}

// out[i] = sum_z part[z*stride + i], z = 0..nsplit-1 in order.
__global__ void k_reduce_splits(const float *__restrict__ part, int nsplit, long stride, int count,
                                float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int z = 0; z < nsplit; z++) s += part[(size_t)z * stride + i];
    out[i] = s;
}

__global__ void k_fill(float *__restrict__ p, float v, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// dO = mask (O - T) / B; per-block partial sums of mask |O - T|^2.
__global__ void __launch_bounds__(256) k_loss(const float *__restrict__ O, const float *__restrict__ T,
                                              const uint8_t *__restrict__ M, int B, float lw0, float lw1,
                                              float *__restrict__ dO, float *__restrict__ lossp)
{
    __shared__ float red[256];
    const int b = blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    if (b < B) {
        const float w = M ? (float)M[b] : 1.f;
        const float e0 = O[2 * b] - T[2 * b], e1 = O[2 * b + 1] - T[2 * b + 1];
        dO[2 * b] = w * lw0 * e0 / (float)B;
        dO[2 * b + 1] = w * lw1 * e1 / (float)B;
        l = w * (lw0 * e0 * e0 + lw1 * e1 * e1);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, int n, float lr, float bc1, float bc2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    p[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// Packs the BF16 A operands and the bias table of harm_mlp.hpp from the
// master parameters: one thread per (chunk, lane).
__global__ void k_pack(const float *__restrict__ P, uint4 *__restrict__ wpack, float *__restrict__ bias)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < MLP_CHUNKS * 64) {
        const int chunk = t / 64, lane = t % 64, row = lane & 15, gq = lane >> 4;
        bf16x8_t v;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            float w = 0.f;
            if (chunk < 4) {
                const int o = 16 * chunk + row;
                if (gq < 2 && i < 4) w = P[MLP_W1 + o * MLP_NIN + 4 * gq + i];
            } else if (chunk < 20) {
                const int base = (chunk < 12) ? MLP_W2 : MLP_W3;
                const int rel = (chunk < 12) ? chunk - 4 : chunk - 12;
                const int mt = rel / 2, c = rel % 2;
                w = P[base + (16 * mt + row) * MLP_NH + mlp_perm(c, gq, i)];
            } else {
                const int c = chunk - 20;
                if (row < MLP_NOUT) w = P[MLP_W4 + row * MLP_NH + mlp_perm(c, gq, i)];
            }
            v[i] = (__bf16)w;
        }
        wpack[t] = __builtin_bit_cast(uint4, v);
    }
    if (t < MLP_BIAS_LEN) {
        float b = 0.f;
        if (t < MLP_NH)
            b = P[MLP_B1 + t];
        else if (t < 2 * MLP_NH)
            b = P[MLP_B2 + t - MLP_NH];
        else if (t < 3 * MLP_NH)
            b = P[MLP_B3 + t - 2 * MLP_NH];
        else if (t < 3 * MLP_NH + MLP_NOUT)
            b = P[MLP_B4 + t - 3 * MLP_NH];
        bias[t] = b;
    }
}

// splitmix64 finalizer: the row drawn for (seed, step, b).
__device__ __forceinline__ unsigned long long mix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void __launch_bounds__(256) k_gather(const double *__restrict__ feat, const double *__restrict__ targ,
                                                const uint8_t *__restrict__ valid, size_t n, const double *norm,
                                                unsigned long long seed, int B, float *__restrict__ X,
                                                float *__restrict__ T, uint8_t *__restrict__ M)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    const size_t r = (size_t)(mix64(seed ^ mix64((unsigned long long)b)) % n);
#pragma unroll
    for (int k = 0; k < MLP_NIN; k++) X[b * MLP_NIN + k] = (float)((feat[k * n + r] - norm[k]) * norm[8 + k]);
    const bool ok = valid ? valid[r] != 0 : true;
    T[2 * b] = ok ? (float)((targ[r] - norm[16]) / norm[18]) : 0.f;
    T[2 * b + 1] = ok ? (float)((targ[n + r] - norm[17]) / norm[19]) : 0.f;
    M[b] = ok ? 1 : 0;
}

// Partial sums over valid rows for the normalization: pass 0 sums x, pass 1
// sums (x - mean)^2; 10 variables (8 features, 2 targets) and the row count.
__global__ void __launch_bounds__(256) k_norm_partials(const double *__restrict__ feat,
                                                       const double *__restrict__ targ,
                                                       const uint8_t *__restrict__ valid, size_t n, int pass,
                                                       const double *__restrict__ mean, double *__restrict__ part)
{
    __shared__ double red[11][256];
    double acc[11];
#pragma unroll
    for (int k = 0; k < 11; k++) acc[k] = 0.;
    for (size_t r = (size_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (size_t)gridDim.x * 256) {
        if (valid && !valid[r]) continue;
#pragma unroll
        for (int k = 0; k < 10; k++) {
            const double x = (k < 8) ? feat[k * n + r] : targ[(k - 8) * n + r];
            const double d = pass ? x - mean[k] : x;
            acc[k] += pass ? d * d : d;
        }
        acc[10] += 1.;
    }
#pragma unroll
    for (int k = 0; k < 11; k++) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int k = 0; k < 11; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 11) part[blockIdx.x * 11 + threadIdx.x] = red[threadIdx.x][0];
}

constexpr int NORM_BLOCKS = 512;

// Single thread: fold the partials in block order into the mean (pass 0) or
// into the normalization block (pass 1).
__global__ void k_norm_finish(const double *__restrict__ part, int pass, double *__restrict__ mean,
                              double *__restrict__ norm)
{
    double s[11];
    for (int k = 0; k < 11; k++) s[k] = 0.;
    for (int b = 0; b < NORM_BLOCKS; b++)
        for (int k = 0; k < 11; k++) s[k] += part[b * 11 + k];
    const double cnt = s[10] > 0. ? s[10] : 1.;
    if (pass == 0) {
        for (int k = 0; k < 10; k++) mean[k] = s[k] / cnt;
        return;
    }
    for (int k = 0; k < 10; k++) {
        double sd = sqrt(s[k] / cnt);
        if (!(sd > 1e-12)) sd = 1.;
        if (k < 8) {
            norm[k] = mean[k];
            norm[8 + k] = 1. / sd;
        } else {
            norm[16 + (k - 8)] = mean[k];
            norm[18 + (k - 8)] = sd;
        }
    }
}

// BF16 inference on standardized inputs (rows beyond n are zero inputs).
__global__ void __launch_bounds__(256) k_forward_bf16(MlpDev m, const float *__restrict__ X, float *__restrict__ Y,
                                                      int n)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    float x[MLP_NIN];
#pragma unroll
    for (int k = 0; k < MLP_NIN; k++) x[k] = (b < n) ? X[b * MLP_NIN + k] : 0.f;
    float y0, y1;
    mlp_forward_wave(m, x, y0, y1);
    if (b < n) {
        Y[2 * b] = y0;
        Y[2 * b + 1] = y1;
    }
}

// Features and seeds of a zone batch (SoA as harm_utoprim2d_batch).
__global__ void __launch_bounds__(256) k_seed(MlpDev m, const double *__restrict__ U, const double *__restrict__ gcov,
                                              const double *__restrict__ gcon, const double *__restrict__ gdet,
                                              size_t n, double *__restrict__ feat, double *__restrict__ W0,
                                              double *__restrict__ vsq0)
{
    const size_t z = (size_t)blockIdx.x * 256 + threadIdx.x;
    const bool act = z < n;
    double f[MLP_NIN], D = 1.;
    bool ok = false;
    if (act) {
        Geom geom;
#pragma unroll
        for (int a = 0; a < 16; a++) {
            geom.gcov[a / 4][a % 4] = gcov[(size_t)a * n + z];
            geom.gcon[a / 4][a % 4] = gcon[(size_t)a * n + z];
        }
        geom.g = gdet[z];
        geom_derive(geom);
        double Uz[NPR];
#pragma unroll
        for (int k = 0; k < NPR; k++) Uz[k] = U[(size_t)k * n + z];
        ok = mlp_features_U(Uz, geom, f);
        if (ok) D = mlp_wref(f);
        if (feat)
#pragma unroll
            for (int k = 0; k < MLP_NIN; k++) feat[(size_t)k * n + z] = f[k];
    }
    if (!W0) return;  // uniform over the grid: no MFMA below
    float x[MLP_NIN];
    if (ok)
        mlp_standardize(m, f, x);
    else
#pragma unroll
        for (int k = 0; k < MLP_NIN; k++) x[k] = 0.f;
    float y0, y1;
    mlp_forward_wave(m, x, y0, y1);
    if (act) {
        double w, v;
        mlp_unstandardize(m, y0, y1, D, w, v);
        W0[z] = ok ? w : 0.;
        if (vsq0) vsq0[z] = ok ? v : 0.;
    }
}

// Training pairs from solved zones: features of U, targets from the converged
// primitives (W = (rho + gam u) (1 + utsq), v^2 = utsq / (1 + utsq)).
__global__ void __launch_bounds__(256) k_pairs(const double *__restrict__ U, const double *__restrict__ gcov,
                                               const double *__restrict__ gcon, const double *__restrict__ gdet,
                                               const double *__restrict__ prim, const int8_t *__restrict__ ret,
                                               size_t n, double gam, double *__restrict__ feat,
                                               double *__restrict__ targ, uint8_t *__restrict__ valid)
{
    const size_t z = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (z >= n) return;
    Geom geom;
#pragma unroll
    for (int a = 0; a < 16; a++) {
        geom.gcov[a / 4][a % 4] = gcov[(size_t)a * n + z];
        geom.gcon[a / 4][a % 4] = gcon[(size_t)a * n + z];
    }
    geom.g = gdet[z];
    geom_derive(geom);
    double Uz[NPR], pr[NPR], f[MLP_NIN];
#pragma unroll
    for (int k = 0; k < NPR; k++) {
        Uz[k] = U[(size_t)k * n + z];
        pr[k] = prim[(size_t)k * n + z];
    }
    const bool ok = mlp_features_U(Uz, geom, f) && ret[z] == 0 && mlp_wref(f) > 0.;
    double t0 = 0., t1 = 0.;
    if (ok) {
        const double utsq = utsq_of(geom, pr);
        t0 = log((pr[RHO] + gam * pr[UU]) * (1. + utsq)) - log(mlp_wref(f));
        t1 = utsq / (1. + utsq);
    }
#pragma unroll
    for (int k = 0; k < MLP_NIN; k++) feat[(size_t)k * n + z] = f[k];
    targ[z] = t0;
    targ[n + z] = t1;
    valid[z] = ok ? 1 : 0;
}

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

template <class T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return 0;
    MCHK(hipMalloc((void **)p, n * sizeof(T)));
    return 0;
}

int launch_gemm(harm_mlp *m, const GemmArgs &a, int splits)
{
    hipLaunchKernelGGL(k_gemm, dim3(cdiv(a.N, 16), cdiv(a.M, 16), splits), dim3(64), 0, m->stream, a);
    MCHK(hipGetLastError());
    return 0;
}

GemmArgs gemm(int M, int N, int K, const float *A, long sam, long sak, const float *B, long sbk, long sbn, float *C,
              long scm, long scn)
{
    GemmArgs a;
    std::memset(&a, 0, sizeof(a));
    a.M = M;
    a.N = N;
    a.K = K;
    a.kchunk = K;
    a.A = A;
    a.sam = sam;
    a.sak = sak;
    a.B = B;
    a.sbk = sbk;
    a.sbn = sbn;
    a.C = C;
    a.scm = scm;
    a.scn = scn;
    return a;
}

// Weight gradient dW[Mo][Ni] = dZ^T H over B rows (dZ [B][Mo], H [B][Ni]) and
// bias gradient db[Mo] = dZ^T 1, split-K then a fixed-order reduction.
int weight_grad(harm_mlp *m, int B, int Mo, int Ni, const float *dZ, const float *H, float *dW, float *db)
{
    const int splits = (int)cdiv(B, KCHUNK);
    GemmArgs a = gemm(Mo, Ni, B, dZ, 1, Mo, H, Ni, 1, m->part, Ni, 1);
    a.kchunk = KCHUNK;
    a.cz = (long)Mo * Ni;
    int rc;
    if ((rc = launch_gemm(m, a, splits))) return rc;
    hipLaunchKernelGGL(k_reduce_splits, dim3(cdiv(Mo * Ni, 256)), dim3(256), 0, m->stream, m->part, splits,
                       (long)Mo * Ni, Mo * Ni, dW);
    MCHK(hipGetLastError());
    GemmArgs b = gemm(Mo, 1, B, dZ, 1, Mo, m->ones, 1, 0, m->part, 1, 1);
    b.kchunk = KCHUNK;
    b.cz = Mo;
    if ((rc = launch_gemm(m, b, splits))) return rc;
    hipLaunchKernelGGL(k_reduce_splits, dim3(1), dim3(256), 0, m->stream, m->part, splits, (long)Mo, Mo, db);
    MCHK(hipGetLastError());
    return 0;
}

int ensure_ws(harm_mlp *m, int B)
{
    if (B <= m->wsB) return 0;
    float **bufs[] = {&m->X, &m->T, &m->H1, &m->H2, &m->H3, &m->O, &m->dO, &m->dZa, &m->dZb, &m->ones, &m->part,
                      &m->lossp};
    for (auto p : bufs) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    if (m->M) (void)hipFree(m->M);
    m->M = nullptr;
    const size_t b = (size_t)B;
    int rc;
    if ((rc = dalloc(&m->X, b * MLP_NIN)) || (rc = dalloc(&m->T, b * MLP_NOUT)) || (rc = dalloc(&m->H1, b * MLP_NH)) ||
        (rc = dalloc(&m->H2, b * MLP_NH)) || (rc = dalloc(&m->H3, b * MLP_NH)) || (rc = dalloc(&m->O, b * MLP_NOUT)) ||
        (rc = dalloc(&m->dO, b * MLP_NOUT)) || (rc = dalloc(&m->dZa, b * MLP_NH)) ||
        (rc = dalloc(&m->dZb, b * MLP_NH)) || (rc = dalloc(&m->ones, b)) ||
        (rc = dalloc(&m->part, (size_t)cdiv(b, KCHUNK) * MLP_NH * MLP_NH)) || (rc = dalloc(&m->lossp, cdiv(b, 256))) ||
        (rc = dalloc(&m->M, b)))
        return rc;
    hipLaunchKernelGGL(k_fill, dim3(cdiv(b, 256)), dim3(256), 0, m->stream, m->ones, 1.f, B);
    MCHK(hipGetLastError());
    m->wsB = B;
    return 0;
}

// Forward of the FP32 training path on X [B][8] into H1..H3, O.
int forward_f32(harm_mlp *m, const float *X, int B)
{
    const float *P = m->params;
    int rc;
    GemmArgs a = gemm(B, MLP_NH, MLP_NIN, X, MLP_NIN, 1, P + MLP_W1, 1, MLP_NIN, m->H1, MLP_NH, 1);
    a.epi = 2;
    a.bias = P + MLP_B1;
    if ((rc = launch_gemm(m, a, 1))) return rc;
    a = gemm(B, MLP_NH, MLP_NH, m->H1, MLP_NH, 1, P + MLP_W2, 1, MLP_NH, m->H2, MLP_NH, 1);
    a.epi = 2;
    a.bias = P + MLP_B2;
    if ((rc = launch_gemm(m, a, 1))) return rc;
    a = gemm(B, MLP_NH, MLP_NH, m->H2, MLP_NH, 1, P + MLP_W3, 1, MLP_NH, m->H3, MLP_NH, 1);
    a.epi = 2;
    a.bias = P + MLP_B3;
    if ((rc = launch_gemm(m, a, 1))) return rc;
    a = gemm(B, MLP_NOUT, MLP_NH, m->H3, MLP_NH, 1, P + MLP_W4, 1, MLP_NH, m->O, MLP_NOUT, 1);
    a.epi = 1;
    a.bias = P + MLP_B4;
    return launch_gemm(m, a, 1);
}

// Forward, loss and gradient into m->grad for B standardized rows.
int loss_grad(harm_mlp *m, const float *X, const float *T, const uint8_t *M, int B)
{
    int rc;
    if ((rc = forward_f32(m, X, B))) return rc;
    hipLaunchKernelGGL(k_loss, dim3(cdiv(B, 256)), dim3(256), 0, m->stream, m->O, T, M, B, m->loss_w[0],
                       m->loss_w[1], m->dO, m->lossp);
    MCHK(hipGetLastError());
    const float *P = m->params;
    float *G = m->grad;
    // layer 4
    if ((rc = weight_grad(m, B, MLP_NOUT, MLP_NH, m->dO, m->H3, G + MLP_W4, G + MLP_B4))) return rc;
    GemmArgs a = gemm(B, MLP_NH, MLP_NOUT, m->dO, MLP_NOUT, 1, P + MLP_W4, MLP_NH, 1, m->dZa, MLP_NH, 1);
    a.epi = 3;
    a.H = m->H3;
    a.shm = MLP_NH;
    a.shn = 1;
    if ((rc = launch_gemm(m, a, 1))) return rc;
    // layer 3
    if ((rc = weight_grad(m, B, MLP_NH, MLP_NH, m->dZa, m->H2, G + MLP_W3, G + MLP_B3))) return rc;
    a = gemm(B, MLP_NH, MLP_NH, m->dZa, MLP_NH, 1, P + MLP_W3, MLP_NH, 1, m->dZb, MLP_NH, 1);
    a.epi = 3;
    a.H = m->H2;
    a.shm = MLP_NH;
    a.shn = 1;
    if ((rc = launch_gemm(m, a, 1))) return rc;
    // layer 2
    if ((rc = weight_grad(m, B, MLP_NH, MLP_NH, m->dZb, m->H1, G + MLP_W2, G + MLP_B2))) return rc;
    a = gemm(B, MLP_NH, MLP_NH, m->dZb, MLP_NH, 1, P + MLP_W2, MLP_NH, 1, m->dZa, MLP_NH, 1);
    a.epi = 3;
    a.H = m->H1;
    a.shm = MLP_NH;
    a.shn = 1;
    if ((rc = launch_gemm(m, a, 1))) return rc;
    // layer 1
    return weight_grad(m, B, MLP_NH, MLP_NIN, m->dZa, X, G + MLP_W1, G + MLP_B1);
}

int adam_step(harm_mlp *m, float lr)
{
    m->adam_t++;
    const float bc1 = (float)(1. - std::pow(0.9, (double)m->adam_t));
    const float bc2 = (float)(1. - std::pow(0.999, (double)m->adam_t));
    hipLaunchKernelGGL(k_adam, dim3(cdiv(MLP_NPARAMS, 256)), dim3(256), 0, m->stream, m->params, m->adam_m, m->adam_v,
                       m->grad, MLP_NPARAMS, lr, bc1, bc2);
    MCHK(hipGetLastError());
    return 0;
}

int repack(harm_mlp *m)
{
    hipLaunchKernelGGL(k_pack, dim3(cdiv(MLP_CHUNKS * 64, 256)), dim3(256), 0, m->stream, m->params, m->wpack,
                       m->bias);
    MCHK(hipGetLastError());
    return 0;
}

double sum_loss(harm_mlp *m, int B, int &rc)
{
    std::vector<float> lp(cdiv(B, 256));
    rc = 0;
    if (hipMemcpyAsync(lp.data(), m->lossp, lp.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream) !=
            hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess) {
        rc = harm_host::set_error("loss copy failed", 2);
        return 0.;
    }
    double s = 0.;
    for (float v : lp) s += v;
    return s / (2. * B);
}

struct DevGuard {
    int prev = -1;
    explicit DevGuard(int d)
    {
        (void)hipGetDevice(&prev);
        if (prev != d) (void)hipSetDevice(d);
    }
    ~DevGuard()
    {
        int cur;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

int check_batch(int n)
{
    if (n <= 0 || n > (1 << 24)) return harm_host::set_error("batch size must be in [1, 2^24]", 4);
    return 0;
}

}  // namespace

harm::MlpDev harm_mlp_view(const harm_mlp *m)
{
    harm::MlpDev d;
    d.wpack = m->wpack;
    d.bias = m->bias;
    d.norm = m->norm;
    d.vsq_from_W = m->vsq_from_W;
    return d;
}
int harm_mlp_device(const harm_mlp *m) { return m->device; }

extern "C" {

harm_mlp *harm_mlp_create(int device)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        harm_host::set_error("no HIP device available: libharmcool.so runs only on a gfx950 GPU", 2);
        return nullptr;
    }
    if (device >= 0 && hipSetDevice(device) != hipSuccess) {
        harm_host::set_error("hipSetDevice failed", 2);
        return nullptr;
    }
    harm_mlp *m = new harm_mlp();
    std::memset(m, 0, sizeof(*m));
    (void)hipGetDevice(&m->device);
    m->vsq_from_W = 1;
    m->loss_w[0] = m->loss_w[1] = 1.f;
    bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && !dalloc(&m->params, MLP_NPARAMS) && !dalloc(&m->adam_m, MLP_NPARAMS) &&
         !dalloc(&m->adam_v, MLP_NPARAMS) && !dalloc(&m->grad, MLP_NPARAMS) && !dalloc(&m->wpack, MLP_CHUNKS * 64) &&
         !dalloc(&m->bias, MLP_BIAS_LEN) && !dalloc(&m->norm, MLP_NORM_LEN) && !dalloc(&m->normp, NORM_BLOCKS * 11 + 16);
    if (!ok) {
        harm_mlp_destroy(m);
        return nullptr;
    }
    std::vector<double> nrm(MLP_NORM_LEN, 0.);
    for (int k = 0; k < 8; k++) nrm[8 + k] = 1.;
    nrm[18] = nrm[19] = 1.;
    if (harm_mlp_set_norm(m, nrm.data()) || harm_mlp_init(m, 1)) {
        harm_mlp_destroy(m);
        return nullptr;
    }
    return m;
}

void harm_mlp_destroy(harm_mlp *m)
{
    if (!m) return;
    DevGuard dg(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void *bufs[] = {m->params, m->adam_m, m->adam_v, m->grad, m->wpack, m->bias, m->norm, m->normp,
                    m->X,      m->T,      m->H1,     m->H2,   m->H3,    m->O,    m->dO,   m->dZa,
                    m->dZb,    m->ones,   m->part,   m->lossp, m->M};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int harm_mlp_init(harm_mlp *m, uint64_t seed)
{
    if (!m) return harm_host::set_error("null network", 1);
    std::vector<float> P(MLP_NPARAMS, 0.f);
    unsigned long long s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int nout, int nin) {
        const double a = std::sqrt(3.0 / nin);
        for (int i = 0; i < nout * nin; i++) P[off + i] = (float)((2. * next() - 1.) * a);
    };
    fill(MLP_W1, MLP_NH, MLP_NIN);
    fill(MLP_W2, MLP_NH, MLP_NH);
    fill(MLP_W3, MLP_NH, MLP_NH);
    fill(MLP_W4, MLP_NOUT, MLP_NH);
    return harm_mlp_set_params(m, P.data());
}

int harm_mlp_set_params(harm_mlp *m, const float *params)
{
    if (!m || !params) return harm_host::set_error("null argument", 1);
    DevGuard dg(m->device);
    MCHK(hipMemcpyAsync(m->params, params, MLP_NPARAMS * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MCHK(hipMemsetAsync(m->adam_m, 0, MLP_NPARAMS * sizeof(float), m->stream));
    MCHK(hipMemsetAsync(m->adam_v, 0, MLP_NPARAMS * sizeof(float), m->stream));
    m->adam_t = 0;
    int rc = repack(m);
    if (rc) return rc;
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int harm_mlp_get_params(harm_mlp *m, float *params)
{
    if (!m || !params) return harm_host::set_error("null argument", 1);
    DevGuard dg(m->device);
    MCHK(hipMemcpyAsync(params, m->params, MLP_NPARAMS * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int harm_mlp_set_norm(harm_mlp *m, const double *norm)
{
    if (!m || !norm) return harm_host::set_error("null argument", 1);
    for (int k = 8; k < 16; k++)
        if (!(norm[k] > 0.) || !std::isfinite(norm[k])) return harm_host::set_error("feature 1/std must be > 0", 4);
    if (!(norm[18] > 0.) || !(norm[19] > 0.)) return harm_host::set_error("output std must be > 0", 4);
    DevGuard dg(m->device);
    MCHK(hipMemcpyAsync(m->norm, norm, MLP_NORM_LEN * sizeof(double), hipMemcpyHostToDevice, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int harm_mlp_get_norm(harm_mlp *m, double *norm)
{
    if (!m || !norm) return harm_host::set_error("null argument", 1);
    DevGuard dg(m->device);
    MCHK(hipMemcpyAsync(norm, m->norm, MLP_NORM_LEN * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int harm_mlp_set_vsq_from_w(harm_mlp *m, int on)
{
    if (!m) return harm_host::set_error("null network", 1);
    m->vsq_from_W = on ? 1 : 0;
    return 0;
}

int harm_mlp_set_loss_weights(harm_mlp *m, float w_logw, float w_vsq)
{
    if (!m) return harm_host::set_error("null network", 1);
    if (!(w_logw >= 0.f) || !(w_vsq >= 0.f)) return harm_host::set_error("loss weights must be >= 0", 4);
    m->loss_w[0] = w_logw;
    m->loss_w[1] = w_vsq;
    return 0;
}

int harm_mlp_fit_norm(harm_mlp *m, const double *feat, const double *targ, const uint8_t *valid, size_t n)
{
    if (!m || !feat || !targ) return harm_host::set_error("null argument", 1);
    if (n == 0) return harm_host::set_error("empty training set", 4);
    DevGuard dg(m->device);
    double *mean = m->normp + NORM_BLOCKS * 11;
    for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(k_norm_partials, dim3(NORM_BLOCKS), dim3(256), 0, m->stream, feat, targ, valid, n, pass,
                           mean, m->normp);
        MCHK(hipGetLastError());
        hipLaunchKernelGGL(k_norm_finish, dim3(1), dim3(1), 0, m->stream, m->normp, pass, mean, m->norm);
        MCHK(hipGetLastError());
    }
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int harm_mlp_grad(harm_mlp *m, const float *X, const float *T, const uint8_t *mask, int n, float *grad, double *loss)
{
    if (!m || !X || !T) return harm_host::set_error("null argument", 1);
    int rc;
    if ((rc = check_batch(n))) return rc;
    DevGuard dg(m->device);
    if ((rc = ensure_ws(m, n))) return rc;
    if ((rc = loss_grad(m, X, T, mask, n))) return rc;
    const double l = sum_loss(m, n, rc);
    if (rc) return rc;
    if (loss) *loss = l;
    if (grad) {
        MCHK(hipMemcpyAsync(grad, m->grad, MLP_NPARAMS * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MCHK(hipStreamSynchronize(m->stream));
    }
    return 0;
}

int harm_mlp_train_batch(harm_mlp *m, const float *X, const float *T, const uint8_t *mask, int n, float lr)
{
    if (!m || !X || !T) return harm_host::set_error("null argument", 1);
    int rc;
    if ((rc = check_batch(n))) return rc;
    DevGuard dg(m->device);
    if ((rc = ensure_ws(m, n)) || (rc = loss_grad(m, X, T, mask, n)) || (rc = adam_step(m, lr)) || (rc = repack(m)))
        return rc;
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int harm_mlp_train(harm_mlp *m, const double *feat, const double *targ, const uint8_t *valid, size_t n, int batch,
                   int steps, float lr0, float lr1, uint64_t seed, float *loss_trace)
{
    if (!m || !feat || !targ) return harm_host::set_error("null argument", 1);
    if (n == 0) return harm_host::set_error("empty training set", 4);
    int rc;
    if ((rc = check_batch(batch))) return rc;
    if (steps < 0 || !(lr0 > 0.f) || !(lr1 > 0.f)) return harm_host::set_error("bad training schedule", 4);
    DevGuard dg(m->device);
    if ((rc = ensure_ws(m, batch))) return rc;
    std::vector<float> lp;
    for (int s = 0; s < steps; s++) {
        const double fr = steps > 1 ? (double)s / (steps - 1) : 0.;
        const float lr = (float)(lr0 * std::pow((double)lr1 / lr0, fr));
        const unsigned long long sd = (unsigned long long)seed * 0x100000001B3ull + (unsigned long long)s;
        hipLaunchKernelGGL(k_gather, dim3(cdiv(batch, 256)), dim3(256), 0, m->stream, feat, targ, valid, n, m->norm,
                           sd, batch, m->X, m->T, m->M);
        MCHK(hipGetLastError());
        if ((rc = loss_grad(m, m->X, m->T, m->M, batch)) || (rc = adam_step(m, lr))) return rc;
        if (loss_trace) {
            loss_trace[s] = (float)sum_loss(m, batch, rc);
            if (rc) return rc;
        }
    }
    if ((rc = repack(m))) return rc;
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int harm_mlp_forward(harm_mlp *m, const float *X, float *Y, int n, int bf16)
{
    if (!m || !X || !Y) return harm_host::set_error("null argument", 1);
    int rc;
    if ((rc = check_batch(n))) return rc;
    DevGuard dg(m->device);
    if (bf16) {
        hipLaunchKernelGGL(k_forward_bf16, dim3(cdiv(n, 256)), dim3(256), 0, m->stream, harm_mlp_view(m), X, Y, n);
        MCHK(hipGetLastError());
    } else {
        if ((rc = ensure_ws(m, n)) || (rc = forward_f32(m, X, n))) return rc;
        MCHK(hipMemcpyAsync(Y, m->O, (size_t)n * MLP_NOUT * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
    }
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int harm_mlp_seed(harm_mlp *m, const double *U, const double *gcov, const double *gcon, const double *gdet, size_t n,
                  double *feat, double *W0, double *vsq0, void *stream)
{
    if (!m || !U || !gcov || !gcon || !gdet) return harm_host::set_error("null argument", 1);
    if (n == 0) return 0;
    if (n > ((size_t)1 << 31)) return harm_host::set_error("batch too large", 4);
    DevGuard dg(m->device);
    hipStream_t st = stream ? (hipStream_t)stream : m->stream;
    hipLaunchKernelGGL(k_seed, dim3(cdiv(n, 256)), dim3(256), 0, st, harm_mlp_view(m), U, gcov, gcon, gdet, n, feat,
                       W0, vsq0);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(st));
    return 0;
}

int harm_mlp_pairs(const double *U, const double *gcov, const double *gcon, const double *gdet, const double *prim,
                   const int8_t *ret, size_t n, double gam, double *feat, double *targ, uint8_t *valid, void *stream)
{
    if (!U || !gcov || !gcon || !gdet || !prim || !ret || !feat || !targ || !valid)
        return harm_host::set_error("null argument", 1);
    if (n == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_pairs, dim3(cdiv(n, 256)), dim3(256), 0, st, U, gcov, gcon, gdet, prim, ret, n, gam, feat,
                       targ, valid);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
