// harm_mlp.hpp -- learned Newton seed for the Noble 2D inversion: input
// features, parameter layout, and the BF16 MFMA forward pass of one wavefront.
//
// No counterpart in the reference: the seed replaces the guess block of
// Utoprim_2d (utoprim_2d.c:268-307) and nothing else; the FP64 Newton solve
// (utoprim_2d.c:442-541) refines it, and the previous-step guess is the
// fallback when the seeded solve fails.
//
// Network: 8 inputs -> 64 tanh -> 64 tanh -> 64 tanh -> 2 outputs
// (log(W/Wref), v^2; Wref = E - B^2/2, mlp_wref), all standardized by
// per-feature mean/std fitted on the training pairs.  FP32 master parameters, row-major W[out][in]:
//   W1 64x8 | b1 64 | W2 64x64 | b2 64 | W3 64x64 | b3 64 | W4 2x64 | b4 2.
//
// Inference runs one wavefront over 64 zones as four 16-zone tiles with
// v_mfma_f32_16x16x32_bf16: activations are kept transposed (rows = neurons,
// columns = zones) so that the D layout of one layer (lane l holds rows
// 4*(l/16)+r of column l%16) is, after tanh and a BF16 cast, the B operand of
// the next one; the contraction order inside each 32-wide chunk is permuted to
// match, and the packed A operands (weights) carry the same permutation.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace harm {

constexpr int MLP_NIN = 8, MLP_NH = 64, MLP_NOUT = 2;
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + MLP_NH * MLP_NIN;
constexpr int MLP_W2 = MLP_B1 + MLP_NH;
constexpr int MLP_B2 = MLP_W2 + MLP_NH * MLP_NH;
constexpr int MLP_W3 = MLP_B2 + MLP_NH;
constexpr int MLP_B3 = MLP_W3 + MLP_NH * MLP_NH;
constexpr int MLP_W4 = MLP_B3 + MLP_NH;
constexpr int MLP_B4 = MLP_W4 + MLP_NOUT * MLP_NH;
constexpr int MLP_NPARAMS = MLP_B4 + MLP_NOUT;  // 9026

// Packed A operands: 64 lanes x 8 BF16 per MFMA.  Chunk order: layer 1 m=0..3;
// layer 2 (m, c) at 4 + 2m + c; layer 3 at 12 + 2m + c; layer 4 c at 20 + c.
constexpr int MLP_CHUNKS = 22;
// Bias table in accumulator order: b1[64] b2[64] b3[64] b4[2] + 14 zeros.
constexpr int MLP_BIAS_LEN = 3 * MLP_NH + 16;
// Normalization block: feature mean[8], 1/std[8], output mean[2], std[2].
constexpr int MLP_NORM_LEN = 20;

struct MlpDev {
    const uint4 *wpack;   // [MLP_CHUNKS][64] x 16 B
    const float *bias;    // [MLP_BIAS_LEN]
    const double *norm;   // [MLP_NORM_LEN]
    int vsq_from_W;       // seed v^2 from W (utoprim_2d.c:395-406) instead of the second output
};

// Input neuron that slot i of lane group gq contracts in chunk c of a 64-wide
// layer: slots 0..3 come from D tile 2c, slots 4..7 from D tile 2c+1.
__host__ __device__ inline int mlp_perm(int c, int gq, int i)
{
    return 32 * c + (i < 4 ? 4 * gq + i : 16 + 4 * gq + (i - 4));
}

// The eight raw inputs of one zone, from the scalars of utoprim_2d.c:240-266
// (s, after the transform of utoprim_2d.c:145-163) and the lapse.
HD void mlp_features(const U2P &s, double lapse, double *f)
{
    const double D = s.D, E = -s.Qdotn;
    f[0] = log(D);
    f[1] = asinh(E / D);
    f[2] = asinh(s.Qtsq / (D * D));
    f[3] = asinh(s.Bsq / D);
    f[4] = asinh(s.QdotBsq / (D * D * D));
    f[5] = lapse;
    f[6] = asinh((E - D - 0.5 * s.Bsq) / D);
    f[7] = asinh(s.Bsq / E);
}

// Reference scale of the first output, Wref = E - B^2/2 = D (sinh f1 - sinh f3 / 2):
// W = Wref + p - B^2 v^2/2 + (Q.B)^2/(2W^2), so log(W/Wref) spans a far
// narrower range than log(W/D) (standard deviations 0.32 and 1.7 on the
// inversion stress distribution).  No seed where it is not positive.
HD double mlp_wref(const double *f) { return exp(f[0]) * (sinh(f[1]) - 0.5 * sinh(f[3])); }

// Raw features of HARM's conserved vector U (x gdet); false when U[0] <= 0.
HD bool mlp_features_U(const double *U, const Geom &geom, double *f)
{
    if (!(U[0] > 0.)) {
#pragma unroll
        for (int k = 0; k < MLP_NIN; k++) f[k] = 0.;
        return false;
    }
    double U_tmp[NPR], Bcon[4], Qcon[4], ncon[4], QdotB;
    u2p_transform(U, geom, U_tmp);
    U2P s;
    s.gam = 0.;
    s.K_atm = 0.;
    u2p_scalars(s, U_tmp, geom, Bcon, Qcon, ncon, QdotB);
    mlp_features(s, geom.alpha, f);
    return true;
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float mlp_tanh(float x)
{
    // tanh(x) = 1 - 2 / (exp(2x) + 1); saturates correctly at both ends.
    const float e = __expf(2.f * x);
    return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}

__device__ __forceinline__ bf16x8_t mlp_ld_chunk(const MlpDev &m, int chunk, int lane)
{
    const uint4 v = m.wpack[chunk * 64 + lane];
    return __builtin_bit_cast(bf16x8_t, v);
}

__device__ __forceinline__ f32x4_t mlp_ld_bias(const MlpDev &m, int off)
{
    const float4 b = *reinterpret_cast<const float4 *>(m.bias + off);
    f32x4_t r = {b.x, b.y, b.z, b.w};
    return r;
}

// One hidden 64 -> 64 layer on a 16-zone tile: acc = W h + b, then tanh and
// repack as the B operand of the next layer.
__device__ __forceinline__ void mlp_hidden(const MlpDev &m, int chunk0, int boff, int lane, bf16x8_t (&h)[2])
{
    const int gq = lane >> 4;
    f32x4_t acc[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        acc[mt] = mlp_ld_bias(m, boff + 16 * mt + 4 * gq);
#pragma unroll
        for (int c = 0; c < 2; c++)
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(mlp_ld_chunk(m, chunk0 + 2 * mt + c, lane), h[c],
                                                              acc[mt], 0, 0, 0);
    }
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            h[c][r] = (__bf16)mlp_tanh(acc[2 * c][r]);
            h[c][4 + r] = (__bf16)mlp_tanh(acc[2 * c + 1][r]);
        }
}

// Forward pass for the 64 zones of this wavefront.  x = standardized inputs
// of this lane's zone; returns the standardized outputs of this lane's zone.
// Every lane of the wavefront must call it (MFMA reads all 64 lanes).
__device__ __forceinline__ void mlp_forward_wave(const MlpDev &m, const float (&x)[MLP_NIN], float &y0, float &y1)
{
    const int lane = threadIdx.x & 63, gq = lane >> 4, n = lane & 15;
    float o0 = 0.f, o1 = 0.f;
#pragma unroll 1
    for (int t = 0; t < 4; t++) {
        const int src = 16 * t + n;
        bf16x8_t b;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float v0 = __shfl(x[i], src, 64);
            const float v1 = __shfl(x[4 + i], src, 64);
            const float v = (gq == 0) ? v0 : ((gq == 1) ? v1 : 0.f);
            b[i] = (__bf16)v;
            b[4 + i] = (__bf16)0.f;
        }
        bf16x8_t h[2];
        {
            f32x4_t acc[4];
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                acc[mt] = mlp_ld_bias(m, 16 * mt + 4 * gq);
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(mlp_ld_chunk(m, mt, lane), b, acc[mt], 0, 0, 0);
            }
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    h[c][r] = (__bf16)mlp_tanh(acc[2 * c][r]);
                    h[c][4 + r] = (__bf16)mlp_tanh(acc[2 * c + 1][r]);
                }
        }
        mlp_hidden(m, 4, MLP_NH, lane, h);
        mlp_hidden(m, 12, 2 * MLP_NH, lane, h);
        f32x4_t out = mlp_ld_bias(m, 3 * MLP_NH + 4 * gq);
#pragma unroll
        for (int c = 0; c < 2; c++)
            out = __builtin_amdgcn_mfma_f32_16x16x32_bf16(mlp_ld_chunk(m, 20 + c, lane), h[c], out, 0, 0, 0);
        // Lanes 0..15 hold rows 0..3 (outputs 0 and 1 in r = 0, 1) of zones 16t..16t+15.
        const float a0 = __shfl(out[0], n, 64);
        const float a1 = __shfl(out[1], n, 64);
        if (gq == t) {
            o0 = a0;
            o1 = a1;
        }
    }
    y0 = o0;
    y1 = o1;
}

// Standardize raw features for the network.
__device__ __forceinline__ void mlp_standardize(const MlpDev &m, const double *f, float (&x)[MLP_NIN])
{
#pragma unroll
    for (int k = 0; k < MLP_NIN; k++) x[k] = (float)((f[k] - m.norm[k]) * m.norm[8 + k]);
}

// Seed (W0, vsq0) from the standardized outputs; wref = mlp_wref(features).
__device__ __forceinline__ void mlp_unstandardize(const MlpDev &m, float y0, float y1, double wref, double &W0,
                                                  double &vsq0)
{
    W0 = (wref > 0.) ? wref * exp((double)y0 * m.norm[18] + m.norm[16]) : 0.;
    vsq0 = (double)y1 * m.norm[19] + m.norm[17];
}

}  // namespace harm

// Device view of a network handle (defined with the training kernels).
harm::MlpDev harm_mlp_view(const harm_mlp *m);
int harm_mlp_device(const harm_mlp *m);
