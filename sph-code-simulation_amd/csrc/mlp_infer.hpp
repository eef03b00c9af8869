// mlp_infer.hpp -- the learned Riemann starting guess as seen by the force kernel.
//
// Network: 6 -> 64 -> 64 -> 64 -> 1 with ReLU, parameters in one FP32 block
// (W1[64][6], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4;
// 8833 values).  The six inputs are the
// left/right states made dimensionless with the pair means:
//   rho/rho_m, p/p_m, (u - u_m)/c_m  with  c_m = sqrt(gamma p_m / rho_m),
// and the output is log(p*/p_m).  The Riemann problem is Galilean and scale
// invariant, so this normalisation loses nothing.
#pragma once

#include <hip/hip_runtime.h>

#include "device_math.hpp"

struct sphgpu_ctx;

namespace sphgpu {

constexpr int kMlpIn = 6;
constexpr int kMlpH = 64;
// parameter block offsets
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpH * kMlpH;
constexpr int kOffW4 = kOffB3 + kMlpH;
constexpr int kOffB4 = kOffW4 + kMlpH;
constexpr int kMlpParams = kOffB4 + 1;

struct MlpWeightsBF16;

struct MLPDev {
    const float *w;                // FP32 parameters (device), or null when untrained
    const MlpWeightsBF16 *img;     // their bf16 image for the MFMA inference
    float gamma;
};

MLPDev mlp_dev(sphgpu_ctx *c);
int riemann_refine64(sphgpu_ctx *c);

// features of one pair, shared by training and inference
SPH_DEV void mlp_features(float gamma, float uL, float rL, float pL, float uR, float rR, float pR, float x[6],
                          float &p_m)
{
    const float r_m = 0.5f * (rL + rR);
    p_m = 0.5f * (pL + pR);
    const float u_m = 0.5f * (uL + uR);
    const float c_m = __builtin_sqrtf(gamma * p_m / r_m);
    const float ic = 1.0f / c_m;
    x[0] = rL / r_m;
    x[1] = pL / p_m;
    x[2] = (uL - u_m) * ic;
    x[3] = rR / r_m;
    x[4] = pR / p_m;
    x[5] = (uR - u_m) * ic;
}

// per-lane evaluation (reference form of the network; FP32)
SPH_DEV float mlp_eval(const float *w, const float x[6])
{
    float h1[kMlpH], h2[kMlpH];
#pragma unroll 8
    for (int o = 0; o < kMlpH; ++o) {
        float s = w[kOffB1 + o];
        for (int k = 0; k < kMlpIn; ++k) s = __builtin_fmaf(w[kOffW1 + o * kMlpIn + k], x[k], s);
        h1[o] = s > 0.f ? s : 0.f;
    }
    for (int o = 0; o < kMlpH; ++o) {
        float s = w[kOffB2 + o];
        for (int k = 0; k < kMlpH; ++k) s = __builtin_fmaf(w[kOffW2 + o * kMlpH + k], h1[k], s);
        h2[o] = s > 0.f ? s : 0.f;
    }
    float out = w[kOffB4];
    for (int o = 0; o < kMlpH; ++o) {
        float s = w[kOffB3 + o];
        for (int k = 0; k < kMlpH; ++k) s = __builtin_fmaf(w[kOffW3 + o * kMlpH + k], h2[k], s);
        out = __builtin_fmaf(w[kOffW4 + o], s > 0.f ? s : 0.f, out);
    }
    return out;
}

SPH_DEV float mlp_guess(const MLPDev &m, float uL, float rL, float pL, float uR, float rR, float pR)
{
    float x[6], p_m;
    mlp_features(m.gamma, uL, rL, pL, uR, rR, pR, x, p_m);
    const float y = mlp_eval(m.w, x);
    return p_m * __expf(y);
}

// ---- wavefront-cooperative bf16 MFMA inference --------------------------------
// The 64 lanes of a wavefront each contribute one pair (row); the network runs
// as four small GEMMs on v_mfma_f32_16x16x32_bf16, one 16-row chunk at a
// time: z1 = X W1^T (K = 32, inputs zero-padded), z2 = relu(z1) W2^T (K = 64),
// z3 = relu(z2) W3^T (K = 64), out = relu(z3) w4 (w4 as column 0 of a 64x16
// B operand).  Tile lane layout
// (CDNA4): A[16][32] -- lane l holds row l%16, k = 8*(l/16)..+7; B[32][16] --
// lane l holds column l%16, same k; D[16][16] -- lane l holds rows
// 4*(l/16)..+3 of column l%16.
typedef __bf16 mlp_bf16;
typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

constexpr int kInfK1 = 32;
constexpr int kInfLd1 = kInfK1 + 8;
constexpr int kInfLd = kMlpH + 8;

struct MlpWeightsBF16 {           // bf16 image of the parameters (global memory)
    mlp_bf16 W1[kMlpH][kInfLd1];  // [out][in], inputs padded to 32
    mlp_bf16 W2[kMlpH][kInfLd];   // [out][in]
    mlp_bf16 W3[kMlpH][kInfLd];   // [out][in]
    mlp_bf16 W4[16][kInfLd];      // row 0 = w4, rows 1..15 zero
    float b1[kMlpH], b2[kMlpH], b3[kMlpH];
    float b4;
};

struct MlpWaveLDS {               // per wavefront scratch
    mlp_bf16 X[64][kInfLd1];      // features, row = lane
    mlp_bf16 H[16][kInfLd];       // relu(z1) or relu(z2) of one 16-row chunk
    float out[64];
};

// conversion of the FP32 parameters into the bf16 image (whole workgroup)
SPH_DEV void mlp_load_weights(MlpWeightsBF16 &L, const float *w)
{
    for (int e = threadIdx.x; e < kMlpH * kInfK1; e += blockDim.x) {
        const int o = e / kInfK1, k = e % kInfK1;
        L.W1[o][k] = (mlp_bf16)(k < kMlpIn ? w[kOffW1 + o * kMlpIn + k] : 0.f);
    }
    for (int e = threadIdx.x; e < kMlpH * kMlpH; e += blockDim.x) {
        const int o = e / kMlpH, k = e % kMlpH;
        L.W2[o][k] = (mlp_bf16)w[kOffW2 + o * kMlpH + k];
        L.W3[o][k] = (mlp_bf16)w[kOffW3 + o * kMlpH + k];
    }
    for (int e = threadIdx.x; e < 16 * kMlpH; e += blockDim.x) {
        const int r = e / kMlpH, k = e % kMlpH;
        L.W4[r][k] = (mlp_bf16)(r == 0 ? w[kOffW4 + k] : 0.f);
    }
    for (int e = threadIdx.x; e < kMlpH; e += blockDim.x) {
        L.b1[e] = w[kOffB1 + e];
        L.b2[e] = w[kOffB2 + e];
        L.b3[e] = w[kOffB3 + e];
    }
    if (threadIdx.x == 0) L.b4 = w[kOffB4];
}

// LDS hand-off between lanes of one wavefront: order the LDS operations and
// keep the compiler from moving accesses across
SPH_DEV void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// whole wavefront: the lanes with `use` set (at most 64) are packed into the
// first rows, only the 16-row chunks that hold them are evaluated, and each
// such lane gets the network output (log(p*/p_m)) for its features
SPH_DEV float mlp_wave_infer(const MlpWeightsBF16 &L, MlpWaveLDS &S, const float x[6], bool use)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long um = __ballot(use);
    const int row = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(um >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)um, 0u));
    const int nrows = __popcll(um);
    if (use)
        for (int k = 0; k < kInfK1; ++k) S.X[row][k] = (mlp_bf16)(k < kMlpIn ? x[k] : 0.f);
    const int r16 = lane & 15, kq = 8 * (lane >> 4);
    const int nchunks = (nrows + 15) >> 4;
    for (int chunk = 0; chunk < nchunks; ++chunk) {
        wave_lds_sync();
        // layer 1: z1 (16 x 64) = X[chunk] (16 x 32) W1^T
        const mlp_bf16x8 a1 = *reinterpret_cast<const mlp_bf16x8 *>(&S.X[chunk * 16 + r16][kq]);
        for (int tn = 0; tn < 4; ++tn) {
            const mlp_bf16x8 b = *reinterpret_cast<const mlp_bf16x8 *>(&L.W1[tn * 16 + r16][kq]);
            mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b, acc, 0, 0, 0);
            const int col = tn * 16 + r16;
            for (int i = 0; i < 4; ++i) S.H[4 * (lane >> 4) + i][col] = (mlp_bf16)fmaxf(acc[i] + L.b1[col], 0.f);
        }
        // layers 2 and 3: z (16 x 64) = H (16 x 64) W^T, relu(z + b) back into H
        for (int layer = 0; layer < 2; ++layer) {
            const mlp_bf16(*W)[kInfLd] = layer == 0 ? L.W2 : L.W3;
            const float *bias = layer == 0 ? L.b2 : L.b3;
            wave_lds_sync();
            mlp_f32x4 acc2[4];
            for (int tn = 0; tn < 4; ++tn) {
                acc2[tn] = mlp_f32x4{0.f, 0.f, 0.f, 0.f};
                for (int k0 = 0; k0 < kMlpH; k0 += 32) {
                    const mlp_bf16x8 a = *reinterpret_cast<const mlp_bf16x8 *>(&S.H[r16][k0 + kq]);
                    const mlp_bf16x8 b = *reinterpret_cast<const mlp_bf16x8 *>(&W[tn * 16 + r16][k0 + kq]);
                    acc2[tn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc2[tn], 0, 0, 0);
                }
            }
            wave_lds_sync();
            for (int tn = 0; tn < 4; ++tn) {
                const int col = tn * 16 + r16;
                for (int i = 0; i < 4; ++i)
                    S.H[4 * (lane >> 4) + i][col] = (mlp_bf16)fmaxf(acc2[tn][i] + bias[col], 0.f);
            }
        }
        wave_lds_sync();
        // output: out (16 x 16, column 0 used) = H (16 x 64) W4^T
        mlp_f32x4 acc3 = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < kMlpH; k0 += 32) {
            const mlp_bf16x8 a = *reinterpret_cast<const mlp_bf16x8 *>(&S.H[r16][k0 + kq]);
            const mlp_bf16x8 b = *reinterpret_cast<const mlp_bf16x8 *>(&L.W4[r16][k0 + kq]);
            acc3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc3, 0, 0, 0);
        }
        if (r16 == 0)
            for (int i = 0; i < 4; ++i) S.out[chunk * 16 + 4 * (lane >> 4) + i] = acc3[i] + L.b4;
    }
    wave_lds_sync();
    const float y = use ? S.out[row] : 0.f;
    wave_lds_sync();
    return y;
}

} // namespace sphgpu
