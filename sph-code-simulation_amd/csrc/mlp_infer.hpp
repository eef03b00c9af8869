// mlp_infer.hpp -- the learned Riemann starting guess as seen by the force kernel.
//
// Network: 6 -> 64 -> 64 -> 1 with ReLU, parameters in one FP32 block
// (W1[64][6], b1[64], W2[64][64], b2[64], w3[64], b3).  The six inputs are the
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
constexpr int kOffB3 = kOffW3 + kMlpH;
constexpr int kMlpParams = kOffB3 + 1;

struct MLPDev {
    const float *w;    // FP32 parameters (device), or null when untrained
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
    float h1[kMlpH];
#pragma unroll 8
    for (int o = 0; o < kMlpH; ++o) {
        float s = w[kOffB1 + o];
        for (int k = 0; k < kMlpIn; ++k) s = __builtin_fmaf(w[kOffW1 + o * kMlpIn + k], x[k], s);
        h1[o] = s > 0.f ? s : 0.f;
    }
    float out = w[kOffB3];
    for (int o = 0; o < kMlpH; ++o) {
        float s = w[kOffB2 + o];
        for (int k = 0; k < kMlpH; ++k) s = __builtin_fmaf(w[kOffW2 + o * kMlpH + k], h1[k], s);
        out = __builtin_fmaf(w[kOffW3 + o], s > 0.f ? s : 0.f, out);
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

} // namespace sphgpu
