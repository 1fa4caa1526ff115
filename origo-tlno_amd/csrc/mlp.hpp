// Learned-screening MLP: 32 -> 64 -> 64 -> 1, ReLU, weights rounded to
// bfloat16 (round to nearest even) for every forward/backward pass, FP32
// accumulation, FP32 master weights and Adam state.  Predicts
// log10 max|(ab|cd)| of a shell quartet from 32 features (features.hpp).
#pragma once

#include <hip/hip_runtime.h>

namespace tlno_amd {

constexpr int MLP_IN = 32, MLP_H = 64;
// parameter layout (floats): W1[64][32], b1[64], W2[64][64], b2[64], w3[64], b3
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + 1;  // 6337
// normalization: x_n = (x - mean) / std, 32 + 32 floats after the parameters
constexpr int MLP_NALL = MLP_NPARAM + 2 * MLP_IN;

__host__ __device__ inline float bf16_round(float f) {
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    __builtin_memcpy(&r, &u, 4);
    return r;
}

// forward pass of one sample with bf16-valued weights w (MLP_NALL floats:
// parameters then mean/std), x raw features; returns the prediction
__device__ inline float mlp_forward(const float* __restrict__ w, const float* x) {
    float xn[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) xn[i] = (x[i] - w[MLP_NPARAM + i]) / w[MLP_NPARAM + MLP_IN + i];
    float h1[MLP_H];
#pragma unroll 4
    for (int j = 0; j < MLP_H; ++j) {
        float s = w[MLP_OFF_B1 + j];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) s = fmaf(w[MLP_OFF_W1 + j * MLP_IN + i], xn[i], s);
        h1[j] = fmaxf(s, 0.0f);
    }
    float out = w[MLP_OFF_B3];
#pragma unroll 2
    for (int j = 0; j < MLP_H; ++j) {
        float s = w[MLP_OFF_B2 + j];
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) s = fmaf(w[MLP_OFF_W2 + j * MLP_H + i], h1[i], s);
        out = fmaf(w[MLP_OFF_W3 + j], fmaxf(s, 0.0f), out);
    }
    return out;
}

}  // namespace tlno_amd
