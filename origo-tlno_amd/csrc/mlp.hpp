// Learned shell-quartet screen: a 32 -> 64 -> 64 -> 1 ReLU network that
// predicts log10 max|(ab|cd)| - log10(Q_ab Q_cd) (<= 0 by Cauchy-Schwarz) from
// 32 features of the quartet.  Weights are bfloat16 values, activations
// feeding a weight product are rounded to bfloat16, every sum is FP32:
//   xb = bf16((x - mean) * inv_std)
//   h1 = bf16(relu(W1 xb + b1))
//   h2 = relu(W2 h1 + b2)
//   y  = w3 . h2 + b3
// Two evaluators of that function share this header: a per-lane FP32 one
// (training, tests) and a wave-level one on the gfx950 bf16 matrix cores
// (v_mfma_f32_32x32x16_bf16) used inside the screening kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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
// feature normalization after the parameters: mean[32], inv_std[32]
constexpr int MLP_OFF_MEAN = MLP_NPARAM;
constexpr int MLP_OFF_ISTD = MLP_NPARAM + MLP_IN;
constexpr int MLP_NALL = MLP_NPARAM + 2 * MLP_IN;  // 6401

// the weight matrices are bf16 values; biases stay FP32
__host__ __device__ inline bool mlp_is_weight(int i) {
    return (i >= MLP_OFF_W1 && i < MLP_OFF_B1) || (i >= MLP_OFF_W2 && i < MLP_OFF_B2) ||
           (i >= MLP_OFF_W3 && i < MLP_OFF_B3);
}

// round to the nearest bfloat16, ties to even (finite inputs)
__host__ __device__ inline float bf16_round(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    __builtin_memcpy(&r, &u, 4);
    return r;
}

__host__ __device__ inline uint32_t bf16_bits(float f) {  // upper half of bf16_round(f)
    const float r = bf16_round(f);
    uint32_t u;
    __builtin_memcpy(&u, &r, 4);
    return u >> 16;
}

// ---------------------------------------------------------------- features --
// Per shell pair (plan order), computed once per plan on the host:
//   f0 = {log10 Q, log10 p_min, log10 p_max, log10 mu}    (mu, P of the primitive pair with p = p_min)
//   f1 = {|AB|, mu |AB|^2, log10 sum|w|, log10 n_prim_pairs}
//   cen = {Px, Py, Pz, p_min}
struct FeatArgs {
    const float4* f0;
    const float4* f1;
    const float4* cen;
};

// The 32 features of quartet (P | Q), P of class X (bra), Q of class Y (ket):
//   0-7 bra pair, 8-15 ket pair, 16 log10(Q_P Q_Q), 17 |P - Q| of the most
//   diffuse primitive pairs, 18 sqrt(rho) |PQ|, 19 log10 rho
//   (rho = p q / (p + q)), 20-25 one-hot X, 26-31 one-hot Y.
__device__ inline void quartet_features(const FeatArgs& f, int P, int Q, int X, int Y, float (&x)[MLP_IN]) {
    const float4 a0 = f.f0[P], a1 = f.f1[P], ac = f.cen[P];
    const float4 b0 = f.f0[Q], b1 = f.f1[Q], bc = f.cen[Q];
    x[0] = a0.x; x[1] = a0.y; x[2] = a0.z; x[3] = a0.w;
    x[4] = a1.x; x[5] = a1.y; x[6] = a1.z; x[7] = a1.w;
    x[8] = b0.x; x[9] = b0.y; x[10] = b0.z; x[11] = b0.w;
    x[12] = b1.x; x[13] = b1.y; x[14] = b1.z; x[15] = b1.w;
    x[16] = a0.x + b0.x;
    const float dx = ac.x - bc.x, dy = ac.y - bc.y, dz = ac.z - bc.z;
    const float r = sqrtf(dx * dx + dy * dy + dz * dz);
    const float rho = ac.w * bc.w / (ac.w + bc.w);
    x[17] = r;
    x[18] = sqrtf(rho) * r;
    x[19] = log10f(rho);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        x[20 + c] = (c == X) ? 1.0f : 0.0f;
        x[26 + c] = (c == Y) ? 1.0f : 0.0f;
    }
}

// ----------------------------------------------------- per-lane evaluator --
// w: MLP_NALL floats whose weight entries are already bf16 values.
__device__ inline float mlp_forward_lane(const float* __restrict__ w, const float* x) {
    float xb[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) xb[i] = bf16_round((x[i] - w[MLP_OFF_MEAN + i]) * w[MLP_OFF_ISTD + i]);
    float h1[MLP_H];
#pragma unroll 4
    for (int j = 0; j < MLP_H; ++j) {
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) s = fmaf(w[MLP_OFF_W1 + j * MLP_IN + i], xb[i], s);
        h1[j] = bf16_round(fmaxf(s + w[MLP_OFF_B1 + j], 0.0f));
    }
    float out = 0.0f;
#pragma unroll 2
    for (int j = 0; j < MLP_H; ++j) {
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) s = fmaf(w[MLP_OFF_W2 + j * MLP_H + i], h1[i], s);
        out = fmaf(w[MLP_OFF_W3 + j], fmaxf(s + w[MLP_OFF_B2 + j], 0.0f), out);
    }
    return out + w[MLP_OFF_B3];
}

// ------------------------------------------------ wave-level MFMA evaluator --
// 64 lanes = 64 quartets, processed as two 32-column tiles (quartets 0-31,
// 32-63).  With v_mfma_f32_32x32x16_bf16, lane l (h = l / 32) supplies
// A[m = l % 32][k-slot (h, e)] and B[k-slot (h, e)][n = l % 32], e = 0..7, and
// receives C[8 j + 4 h + i][l % 32] in register 4 j + i.
//   layer 1: Z1 (64 hidden x 32 quartets) = W1 . Xb, A = W1 rows, k-slot
//            (h, e) of step ks = feature 16 ks + 8 h + e;
//   layer 2: the hidden index of k-slot (h, e) of step ks is
//            r = 16 ks + 8 (e / 4) + 4 h + e % 4, which is exactly where layer
//            1 leaves that hidden unit in this lane's C registers, so h1 goes
//            from C to B with no data movement;
//   layer 3: each lane dots its 32 rows with w3, and adds the other half-wave.
// Packed operands (MlpPacked, built on the host from bf16 weights):
//   frag[(f * 64 + lane)] = 8 bf16 of A fragment f for that lane,
//     f = 2 mt + ks (W1, mt, ks in 0..1), 4 + 4 mt + ks (W2, ks in 0..3);
//   vec[(kind * 2 + h) * 32 + 16 mt + 4 j + i] = b1 / b2 / w3 of row
//     32 mt + 8 j + 4 h + i (kind 0, 1, 2);
//   norm = mean[32], inv_std[32]; b3.
constexpr int MLP_NFRAG = 12;
struct MlpPacked {
    const uint4* frag;
    const float* vec;
    const float* norm;
    float b3;
};

typedef __bf16 mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f32x16 __attribute__((ext_vector_type(16)));

__device__ inline mlp_bf16x8 mlp_as_bf16x8(uint4 v) {
    mlp_bf16x8 r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}

typedef __bf16 mlp_bf16x2 __attribute__((ext_vector_type(2)));
typedef float mlp_f32x2 __attribute__((ext_vector_type(2)));

// Two values rounded to bf16 (nearest even, as bf16_round for finite values)
// and packed low / high: one v_cvt_pk_bf16_f32 on gfx950.
__device__ inline uint32_t mlp_pack2(float lo, float hi) {
    const mlp_bf16x2 b = __builtin_convertvector(mlp_f32x2{lo, hi}, mlp_bf16x2);
    uint32_t u;
    __builtin_memcpy(&u, &b, 4);
    return u;
}

__device__ inline uint4 mlp_pack8(const float* v) {
    return make_uint4(mlp_pack2(v[0], v[1]), mlp_pack2(v[2], v[3]), mlp_pack2(v[4], v[5]), mlp_pack2(v[6], v[7]));
}

__device__ inline uint4 mlp_select(bool c, uint4 a, uint4 b) {
    return make_uint4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w);
}

__device__ inline uint4 mlp_shfl32(uint4 v) {
    return make_uint4(__shfl_xor(v.x, 32, 64), __shfl_xor(v.y, 32, 64), __shfl_xor(v.z, 32, 64),
                      __shfl_xor(v.w, 32, 64));
}

// Host-side packing of the operands above from MLP_NALL parameters (weights
// are rounded to bf16 here).  frag: MLP_NFRAG * 64 entries, vec: 192, norm: 64.
inline void mlp_pack(const float* w, uint4* frag, float* vec, float* norm, float* b3) {
    for (int f = 0; f < MLP_NFRAG; ++f)
        for (int lane = 0; lane < 64; ++lane) {
            const int h = lane >> 5, ml = lane & 31;
            uint32_t bits[8];
            for (int e = 0; e < 8; ++e) {
                float v;
                if (f < 4) {
                    const int mt = f >> 1, ks = f & 1;
                    v = w[MLP_OFF_W1 + (32 * mt + ml) * MLP_IN + 16 * ks + 8 * h + e];
                } else {
                    const int g = f - 4, mt = g >> 2, ks = g & 3;
                    const int r = 16 * ks + 8 * (e >> 2) + 4 * h + (e & 3);
                    v = w[MLP_OFF_W2 + (32 * mt + ml) * MLP_H + r];
                }
                bits[e] = bf16_bits(v);
            }
            frag[f * 64 + lane] = make_uint4(bits[0] | (bits[1] << 16), bits[2] | (bits[3] << 16),
                                             bits[4] | (bits[5] << 16), bits[6] | (bits[7] << 16));
        }
    for (int kind = 0; kind < 3; ++kind)
        for (int h = 0; h < 2; ++h)
            for (int mt = 0; mt < 2; ++mt)
                for (int j = 0; j < 4; ++j)
                    for (int i = 0; i < 4; ++i) {
                        const int row = 32 * mt + 8 * j + 4 * h + i;
                        const float v = kind == 0 ? w[MLP_OFF_B1 + row]
                                                  : (kind == 1 ? w[MLP_OFF_B2 + row] : bf16_round(w[MLP_OFF_W3 + row]));
                        vec[(kind * 2 + h) * 32 + 16 * mt + 4 * j + i] = v;
                    }
    for (int i = 0; i < 2 * MLP_IN; ++i) norm[i] = w[MLP_OFF_MEAN + i];
    *b3 = w[MLP_OFF_B3];
}

// Must be reached by all 64 lanes of the wave (uniform control flow).
// x: this lane's raw features; returns the network output for this lane's quartet.
__device__ inline float mlp_forward_wave(const MlpPacked& m, const float (&x)[MLP_IN]) {
    const int lane = threadIdx.x & 63;
    const int h = lane >> 5;
    float xn[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) xn[i] = (x[i] - m.norm[i]) * m.norm[MLP_IN + i];
    // layer-1 B operands: own quartet feeds tile h, the partner lane's quartet tile 1 - h
    uint4 bown[2], bother[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        const uint4 lo = mlp_pack8(xn + 16 * ks), hi = mlp_pack8(xn + 16 * ks + 8);
        bown[ks] = mlp_select(h != 0, hi, lo);
        bother[ks] = mlp_shfl32(mlp_select(h != 0, lo, hi));
    }
    float y = 0.0f;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const bool own = nt == h;
        mlp_f32x16 c1[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
            for (int r = 0; r < 16; ++r) c1[mt][r] = m.vec[(0 * 2 + h) * 32 + 16 * mt + r];
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                c1[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    mlp_as_bf16x8(m.frag[(2 * mt + ks) * 64 + lane]),
                    mlp_as_bf16x8(mlp_select(own, bown[ks], bother[ks])), c1[mt], 0, 0, 0);
        }
        uint4 b2[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = fmaxf(c1[ks >> 1][8 * (ks & 1) + e], 0.0f);
            b2[ks] = mlp_pack8(v);
        }
        float part = 0.0f;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            mlp_f32x16 c2;
#pragma unroll
            for (int r = 0; r < 16; ++r) c2[r] = m.vec[(1 * 2 + h) * 32 + 16 * mt + r];
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                c2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(mlp_as_bf16x8(m.frag[(4 + 4 * mt + ks) * 64 + lane]),
                                                             mlp_as_bf16x8(b2[ks]), c2, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) part = fmaf(m.vec[(2 * 2 + h) * 32 + 16 * mt + r], fmaxf(c2[r], 0.0f), part);
        }
        part += __shfl_xor(part, 32, 64);
        if (own) y = part + m.b3;
    }
    return y;
}

}  // namespace tlno_amd
