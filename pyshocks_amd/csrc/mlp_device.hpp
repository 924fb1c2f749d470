// Learned WENO weights: a 5 -> 32 -> 32 -> 3 perceptron (ReLU, softmax)
// evaluated with BF16 matrix-core instructions and FP32 accumulation.
//
// Parameters are kept as FP32 in one packed array (offsets below); the
// kernels round them to BF16 when they build their operand registers.
//
// Semantics for one stencil v = (v0, ..., v4) of the right reconstruction at
// x_{i+1/2} (v = u[i-2 .. i+2]; the left reconstruction passes the mirrored
// stencil, as the reference's array reversal does):
//
//   mean = (((v0 + v4) + (v1 + v3)) + v2) * 0.2,  s = max_k |v_k - mean|
//   x_k  = (v_k - mean) / s   (0 when s == 0)
//   h1   = relu(W1 bf16(x) + b1),       rounded to BF16
//   h2   = relu(W2 h1 + b2),            rounded to BF16
//   z    = W3 h2 + (b3 + ln d),         d = (1/10, 6/10, 3/10)
//   omega = softmax(z)
//
// with W1, W2, W3 rounded to BF16 and every sum accumulated in FP32, so that
// z = 0 gives the ideal linear weights d. The mean is summed symmetrically so
// that the inputs of the mirrored stencil are exactly the mirrored inputs: a
// face shared by two tiles then gets the same flux from both.
//
// Matrix-core layout (v_mfma_f32_16x16x32_bf16, one wave): the products are
// formed transposed, H^T = W X^T, so that the samples are the 16 columns.
// Lane l holds column l % 16 of the B operand and the 8 k-slots
// (q = l / 16, j = 0..7) of that column; the A operand holds row l % 16 and
// the same k-slots; the 16x16 FP32 result holds rows 4q + v (v = 0..3) of
// column l % 16. The result of a 32-row layer (two 16-row blocks) therefore
// already sits where the next layer's B operand needs it, with k-slot
// (q, j) standing for hidden unit pi(q, j) = 4q + j (j < 4) or
// 16 + 4q + (j - 4) (j >= 4); the A operands of the next layer are built with
// the same slot -> unit map, so no data moves between layers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "weno5_device.hpp"

namespace psa {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int MLP_IN = 5;
constexpr int MLP_HID = 32;
constexpr int MLP_OUT = 3;
constexpr int MLP_W1 = 0;      // [32][5]
constexpr int MLP_B1 = 160;    // [32]
constexpr int MLP_W2 = 192;    // [32][32]
constexpr int MLP_B2 = 1216;   // [32]
constexpr int MLP_W3 = 1248;   // [3][32]
constexpr int MLP_B3 = 1344;   // [3]
constexpr int MLP_NPARAM = 1347;

// ln of the ideal weights, in FP32
constexpr float MLP_LN_D0 = -2.302585092994046f;   // ln(1/10)
constexpr float MLP_LN_D1 = -0.5108256237659907f;  // ln(6/10)
constexpr float MLP_LN_D2 = -1.2039728043259361f;  // ln(3/10)

__device__ __forceinline__ int mlp_slot_unit(int q, int j) {
  return j < 4 ? 4 * q + j : 16 + 4 * q + (j - 4);
}

__device__ __forceinline__ f32x4_t mlp_mfma(bf16x8_t a, bf16x8_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Operand registers of one wave.
struct MlpWave {
  bf16x8_t w1[2];  // layer 1, rows 16b + l%16; slot (0, j < 5) -> input j
  bf16x8_t w2[2];  // layer 2, rows 16b + l%16; slot (q, j) -> unit pi(q, j)
  bf16x8_t w3;     // layer 3, rows l%16 < 3
  f32x4_t b1[2];   // accumulator init: bias of rows 16b + 4q + v
  f32x4_t b2[2];
  f32x4_t b3;      // rows 4q + v < 3: b3 + ln d
};

__device__ __forceinline__ void mlp_load(MlpWave &W, const float *__restrict__ P) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4;
  const int i = lane & 15;
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int row = 16 * b + i;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float w1 = (q == 0 && j < MLP_IN) ? P[MLP_W1 + row * MLP_IN + j] : 0.0f;
      W.w1[b][j] = static_cast<__bf16>(w1);
      W.w2[b][j] = static_cast<__bf16>(P[MLP_W2 + row * MLP_HID + mlp_slot_unit(q, j)]);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      W.b1[b][v] = P[MLP_B1 + 16 * b + 4 * q + v];
      W.b2[b][v] = P[MLP_B2 + 16 * b + 4 * q + v];
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const float w3 = i < MLP_OUT ? P[MLP_W3 + i * MLP_HID + mlp_slot_unit(q, j)] : 0.0f;
    W.w3[j] = static_cast<__bf16>(w3);
  }
  const float lnd[3] = {MLP_LN_D0, MLP_LN_D1, MLP_LN_D2};
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = 4 * q + v;
    W.b3[v] = r < MLP_OUT ? P[MLP_B3 + r] + lnd[r < MLP_OUT ? r : 0] : 0.0f;
  }
}

// B operand of the next layer from the two 16-row blocks of this one:
// relu, then BF16.
__device__ __forceinline__ bf16x8_t mlp_relu_pack(f32x4_t h0, f32x4_t h1) {
  bf16x8_t r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    r[j] = static_cast<__bf16>(fmaxf(h0[j], 0.0f));
    r[4 + j] = static_cast<__bf16>(fmaxf(h1[j], 0.0f));
  }
  return r;
}

// Inputs of one stencil (see the header comment). T is the state type.
template <typename T>
__device__ __forceinline__ void mlp_features(T v0, T v1, T v2, T v3, T v4,
                                             float x[MLP_IN]) {
  const T mean = (((v0 + v4) + (v1 + v3)) + v2) * T(0.2);
  const T d0 = v0 - mean, d1 = v1 - mean, d2 = v2 - mean, d3 = v3 - mean,
          d4 = v4 - mean;
  const T s = fmax(fmax(fmax(fabs(d0), fabs(d1)), fmax(fabs(d2), fabs(d3))),
                   fabs(d4));
  if (s > T(0)) {
    x[0] = static_cast<float>(d0 / s);
    x[1] = static_cast<float>(d1 / s);
    x[2] = static_cast<float>(d2 / s);
    x[3] = static_cast<float>(d3 / s);
    x[4] = static_cast<float>(d4 / s);
  } else {
    x[0] = x[1] = x[2] = x[3] = x[4] = 0.0f;
  }
}

// Logits of one sample per lane (all 64 lanes of the wave must call it).
// The 64 samples run as 4 chains of 16: chain c takes the samples of lanes
// 16c .. 16c + 15, whose inputs move to lanes 0..15 of its B operand.
__device__ __forceinline__ void mlp_logits(const MlpWave &W, const float x[MLP_IN],
                                           float z[MLP_OUT]) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4;
  const int col = lane & 15;
  z[0] = z[1] = z[2] = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int src = 16 * c + col;
    bf16x8_t B;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float xs = j < MLP_IN ? __shfl(x[j < MLP_IN ? j : 0], src, 64) : 0.0f;
      B[j] = static_cast<__bf16>(q == 0 ? xs : 0.0f);
    }
    f32x4_t h0 = mlp_mfma(W.w1[0], B, W.b1[0]);
    f32x4_t h1 = mlp_mfma(W.w1[1], B, W.b1[1]);
    B = mlp_relu_pack(h0, h1);
    h0 = mlp_mfma(W.w2[0], B, W.b2[0]);
    h1 = mlp_mfma(W.w2[1], B, W.b2[1]);
    B = mlp_relu_pack(h0, h1);
    const f32x4_t o = mlp_mfma(W.w3, B, W.b3);
    // lanes 0..15 hold the logits of chain c's samples in rows 0..2
    const float t0 = __shfl(o[0], col, 64);
    const float t1 = __shfl(o[1], col, 64);
    const float t2 = __shfl(o[2], col, 64);
    if (q == c) {
      z[0] = t0;
      z[1] = t1;
      z[2] = t2;
    }
  }
}

__device__ __forceinline__ void mlp_softmax(const float z[MLP_OUT], float w[MLP_OUT]) {
  const float zm = fmaxf(fmaxf(z[0], z[1]), z[2]);
  const float e0 = __expf(z[0] - zm), e1 = __expf(z[1] - zm), e2 = __expf(z[2] - zm);
  const float inv = 1.0f / ((e0 + e1) + e2);
  w[0] = e0 * inv;
  w[1] = e1 * inv;
  w[2] = e2 * inv;
}

// Shock flag of a cell from its right-side smoothness indicators (the left
// side's are the same three): max beta > tau (min beta + eps).
template <typename T>
__device__ __forceinline__ bool mlp_shock(T b0, T b1, T b2, T eps, T tau) {
  const T bmax = fmax(fmax(b0, b1), b2);
  const T bmin = fmin(fmin(b0, b1), b2);
  return bmax > tau * (bmin + eps);
}

// Both reconstructions of one cell per lane (wave-collective). ulJS, urJS
// are the WENO-JS values of the cell, used where the cell is flagged.
template <typename T>
__device__ __forceinline__ void learned_pair(const MlpWave &W, T um2, T um1, T u0,
                                             T up1, T up2, T b0, T b1, T b2,
                                             T eps, T tau, T ulJS, T urJS,
                                             T &ul, T &ur) {
  float x[MLP_IN], xm[MLP_IN], zr[MLP_OUT], zl[MLP_OUT], wr[MLP_OUT], wl[MLP_OUT];
  mlp_features(um2, um1, u0, up1, up2, x);
  xm[0] = x[4];
  xm[1] = x[3];
  xm[2] = x[2];
  xm[3] = x[1];
  xm[4] = x[0];
  mlp_logits(W, x, zr);
  mlp_logits(W, xm, zl);
  mlp_softmax(zr, wr);
  mlp_softmax(zl, wl);
  T hr0, hr1, hr2, hl0, hl1, hl2;
  weno_uhat(um2, um1, u0, up1, up2, hr0, hr1, hr2);
  weno_uhat(up2, up1, u0, um1, um2, hl0, hl1, hl2);
  const T lr = (T(wr[0]) * hr0 + T(wr[1]) * hr1) + T(wr[2]) * hr2;
  const T ll = (T(wl[0]) * hl0 + T(wl[1]) * hl1) + T(wl[2]) * hl2;
  const bool shock = mlp_shock(b0, b1, b2, eps, tau);
  ur = shock ? urJS : lr;
  ul = shock ? ulJS : ll;
}

// Both reconstructions of one cell per lane from its stencil alone
// (wave-collective): indicators, WENO-JS values and the learned values.
template <typename T>
__device__ __forceinline__ void learned_cell(const MlpWave &W, T um2, T um1, T u0,
                                             T up1, T up2, T eps, T tau, T &ul,
                                             T &ur) {
  T b0, b1, b2, ulJS, urJS;
  weno_beta(um2, um1, u0, up1, up2, b0, b1, b2);
  weno_pair_from_beta(um2, um1, u0, up1, up2, b0, b1, b2, eps, ulJS, urJS);
  learned_pair(W, um2, um1, u0, up1, up2, b0, b1, b2, eps, tau, ulJS, urJS, ul,
               ur);
}

}  // namespace psa
