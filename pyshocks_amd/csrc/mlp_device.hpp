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
//   x_k  = (v_k - mean) * (1 / s)   (0 when s == 0)
//   h1   = relu(W1 bf16(x) + b1),       rounded to BF16
//   h2   = relu(W2 h1 + b2),            rounded to BF16
//   z    = W3 h2 + (b3 + ln d),         d = (1/10, 6/10, 3/10)
//   omega = softmax(z),  value = h1 + omega0 (h0 - h1) + omega2 (h2 - h1)
//
// Cells in the shock regime keep WENO-JS, cells in the smooth regime take
// the ideal weights d without evaluating the network (mlp_regime).
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
//
// A wave evaluates its 64 samples (one per lane) as four chains of 16: chain
// c takes the samples of lanes 16c .. 16c + 15, which are columns 0..15.
// W1 is replicated in all four k-slot groups of its A operand, and chain c's
// B operand carries the inputs in k-slot group c only (lanes 16c .. 16c + 15
// hold their own inputs, the other lanes zeros), so the product picks W1
// once. W3 is replicated in the four row groups of its A operand, so every
// lane group holds the logits of its column: the lanes of group c read
// chain c's logits from their own registers. No lane exchanges are needed.
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
      const float w1 = j < MLP_IN ? P[MLP_W1 + row * MLP_IN + (j < MLP_IN ? j : 0)] : 0.0f;
      W.w1[b][j] = static_cast<__bf16>(w1);
      W.w2[b][j] = static_cast<__bf16>(P[MLP_W2 + row * MLP_HID + mlp_slot_unit(q, j)]);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      W.b1[b][v] = P[MLP_B1 + 16 * b + 4 * q + v];
      W.b2[b][v] = P[MLP_B2 + 16 * b + 4 * q + v];
    }
  }
  // row 4g + r of layer 3 is output r (r < 3) for every row group g
  const int r3 = i & 3;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const float w3 = r3 < MLP_OUT ? P[MLP_W3 + (r3 < MLP_OUT ? r3 : 0) * MLP_HID +
                                      mlp_slot_unit(q, j)]
                                  : 0.0f;
    W.w3[j] = static_cast<__bf16>(w3);
  }
  const float lnd[3] = {MLP_LN_D0, MLP_LN_D1, MLP_LN_D2};
#pragma unroll
  for (int v = 0; v < 4; ++v) W.b3[v] = v < MLP_OUT ? P[MLP_B3 + v] + lnd[v] : 0.0f;
}

// The operand registers of every lane of a wave, kept in LDS by persistent
// kernels (they depend on the lane only): a lane fetches its 160 bytes when
// the network runs instead of holding 40 registers throughout.
struct MlpWaveLds {
  bf16x8_t w1[2][64];
  bf16x8_t w2[2][64];
  bf16x8_t w3[64];
  f32x4_t b1[2][64];
  f32x4_t b2[2][64];
  f32x4_t b3[64];
};

// LDS room for MlpWaveLds only in the kernels that use it.
template <bool USED>
struct MlpLdsSlot {
  MlpWaveLds w;
};
template <>
struct MlpLdsSlot<false> {
  int unused;
};

// One wave stores its operand registers; a barrier must follow.
__device__ __forceinline__ void mlp_store_lds(const MlpWave &W, MlpWaveLds &L) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    L.w1[b][lane] = W.w1[b];
    L.w2[b][lane] = W.w2[b];
    L.b1[b][lane] = W.b1[b];
    L.b2[b][lane] = W.b2[b];
  }
  L.w3[lane] = W.w3;
  L.b3[lane] = W.b3;
}

__device__ __forceinline__ void mlp_load_lds(MlpWave &W, const MlpWaveLds &L) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    W.w1[b] = L.w1[b][lane];
    W.w2[b] = L.w2[b][lane];
    W.b1[b] = L.b1[b][lane];
    W.b2[b] = L.b2[b][lane];
  }
  W.w3 = L.w3[lane];
  W.b3 = L.b3[lane];
}

// Where a kernel finds the network: P in global memory, or L in LDS if set.
struct MlpSource {
  const float *P;
  const MlpWaveLds *L;
};

__device__ __forceinline__ void mlp_fetch(MlpWave &W, const MlpSource &src) {
  if (src.L != nullptr) {
    mlp_load_lds(W, *src.L);
  } else {
    mlp_load(W, src.P);
  }
}

typedef short mlp_s16x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 mlp_bf16x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int mlp_u32x4_t __attribute__((ext_vector_type(4)));

typedef float mlp_f32x2_t __attribute__((ext_vector_type(2)));

// Two FP32 values -> packed BF16 (round to nearest even, one conversion).
__device__ __forceinline__ unsigned int mlp_pack(float a, float b) {
  const mlp_f32x2_t f = {a, b};
  return __builtin_bit_cast(unsigned int, __builtin_convertvector(f, mlp_bf16x2_t));
}

// relu of packed BF16 values: a negative BF16 (sign bit set, -0 included) is
// a negative 16-bit integer, so an integer max with 0 clears it and keeps
// every non-negative value.
__device__ __forceinline__ unsigned int mlp_relu2(unsigned int x) {
  const mlp_s16x2_t zero = {0, 0};
  const mlp_s16x2_t r = __builtin_elementwise_max(__builtin_bit_cast(mlp_s16x2_t, x), zero);
  return __builtin_bit_cast(unsigned int, r);
}

// B operand of the next layer from the two 16-row blocks of this one:
// BF16, then relu (the same values as relu, then BF16).
__device__ __forceinline__ bf16x8_t mlp_relu_pack(f32x4_t h0, f32x4_t h1) {
  const mlp_u32x4_t r = {mlp_relu2(mlp_pack(h0[0], h0[1])), mlp_relu2(mlp_pack(h0[2], h0[3])),
                         mlp_relu2(mlp_pack(h1[0], h1[1])), mlp_relu2(mlp_pack(h1[2], h1[3]))};
  return __builtin_bit_cast(bf16x8_t, r);
}

// Inputs of one stencil (see the header comment). T is the state type.
template <typename T>
__device__ __forceinline__ void mlp_features(T v0, T v1, T v2, T v3, T v4,
                                             float x[MLP_IN]) {
  // every operation rounded as written (no multiply-add contraction), so the
  // inputs are bit-identical to the same formula evaluated elsewhere
#pragma clang fp contract(off)
  const T mean = (((v0 + v4) + (v1 + v3)) + v2) * T(0.2);
  const T d0 = v0 - mean, d1 = v1 - mean, d2 = v2 - mean, d3 = v3 - mean,
          d4 = v4 - mean;
  const T s = fmax(fmax(fmax(fabs(d0), fabs(d1)), fmax(fabs(d2), fabs(d3))),
                   fabs(d4));
  if (s > T(0)) {
    const T inv = T(1) / s;
    x[0] = static_cast<float>(d0 * inv);
    x[1] = static_cast<float>(d1 * inv);
    x[2] = static_cast<float>(d2 * inv);
    x[3] = static_cast<float>(d3 * inv);
    x[4] = static_cast<float>(d4 * inv);
  } else {
    x[0] = x[1] = x[2] = x[3] = x[4] = 0.0f;
  }
}

// Logits of one sample per lane (all 64 lanes of the wave must call it).
// With CHAINS = 1 only the lanes 0..15 are evaluated (the others' results
// are undefined).
template <int CHAINS = 4>
__device__ __forceinline__ void mlp_logits(const MlpWave &W, const float x[MLP_IN],
                                           float z[MLP_OUT]) {
  const int q = (threadIdx.x & 63) >> 4;
  const mlp_u32x4_t own = {mlp_pack(x[0], x[1]), mlp_pack(x[2], x[3]),
                           mlp_pack(x[4], 0.0f), 0u};
  const mlp_u32x4_t none = {0u, 0u, 0u, 0u};
  z[0] = z[1] = z[2] = 0.0f;
#pragma unroll
  for (int c = 0; c < CHAINS; ++c) {
    bf16x8_t B = __builtin_bit_cast(bf16x8_t, q == c ? own : none);
    f32x4_t h0 = mlp_mfma(W.w1[0], B, W.b1[0]);
    f32x4_t h1 = mlp_mfma(W.w1[1], B, W.b1[1]);
    B = mlp_relu_pack(h0, h1);
    h0 = mlp_mfma(W.w2[0], B, W.b2[0]);
    h1 = mlp_mfma(W.w2[1], B, W.b2[1]);
    B = mlp_relu_pack(h0, h1);
    const f32x4_t o = mlp_mfma(W.w3, B, W.b3);
    if (q == c) {
      z[0] = o[0];
      z[1] = o[1];
      z[2] = o[2];
    }
  }
}

// mlp_logits for inputs already packed to BF16 (own = {x0 x1, x2 x3, x4 0, 0})
// and only the first nch chains evaluated (nch wave-uniform).
__device__ __forceinline__ void mlp_logits_packed(const MlpWave &W, mlp_u32x4_t own,
                                                  int nch, float z[MLP_OUT]) {
  const int q = (threadIdx.x & 63) >> 4;
  const mlp_u32x4_t none = {0u, 0u, 0u, 0u};
  z[0] = z[1] = z[2] = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (c < nch) {
      bf16x8_t B = __builtin_bit_cast(bf16x8_t, q == c ? own : none);
      f32x4_t h0 = mlp_mfma(W.w1[0], B, W.b1[0]);
      f32x4_t h1 = mlp_mfma(W.w1[1], B, W.b1[1]);
      B = mlp_relu_pack(h0, h1);
      h0 = mlp_mfma(W.w2[0], B, W.b2[0]);
      h1 = mlp_mfma(W.w2[1], B, W.b2[1]);
      B = mlp_relu_pack(h0, h1);
      const f32x4_t o = mlp_mfma(W.w3, B, W.b3);
      if (q == c) {
        z[0] = o[0];
        z[1] = o[1];
        z[2] = o[2];
      }
    }
  }
}

__device__ __forceinline__ mlp_u32x4_t mlp_pack_inputs(const float x[MLP_IN]) {
  const mlp_u32x4_t own = {mlp_pack(x[0], x[1]), mlp_pack(x[2], x[3]), mlp_pack(x[4], 0.0f),
                           0u};
  return own;
}

__device__ __forceinline__ void mlp_softmax(const float z[MLP_OUT], float w[MLP_OUT]) {
  const float zm = fmaxf(fmaxf(z[0], z[1]), z[2]);
  const float e0 = __expf(z[0] - zm), e1 = __expf(z[1] - zm), e2 = __expf(z[2] - zm);
  const float inv = 1.0f / ((e0 + e1) + e2);
  w[0] = e0 * inv;
  w[1] = e1 * inv;
  w[2] = e2 * inv;
}

// sum_k w_k h_k written as h1 + w0 (h0 - h1) + w2 (h2 - h1): the FP32
// weights sum to 1 only to rounding, and this form keeps that rounding
// relative to the spread of the candidates instead of to their size.
template <typename T>
__device__ __forceinline__ T mlp_combine(const float w[MLP_OUT], T h0, T h1, T h2) {
  return (h1 + T(w[0]) * (h0 - h1)) + T(w[2]) * (h2 - h1);
}

// The ideal weights d = (1/10, 6/10, 3/10) in the same form.
template <typename T>
__device__ __forceinline__ T mlp_combine_ideal(T h0, T h1, T h2) {
  return (h1 + T(0.1) * (h0 - h1)) + T(0.3) * (h2 - h1);
}

// Regime of a cell from its right-side smoothness indicators (the left
// side's are the same three):
//   0 (shock, WENO-JS)     max beta >  tau (min beta + eps)
//   1 (smooth, ideal d)    max beta <= sigma min beta   (scale-free)
//   2 (learned)            otherwise.
template <typename T>
__device__ __forceinline__ int mlp_regime(T b0, T b1, T b2, T eps, T tau, T sigma) {
  const T bmax = fmax(fmax(b0, b1), b2);
  const T bmin = fmin(fmin(b0, b1), b2);
  if (bmax > tau * (bmin + eps)) return 0;
  return bmax <= sigma * bmin ? 1 : 2;
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
// are the WENO-JS values of the cell, used in the shock regime; the smooth
// regime takes the ideal weights, the learned one the network's.
template <typename T>
__device__ __forceinline__ void learned_pair(const MlpWave &W, T um2, T um1, T u0,
                                             T up1, T up2, T b0, T b1, T b2,
                                             T eps, T tau, T sigma, T ulJS, T urJS,
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
  const int regime = mlp_regime(b0, b1, b2, eps, tau, sigma);
  if (regime == 0) {
    ur = urJS;
    ul = ulJS;
  } else if (regime == 1) {
    ur = mlp_combine_ideal(hr0, hr1, hr2);
    ul = mlp_combine_ideal(hl0, hl1, hl2);
  } else {
    ur = mlp_combine(wr, hr0, hr1, hr2);
    ul = mlp_combine(wl, hl0, hl1, hl2);
  }
}

// Learned right reconstruction of the stencils of lanes 0 and 1 (wave-
// collective; the other lanes' results are undefined): the network (fetched
// from src only then) runs as one 16-sample chain, and only if one of the two
// cells is in the learned regime; js reports the shock regime (the WENO-JS
// value is then to be used instead).
template <typename T>
__device__ __forceinline__ T learned_right16(const MlpSource &src, T um2, T um1,
                                             T u0, T up1, T up2, T eps, T tau, T sigma,
                                             bool &js) {
  // the indicators and candidates of the fused kernels (reconstruct_cells),
  // expression for expression, so that a cell gets the same regime and
  // values here (at a tile edge, or from its mirrored stencil)
  using C = Weno5Coeffs<T>;
  const T D0 = um1 - um2, D1 = u0 - um1, D2 = up1 - u0, D3 = up2 - up1;
  const T s0 = D1 - D0, s1 = D2 - D1, s2 = D3 - D2;
  const T E0 = C::a0x4 * (s0 * s0), E1 = C::a0x4 * (s1 * s1), E2 = C::a0x4 * (s2 * s2);
  const T t01 = fma(T(2), D1, s0);
  const T t11 = D2 + D1;
  const T t21 = fma(T(-2), D2, s2);
  const T b0 = fma(t01, t01, E0);
  const T b1 = fma(t11, t11, E1);
  const T b2 = fma(t21, t21, E2);
  const int regime = mlp_regime(b0, b1, b2, T(4) * eps, tau, sigma);
  js = regime == 0;
  T h0, h1, h2, g0, g1, g2;
  weno_uhat_diff(u0, D0, D1, D2, D3, h0, h1, h2, g0, g1, g2);
  T value = mlp_combine_ideal(h0, h1, h2);
  const int lane = threadIdx.x & 63;
  if (__ballot(lane < 2 && regime == 2) != 0) {  // wave-uniform
    MlpWave W;
    mlp_fetch(W, src);
    float x[MLP_IN], z[MLP_OUT], w[MLP_OUT];
    mlp_features(um2, um1, u0, up1, up2, x);
    mlp_logits<1>(W, x, z);
    mlp_softmax(z, w);
    if (regime == 2) value = mlp_combine(w, h0, h1, h2);
  }
  return value;
}

// Both reconstructions of one cell per lane from its stencil alone
// (wave-collective): indicators, WENO-JS values and the learned values.
template <typename T>
__device__ __forceinline__ void learned_cell(const MlpWave &W, T um2, T um1, T u0,
                                             T up1, T up2, T eps, T tau, T sigma,
                                             T &ul, T &ur) {
  T b0, b1, b2, ulJS, urJS;
  weno_beta(um2, um1, u0, up1, up2, b0, b1, b2);
  weno_pair_from_beta(um2, um1, u0, up1, up2, b0, b1, b2, eps, ulJS, urJS);
  learned_pair(W, um2, um1, u0, up1, up2, b0, b1, b2, eps, tau, sigma, ulJS, urJS, ul,
               ur);
}

}  // namespace psa
