// Discrete adjoint (vector-Jacobian products) of the fused SSP-RK3 step.
//
// The reference differentiates one step with a dense forward-mode Jacobian
// (timestepping.py:155-215, jax.jacfwd of advance). Here the transposed
// Jacobian is applied matrix-free, one reverse kernel per Runge-Kutta stage:
//
//   forward  k1 = x + dt L(x)
//            k2 = 3/4 x + 1/4 (k1 + dt L(k1))
//            x+ = 1/3 x + 2/3 (k2 + dt L(k2))
//   reverse  lk2 = 2/3 (l + dt JL(k2)^T l)
//            lk1 = 1/4 (lk2 + dt JL(k1)^T lk2)
//            lx  = 1/3 l + 3/4 lk2 + lk1 + dt JL(x)^T lk1
//
// JL^T mu is formed per cell from the closed-form gradient of the two WENO-JS5
// reconstructions of the cell and the partial derivatives of the face flux.
// Derivative conventions at the non-smooth points follow jax: jnp.maximum
// splits a tie 0.5 / 0.5, d|u|/du = sign(u) with sign(0) = 0, and the upwind
// selection (jnp.where) passes the derivative of the selected branch only.
// dt is a constant (adjoint_step recovers it as t_{m+1} - t_m).
//
// The state and the cotangents are periodic: every index is mapped to its
// interior image, which also covers the 5-cell halo the transpose needs.
//
// Learned weights (LEARNED kernels): each cell takes its forward regime (from
// the same fused indicator arithmetic as the forward kernels); shock cells
// differentiate WENO-JS, smooth cells the ideal weights, learned cells the
// network. The network is differentiated as the smooth function it rounds:
// BF16 parameters (as the forward's operands), inputs and activations
// unrounded, in the state dtype -- the adjoint of that surrogate along the
// forward's trajectory.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/pyshocks_amd.h"
#include "mlp_device.hpp"
#include "psa_internal.hpp"
#include "weno5_device.hpp"

namespace psa {
namespace {

constexpr int kVjpBlock = 256;              // computing cells per workgroup
constexpr int kVjpThreads = 2 * kVjpBlock;  // split form: a thread per cell and side
constexpr int kVjpOut = kVjpBlock - 6;  // output cells per workgroup

// cell gm of a periodic array with n interior cells and g ghosts
template <typename T>
__device__ __forceinline__ T load_periodic(const T *__restrict__ u, int64_t gm,
                                           int64_t n, int g) {
  int64_t k = (gm - g) % n;
  if (k < 0) k += n;
  return u[g + k];
}

// 1 / x for a positive normal FP64 x: the hardware reciprocal refined by
// two Newton steps.
__device__ __forceinline__ double rcp_refined(double x) {
  double inv = __builtin_amdgcn_rcp(x);
  double e = fma(-x, inv, 1.0);
  inv = fma(inv, e, inv);
  e = fma(-x, inv, 1.0);
  return fma(inv, e, inv);
}

// The normalised WENO-JS weights w_k of a_k = d_k / e_k^2 and the
// reciprocals ie_k = 1 / e_k of the indicators e_k = eps + beta_k. FP64: one
// refined reciprocal of e0 e1 e2 and one of the sum of the a_k when both are
// positive normal numbers (6 divisions fewer per side, a few ulps from
// them), else the divisions; FP32: the divisions.
template <typename T>
__device__ __forceinline__ void weno_weights_inv(T e0, T e1, T e2, T &w0, T &w1, T &w2,
                                                 T &ie0, T &ie1, T &ie2) {
  using C = Weno5Coeffs<T>;
  if constexpr (std::is_same<T, double>::value) {
    const double e01 = e0 * e1;
    const double den = e01 * e2;
    if (__builtin_amdgcn_class(den, 0x100)) {
      const double inv = rcp_refined(den);
      ie0 = (e1 * e2) * inv;
      ie1 = (e0 * e2) * inv;
      ie2 = e01 * inv;
      const double a0 = C::d0 * (ie0 * ie0);
      const double a1 = C::d1 * (ie1 * ie1);
      const double a2 = C::d2 * (ie2 * ie2);
      const double s = (a0 + a1) + a2;
      if (__builtin_amdgcn_class(s, 0x100)) {
        const double is = rcp_refined(s);
        w0 = a0 * is;
        w1 = a1 * is;
        w2 = a2 * is;
        return;
      }
    }
  }
  const T a0 = C::d0 / (e0 * e0);
  const T a1 = C::d1 / (e1 * e1);
  const T a2 = C::d2 / (e2 * e2);
  const T s = (a0 + a1) + a2;
  w0 = a0 / s;
  w1 = a1 / s;
  w2 = a2 / s;
  ie0 = T(1) / e0;
  ie1 = T(1) / e1;
  ie2 = T(1) / e2;
}

// Right reconstruction (value at x_{i+1/2}) of a cell from
// (u[i-2], ..., u[i+2]) and its gradient with respect to those five values.
template <typename T>
__device__ __forceinline__ T weno_side_grad(T um2, T um1, T u0, T up1, T up2,
                                            T eps, T G[5]) {
  using C = Weno5Coeffs<T>;
  const T t00 = (u0 - T(2) * um1) + um2;
  const T t01 = (T(3) * u0 - T(4) * um1) + um2;
  const T t10 = (up1 - T(2) * u0) + um1;
  const T t11 = up1 - um1;
  const T t20 = (up2 - T(2) * up1) + u0;
  const T t21 = (up2 - T(4) * up1) + T(3) * u0;
  const T b0 = C::a0 * (t00 * t00) + C::a1 * (t01 * t01);
  const T b1 = C::a0 * (t10 * t10) + C::a1 * (t11 * t11);
  const T b2 = C::a0 * (t20 * t20) + C::a1 * (t21 * t21);
  T h0, h1, h2;
  weno_uhat(um2, um1, u0, up1, up2, h0, h1, h2);

  const T e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  T w0, w1, w2, ie0, ie1, ie2;
  weno_weights_inv(e0, e1, e2, w0, w1, w2, ie0, ie1, ie2);
  const T r = (w0 * h0 + w1 * h1) + w2 * h2;

  // dr/dbeta_k = -2 omega_k (uhat_k - r) / (eps + beta_k)
  const T q0 = T(-2) * w0 * (h0 - r) * ie0;
  const T q1 = T(-2) * w1 * (h1 - r) * ie1;
  const T q2 = T(-2) * w2 * (h2 - r) * ie2;
  // dr/dt_kj = dr/dbeta_k * 2 a_j t_kj
  const T p00 = q0 * (T(2) * C::a0) * t00, p01 = q0 * (T(2) * C::a1) * t01;
  const T p10 = q1 * (T(2) * C::a0) * t10, p11 = q1 * (T(2) * C::a1) * t11;
  const T p20 = q2 * (T(2) * C::a0) * t20, p21 = q2 * (T(2) * C::a1) * t21;

  G[0] = (p00 + p01) + w0 * C::c2;
  G[1] = (T(-2) * p00 - T(4) * p01 + p10 - p11) - w0 * C::c7 - w1 * C::c1;
  G[2] = (p00 + T(3) * p01 - T(2) * p10 + p20 + T(3) * p21) + w0 * C::c11 +
         w1 * C::c5 + w2 * C::c2;
  G[3] = (p10 + p11 - T(2) * p20 - T(4) * p21) + w1 * C::c2 + w2 * C::c5;
  G[4] = (p20 + p21) - w2 * C::c1;
  return r;
}

template <typename T>
__device__ __forceinline__ T sgn(T x) {
  return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0));
}

template <typename T>
struct VjpArgs {
  const T *state;   // stencil input of the forward stage
  const T *mu;      // cotangent of the stage operator output
  const T *lam;     // RSTAGE 1: cotangent of the step output
  const T *lamk2;   // RSTAGE 1: cotangent of k2
  T *out;
  const T *dt_ptr;
  int64_t m;
  int64_t n;
  int g;
  T inv_dx;
  T eps;
  T vconst;
  const float *mlp;   // LEARNED: the network's parameters (PSA_MLP_NPARAM FP32)
  T eps4;             // LEARNED: 4 eps (the fused indicators carry 4 beta)
  T tau;              // LEARNED: shock threshold
  T sigma;            // LEARNED: smooth threshold
  T alpha;            // Lax-Friedrichs: the global speed max|state| (in-kernel)
};

// Regime of a cell (mlp_regime) from its right stencil with the indicator
// arithmetic of the forward kernels (reconstruct_cells: shared first and
// second differences, 4 beta, one multiply-add each), so that the reverse
// kernels see the regime the forward took.
template <typename T>
__device__ __forceinline__ int learned_regime5(T v0, T v1, T v2, T v3, T v4, T eps4,
                                               T tau, T sigma) {
  using C = Weno5Coeffs<T>;
  const T D0 = v1 - v0, D1 = v2 - v1, D2 = v3 - v2, D3 = v4 - v3;
  const T s0 = D1 - D0, s1 = D2 - D1, s2 = D3 - D2;
  const T E0 = C::a0x4 * (s0 * s0), E1 = C::a0x4 * (s1 * s1), E2 = C::a0x4 * (s2 * s2);
  const T t01 = fma(T(2), D1, s0);
  const T t11 = D2 + D1;
  const T t21 = fma(T(-2), D2, s2);
  return mlp_regime(fma(t01, t01, E0), fma(t11, t11, E1), fma(t21, t21, E2), eps4, tau,
                    sigma);
}

// Gradient of the candidates with respect to the stencil: G += sum_k c_k
// d uhat_k / dv (weno.py:231-239 taps in the order of weno_side_grad).
template <typename T>
__device__ __forceinline__ void uhat_grad(T c0, T c1, T c2, T G[5]) {
  using C = Weno5Coeffs<T>;
  G[0] += c0 * C::c2;
  G[1] += -(c0 * C::c7) - c1 * C::c1;
  G[2] += (c0 * C::c11 + c1 * C::c5) + c2 * C::c2;
  G[3] += c1 * C::c2 + c2 * C::c5;
  G[4] += -(c2 * C::c1);
}

// Ideal-weight reconstruction (mlp_combine_ideal) and its stencil gradient.
template <typename T>
__device__ __forceinline__ T ideal_side_grad(T v0, T v1, T v2, T v3, T v4, T G[5]) {
  T h0, h1, h2;
  weno_uhat(v0, v1, v2, v3, v4, h0, h1, h2);
#pragma unroll
  for (int k = 0; k < 5; ++k) G[k] = T(0);
  uhat_grad(T(0.1), T(1) - T(0.1) - T(0.3), T(0.3), G);
  return mlp_combine_ideal(h0, h1, h2);
}

// A weight as the forward kernels' BF16 operand (round to nearest even, on
// the bit pattern: the parameters are uniform, so this stays in scalar
// registers), as the state dtype.
template <typename T>
__device__ __forceinline__ T mlp_weight(const float *__restrict__ P, int i);

template <>
__device__ __forceinline__ float mlp_weight<float>(const float *__restrict__ P, int i) {
  unsigned int b = __float_as_uint(P[i]);
  b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
  return __uint_as_float(b);
}

// FP64: the BF16 value's exact FP64 bits built from its FP32 bits (sign,
// exponent rebiased by 1023 - 127, the 7 mantissa bits; BF16 values below
// the FP32 normal range are taken as zero), so that the weight stays a
// scalar operand of the multiply-adds.
template <>
__device__ __forceinline__ double mlp_weight<double>(const float *__restrict__ P, int i) {
  unsigned int b = __float_as_uint(P[i]);
  b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
  const unsigned int e = (b >> 23) & 0xFFu;
  const unsigned int hi =
      e == 0u ? (b & 0x80000000u)
              : ((b & 0x80000000u) | ((e + 896u) << 20) | ((b >> 3) & 0x000FFFFFu));
  return __hiloint2double(static_cast<int>(hi), 0);
}

// Network reconstruction of one side, r = h1 + w0 (h0 - h1) + w2 (h2 - h1)
// with w = softmax(z) of the network on x = (v - mean) / max|v - mean|, and
// its stencil gradient: the candidates' part with (w0, 1 - w0 - w2, w2), the
// weights' part back through softmax, the three layers (ReLU masks of this
// evaluation) and the normalisation. P: the parameters (weights rounded as
// mlp_weight, b3 + ln d summed in FP32 as the forward's accumulator start).
// Layer 2 runs row by row, so 32 values of T are live.
template <typename T>
__device__ __forceinline__ T learned_side_grad(const float *__restrict__ P, T v0, T v1,
                                               T v2, T v3, T v4, T G[5]) {
  const T v[5] = {v0, v1, v2, v3, v4};
  const T mean = (((v0 + v4) + (v1 + v3)) + v2) * T(0.2);
  T d[5], x[5];
  int km = 0;
  T s = T(0);
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    d[k] = v[k] - mean;
    if (fabs(d[k]) > s) {
      s = fabs(d[k]);
      km = k;
    }
  }
  const T inv = s > T(0) ? T(1) / s : T(0);
#pragma unroll
  for (int k = 0; k < 5; ++k) x[k] = d[k] * inv;

  // forward: h1 = relu(W1 x + b1), z = W3 relu(W2 h1 + b2) + b3 + ln d
  T h1[MLP_HID];
  unsigned int m1 = 0, m2 = 0;
#pragma unroll
  for (int j = 0; j < MLP_HID; ++j) {
    T a = static_cast<T>(P[MLP_B1 + j]);
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) a = fma(mlp_weight<T>(P, MLP_W1 + j * MLP_IN + k), x[k], a);
    m1 |= a > T(0) ? (1u << j) : 0u;
    h1[j] = a > T(0) ? a : T(0);
  }
  T z[MLP_OUT];
  const float lnd[3] = {MLP_LN_D0, MLP_LN_D1, MLP_LN_D2};
#pragma unroll
  for (int o = 0; o < MLP_OUT; ++o) z[o] = static_cast<T>(P[MLP_B3 + o] + lnd[o]);
  // rows of layer 2 one at a time (a rolled loop: the row's 32 weights are
  // loaded when it runs)
#pragma unroll 1
  for (int i = 0; i < MLP_HID; ++i) {
    T a = static_cast<T>(P[MLP_B2 + i]);
#pragma unroll
    for (int j = 0; j < MLP_HID; ++j) a = fma(mlp_weight<T>(P, MLP_W2 + i * MLP_HID + j), h1[j], a);
    if (a > T(0)) {
      m2 |= 1u << i;
#pragma unroll
      for (int o = 0; o < MLP_OUT; ++o)
        z[o] = fma(mlp_weight<T>(P, MLP_W3 + o * MLP_HID + i), a, z[o]);
    }
  }
  const T zm = fmax(fmax(z[0], z[1]), z[2]);
  const T e0 = exp(z[0] - zm), e1 = exp(z[1] - zm), e2 = exp(z[2] - zm);
  const T es = (e0 + e1) + e2;
  const T w0 = e0 / es, w1 = e1 / es, w2 = e2 / es;
  T h0c, h1c, h2c;
  weno_uhat(v0, v1, v2, v3, v4, h0c, h1c, h2c);
  const T r = (h1c + w0 * (h0c - h1c)) + w2 * (h2c - h1c);

#pragma unroll
  for (int k = 0; k < 5; ++k) G[k] = T(0);
  uhat_grad(w0, (T(1) - w0) - w2, w2, G);

  // backward: dr/dw = (h0 - h1, 0, h2 - h1), through softmax
  const T c0 = h0c - h1c, c2 = h2c - h1c;
  const T sw = c0 * w0 + c2 * w2;
  const T zb[MLP_OUT] = {w0 * (c0 - sw), w1 * (T(0) - sw), w2 * (c2 - sw)};
  // h1b = W2^T (mask2 * W3^T zb), then xb = W1^T (mask1 * h1b)
  T h1b[MLP_HID];
#pragma unroll
  for (int j = 0; j < MLP_HID; ++j) h1b[j] = T(0);
#pragma unroll 1
  for (int i = 0; i < MLP_HID; ++i) {
    if (m2 & (1u << i)) {
      T h2b = T(0);
#pragma unroll
      for (int o = 0; o < MLP_OUT; ++o)
        h2b = fma(mlp_weight<T>(P, MLP_W3 + o * MLP_HID + i), zb[o], h2b);
#pragma unroll
      for (int j = 0; j < MLP_HID; ++j)
        h1b[j] = fma(mlp_weight<T>(P, MLP_W2 + i * MLP_HID + j), h2b, h1b[j]);
    }
  }
  T xb[MLP_IN] = {T(0), T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int j = 0; j < MLP_HID; ++j) {
    if (m1 & (1u << j)) {
#pragma unroll
      for (int k = 0; k < MLP_IN; ++k)
        xb[k] = fma(mlp_weight<T>(P, MLP_W1 + j * MLP_IN + k), h1b[j], xb[k]);
    }
  }
  if (s > T(0)) {
    // x_k = d_k / s: dx_k/dv_j = (delta_kj - 1/5) / s - x_k sgn(d_km)
    // (delta_km,j - 1/5) / s
    T sxb = T(0), dot = T(0);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      sxb += xb[k];
      dot += xb[k] * x[k];
    }
    const T sg = d[km] > T(0) ? T(1) : T(-1);
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const T ds = sg * ((j == km ? T(1) : T(0)) - T(0.2));
      G[j] += inv * ((xb[j] - T(0.2) * sxb) - dot * ds);
    }
  }
  return r;
}

// Scratch of one reverse stage over a block of B computing cells.
template <typename T, bool SPLIT = true>
struct VjpScratch {
  T ur[kVjpBlock], ul[kVjpBlock];
  T cR[kVjpBlock], cL[kVjpBlock], sL[kVjpBlock], sR[kVjpBlock];
  T P[5][kVjpBlock];
  // gradients of the left reconstructions (computed by the other half of a
  // split block)
  T GL[SPLIT ? 5 : 1][SPLIT ? kVjpBlock : 1];
};

// JL(state)^T mu for the output cells Z .. Z + B - 7 of one block: s_u holds
// the state at cells Z - 5 .. Z + B - 2, s_mu the cotangent at Z - 3 ..
// Z + B - 3 (both in LDS, visible to the block). SPLIT (2 B threads): thread
// t < B computes the right reconstruction of cell Z - 3 + t and its
// gradient, thread B + t the left one (the two chains side by side, for
// grids too small to fill the device); otherwise (B threads) thread t
// computes both. Thread t returns the gathered sum of output cell Z + t
// (valid for t < B - 6; threads B .. 2 B - 1 return 0). Both forms compute
// the same values with the same roundings. s_u is read only before the
// second barrier (a caller may overwrite it once this returns). Ends with
// the scratch in use.
template <typename T, int FLUX, bool LEARNED = false, bool SPLIT = true>
__device__ __forceinline__ T vjp_block(const VjpArgs<T> &p, const T *s_u,
                                       const T *s_mu, VjpScratch<T, SPLIT> &sc) {
  constexpr int B = kVjpBlock;
  const int tid = threadIdx.x;
  const bool left = SPLIT && tid >= B;
  const int c = left ? tid - B : tid;

  // ---- one reconstruction of computing cell Z - 3 + c and its gradient:
  // the right one from v0 .. v4, the left one from the mirrored stencil
  T G[5];
  const T v0 = s_u[c], v1 = s_u[c + 1], v2 = s_u[c + 2], v3 = s_u[c + 3], v4 = s_u[c + 4];
  const T w0 = left ? v4 : v0, w1 = left ? v3 : v1, w3 = left ? v1 : v3,
          w4 = left ? v0 : v4;
  T rec;
  if constexpr (LEARNED) {
    const int regime = learned_regime5(v0, v1, v2, v3, v4, p.eps4, p.tau, p.sigma);
    if (regime == 0) {
      rec = weno_side_grad(w0, w1, v2, w3, w4, p.eps, G);
    } else if (regime == 1) {
      rec = ideal_side_grad(w0, w1, v2, w3, w4, G);
    } else {
      rec = learned_side_grad(p.mlp, w0, w1, v2, w3, w4, G);
    }
  } else {
    rec = weno_side_grad(w0, w1, v2, w3, w4, p.eps, G);
  }
  if constexpr (SPLIT) {
    if (left) {
      sc.ul[c] = rec;
#pragma unroll
      for (int k = 0; k < 5; ++k) sc.GL[k][c] = G[k];
    } else {
      sc.ur[c] = rec;
    }
  } else {
    sc.ur[c] = rec;
  }
  T GLm[5];
  if constexpr (!SPLIT) {
    // the left reconstruction in the same thread, kept in registers
    T recl;
    if constexpr (LEARNED) {
      const int regime = learned_regime5(v0, v1, v2, v3, v4, p.eps4, p.tau, p.sigma);
      if (regime == 0) {
        recl = weno_side_grad(v4, v3, v2, v1, v0, p.eps, GLm);
      } else if (regime == 1) {
        recl = ideal_side_grad(v4, v3, v2, v1, v0, GLm);
      } else {
        recl = learned_side_grad(p.mlp, v4, v3, v2, v1, v0, GLm);
      }
    } else {
      recl = weno_side_grad(v4, v3, v2, v1, v0, p.eps, GLm);
    }
    sc.ul[c] = recl;
  }
  if (tid == 0) {
    sc.cL[0] = T(0);
    sc.sR[0] = T(0);
  }
  __syncthreads();

  // ---- face c + 1/2: cotangent of the flux and its partial derivatives
  if (tid < B - 1) {
    const T urL = sc.ur[tid], ulR = sc.ul[tid + 1];
    const T uL = s_u[tid + 2], uR = s_u[tid + 3];
    T gF = (s_mu[tid + 1] - s_mu[tid]) * p.inv_dx;
    T dur = T(0), dul = T(0), duL = T(0), duR = T(0);
    if constexpr (FLUX == PSA_FLUX_RUSANOV) {
      // F = 0.5 (ul^2/2 + ur^2/2) - 0.5 a (ul - ur), a = max(|uR|, |uL|)
      const T aR = fabs(uR), aL = fabs(uL);
      const T a = aR > aL ? aR : aL;
      dur = T(0.5) * urL + T(0.5) * a;
      dul = T(0.5) * ulR - T(0.5) * a;
      const T dFda = T(-0.5) * (ulR - urL);
      const T wR = aR > aL ? T(1) : (aR < aL ? T(0) : T(0.5));
      duR = dFda * wR * sgn(uR);
      duL = dFda * (T(1) - wR) * sgn(uL);
    } else if constexpr (FLUX == PSA_FLUX_LAX_FRIEDRICHS) {
      // F = 0.5 (ul^2/2 + ur^2/2) - 0.5 alpha (ul - ur) with the global speed
      // alpha, whose own derivative k_lf_speed_vjp adds
      dur = T(0.5) * urL + T(0.5) * p.alpha;
      dul = T(0.5) * ulR - T(0.5) * p.alpha;
    } else if constexpr (FLUX == PSA_FLUX_ENGQUIST_OSHER) {
      // F = max(ur, 0)^2 / 2 + min(ul, 0)^2 / 2
      dur = urL > T(0) ? urL : T(0);
      dul = ulR < T(0) ? ulR : T(0);
    } else if constexpr (FLUX == PSA_FLUX_BURGERS_UPWIND) {
      const T aavg = (urL + ulR) / T(2);
      if (aavg > T(0)) {
        dur = urL;
      } else {
        dul = ulR;
      }
    } else {
      // constant velocity a: L = -a (F[1:] - F[:-1]) / dx, F = ur if a > 0
      gF *= p.vconst;
      if (p.vconst > T(0)) {
        dur = T(1);
      } else {
        dul = T(1);
      }
    }
    sc.cR[tid] = gF * dur;
    sc.cL[tid + 1] = gF * dul;
    sc.sL[tid] = gF * duL;
    sc.sR[tid + 1] = gF * duR;
  }
  __syncthreads();

  // ---- scatter: contribution of cell c's reconstructions to c - 2 .. c + 2
  if (tid >= 1 && tid <= B - 2) {
    const T cR = sc.cR[tid], cL = sc.cL[tid];
    // (one explicit multiply-add, so that both forms round alike)
#pragma unroll
    for (int r = 0; r < 5; ++r) {
      T gl;
      if constexpr (SPLIT) {
        gl = sc.GL[4 - r][tid];
      } else {
        gl = GLm[4 - r];
      }
      sc.P[r][tid] = fma(cR, G[r], cL * gl);
    }
  }
  __syncthreads();

  // ---- gather for output cell Z + tid (local computing index q)
  T gsum = T(0);
  if (tid < B - 6) {
    const int q = tid + 3;
    gsum = (sc.sL[q] + sc.sR[q]);
#pragma unroll
    for (int r = 0; r < 5; ++r) gsum += sc.P[r][q + 2 - r];
  }
  return gsum;
}

// The reverse stage combinations from mu (the cotangent the stage received)
// and the gathered JL^T mu.
template <typename T, int RSTAGE>
__device__ __forceinline__ T vjp_value(T mu, T gsum, T dt, T lam, T lamk2) {
  if constexpr (RSTAGE == 0) {
    return gsum;
  } else if constexpr (RSTAGE == 3) {
    return T(2.0 / 3.0) * (mu + dt * gsum);
  } else if constexpr (RSTAGE == 2) {
    return T(1.0 / 4.0) * (mu + dt * gsum);
  } else {
    return (T(1.0 / 3.0) * lam + T(3.0 / 4.0) * lamk2) + (mu + dt * gsum);
  }
}

// RSTAGE 0: out = JL(state)^T mu; 3, 2, 1: the reverse stages above, with
// mu = l, lk2, lk1 respectively.
template <typename T, int FLUX, int RSTAGE, bool LEARNED = false>
__global__ __launch_bounds__(kVjpThreads) void k_stage_vjp(VjpArgs<T> p) {
  constexpr int B = kVjpBlock;
  __shared__ T s_u[B + 4];   // cells S - 5 ..
  __shared__ T s_mu[B + 1];  // cells S - 3 ..
  __shared__ VjpScratch<T> sc;

  const int tid = threadIdx.x;
  const int64_t n = p.n;
  const int g = p.g;
  const int64_t S = g + static_cast<int64_t>(blockIdx.x) * kVjpOut;

  if constexpr (FLUX == PSA_FLUX_LAX_FRIEDRICHS) {
    // the speed k_lf_alpha left in the output's first ghost entry (the
    // ghosts are cleared by k_lf_speed_vjp after this launch)
    p.alpha = p.out[0];
  } else if (blockIdx.x == 0 && tid < g) {
    // ghost entries of a cotangent are zero
    p.out[tid] = T(0);
    p.out[p.m - g + tid] = T(0);
  }

  for (int j = tid; j < B + 4; j += kVjpThreads)
    s_u[j] = load_periodic(p.state, S - 5 + j, n, g);
  for (int j = tid; j < B + 1; j += kVjpThreads)
    s_mu[j] = load_periodic(p.mu, S - 3 + j, n, g);
  __syncthreads();

  const T gsum = vjp_block<T, FLUX, LEARNED>(p, s_u, s_mu, sc);
  if (tid < kVjpOut) {
    const int64_t j = S + tid;
    if (j < p.m - g) {
      const T dt = RSTAGE == 0 ? T(0) : *p.dt_ptr;
      const T lam = RSTAGE == 1 ? p.lam[j] : T(0);
      const T lamk2 = RSTAGE == 1 ? p.lamk2[j] : T(0);
      p.out[j] = vjp_value<T, RSTAGE>(s_mu[tid + 3], gsum, dt, lam, lamk2);
    }
  }
}

// Arguments of the fused reverse step.
template <typename T>
struct StepVjpArgs {
  VjpArgs<T> v;     // eps, inv_dx, vconst, dt_ptr, m, n, g
  const T *x;       // the step's input state
  const T *k1;      // its stage states
  const T *k2;
  const T *lam;     // cotangent of the step output
  T *out;           // cotangent of the step input
  int64_t bstride;  // a batch: instance blockIdx.y at offset blockIdx.y *
                    //   bstride of every array, its dt at dt_ptr[blockIdx.y]
};

constexpr int kStepVjpOut = kVjpBlock - 18;  // output cells per workgroup

// The three reverse stages of one step in one pass: a workgroup loads k2,
// k1, x and the incoming cotangent with halos of 11, 8, 5 and 9 cells and runs
// reverse stages 3, 2 and 1 in LDS (vjp_block each): lk2 is exact on
// [S - 6, S + B - 12), lk1 on [S - 3, S + B - 15), the step's cotangent on the
// output cells [S, S + B - 18). The same arithmetic as k_stage_vjp.
template <typename T, int FLUX, bool LEARNED = false, bool SPLIT = true>
__global__ __launch_bounds__(SPLIT ? kVjpThreads : kVjpBlock)
__attribute__((amdgpu_waves_per_eu(SPLIT || LEARNED ? 1 : 5))) void k_step_vjp(
    StepVjpArgs<T> a) {
  constexpr int B = kVjpBlock;
  constexpr int NT = SPLIT ? kVjpThreads : kVjpBlock;
  __shared__ T sK2[B + 4], sK1[B + 4], sX[B + 4];
  __shared__ T sLam[B + 4];  // cells S - 9 ..
  __shared__ VjpScratch<T, SPLIT> sc;
  // lk2 (cells S - 6 ..) takes k2's buffer and lk1 (cells S - 3 ..) k1's:
  // each is written once the reverse stage that reads that state has
  // returned, and vjp_block reads its state only before its second barrier
  // (4 KB less LDS: one more workgroup per CU for the one-thread form)
  T *const sLk2 = sK2;
  T *const sLk1 = sK1;
  const VjpArgs<T> &p = a.v;
  const int tid = threadIdx.x;
  const int64_t n = p.n;
  const int g = p.g;
  const int64_t S = g + static_cast<int64_t>(blockIdx.x) * kStepVjpOut;
  const int64_t off = static_cast<int64_t>(blockIdx.y) * a.bstride;
  const T dt = p.dt_ptr[blockIdx.y];
  const T *xk2 = a.k2 + off;
  const T *xk1 = a.k1 + off;
  const T *xx = a.x + off;
  const T *xlam = a.lam + off;
  T *xout = a.out + off;

  if (blockIdx.x == 0 && tid < g) {
    xout[tid] = T(0);
    xout[p.m - g + tid] = T(0);
  }
  for (int j = tid; j < B + 4; j += NT) {
    sK2[j] = load_periodic(xk2, S - 11 + j, n, g);
    sK1[j] = load_periodic(xk1, S - 8 + j, n, g);
    sX[j] = load_periodic(xx, S - 5 + j, n, g);
    sLam[j] = load_periodic(xlam, S - 9 + j, n, g);
  }
  __syncthreads();

  // reverse stage 3 on the output cells S - 6 + t: lk2 (with zeros in the
  // entries past the exact cotangents, read only for cells outside the exact
  // ranges: finite values)
  T gsum = vjp_block<T, FLUX, LEARNED, SPLIT>(p, sK2, sLam, sc);
  if (tid < B - 6) sLk2[tid] = vjp_value<T, 3>(sLam[tid + 3], gsum, dt, T(0), T(0));
  for (int j = B - 6 + tid; j < B + 1; j += B) sLk2[j] = T(0);
  __syncthreads();
  // reverse stage 2 on S - 3 + t: lk1
  gsum = vjp_block<T, FLUX, LEARNED, SPLIT>(p, sK1, sLk2, sc);
  if (tid < B - 6) sLk1[tid] = vjp_value<T, 2>(sLk2[tid + 3], gsum, dt, T(0), T(0));
  for (int j = B - 6 + tid; j < B + 1; j += B) sLk1[j] = T(0);
  __syncthreads();
  // reverse stage 1 on S + t: the cotangent of the step input
  gsum = vjp_block<T, FLUX, LEARNED, SPLIT>(p, sX, sLk1, sc);
  if (tid < kStepVjpOut) {
    const int64_t j = S + tid;
    if (j < p.m - g)
      xout[j] = vjp_value<T, 1>(sLk1[tid + 3], gsum, dt, sLam[tid + 9], sLk2[tid + 6]);
  }
}

template <typename T>
__global__ void k_zero_ghosts(T *u, int64_t m, int g, int64_t ld) {
  const int t = threadIdx.x;
  T *v = u + static_cast<int64_t>(blockIdx.x) * ld;
  if (t < g) {
    v[t] = T(0);
    v[m - g + t] = T(0);
  }
}

// ---------------------------------------------------------------------------
// Lax-Friedrichs: the global speed alpha = max |state| over the ghost-filled
// array (scalar.py:276-277) couples every face to the cells where |state|
// is largest. One workgroup each: k_lf_alpha puts alpha into out[0] (a
// ghost entry) for k_stage_vjp, which differentiates the faces at fixed
// alpha; k_lf_speed_vjp then adds the derivative through alpha, G =
// sum over the faces of gF dF/dalpha, dF/dalpha = -(ul - ur) / 2, shared by
// the entries of the ghost-filled array where |state| = alpha (jax's max:
// ties share equally; d|x|/dx = sign(x)), mapped to their cells through the
// periodic fill, and clears the ghost entries.

constexpr int kLfBlock = 256;

template <typename T>
__device__ __forceinline__ T lf_block_sum(T v, T *s_red) {
  const int tid = threadIdx.x;
  s_red[tid] = v;
  __syncthreads();
  for (int w = kLfBlock / 2; w > 0; w >>= 1) {
    if (tid < w) s_red[tid] = s_red[tid] + s_red[tid + w];
    __syncthreads();
  }
  const T r = s_red[0];
  __syncthreads();
  return r;
}

template <typename T>
__global__ __launch_bounds__(kLfBlock) void k_lf_alpha(VjpArgs<T> p) {
  __shared__ T s_red[kLfBlock];
  const int tid = threadIdx.x;
  T mx = T(0);
  bool nan = false;
  for (int64_t c = p.g + tid; c < p.m - p.g; c += kLfBlock) {
    const T v = p.state[c];
    mx = fmax(mx, fabs(v));
    nan |= isnan(v);
  }
  s_red[tid] = nan ? T(NAN) : mx;
  __syncthreads();
  for (int w = kLfBlock / 2; w > 0; w >>= 1) {
    if (tid < w) {
      const T a = s_red[tid], b = s_red[tid + w];
      s_red[tid] = (isnan(a) || isnan(b)) ? T(NAN) : fmax(a, b);
    }
    __syncthreads();
  }
  if (tid == 0) p.out[0] = s_red[0];
}

template <typename T, int RSTAGE>
__global__ __launch_bounds__(kLfBlock) void k_lf_speed_vjp(VjpArgs<T> p) {
  __shared__ T s_red[kLfBlock];
  const int tid = threadIdx.x;
  const int64_t n = p.n, m = p.m;
  const int g = p.g;
  const T alpha = p.out[0];
  // G over the n faces of the periodic ring (face j between cells j - 1, j)
  T acc = T(0);
  for (int64_t j = g + tid; j < m - g; j += kLfBlock) {
    T v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = load_periodic(p.state, j - 3 + k, n, g);
    T G[5];
    const T urL = weno_side_grad(v[0], v[1], v[2], v[3], v[4], p.eps, G);
    const T ulR = weno_side_grad(v[5], v[4], v[3], v[2], v[1], p.eps, G);
    const T gF = (load_periodic(p.mu, j, n, g) - load_periodic(p.mu, j - 1, n, g)) * p.inv_dx;
    acc += gF * (T(-0.5) * (ulR - urL));
  }
  const T G = lf_block_sum(acc, s_red);
  // entries of the ghost-filled array that hold cell c: c itself, the left
  // ghost c - n for cells n .. n + g - 1 and the right ghost c + n for cells
  // g .. 2g - 1
  auto mult = [&](int64_t c) {
    return 1 + ((c >= n && c < n + g) ? 1 : 0) + ((c >= g && c < 2 * g) ? 1 : 0);
  };
  T cnt = T(0);
  for (int64_t c = g + tid; c < m - g; c += kLfBlock)
    if (fabs(p.state[c]) == alpha) cnt += T(mult(c));
  const T count = lf_block_sum(cnt, s_red);
  T coef = T(1);
  if constexpr (RSTAGE == 3) {
    coef = T(2.0 / 3.0) * *p.dt_ptr;
  } else if constexpr (RSTAGE == 2) {
    coef = T(1.0 / 4.0) * *p.dt_ptr;
  } else if constexpr (RSTAGE == 1) {
    coef = *p.dt_ptr;
  }
  if (count > T(0)) {
    const T share = coef * G / count;
    for (int64_t c = g + tid; c < m - g; c += kLfBlock) {
      const T v = p.state[c];
      if (fabs(v) == alpha) p.out[c] += share * T(mult(c)) * sgn(v);
    }
  }
  // the ghost entries (out[0] carried alpha; every thread read it above)
  if (tid < g) {
    p.out[tid] = T(0);
    p.out[m - g + tid] = T(0);
  }
}

template <typename T, int FLUX, bool LEARNED = false>
int launch_vjp_flux(int rstage, const VjpArgs<T> &a, hipStream_t s) {
  const unsigned int blocks =
      static_cast<unsigned int>((a.n + kVjpOut - 1) / kVjpOut);
  switch (rstage) {
    case 0:
      hipLaunchKernelGGL((k_stage_vjp<T, FLUX, 0, LEARNED>), dim3(blocks),
                         dim3(kVjpThreads), 0, s, a);
      break;
    case 1:
      hipLaunchKernelGGL((k_stage_vjp<T, FLUX, 1, LEARNED>), dim3(blocks),
                         dim3(kVjpThreads), 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((k_stage_vjp<T, FLUX, 2, LEARNED>), dim3(blocks),
                         dim3(kVjpThreads), 0, s, a);
      break;
    default:
      hipLaunchKernelGGL((k_stage_vjp<T, FLUX, 3, LEARNED>), dim3(blocks),
                         dim3(kVjpThreads), 0, s, a);
      break;
  }
  return check_launch("stage vjp");
}

template <typename T>
int launch_vjp(const psa_solver_desc *d, int rstage, const void *state,
               const void *mu, const void *lam, const void *lamk2, void *out,
               const void *dt, hipStream_t s) {
  VjpArgs<T> a{};
  a.state = static_cast<const T *>(state);
  a.mu = static_cast<const T *>(mu);
  a.lam = static_cast<const T *>(lam);
  a.lamk2 = static_cast<const T *>(lamk2);
  a.out = static_cast<T *>(out);
  a.dt_ptr = static_cast<const T *>(dt);
  a.n = d->n;
  a.g = d->g;
  a.m = d->n + 2 * static_cast<int64_t>(d->g);
  a.inv_dx = T(1) / static_cast<T>(d->dx);
  a.eps = static_cast<T>(d->eps);
  a.vconst = static_cast<T>(d->velocity_const);
  a.mlp = d->mlp;
  a.eps4 = T(4) * a.eps;
  a.tau = static_cast<T>(d->mlp_tau);
  a.sigma = static_cast<T>(d->mlp_sigma);
  if (d->mlp != nullptr) {
    if (d->flux == PSA_FLUX_RUSANOV)
      return launch_vjp_flux<T, PSA_FLUX_RUSANOV, true>(rstage, a, s);
    if (d->flux == PSA_FLUX_ENGQUIST_OSHER)
      return launch_vjp_flux<T, PSA_FLUX_ENGQUIST_OSHER, true>(rstage, a, s);
    return launch_vjp_flux<T, PSA_FLUX_BURGERS_UPWIND, true>(rstage, a, s);
  }
  if (d->flux == PSA_FLUX_LAX_FRIEDRICHS) {
    hipLaunchKernelGGL(k_lf_alpha<T>, dim3(1), dim3(kLfBlock), 0, s, a);
    int st = check_launch("speed of the stage");
    if (st != PSA_OK) return st;
    st = launch_vjp_flux<T, PSA_FLUX_LAX_FRIEDRICHS>(rstage, a, s);
    if (st != PSA_OK) return st;
    switch (rstage) {
      case 0:
        hipLaunchKernelGGL((k_lf_speed_vjp<T, 0>), dim3(1), dim3(kLfBlock), 0, s, a);
        break;
      case 1:
        hipLaunchKernelGGL((k_lf_speed_vjp<T, 1>), dim3(1), dim3(kLfBlock), 0, s, a);
        break;
      case 2:
        hipLaunchKernelGGL((k_lf_speed_vjp<T, 2>), dim3(1), dim3(kLfBlock), 0, s, a);
        break;
      default:
        hipLaunchKernelGGL((k_lf_speed_vjp<T, 3>), dim3(1), dim3(kLfBlock), 0, s, a);
        break;
    }
    return check_launch("speed vjp");
  }
  if (d->flux == PSA_FLUX_RUSANOV)
    return launch_vjp_flux<T, PSA_FLUX_RUSANOV>(rstage, a, s);
  if (d->flux == PSA_FLUX_BURGERS_UPWIND)
    return launch_vjp_flux<T, PSA_FLUX_BURGERS_UPWIND>(rstage, a, s);
  if (d->flux == PSA_FLUX_ENGQUIST_OSHER)
    return launch_vjp_flux<T, PSA_FLUX_ENGQUIST_OSHER>(rstage, a, s);
  return launch_vjp_flux<T, PSA_FLUX_ADVECTION_UPWIND>(rstage, a, s);
}

// Grids of at most this many workgroups run the split reverse step.
constexpr long long kVjpSplitMaxBlocks = 512;

template <typename T, bool SPLIT>
int launch_step_vjp_plain(const psa_solver_desc *d, dim3 grid, const StepVjpArgs<T> &a,
                          hipStream_t s) {
  const dim3 block(SPLIT ? kVjpThreads : kVjpBlock);
  if (d->flux == PSA_FLUX_RUSANOV) {
    hipLaunchKernelGGL((k_step_vjp<T, PSA_FLUX_RUSANOV, false, SPLIT>), grid, block, 0, s,
                       a);
  } else if (d->flux == PSA_FLUX_BURGERS_UPWIND) {
    hipLaunchKernelGGL((k_step_vjp<T, PSA_FLUX_BURGERS_UPWIND, false, SPLIT>), grid, block,
                       0, s, a);
  } else if (d->flux == PSA_FLUX_ENGQUIST_OSHER) {
    hipLaunchKernelGGL((k_step_vjp<T, PSA_FLUX_ENGQUIST_OSHER, false, SPLIT>), grid, block,
                       0, s, a);
  } else {
    hipLaunchKernelGGL((k_step_vjp<T, PSA_FLUX_ADVECTION_UPWIND, false, SPLIT>), grid,
                       block, 0, s, a);
  }
  return check_launch("step vjp");
}

// The three reverse stages of one step as one launch (k_step_vjp).
template <typename T>
int launch_step_vjp(const psa_solver_desc *d, const void *x, const void *k1,
                    const void *k2, const void *lam, void *out, const void *dt,
                    hipStream_t s, long long batch = 1, int64_t ld = 0) {
  StepVjpArgs<T> a{};
  a.bstride = ld;
  a.v.mlp = d->mlp;
  a.v.eps4 = T(4) * static_cast<T>(d->eps);
  a.v.tau = static_cast<T>(d->mlp_tau);
  a.v.sigma = static_cast<T>(d->mlp_sigma);
  a.v.dt_ptr = static_cast<const T *>(dt);
  a.v.n = d->n;
  a.v.g = d->g;
  a.v.m = d->n + 2 * static_cast<int64_t>(d->g);
  a.v.inv_dx = T(1) / static_cast<T>(d->dx);
  a.v.eps = static_cast<T>(d->eps);
  a.v.vconst = static_cast<T>(d->velocity_const);
  a.x = static_cast<const T *>(x);
  a.k1 = static_cast<const T *>(k1);
  a.k2 = static_cast<const T *>(k2);
  a.lam = static_cast<const T *>(lam);
  a.out = static_cast<T *>(out);
  const dim3 grid(static_cast<unsigned int>((d->n + kStepVjpOut - 1) / kStepVjpOut),
                  static_cast<unsigned int>(batch));
  if (d->mlp != nullptr) {
    if (d->flux == PSA_FLUX_RUSANOV) {
      hipLaunchKernelGGL((k_step_vjp<T, PSA_FLUX_RUSANOV, true>), grid, dim3(kVjpThreads), 0,
                         s, a);
    } else if (d->flux == PSA_FLUX_ENGQUIST_OSHER) {
      hipLaunchKernelGGL((k_step_vjp<T, PSA_FLUX_ENGQUIST_OSHER, true>), grid,
                         dim3(kVjpThreads), 0, s, a);
    } else {
      hipLaunchKernelGGL((k_step_vjp<T, PSA_FLUX_BURGERS_UPWIND, true>), grid,
                         dim3(kVjpThreads), 0, s, a);
    }
    return check_launch("learned step vjp");
  }
  // a grid that leaves the device short of waves splits each cell's two
  // reconstructions between two threads (latency); a larger one keeps them
  // in one thread (throughput)
  if (static_cast<long long>(grid.x) * grid.y <= kVjpSplitMaxBlocks)
    return launch_step_vjp_plain<T, true>(d, grid, a, s);
  return launch_step_vjp_plain<T, false>(d, grid, a, s);
}

int check_vjp_desc(const psa_solver_desc *d) {
  if (d == nullptr) return set_error(PSA_ERR_INVALID_ARGUMENT, "null desc");
  if (d->dtype != PSA_FLOAT64 && d->dtype != PSA_FLOAT32)
    return set_error(PSA_ERR_NOT_IMPLEMENTED, "unsupported dtype");
  if (d->rec != PSA_REC_WENOJS53)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the adjoint supports the WENO-JS5 reconstruction only");
  if (d->bc != PSA_BC_PERIODIC)
    return set_error(PSA_ERR_NOT_IMPLEMENTED, "the adjoint is periodic");
  if (d->g < 3 || d->n < d->g)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad grid size");
  if (d->flux < 0 || d->flux > PSA_FLUX_LAX_FRIEDRICHS)
    return set_error(PSA_ERR_NOT_IMPLEMENTED, "unsupported flux");
  if ((d->flux == PSA_FLUX_RUSANOV || d->flux == PSA_FLUX_LAX_FRIEDRICHS) &&
      d->nu != nullptr)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the adjoint supports the Rusanov and Lax-Friedrichs fluxes "
                     "with alpha = 1");
  if (d->flux == PSA_FLUX_LAX_FRIEDRICHS && d->mlp != nullptr)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the adjoint of learned weights takes Rusanov, Godunov and "
                     "Engquist-Osher");
  if (d->flux == PSA_FLUX_ADVECTION_UPWIND && !d->velocity_is_const)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the adjoint supports advection with a constant velocity");
  if (d->mlp != nullptr && d->flux == PSA_FLUX_ADVECTION_UPWIND)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the adjoint of learned weights takes the Burgers fluxes");
  if (d->mlp != nullptr && (!(d->mlp_tau >= 0.0) || !(d->mlp_sigma >= 0.0)))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "mlp_tau and mlp_sigma must be >= 0");
  if (!(d->dx > 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "dx must be positive");
  return PSA_OK;
}

size_t elem_size(const psa_solver_desc *d) {
  return d->dtype == PSA_FLOAT64 ? sizeof(double) : sizeof(float);
}

int copy_async(void *dst, const void *src, size_t bytes, hipStream_t s) {
  hipError_t err = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
  if (err != hipSuccess) return set_error(PSA_ERR_HIP, hipGetErrorString(err));
  return PSA_OK;
}

struct AdjointLayout {
  long long ncp;     // checkpoints at steps 0, every, 2 every, ...
  long long nwin;    // arrays of the recompute window
  size_t array_bytes;
  size_t bytes;
};

// batch instances at a stride of ld elements per array (ld = m for one)
AdjointLayout adjoint_layout(const psa_solver_desc *d, long long nsteps,
                             int every, long long batch = 1, int64_t ld = 0) {
  AdjointLayout L{};
  const size_t m = static_cast<size_t>(d->n + 2 * static_cast<int64_t>(d->g));
  const size_t cells = batch > 1 ? static_cast<size_t>(batch) * static_cast<size_t>(ld) : m;
  L.array_bytes = ((cells * elem_size(d) + 255) / 256) * 256;
  L.ncp = nsteps / every + 1;
  L.nwin = 3LL * every;
  // checkpoints, window (x, k1, k2 per step), running state, 4 cotangents,
  // and 256 bytes for the Lax-Friedrichs stage speed
  L.bytes = static_cast<size_t>(L.ncp + L.nwin + 1 + 4) * L.array_bytes + 256;
  return L;
}

// The device scalar of the Lax-Friedrichs stage speed in the workspace.
void *lf_speed_slot(void *workspace, const AdjointLayout &L) {
  return static_cast<char *>(workspace) + (L.bytes - 256);
}

}  // namespace
}  // namespace psa

// defined in forward.hip and reductions.hip
extern "C" int psa_ssprk33_stage(const psa_solver_desc *desc, int stage,
                                 const void *un, const void *uk, void *out,
                                 void *control, const void *dt, void *stream);
extern "C" int psa_max_abs_interior(int dtype, const void *u, void *out, int64_t m,
                                    int g, void *stream);

using namespace psa;

namespace {

// One fixed-dt stage (psa_ssprk33_stage without a control block); for
// Lax-Friedrichs the flux speed max|uk| of the stage input first goes to the
// device scalar speed (scalar.py:276-277: the interior maximum, the ghosts
// being images).
int forward_stage(const psa_solver_desc *d, int stage, const void *un, const void *uk,
                  void *out, const void *dt, void *speed, void *stream) {
  if (d->flux != PSA_FLUX_LAX_FRIEDRICHS)
    return psa_ssprk33_stage(d, stage, un, uk, out, nullptr, dt, stream);
  const int64_t m = d->n + 2 * static_cast<int64_t>(d->g);
  const int st = psa_max_abs_interior(d->dtype, uk, speed, m, d->g, stream);
  if (st != PSA_OK) return st;
  psa_solver_desc dl = *d;
  dl.velocity = speed;
  return psa_ssprk33_stage(&dl, stage, un, uk, out, nullptr, dt, stream);
}

}  // namespace

extern "C" {

int psa_ssprk33_stage_vjp(const psa_solver_desc *desc, int stage,
                          const void *state, const void *mu, const void *lam,
                          const void *lam_k2, void *out, const void *dt,
                          void *stream) {
  int st = check_vjp_desc(desc);
  if (st != PSA_OK) return st;
  if (stage < 0 || stage > 3)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "stage must be 0..3");
  if (state == nullptr || mu == nullptr || out == nullptr ||
      (stage > 0 && dt == nullptr) ||
      (stage == 1 && (lam == nullptr || lam_k2 == nullptr)))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (out == mu || out == state || (stage == 1 && (out == lam || out == lam_k2)))
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "the output must not alias an input");
  auto s = static_cast<hipStream_t>(stream);
  if (desc->dtype == PSA_FLOAT64)
    return launch_vjp<double>(desc, stage, state, mu, lam, lam_k2, out, dt, s);
  return launch_vjp<float>(desc, stage, state, mu, lam, lam_k2, out, dt, s);
}

int psa_ssprk33_step_vjp(const psa_solver_desc *desc, const void *u,
                         const void *lam, void *lam_out, void *work,
                         size_t work_bytes, const void *dt, void *stream) {
  int st = check_vjp_desc(desc);
  if (st != PSA_OK) return st;
  if (u == nullptr || lam == nullptr || lam_out == nullptr || work == nullptr ||
      dt == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  const size_t m = static_cast<size_t>(desc->n + 2 * static_cast<int64_t>(desc->g));
  const size_t ab = ((m * elem_size(desc) + 255) / 256) * 256;
  if (work_bytes < 4 * ab)
    return set_error(PSA_ERR_WORKSPACE, "step vjp needs 4 arrays of workspace");
  if (lam_out == lam || lam_out == u)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "lam_out must not alias");
  char *w = static_cast<char *>(work);
  void *k1 = w, *k2 = w + ab, *lk2 = w + 2 * ab, *lk1 = w + 3 * ab;
  // recompute the stages of the forward step (a Lax-Friedrichs stage speed
  // in lk1's storage, which the reverse stages write later)
  st = forward_stage(desc, 1, u, u, k1, dt, lk1, stream);
  if (st != PSA_OK) return st;
  st = forward_stage(desc, 2, u, k1, k2, dt, lk1, stream);
  if (st != PSA_OK) return st;

  st = psa_ssprk33_stage_vjp(desc, 3, k2, lam, nullptr, nullptr, lk2, dt, stream);
  if (st != PSA_OK) return st;
  st = psa_ssprk33_stage_vjp(desc, 2, k1, lk2, nullptr, nullptr, lk1, dt, stream);
  if (st != PSA_OK) return st;
  return psa_ssprk33_stage_vjp(desc, 1, u, lk1, lam, lk2, lam_out, dt, stream);
}

}  // extern "C"

namespace {

// The batch arguments: batch >= 1 instances at ld >= m elements apart (ld a
// multiple of 16 bytes), whole-step kernels only when batch > 1.
int check_batch(const psa_solver_desc *d, long long batch, long long ld) {
  const int64_t m = d->n + 2 * static_cast<int64_t>(d->g);
  if (batch < 1 || batch > 65535)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "batch must be in [1, 65535]");
  if (batch > 1 && (ld < m || (ld * static_cast<long long>(elem_size(d))) % 16 != 0))
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "ld must be >= m and a multiple of 16 bytes");
  return PSA_OK;
}

// With one window over the whole horizon the forward pass keeps the window's
// states itself (the window's 3 every arrays hold all of them, X(nsteps)
// being at most the running-state array X(every)).
bool keeps_window(long long nsteps, int every) { return nsteps > 0 && every >= nsteps; }

int forward_checkpointed(const psa_solver_desc *desc, long long batch, int64_t ld,
                         const void *u0, void *u_final, long long nsteps,
                         const void *dt, int every, void *workspace,
                         size_t workspace_bytes, hipStream_t s) {
  int st = check_vjp_desc(desc);
  if (st != PSA_OK) return st;
  st = check_batch(desc, batch, ld);
  if (st != PSA_OK) return st;
  if (u0 == nullptr || u_final == nullptr || dt == nullptr ||
      workspace == nullptr || nsteps < 0 || every <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad arguments");
  const AdjointLayout L = adjoint_layout(desc, nsteps, every, batch, ld);
  if (workspace_bytes < L.bytes)
    return set_error(PSA_ERR_WORKSPACE, "adjoint workspace too small");
  const size_t m = static_cast<size_t>(desc->n + 2 * static_cast<int64_t>(desc->g));
  const size_t bytes =
      (batch > 1 ? static_cast<size_t>(batch) * static_cast<size_t>(ld) : m) *
      elem_size(desc);
  char *base = static_cast<char *>(workspace);
  auto cp = [&](long long k) { return base + k * L.array_bytes; };
  char *win = cp(L.ncp);
  char *cur = win + L.nwin * L.array_bytes;
  void *k1 = win, *k2 = win + L.array_bytes;

  st = copy_async(cp(0), u0, bytes, s);
  if (st != PSA_OK) return st;
  if (keeps_window(nsteps, every)) {
    // one window over the whole horizon: keep every step's states (x, k1,
    // k2) in it now, so that the reverse sweep recomputes nothing
    auto X = [&](long long i) { return win + (3 * i) * L.array_bytes; };
    auto K1 = [&](long long i) { return win + (3 * i + 1) * L.array_bytes; };
    auto K2 = [&](long long i) { return win + (3 * i + 2) * L.array_bytes; };
    st = copy_async(X(0), u0, bytes, s);
    if (st != PSA_OK) return st;
    for (long long i = 0; i < nsteps; ++i) {
      st = step_fixed_dt(desc, X(i), X(i + 1), K1(i), K2(i), dt, s, batch, ld);
      if (st == PSA_OK) continue;
      if (st != PSA_ERR_NOT_IMPLEMENTED) return st;
      if (batch > 1)
        return set_error(PSA_ERR_NOT_IMPLEMENTED,
                         "a batch runs on the whole-step kernel (default or step "
                         "schedule, tiles 256 x 4 or 256 x 2, not Lax-Friedrichs)");
      void *speed = lf_speed_slot(workspace, L);
      st = forward_stage(desc, 1, X(i), X(i), K1(i), dt, speed, s);
      if (st != PSA_OK) return st;
      st = forward_stage(desc, 2, X(i), K1(i), K2(i), dt, speed, s);
      if (st != PSA_OK) return st;
      st = forward_stage(desc, 3, X(i), K2(i), X(i + 1), dt, speed, s);
      if (st != PSA_OK) return st;
    }
    if (nsteps % every == 0) {
      st = copy_async(cp(1), X(nsteps), bytes, s);
      if (st != PSA_OK) return st;
    }
    return copy_async(u_final, X(nsteps), bytes, s);
  }
  st = copy_async(cur, u0, bytes, s);
  if (st != PSA_OK) return st;
  // whole steps, the state alternating between cur and the window's first
  // array, where the whole-step kernel applies
  char *alt = win;
  for (long long step = 0; step < nsteps; ++step) {
    st = step_fixed_dt(desc, cur, alt, nullptr, nullptr, dt, s, batch, ld);
    if (st == PSA_OK) {
      char *tmp = cur;
      cur = alt;
      alt = tmp;
      if ((step + 1) % every == 0) {
        st = copy_async(cp((step + 1) / every), cur, bytes, s);
        if (st != PSA_OK) return st;
      }
      continue;
    }
    if (st != PSA_ERR_NOT_IMPLEMENTED) return st;
    if (batch > 1)
      return set_error(PSA_ERR_NOT_IMPLEMENTED,
                       "a batch runs on the whole-step kernel (default or step "
                       "schedule, tiles 256 x 4 or 256 x 2, not Lax-Friedrichs)");
    void *speed = lf_speed_slot(workspace, L);
    st = forward_stage(desc, 1, cur, cur, k1, dt, speed, s);
    if (st != PSA_OK) return st;
    st = forward_stage(desc, 2, cur, k1, k2, dt, speed, s);
    if (st != PSA_OK) return st;
    st = forward_stage(desc, 3, cur, k2, cur, dt, speed, s);
    if (st != PSA_OK) return st;
    if ((step + 1) % every == 0) {
      st = copy_async(cp((step + 1) / every), cur, bytes, s);
      if (st != PSA_OK) return st;
    }
  }
  return copy_async(u_final, cur, bytes, s);
}

int adjoint(const psa_solver_desc *desc, long long batch, int64_t ld, long long nsteps,
            const void *dt, int every, const void *lam_final, void *grad_u0,
            void *workspace, size_t workspace_bytes, hipStream_t s) {
  int st = check_vjp_desc(desc);
  if (st != PSA_OK) return st;
  st = check_batch(desc, batch, ld);
  if (st != PSA_OK) return st;
  if (lam_final == nullptr || grad_u0 == nullptr || dt == nullptr ||
      workspace == nullptr || nsteps < 0 || every <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad arguments");
  const AdjointLayout L = adjoint_layout(desc, nsteps, every, batch, ld);
  if (workspace_bytes < L.bytes)
    return set_error(PSA_ERR_WORKSPACE, "adjoint workspace too small");
  const size_t m = static_cast<size_t>(desc->n + 2 * static_cast<int64_t>(desc->g));
  const size_t bytes =
      (batch > 1 ? static_cast<size_t>(batch) * static_cast<size_t>(ld) : m) *
      elem_size(desc);
  char *base = static_cast<char *>(workspace);
  auto cp = [&](long long k) { return base + k * L.array_bytes; };
  char *win = cp(L.ncp);
  auto X = [&](long long i) { return win + (3 * i) * L.array_bytes; };
  auto K1 = [&](long long i) { return win + (3 * i + 1) * L.array_bytes; };
  auto K2 = [&](long long i) { return win + (3 * i + 2) * L.array_bytes; };
  char *lam_a = win + L.nwin * L.array_bytes + L.array_bytes;
  char *lam_b = lam_a + L.array_bytes;
  char *lk2 = lam_b + L.array_bytes;
  char *lk1 = lk2 + L.array_bytes;
  // (Lax-Friedrichs: the stage-wise reverse, whose global speed term needs a
  // reduction between the stages)
  if (desc->flux == PSA_FLUX_LAX_FRIEDRICHS && batch > 1)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the batched adjoint does not take Lax-Friedrichs");
  const bool fused_reverse =
      desc->flux != PSA_FLUX_LAX_FRIEDRICHS &&
      (batch > 1 || (desc->schedule != PSA_SCHEDULE_TILE &&
                     desc->schedule != PSA_SCHEDULE_PIPELINE));

  st = copy_async(lam_a, lam_final, bytes, s);
  if (st != PSA_OK) return st;
  char *lam = lam_a, *lam_next = lam_b;
  const long long nwin = (nsteps + every - 1) / every;
  for (long long w = nwin - 1; w >= 0; --w) {
    const long long s0 = w * every;
    const long long s1 = (s0 + every < nsteps) ? s0 + every : nsteps;
    // recompute the window from its checkpoint, keeping every stage state
    // (kept by the forward pass already when it spans the horizon)
    const bool kept = keeps_window(nsteps, every);
    if (!kept) {
      st = copy_async(X(0), cp(w), bytes, s);
      if (st != PSA_OK) return st;
    }
    for (long long step = s0; step < s1 && !kept; ++step) {
      const long long i = step - s0;
      // one whole step writing k1, k2 and the next state (X(every) is the
      // running-state array, free in this pass)
      st = step_fixed_dt(desc, X(i), X(i + 1), K1(i), K2(i), dt, s, batch, ld);
      if (st == PSA_OK) continue;
      if (st != PSA_ERR_NOT_IMPLEMENTED) return st;
      if (batch > 1)
        return set_error(PSA_ERR_NOT_IMPLEMENTED,
                         "a batch runs on the whole-step kernel (default or step "
                         "schedule, tiles 256 x 4 or 256 x 2, not Lax-Friedrichs)");
      void *speed = lf_speed_slot(workspace, L);
      st = forward_stage(desc, 1, X(i), X(i), K1(i), dt, speed, s);
      if (st != PSA_OK) return st;
      st = forward_stage(desc, 2, X(i), K1(i), K2(i), dt, speed, s);
      if (st != PSA_OK) return st;
      if (step + 1 < s1) {
        st = forward_stage(desc, 3, X(i), K2(i), X(i + 1), dt, speed, s);
        if (st != PSA_OK) return st;
      }
    }
    // reverse sweep through the window
    for (long long step = s1 - 1; step >= s0; --step) {
      const long long i = step - s0;
      if (fused_reverse) {
        // the three reverse stages in one launch
        st = desc->dtype == PSA_FLOAT64
                 ? launch_step_vjp<double>(desc, X(i), K1(i), K2(i), lam, lam_next,
                                           dt, s, batch, ld)
                 : launch_step_vjp<float>(desc, X(i), K1(i), K2(i), lam, lam_next,
                                          dt, s, batch, ld);
        if (st != PSA_OK) return st;
        char *tmp = lam;
        lam = lam_next;
        lam_next = tmp;
        continue;
      }
      st = psa_ssprk33_stage_vjp(desc, 3, K2(i), lam, nullptr, nullptr, lk2, dt, s);
      if (st != PSA_OK) return st;
      st = psa_ssprk33_stage_vjp(desc, 2, K1(i), lk2, nullptr, nullptr, lk1, dt, s);
      if (st != PSA_OK) return st;
      st = psa_ssprk33_stage_vjp(desc, 1, X(i), lk1, lam, lk2, lam_next, dt, s);
      if (st != PSA_OK) return st;
      char *tmp = lam;
      lam = lam_next;
      lam_next = tmp;
    }
  }
  st = copy_async(grad_u0, lam, bytes, s);
  if (st != PSA_OK) return st;
  if (nsteps == 0) {
    // no step: the gradient is the incoming cotangent with zero ghosts
    const int64_t mm = static_cast<int64_t>(m);
    const dim3 grid(static_cast<unsigned int>(batch));
    if (desc->dtype == PSA_FLOAT64) {
      hipLaunchKernelGGL(k_zero_ghosts<double>, grid, dim3(64), 0, s,
                         static_cast<double *>(grad_u0), mm, desc->g, ld);
    } else {
      hipLaunchKernelGGL(k_zero_ghosts<float>, grid, dim3(64), 0, s,
                         static_cast<float *>(grad_u0), mm, desc->g, ld);
    }
    return check_launch("zero ghosts");
  }
  return PSA_OK;
}

}  // namespace

extern "C" {

int psa_adjoint_workspace_bytes(const psa_solver_desc *desc, long long nsteps,
                                int every, size_t *bytes) {
  return psa_adjoint_workspace_bytes_batch(desc, 1, 0, nsteps, every, bytes);
}

int psa_adjoint_workspace_bytes_batch(const psa_solver_desc *desc, long long batch,
                                      long long ld, long long nsteps, int every,
                                      size_t *bytes) {
  int st = check_vjp_desc(desc);
  if (st != PSA_OK) return st;
  st = check_batch(desc, batch, ld);
  if (st != PSA_OK) return st;
  if (bytes == nullptr || nsteps < 0 || every <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad arguments");
  *bytes = adjoint_layout(desc, nsteps, every, batch, ld).bytes;
  return PSA_OK;
}

int psa_ssprk33_forward_checkpointed(const psa_solver_desc *desc,
                                     const void *u0, void *u_final,
                                     long long nsteps, const void *dt,
                                     int every, void *workspace,
                                     size_t workspace_bytes, void *stream) {
  return forward_checkpointed(desc, 1, 0, u0, u_final, nsteps, dt, every, workspace,
                              workspace_bytes, static_cast<hipStream_t>(stream));
}

int psa_ssprk33_forward_checkpointed_batch(const psa_solver_desc *desc, long long batch,
                                           long long ld, const void *u0, void *u_final,
                                           long long nsteps, const void *dt, int every,
                                           void *workspace, size_t workspace_bytes,
                                           void *stream) {
  return forward_checkpointed(desc, batch, ld, u0, u_final, nsteps, dt, every, workspace,
                              workspace_bytes, static_cast<hipStream_t>(stream));
}

int psa_ssprk33_adjoint(const psa_solver_desc *desc, long long nsteps,
                        const void *dt, int every, const void *lam_final,
                        void *grad_u0, void *workspace, size_t workspace_bytes,
                        void *stream) {
  return adjoint(desc, 1, 0, nsteps, dt, every, lam_final, grad_u0, workspace,
                 workspace_bytes, static_cast<hipStream_t>(stream));
}

int psa_ssprk33_adjoint_batch(const psa_solver_desc *desc, long long batch, long long ld,
                              long long nsteps, const void *dt, int every,
                              const void *lam_final, void *grad_u0, void *workspace,
                              size_t workspace_bytes, void *stream) {
  return adjoint(desc, batch, ld, nsteps, dt, every, lam_final, grad_u0, workspace,
                 workspace_bytes, static_cast<hipStream_t>(stream));
}

}  // extern "C"
