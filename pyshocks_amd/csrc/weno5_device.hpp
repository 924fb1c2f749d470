// Device-side arithmetic of the WENO-JS5 / Rusanov / upwind hot path.
//
// Two forms of the reconstruction are provided:
//
// * ``weno_side_ref``: the reference's expression order for one side
//   (weno.py:114-157,247-256 and reconstruction.py:351-355): beta from the
//   (a, b) coefficients, alpha = d / (eps + beta)^2, omega = alpha / sum(alpha),
//   sum(omega * uhat). Used by the standalone weno_* / reconstruct kernels.
//
// * ``weno_pair``: both sides of one cell at once for the fused kernels. The
//   mirrored left reconstruction (reconstruction.py:375 reverses the array)
//   uses the same three smoothness indicators as the right one with beta0 and
//   beta2 swapped, so they are computed once. In FP64 the six weight
//   divisions and the two normalisations are replaced by one division: with
//   q_k = (eps + beta_k)^2, omega_k = (d_k prod_{j != k} q_j) / sum(...), and
//   both sides share the pairwise products q_i q_j. A guard falls back to the
//   reference form when the product leaves the normal range.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <type_traits>

#include "../../include/pyshocks_amd.h"

namespace psa {

template <typename T>
struct Weno5Coeffs {
  // smoothness (weno.py:218-228): a = [13/12, 1/4]
  static constexpr T a0 = T(13.0 / 12.0);
  static constexpr T a1 = T(1.0 / 4.0);
  // interpolation (weno.py:231-239)
  static constexpr T c11 = T(11.0 / 6.0);
  static constexpr T c7 = T(7.0 / 6.0);
  static constexpr T c2 = T(2.0 / 6.0);
  static constexpr T c5 = T(5.0 / 6.0);
  static constexpr T c1 = T(1.0 / 6.0);
  // the fused kernels carry 4 beta_k = (13/3) t_k0^2 + t_k1^2 with 4 eps: a
  // power-of-two scale of every eps + beta_k, which the normalised weights
  // (ratios of products of (eps + beta)^2 of equal degree) do not see
  static constexpr T a0x4 = T(13.0 / 3.0);
  // ideal weights (weno.py:242)
  static constexpr T d0 = T(1.0 / 10.0);
  static constexpr T d1 = T(6.0 / 10.0);
  static constexpr T d2 = T(3.0 / 10.0);
};

template <typename T>
struct Limits;
template <>
struct Limits<double> {
  static constexpr double min_normal = DBL_MIN;
  static constexpr double max_finite = DBL_MAX;
};
template <>
struct Limits<float> {
  static constexpr float min_normal = FLT_MIN;
  static constexpr float max_finite = FLT_MAX;
};

// Smoothness indicators of the right (left-biased) reconstruction at cell i
// from v = (u[i-2], u[i-1], u[i], u[i+1], u[i+2]) (weno.py:114-140 with the
// taps of weno.py:219-226; each row convolves as sum_k b[k] u[i+2-k]).
template <typename T>
__device__ __forceinline__ void weno_beta(T um2, T um1, T u0, T up1, T up2,
                                          T &b0, T &b1, T &b2) {
  using C = Weno5Coeffs<T>;
  const T t00 = (u0 - T(2) * um1) + um2;
  const T t01 = (T(3) * u0 - T(4) * um1) + um2;
  const T t10 = (up1 - T(2) * u0) + um1;
  const T t11 = up1 - um1;
  const T t20 = (up2 - T(2) * up1) + u0;
  const T t21 = (up2 - T(4) * up1) + T(3) * u0;
  b0 = C::a0 * (t00 * t00) + C::a1 * (t01 * t01);
  b1 = C::a0 * (t10 * t10) + C::a1 * (t11 * t11);
  b2 = C::a0 * (t20 * t20) + C::a1 * (t21 * t21);
}

// Candidate interpolants of the right reconstruction (weno.py:143-157).
template <typename T>
__device__ __forceinline__ void weno_uhat(T um2, T um1, T u0, T up1, T up2,
                                          T &h0, T &h1, T &h2) {
  using C = Weno5Coeffs<T>;
  h0 = (C::c11 * u0 - C::c7 * um1) + C::c2 * um2;
  h1 = (C::c2 * up1 + C::c5 * u0) - C::c1 * um1;
  h2 = (-C::c1 * up2 + C::c5 * up1) + C::c2 * u0;
}

// One side in the reference's expression order (weno.py:247-256,
// reconstruction.py:351-355). For the right value at x_{i+1/2} pass
// (u[i-2], ..., u[i+2]); for the left value at x_{i-1/2} pass the mirrored
// stencil (u[i+2], ..., u[i-2]).
template <typename T>
__device__ __forceinline__ T weno_side_ref(T um2, T um1, T u0, T up1, T up2,
                                           T eps) {
  using C = Weno5Coeffs<T>;
  T b0, b1, b2, h0, h1, h2;
  weno_beta(um2, um1, u0, up1, up2, b0, b1, b2);
  weno_uhat(um2, um1, u0, up1, up2, h0, h1, h2);
  const T e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  const T a0 = C::d0 / (e0 * e0);
  const T a1 = C::d1 / (e1 * e1);
  const T a2 = C::d2 / (e2 * e2);
  const T s = (a0 + a1) + a2;
  return ((a0 / s) * h0 + (a1 / s) * h1) + (a2 / s) * h2;
}

// Right and left reconstructions of one cell from shared indicators
// (weno_pair_from_beta below; b0, b1, b2 are the right-side indicators).

// Both sides' weights and values from the right-side indicators and the six
// candidate values (right hr*, left hl*). FP64: the product form with one
// reciprocal (see the top of this file) where the shared denominator is a
// positive normal number, else the reference form.
template <typename T>
__device__ __forceinline__ void weno_combine_pair(T hr0, T hr1, T hr2, T hl0,
                                                  T hl1, T hl2, T b0, T b1,
                                                  T b2, T eps, T &ul, T &ur);

// The product form of both sides (FP64): ur = nr / sr, ul = nl / sl.
struct WenoPairSums {
  double sr, sl, nr, nl;
};

__device__ __forceinline__ WenoPairSums weno_pair_sums(double hr0, double hr1,
                                                       double hr2, double hl0,
                                                       double hl1, double hl2,
                                                       double b0, double b1,
                                                       double b2, double eps) {
  const double e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  const double q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
  const double p01 = q0 * q1, p02 = q0 * q2, p12 = q1 * q2;
  // right: alpha_k ~ d_k / q_k; left: d_0 / q_2, d_1 / q_1, d_2 / q_0; the
  // ideal weights 1/10, 6/10, 3/10 enter in their ratio 1 : 6 : 3, which
  // the normalisation does not see
  const double w1 = 6.0 * p02;
  const double wr2 = 3.0 * p01, wl2 = 3.0 * p12;
  WenoPairSums r;
  r.sr = (p12 + w1) + wr2;
  r.sl = (p01 + w1) + wl2;
  r.nr = (p12 * hr0 + w1 * hr1) + wr2 * hr2;
  r.nl = (p01 * hl0 + w1 * hl1) + wl2 * hl2;
  return r;
}

// Whether the shared reciprocal applies: sr sl a positive normal number.
__device__ __forceinline__ bool weno_pair_fast(const WenoPairSums &s) {
  return __builtin_amdgcn_class(s.sr * s.sl, 0x100);
}

// ur, ul by the shared reciprocal: the hardware reciprocal of sr sl refined
// by two Newton steps (each squares the relative error; within an ulp or two
// of 1 / (sr sl)).
__device__ __forceinline__ void weno_pair_divide(const WenoPairSums &s, double &ul,
                                                 double &ur) {
  const double den = s.sr * s.sl;
  double inv = __builtin_amdgcn_rcp(den);
  double e = fma(-den, inv, 1.0);
  inv = fma(inv, e, inv);
  e = fma(-den, inv, 1.0);
  inv = fma(inv, e, inv);
  ur = (s.nr * s.sl) * inv;
  ul = (s.nl * s.sr) * inv;
}

// Third-difference form of both sides (FP64, the fused kernels). With the
// first differences D(c) = u(c) - u(c-1), the second d2(c) = D(c+1) - D(c)
// and the third T0 = d2(c) - d2(c-1), T1 = d2(c+1) - d2(c), the candidates of
// weno.py:171-202 differ from the middle ones
//   hr1 = u(c) + (D(c) + 2 D(c+1)) / 6,   hl1 = u(c) - (2 D(c) + D(c+1)) / 6
// by hr0 - hr1 = -T0/3, hr2 - hr1 = -T1/6, hl0 - hl1 = T1/3, hl2 - hl1 = T0/6.
// With the weights of weno_pair_sums (right p12 : 6 p02 : 3 p01, left
// p01 : 6 p02 : 3 p12 over their sums sr, sl):
//   ur = hr1 - (p12 T0/3 + p01 T1/2) / sr,  ul = hl1 + (p01 T1/3 + p12 T0/2) / sl,
// both by the one reciprocal of sr sl. e_k = 4 (eps + beta_k) (the scale
// cancels).
struct WenoDiffSums {
  double sr, sl, nr, nl;
};

__device__ __forceinline__ WenoDiffSums weno_diff_sums(double e0, double e1, double e2,
                                                       double t0_3, double t0_2,
                                                       double t1_3, double t1_2) {
  const double q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
  const double p01 = q0 * q1, p02 = q0 * q2, p12 = q1 * q2;
  WenoDiffSums r;
  r.sr = fma(3.0, p01, fma(6.0, p02, p12));
  r.sl = fma(3.0, p12, fma(6.0, p02, p01));
  r.nr = fma(p12, t0_3, p01 * t1_2);
  r.nl = fma(p01, t1_3, p12 * t0_2);
  return r;
}

// Whether the shared reciprocal applies: sr sl a positive normal number.
__device__ __forceinline__ bool weno_diff_fast(const WenoDiffSums &s) {
  return __builtin_amdgcn_class(s.sr * s.sl, 0x100);
}

// ur, ul from the middle candidates by the shared reciprocal (the hardware
// reciprocal of sr sl refined by two Newton steps, as in weno_pair_divide).
__device__ __forceinline__ void weno_diff_divide(const WenoDiffSums &s, double hr1,
                                                 double hl1, double &ul, double &ur) {
  const double den = s.sr * s.sl;
  double inv = __builtin_amdgcn_rcp(den);
  double e = fma(-den, inv, 1.0);
  inv = fma(inv, e, inv);
  e = fma(-den, inv, 1.0);
  inv = fma(inv, e, inv);
  ur = fma(-(s.nr * s.sl), inv, hr1);
  ul = fma(s.nl * s.sr, inv, hl1);
}

// The third-difference form in FP32 (fused kernels): the reciprocal weights
// r_k = 1 / e_k^2 (the product form would underflow), right p12 : 6 p02 :
// 3 p01 becoming r0 : 6 r1 : 3 r2 and left r2 : 6 r1 : 3 r0, so that
//   ur = hr1 - (r0 T0/3 + r2 T1/2) / sr,  ul = hl1 + (r2 T1/3 + r0 T0/2) / sl.
__device__ __forceinline__ void weno_diff_pair_f32(float e0, float e1, float e2,
                                                   float t0_3, float t0_2, float t1_3,
                                                   float t1_2, float hr1, float hl1,
                                                   float &ul, float &ur) {
  const float r0 = __builtin_amdgcn_rcpf(e0 * e0);
  const float r1 = __builtin_amdgcn_rcpf(e1 * e1);
  const float r2 = __builtin_amdgcn_rcpf(e2 * e2);
  const float a1 = 6.0f * r1;
  const float sr = fmaf(3.0f, r2, r0 + a1);
  const float sl = fmaf(3.0f, r0, r2 + a1);
  const float nr = fmaf(r0, t0_3, r2 * t1_2);
  const float nl = fmaf(r2, t1_3, r0 * t0_2);
  ur = fmaf(-nr, __builtin_amdgcn_rcpf(sr), hr1);
  ul = fmaf(nl, __builtin_amdgcn_rcpf(sl), hl1);
}

// The reference form from q_k = (eps + beta_k)^2: safe over the whole
// exponent range.
__device__ __forceinline__ void weno_pair_ref_q(double hr0, double hr1, double hr2,
                                                double hl0, double hl1, double hl2,
                                                double q0, double q1, double q2,
                                                double &ul, double &ur) {
  using C = Weno5Coeffs<double>;
  const double a0 = C::d0 / q0, a1 = C::d1 / q1, a2 = C::d2 / q2;
  const double s = (a0 + a1) + a2;
  ur = ((a0 / s) * hr0 + (a1 / s) * hr1) + (a2 / s) * hr2;
  const double l0 = C::d0 / q2, l2 = C::d2 / q0;
  const double sll = (l0 + a1) + l2;
  ul = ((l0 / sll) * hl0 + (a1 / sll) * hl1) + (l2 / sll) * hl2;
}

template <>
__device__ __forceinline__ void weno_combine_pair<double>(
    double hr0, double hr1, double hr2, double hl0, double hl1, double hl2,
    double b0, double b1, double b2, double eps, double &ul, double &ur) {
  const WenoPairSums s = weno_pair_sums(hr0, hr1, hr2, hl0, hl1, hl2, b0, b1, b2, eps);
  if (weno_pair_fast(s)) {
    weno_pair_divide(s, ul, ur);
  } else {
    const double e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
    weno_pair_ref_q(hr0, hr1, hr2, hl0, hl1, hl2, e0 * e0, e1 * e1, e2 * e2, ul, ur);
  }
}

template <>
__device__ __forceinline__ void weno_combine_pair<float>(
    float hr0, float hr1, float hr2, float hl0, float hl1, float hl2, float b0,
    float b1, float b2, float eps, float &ul, float &ur) {
  const float e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  const float r0 = __builtin_amdgcn_rcpf(e0 * e0);
  const float r1 = __builtin_amdgcn_rcpf(e1 * e1);
  const float r2 = __builtin_amdgcn_rcpf(e2 * e2);
  // the ideal weights in their ratio 1 : 6 : 3, as in the FP64 form
  const float a1 = 6.0f * r1;
  const float ar2 = 3.0f * r2, al2 = 3.0f * r0;
  const float sr = (r0 + a1) + ar2;
  const float sl = (r2 + a1) + al2;
  ur = ((r0 * hr0 + a1 * hr1) + ar2 * hr2) * __builtin_amdgcn_rcpf(sr);
  ul = ((r2 * hl0 + a1 * hl1) + al2 * hl2) * __builtin_amdgcn_rcpf(sl);
}

template <typename T>
__device__ __forceinline__ void weno_pair_from_beta(T um2, T um1, T u0, T up1,
                                                    T up2, T b0, T b1, T b2,
                                                    T eps, T &ul, T &ur) {
  T hr0, hr1, hr2, hl0, hl1, hl2;
  weno_uhat(um2, um1, u0, up1, up2, hr0, hr1, hr2);
  weno_uhat(up2, up1, u0, um1, um2, hl0, hl1, hl2);
  weno_combine_pair(hr0, hr1, hr2, hl0, hl1, hl2, b0, b1, b2, eps, ul, ur);
}

// The six candidate values of a cell (weno.py:143-157, right and mirrored
// left) from u0 and the first differences dm1 = u[i-1] - u[i-2],
// d0 = u[i] - u[i-1], dp1 = u[i+1] - u[i], dp2 = u[i+2] - u[i+1]: the same
// quadratic interpolants written as u0 plus differences, two multiply-adds
// each (hr0 = u0 + 5/6 d0 - 1/3 dm1 = 11/6 u0 - 7/6 u[i-1] + 1/3 u[i-2], ...).
template <typename T>
__device__ __forceinline__ void weno_uhat_diff(T u0, T dm1, T d0, T dp1, T dp2,
                                               T &hr0, T &hr1, T &hr2, T &hl0,
                                               T &hl1, T &hl2) {
  hr0 = fma(T(-1.0 / 3.0), dm1, fma(T(5.0 / 6.0), d0, u0));
  hr1 = fma(T(1.0 / 6.0), d0, fma(T(1.0 / 3.0), dp1, u0));
  hr2 = fma(T(-1.0 / 6.0), dp2, fma(T(2.0 / 3.0), dp1, u0));
  hl0 = fma(T(1.0 / 3.0), dp2, fma(T(-5.0 / 6.0), dp1, u0));
  hl1 = fma(T(-1.0 / 6.0), dp1, fma(T(-1.0 / 3.0), d0, u0));
  hl2 = fma(T(1.0 / 6.0), dm1, fma(T(-2.0 / 3.0), d0, u0));
}

template <typename T>
__device__ __forceinline__ void weno_pair(T um2, T um1, T u0, T up1, T up2,
                                          T eps, T &ul, T &ur) {
  T b0, b1, b2;
  weno_beta(um2, um1, u0, up1, up2, b0, b1, b2);
  weno_pair_from_beta(um2, um1, u0, up1, up2, b0, b1, b2, eps, ul, ur);
}

// One-sided fast reconstruction (right value at x_{i+1/2}); the left value
// is obtained by passing the mirrored stencil.
template <typename T>
__device__ __forceinline__ T weno_right(T um2, T um1, T u0, T up1, T up2,
                                        T eps) {
  T ul, ur;
  weno_pair(um2, um1, u0, up1, up2, eps, ul, ur);
  return ur;
}

// ---------------------------------------------------------------------------
// Face fluxes. "l" is the cell left of the face (its right reconstruction
// urL), "r" the cell right of it (its left reconstruction ulR).

// scalar.py:238-248 with flux = u^2 / 2 (burgers/schemes.py:37-39):
// F = 0.5 (fl + fr) - 0.5 a nu (ul - ur), a = max(|u_l|, |u_r|).
template <typename T>
__device__ __forceinline__ T flux_rusanov(T urL, T ulR, T uL, T uR, T nu) {
  const T fl = (ulR * ulR) / T(2);
  const T fr = (urL * urL) / T(2);
  const T a = fmax(fabs(uR), fabs(uL));
  return T(0.5) * (fl + fr) - ((T(0.5) * a) * nu) * (ulR - urL);
}

// The same flux with nu = 1 and the two flux halves combined,
// 0.5 (ul^2 / 2 + ur^2 / 2) = 0.25 (ul^2 + ur^2); it differs from the
// reference's expression order by rounding only.
template <typename T>
__device__ __forceinline__ T flux_rusanov_nu1(T urL, T ulR, T uL, T uR) {
  const T a = fmax(fabs(uR), fabs(uL));
  return T(0.25) * fma(ulR, ulR, urL * urL) - (T(0.5) * a) * (ulR - urL);
}

// Twice flux_rusanov_nu1, 0.5 (ul^2 + ur^2) - a (ul - ur): the stage kernels
// fold the factor 1/2 into the update.
template <typename T>
__device__ __forceinline__ T flux_rusanov_nu1_x2(T urL, T ulR, T uL, T uR) {
  const T a = fmax(fabs(uR), fabs(uL));
  return T(0.5) * fma(ulR, ulR, urL * urL) - a * (ulR - urL);
}

// scalar.py:276-277 with scalar.py:238-248: the Rusanov flux with the global
// speed a = max|a_i| for every face (Lax-Friedrichs).
template <typename T>
__device__ __forceinline__ T flux_lax_friedrichs(T urL, T ulR, T a, T nu) {
  const T fl = (ulR * ulR) / T(2);
  const T fr = (urL * urL) / T(2);
  return T(0.5) * (fl + fr) - ((T(0.5) * a) * nu) * (ulR - urL);
}

// Twice flux_lax_friedrichs with nu = 1, 0.5 (ul^2 + ur^2) - a (ul - ur)
// (the form of flux_rusanov_nu1_x2).
template <typename T>
__device__ __forceinline__ T flux_lax_friedrichs_nu1_x2(T urL, T ulR, T a) {
  return T(0.5) * fma(ulR, ulR, urL * urL) - a * (ulR - urL);
}

// scalar.py:309-319 with the Burgers flux and omega = 0
// (burgers/schemes.py:184): F = f(max(ur_l, 0)) + f(min(ul_r, 0)) - f(0),
// f(0) = 0. The comparisons pass a NaN state through, as jnp.maximum and
// jnp.minimum do.
template <typename T>
__device__ __forceinline__ T flux_engquist_osher(T urL, T ulR) {
  const T p = urL < T(0) ? T(0) : urL;
  const T q = ulR > T(0) ? T(0) : ulR;
  return (p * p) / T(2) + (q * q) / T(2);
}

// Twice flux_engquist_osher, p^2 + q^2: the stage kernels fold the factor
// 1/2 into the update.
template <typename T>
__device__ __forceinline__ T flux_engquist_osher_x2(T urL, T ulR) {
  const T p = urL < T(0) ? T(0) : urL;
  const T q = ulR > T(0) ? T(0) : ulR;
  return fma(p, p, q * q);
}

// scalar.py:125-130: aavg = (ar + al) / 2 with a = u reconstructed the same
// way as u, F = fr if aavg > 0 else fl.
template <typename T>
__device__ __forceinline__ T flux_burgers_upwind(T urL, T ulR) {
  const T aavg = (urL + ulR) / T(2);
  return aavg > T(0) ? (urL * urL) / T(2) : (ulR * ulR) / T(2);
}

// advection/schemes.py:110-112: F = ur if (ar + al) / 2 > 0 else ul.
template <typename T>
__device__ __forceinline__ T flux_advection_upwind(T urL, T ulR, T arL,
                                                   T alR) {
  const T aavg = (arL + alR) / T(2);
  return aavg > T(0) ? urL : ulR;
}

// ---------------------------------------------------------------------------
// Third-order reconstructions: WENO-JS3 (weno.py:166-203 coefficients,
// weno.py:247-256 weights, reconstruction.py:320-333) and the ES-WENO weights
// of the same stencils (weno.py:285-296, reconstruction.py:387-434):
//   beta0 = (u_i - u_{i-1})^2, beta1 = (u_{i+1} - u_i)^2,
//   uhat0 = 3/2 u_i - 1/2 u_{i-1}, uhat1 = 1/2 u_{i+1} + 1/2 u_i, d = (1/3, 2/3),
//   JS: alpha_k = d_k / (eps + beta_k)^2,
//   ES: alpha_k = d_k (1 + tau / (eps + beta_k)), tau = (u_{i+1} - 2 u_i + u_{i-1})^2
//       (zero for the first and last cell of the array, jnp.pad(.., 1)).
// The left state reverses the array, which swaps u_{i-1} and u_{i+1}.

template <typename T>
struct Weno3Coeffs {
  static constexpr T d0 = T(1.0 / 3.0);
  static constexpr T d1 = T(2.0 / 3.0);
  static constexpr T c00 = T(3.0 / 2.0);
  static constexpr T c01 = T(-1.0 / 2.0);
  static constexpr T c1 = T(1.0 / 2.0);
};

// One side in the reference's expression order (standalone kernels): the
// right state of cell i from (u_{i-1}, u_i, u_{i+1}), or, with the arguments
// reversed, its left state. es selects the ES-WENO weights with this tau.
template <typename T>
__device__ __forceinline__ T weno3_side_ref(T um1, T u0, T up1, T eps, bool es,
                                            T tau) {
  using C = Weno3Coeffs<T>;
  const T t0 = u0 - um1;
  const T t1 = up1 - u0;
  const T b0 = t0 * t0;
  const T b1 = t1 * t1;
  const T h0 = C::c00 * u0 + C::c01 * um1;
  const T h1 = C::c1 * up1 + C::c1 * u0;
  T a0, a1;
  if (es) {
    a0 = C::d0 * (T(1) + tau / (eps + b0));
    a1 = C::d1 * (T(1) + tau / (eps + b1));
  } else {
    const T e0 = eps + b0, e1 = eps + b1;
    a0 = C::d0 / (e0 * e0);
    a1 = C::d1 / (e1 * e1);
  }
  const T s = a0 + a1;
  return (a0 / s) * h0 + (a1 / s) * h1;
}

template <typename T>
__device__ __forceinline__ T weno3_tau_ref(T um1, T u0, T up1);  // below

// The ES-WENO weight of stencil K (0: d = 1/3 on cells i - 1, i; 1: d = 2/3
// on i, i + 1) of cell i from the cells i - 1 .. i + 1 (weno.py:284-296;
// end: tau zero, the array's first or last cell).
template <typename T, int K>
__device__ __forceinline__ T es_weight(T um1, T u0, T up1, T eps, bool end) {
  using C = Weno3Coeffs<T>;
  const T t0 = u0 - um1;
  const T t1 = up1 - u0;
  const T tau = end ? T(0) : weno3_tau_ref(um1, u0, up1);
  const T a0 = C::d0 * (T(1) + tau / (eps + t0 * t0));
  const T a1 = C::d1 * (T(1) + tau / (eps + t1 * t1));
  return (K == 0 ? a0 : a1) / (a0 + a1);
}

// sqrt of a positive normal FP64 value: the hardware reciprocal square root
// (about 24 bits) refined by one Goldschmidt step (s = x y, h = y / 2,
// r = 1/2 - s h, s += s r, h += h r: ~48 bits) and two corrections
// s += (x - s^2) h, each squaring the relative error; anything else goes to
// the library's sqrt.
__device__ __forceinline__ double sqrt_pos(double x) {
  if (!__builtin_amdgcn_class(x, 0x100)) return sqrt(x);
  const double y = __builtin_amdgcn_rsq(x);
  double s = x * y;
  double h = 0.5 * y;
  const double r = fma(-s, h, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  s = fma(fma(-s, s, x), h, s);
  return fma(fma(-s, s, x), h, s);
}
__device__ __forceinline__ float sqrt_pos(float x) { return sqrtf(x); }

// The dissipative flux from the two cells' weights wl, wr and values v1, v2.
template <typename T>
__device__ __forceinline__ T es_dissipation_w(T wl, T wr, T v1, T v2, T delta) {
  const T dw = wr - wl;
  const T mu = sqrt_pos(fma(dw, dw, delta * delta)) * T(0.125);
  return -(mu + dw * T(0.125)) * (v2 - v1);
}

// The dissipative flux of ES-WENO at the face between cells v1 and v2 from
// the cells v0 .. v3 (burgers/schemes.py:240-252): with dw the difference of
// the two cells' weights, -(sqrt(dw^2 + delta^2) / 8 + dw / 8) (v2 - v1).
template <typename T, int K>
__device__ __forceinline__ T es_dissipation(T v0, T v1, T v2, T v3, T eps, T delta,
                                            bool end_l, bool end_r) {
  const T wl = es_weight<T, K>(v0, v1, v2, eps, end_l);
  const T wr = es_weight<T, K>(v1, v2, v3, eps, end_r);
  const T dw = wr - wl;
  const T mu = sqrt(dw * dw + delta * delta) / T(8);
  return -(mu + dw / T(8)) * (v2 - v1);
}

// tau of ES-WENO for cell i (weno.py:293), (u_{i+1} - 2 u_i) + u_{i-1} squared.
template <typename T>
__device__ __forceinline__ T weno3_tau_ref(T um1, T u0, T up1) {
  const T t = (up1 - T(2) * u0) + um1;
  return t * t;
}

// ur = nr / dr and ul = nl / dl. FP64: one reciprocal of dr dl refined by two
// Newton steps (as in weno_combine_pair), guarded by the class of dr dl (a
// positive normal number), else two divisions; FP32: the hardware
// reciprocal of each denominator.
template <typename T>
__device__ __forceinline__ void weno3_divide(T nr, T dr, T nl, T dl, T &ur,
                                             T &ul);

template <>
__device__ __forceinline__ void weno3_divide<double>(double nr, double dr,
                                                     double nl, double dl,
                                                     double &ur, double &ul) {
  const double den = dr * dl;
  if (__builtin_amdgcn_class(den, 0x100)) {
    double inv = __builtin_amdgcn_rcp(den);
    double e = fma(-den, inv, 1.0);
    inv = fma(inv, e, inv);
    e = fma(-den, inv, 1.0);
    inv = fma(inv, e, inv);
    ur = (nr * dl) * inv;
    ul = (nl * dr) * inv;
  } else {
    ur = nr / dr;
    ul = nl / dl;
  }
}

template <>
__device__ __forceinline__ void weno3_divide<float>(float nr, float dr, float nl,
                                                    float dl, float &ur,
                                                    float &ul) {
  ur = nr * __builtin_amdgcn_rcpf(dr);
  ul = nl * __builtin_amdgcn_rcpf(dl);
}

// The right state's first weight w0r / dr of weno3_pair's ES-WENO branch,
// with the same arithmetic (the shared reciprocal of dr dl in FP64), so that
// a cell's weight has the same bits whichever thread forms it.
template <typename T>
__device__ __forceinline__ T weno3_omega0(T w0r, T dr, T dl);

template <>
__device__ __forceinline__ double weno3_omega0<double>(double w0r, double dr, double dl) {
  const double den = dr * dl;
  if (__builtin_amdgcn_class(den, 0x100)) {
    double inv = __builtin_amdgcn_rcp(den);
    double e = fma(-den, inv, 1.0);
    inv = fma(inv, e, inv);
    e = fma(-den, inv, 1.0);
    inv = fma(inv, e, inv);
    return (w0r * dl) * inv;
  }
  return w0r / dr;
}

template <>
__device__ __forceinline__ float weno3_omega0<float>(float w0r, float dr, float) {
  return w0r * __builtin_amdgcn_rcpf(dr);
}

// Both states of one cell for the fused kernels from the first differences
// Dm = u_i - u_{i-1}, Dp = u_{i+1} - u_i: with q_k the JS denominators
// (eps + beta_k)^2, omega_0 = q1 / (q1 + 2 q0) on the right and
// q0 / (q0 + 2 q1) on the left (beta0 and beta1 swap); with ES-WENO, r_k =
// (eps + beta_k) and p_k = r_k + tau, omega_0 = p0 r1 / (p0 r1 + 2 p1 r0) and
// mirrored. The candidates are u_i plus half a difference.
template <typename T>
__device__ __forceinline__ void weno3_pair(T u0, T Dm, T Dp, T eps, bool es,
                                           bool tau_zero, T &ul, T &ur,
                                           T *omega0 = nullptr) {
  const T r0 = fma(Dm, Dm, eps);  // eps + beta0
  const T r1 = fma(Dp, Dp, eps);  // eps + beta1
  // right: uhat0 = u + Dm / 2, uhat1 = u + Dp / 2; left: u - Dp / 2, u - Dm / 2
  const T hr0 = fma(T(0.5), Dm, u0), hr1 = fma(T(0.5), Dp, u0);
  const T hl0 = fma(T(-0.5), Dp, u0), hl1 = fma(T(-0.5), Dm, u0);
  T w0r, w1r, w0l, w1l;  // unnormalised weights of the right and left states
  if (es) {
    const T dd = Dp - Dm;
    const T tau = tau_zero ? T(0) : dd * dd;
    if constexpr (std::is_same<T, float>::value) {
      // FP32: the reference's scale-free a_k = 1 + tau / r_k (the products
      // below scale as the fourth power of the differences and would
      // overflow above differences of ~1e9); hardware reciprocals
      const float a0 = fmaf(tau, __builtin_amdgcn_rcpf(r0), 1.0f);
      const float a1 = fmaf(tau, __builtin_amdgcn_rcpf(r1), 1.0f);
      w0r = a0;
      w1r = 2.0f * a1;
      w0l = a1;
      w1l = 2.0f * a0;
    } else {
      // a_k = (r_k + tau) / r_k, both sides times r0 r1
      const T p0 = r0 + tau, p1 = r1 + tau;
      w0r = p0 * r1;
      w1r = T(2) * (p1 * r0);
      w0l = p1 * r0;
      w1l = T(2) * (p0 * r1);
    }
  } else {
    const T q0 = r0 * r0, q1 = r1 * r1;
    w0r = q1;
    w1r = T(2) * q0;
    w0l = q0;
    w1l = T(2) * q1;
  }
  const T nr = fma(w0r, hr0, w1r * hr1), dr = w0r + w1r;
  const T nl = fma(w0l, hl0, w1l * hl1), dl = w0l + w1l;
  if constexpr (std::is_same<T, double>::value) {
    if (omega0 != nullptr) {
      // one reciprocal of dr dl for both states and the weight, with
      // weno3_omega0's arithmetic
      const double den = dr * dl;
      if (__builtin_amdgcn_class(den, 0x100)) {
        double inv = __builtin_amdgcn_rcp(den);
        double e = fma(-den, inv, 1.0);
        inv = fma(inv, e, inv);
        e = fma(-den, inv, 1.0);
        inv = fma(inv, e, inv);
        ur = (nr * dl) * inv;
        ul = (nl * dr) * inv;
        *omega0 = (w0r * dl) * inv;
      } else {
        ur = nr / dr;
        ul = nl / dl;
        *omega0 = w0r / dr;
      }
      return;
    }
  }
  weno3_divide(nr, dr, nl, dl, ur, ul);
  if (omega0 != nullptr) *omega0 = weno3_omega0(w0r, dr, dl);
}

// The first ES-WENO weight of the right state of cell i (the weight of the
// dissipative flux, burgers/schemes.py:240-252), from Dm = u_i - u_{i-1},
// Dp = u_{i+1} - u_i, with weno3_pair's arithmetic.
template <typename T>
__device__ __forceinline__ T es_omega0_pair(T Dm, T Dp, T eps, bool tau_zero) {
  const T r0 = fma(Dm, Dm, eps);
  const T r1 = fma(Dp, Dp, eps);
  const T dd = Dp - Dm;
  const T tau = tau_zero ? T(0) : dd * dd;
  if constexpr (std::is_same<T, float>::value) {
    // the products in statements of their own, as in weno3_pair (a
    // multiply-add fuses within an expression only)
    const float a0 = fmaf(tau, __builtin_amdgcn_rcpf(r0), 1.0f);
    const float a1 = fmaf(tau, __builtin_amdgcn_rcpf(r1), 1.0f);
    const float w1r = 2.0f * a1, w1l = 2.0f * a0;
    return weno3_omega0(a0, a0 + w1r, a1 + w1l);
  } else {
    const T p0 = r0 + tau, p1 = r1 + tau;
    const T w0r = p0 * r1, w1r = T(2) * (p1 * r0);
    const T w0l = p1 * r0, w1l = T(2) * (p0 * r1);
    return weno3_omega0(w0r, w0r + w1r, w0l + w1l);
  }
}

// ---------------------------------------------------------------------------
// MUSCL with a flux limiter (reconstruction.py:172-252, limiters.py:110-408).

// phi(r) of limiters.py:194-399 (evaluate); theta is the minmod parameter.
template <typename T>
__device__ __forceinline__ T limiter_phi(int lm, T theta, T r) {
  switch (lm) {
    case PSA_LIMITER_MINMOD:
      return fmax(T(0), fmin(fmin(theta, theta * r), (T(1) + r) / T(2)));
    case PSA_LIMITER_MC:
      return fmax(T(0), fmin(fmin(T(2), T(2) * r), (T(1) + r) / T(2)));
    case PSA_LIMITER_SUPERBEE:
      return fmax(T(0), fmax(fmin(T(1), T(2) * r), fmin(T(2), r)));
    case PSA_LIMITER_VANALBADA1:
      return fmax(T(0), (r * r + r) / (r * r + T(1)));
    case PSA_LIMITER_VANALBADA2:
      return fmax(T(0), (T(2) * r) / (r * r + T(1)));
    case PSA_LIMITER_VANLEER: {
      const T a = fabs(r);
      return fmax(T(0), (r + a) / (T(1) + a));
    }
    case PSA_LIMITER_KOREN:
      return fmax(T(0), fmin(fmin(T(2), T(2) * r), (T(1) + T(2) * r) / T(3)));
    default:  // PSA_LIMITER_NONE
      return T(1);
  }
}

// MUSCL's limiter and its parameters (standalone kernels); dx is the cell
// size of the slope-limited form.
template <typename T>
struct LimParams {
  int lm;
  T theta;
  T atol;
  T dx;
};

// limiters.py:158-163: 0 where a b < 0, else the one of smaller modulus.
template <typename T>
__device__ __forceinline__ T minmod2(T a, T b) {
  return a * b < T(0) ? T(0) : (fabs(a) < fabs(b) ? a : b);
}

// slope_limit of limiters.py:148-371 from the one-sided slopes sl, sr and
// the central slope sc of cell i (the unlimited slope is the reference's
// (u[i+1] + u[i-1]) / (x[i+1] - x[i-1]), given as sc_sum).
template <typename T>
__device__ __forceinline__ T slope_limit(int lm, T sl, T sr, T sc, T sc_sum) {
  switch (lm) {
    case PSA_LIMITER_MINMOD:
      return minmod2(sl, sr);
    case PSA_LIMITER_MC:
      return minmod2(minmod2(sl, sr), sc);
    case PSA_LIMITER_SUPERBEE:
      return fmax(minmod2(sl, T(2) * sr), minmod2(T(2) * sl, sr));
    case PSA_LIMITER_VANALBADA1:
      return sl * sr <= T(0) ? T(0) : sl * sr * (sl + sr) / (sl * sl + sr * sr);
    case PSA_LIMITER_VANALBADA2:
      return sl * sr <= T(0) ? T(0) : T(2) * sl * sr / (sl * sl + sr * sr);
    case PSA_LIMITER_VANLEER:
      return sl * sr <= T(0) ? T(0) : T(2) * sl * sr / (sl + sr);
    default:  // PSA_LIMITER_NONE
      return sc_sum;
  }
}

// Both states of cell i of the slope-limited MUSCL (reconstruction.py:
// 279-296) from Dm = u_i - u_{i-1}, Dp = u_{i+1} - u_i, Ds = u_{i+1} + u_{i-1}
// on a uniform grid: ur, ul = u +- dx/2 du, du = 0 at the ends of the array.
template <typename T>
__device__ __forceinline__ void muscls_pair(T u0, T Dm, T Dp, T Ds, int lm, T dx,
                                            bool edge, T &ul, T &ur) {
  const T inv_dx = T(1) / dx;
  const T sl = Dm * inv_dx, sr = Dp * inv_dx;
  const T sc = (Dm + Dp) * (T(0.5) * inv_dx);
  const T du = edge ? T(0) : slope_limit(lm, sl, sr, sc, Ds * (T(0.5) * inv_dx));
  const T h = T(0.5) * dx * du;
  ur = u0 + h;
  ul = u0 - h;
}

// Limiters with phi(1/r) = phi(r) / r (their is_symmetric property).
__device__ __forceinline__ bool limiter_symmetric(int lm) {
  return lm == PSA_LIMITER_MINMOD || lm == PSA_LIMITER_MC ||
         lm == PSA_LIMITER_SUPERBEE || lm == PSA_LIMITER_VANALBADA1 ||
         lm == PSA_LIMITER_VANLEER;
}

// r = Dm / Dp and its inverse ir = Dp / Dm. FP64: one reciprocal of Dm Dp
// refined by two Newton steps, r = Dm^2 / (Dm Dp), ir = Dp^2 / (Dm Dp),
// where Dm Dp is a normal number (either sign), else two divisions; FP32:
// the hardware reciprocal of each difference. ir is used only where r is
// not below the cut-off (it is not finite where Dm = 0).
template <typename T>
__device__ __forceinline__ void muscl_ratios(T Dm, T Dp, T &r, T &ir);

template <>
__device__ __forceinline__ void muscl_ratios<double>(double Dm, double Dp,
                                                     double &r, double &ir) {
  const double den = Dm * Dp;
  if (__builtin_amdgcn_class(den, 0x108)) {  // -normal or +normal
    double inv = __builtin_amdgcn_rcp(den);
    double e = fma(-den, inv, 1.0);
    inv = fma(inv, e, inv);
    e = fma(-den, inv, 1.0);
    inv = fma(inv, e, inv);
    r = (Dm * Dm) * inv;
    ir = (Dp * Dp) * inv;
  } else {
    r = Dm / Dp;
    ir = Dp / Dm;
  }
}

template <>
__device__ __forceinline__ void muscl_ratios<float>(float Dm, float Dp, float &r,
                                                    float &ir) {
  r = Dm * __builtin_amdgcn_rcpf(Dp);
  ir = Dp * __builtin_amdgcn_rcpf(Dm);
}

// Both states of cell i from Dm = u_i - u_{i-1} and Dp = u_{i+1} - u_i:
// r = Dm / Dp (0 where |Dp| < atol, limiters.py:119-129, and 0 at the ends
// of the array, jnp.pad), ur = u + phi(r) / 2 Dp, ul = u - phi(1/r) / 2 Dm
// (phi(r) / r for a symmetric limiter, 0 where |r| < atol; else phi of 1/r,
// atol where |r| < atol); ur of the last cell and ul of the first are 0.
template <typename T>
__device__ __forceinline__ void muscl_pair(T u0, T Dm, T Dp, int lm, T theta,
                                           T atol, bool lo, bool hi, T &ul,
                                           T &ur) {
  T q, iq;
  muscl_ratios(Dm, Dp, q, iq);
  T r = fabs(Dp) < atol ? T(0) : q;
  if (lo || hi) r = T(0);
  const T phi = limiter_phi(lm, theta, r);
  const bool small = fabs(r) < atol;
  T phi_inv;
  if (limiter_symmetric(lm)) {
    // phi(r) / r = phi(r) (Dp / Dm)
    phi_inv = small ? T(0) : phi * iq;
  } else {
    phi_inv = limiter_phi(lm, theta, small ? atol : iq);
  }
  ur = hi ? T(0) : u0 + (T(0.5) * phi) * Dp;
  ul = lo ? T(0) : u0 - (T(0.5) * phi_inv) * Dm;
}

}  // namespace psa
