// Device building blocks of the fused stage computation, shared by the
// single-instance stage kernels (forward.hip) and the batched solver
// (batched.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "../../include/pyshocks_amd.h"
#include "mlp_device.hpp"
#include "weno5_device.hpp"

namespace psa {

// ---------------------------------------------------------------------------
// loads

// Value of cell gm of the periodic ghost-filled array (scalar.py:529-540):
// ghost cells read their interior image; outside [0, m) the 'same'
// convolution pads with zeros (convolve.py:113-114).
template <typename T>
__device__ __forceinline__ T load_filled(const T *__restrict__ u, int64_t gm,
                                         int64_t m, int64_t n, int g) {
  if (gm < 0 || gm >= m) return T(0);
  int64_t idx = gm;
  if (idx < g) {
    idx += n;
  } else if (idx >= m - g) {
    idx -= n;
  }
  return u[idx];
}

// Value of cell gm of an array used as given, zero outside [0, m).
template <typename T>
__device__ __forceinline__ T load_raw(const T *__restrict__ u, int64_t gm,
                                      int64_t m) {
  return (gm >= 0 && gm < m) ? u[gm] : T(0);
}

// Non-negative value -> order-preserving bit pattern (NaN sorts above inf).
__device__ __forceinline__ unsigned long long abs_bits(double v) {
  return static_cast<unsigned long long>(__double_as_longlong(fabs(v)));
}
__device__ __forceinline__ unsigned long long abs_bits(float v) {
  return static_cast<unsigned long long>(
      static_cast<unsigned int>(__float_as_uint(fabsf(v))));
}
// A NaN's pattern (above every other abs_bits value).
template <typename T>
__device__ __forceinline__ unsigned long long nan_bits() {
  return sizeof(T) == 8 ? 0x7FF8000000000000ull : 0x7FC00000ull;
}
template <typename T>
__device__ __forceinline__ T from_bits(unsigned long long b);
template <>
__device__ __forceinline__ double from_bits<double>(unsigned long long b) {
  return __longlong_as_double(static_cast<long long>(b));
}
template <>
__device__ __forceinline__ float from_bits<float>(unsigned long long b) {
  return __uint_as_float(static_cast<unsigned int>(b));
}

// Maximum of a 32-bit unsigned value over the wave (returned to every lane):
// DPP row reductions (quad permutations, row rotations by 4 and 8) give each
// row of 16 lanes its maximum, the row broadcasts of lanes 15 and 31 carry
// them into the last row, and lane 63 holds the result. Lanes a DPP move
// does not write read 0, the identity of the maximum.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ unsigned int dpp_max_u32(unsigned int v) {
  const unsigned int o = static_cast<unsigned int>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, ROWS, 0xF, true));
  return o > v ? o : v;
}

__device__ __forceinline__ unsigned int wave_max_u32(unsigned int v) {
  v = dpp_max_u32<0xB1>(v);        // quad_perm [1, 0, 3, 2]
  v = dpp_max_u32<0x4E>(v);        // quad_perm [2, 3, 0, 1]
  v = dpp_max_u32<0x124>(v);       // row_ror 4
  v = dpp_max_u32<0x128>(v);       // row_ror 8
  v = dpp_max_u32<0x142, 0xA>(v);  // row_bcast 15 into rows 1 and 3
  v = dpp_max_u32<0x143, 0xC>(v);  // row_bcast 31 into rows 2 and 3
  return static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// Maximum of a 64-bit unsigned value over the wave (to every lane): the high
// words first, then the low words of the lanes that hold the high maximum.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  const unsigned int hi = wave_max_u32(static_cast<unsigned int>(v >> 32));
  const unsigned int lo = wave_max_u32(
      static_cast<unsigned int>(v >> 32) == hi ? static_cast<unsigned int>(v) : 0u);
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

template <int BLOCK>
__device__ __forceinline__ unsigned long long block_max_bits(
    unsigned long long v, unsigned long long *s_red) {
  v = wave_max_u64(v);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) s_red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLOCK / 64; ++w) v = s_red[w] > v ? s_red[w] : v;
  }
  return v;
}

// max of PSA_UMAX_SLOTS partials (lane k of every wave reads the k-th and
// the wave reduces them, so each wave holds the result without a barrier)
__device__ __forceinline__ unsigned long long wave_max_bits(unsigned long long v) {
  return wave_max_u64(v);
}

// max of the PSA_UMAX_SLOTS partials in word 0 of `slots` (wave_max_bits).
__device__ __forceinline__ unsigned long long gather_slots(
    const unsigned long long (*slots)[PSA_UMAX_SLOT_STRIDE]) {
  static_assert(PSA_UMAX_SLOTS == 64, "one slot per lane");
  return wave_max_bits(slots[threadIdx.x & 63][0]);
}

// ---------------------------------------------------------------------------
// fused tiled stage kernel

template <typename T>
struct StageArgs {
  const T *un;      // u^n (stages 2 and 3)
  const T *uk;      // stage input: the stencil is evaluated on it
  T *out;           // stage output
  const T *vel;     // advection velocity (VEL == 2)
  const T *nu;      // Rusanov df^(alpha-1) per face j - 1, or null
  psa_control *ctrl;
  const T *dt_ptr;  // used when ctrl is null
  int64_t m;
  int64_t n;
  int g;
  int aligned;      // uk, un, out, vel are 16-byte aligned
  T dx;
  T inv_dx;
  T eps;
  T vconst;         // advection: the constant velocity; Lax-Friedrichs: the
                    //   global flux speed of the stage (set by the kernel)
  const T *speed;   // Lax-Friedrichs without a control block: the device
                    //   scalar max|a| of the stage input
  // Lax-Friedrichs stage kernels: the slots whose maximum is the flux speed,
  // gathered once the tile's loads are in flight (null: vconst holds it)
  const unsigned long long (*speed_slots)[PSA_UMAX_SLOT_STRIDE];
  int es;            // third-order reconstruction: ES-WENO weights (1) or
                     //   WENO-JS (0)
  int rec3;          // host side: the kernels take the third-order
                     //   reconstruction (REC = 1)
  int muscl;         // host side: the kernels take MUSCL (REC = 2)
  int muscls;        // host side: the slope-limited MUSCL (REC = 3)
  T es_delta;        // PSA_FLUX_ESWENO: delta of the dissipative flux
  int nowrap;        // operator (STAGE 0): the ghost cells of uk as given
  int limiter;       // MUSCL: PSA_LIMITER_*
  T lim_theta;       // MUSCL: the minmod parameter
  T lim_atol;        // MUSCL: the slope cut-off
  const float *mlp;  // learned weights (LEARNED kernels)
  T tau;             // their shock threshold
  T sigma;           // their smooth threshold
  T *k1out;          // whole-step kernel: the stage states k1, k2 as well,
  T *k2out;          //   or null
  int64_t bstride;   // fixed-dt whole steps of a batch: instance blockIdx.y
                     //   at offset blockIdx.y * bstride of uk, out, k1out,
                     //   k2out, its dt at dt_ptr[blockIdx.y]
  int batch;         // host side: instances of the launch (grid.y)
  int step_seq;      // whole-step launch s of a controlled sequence: s mod 6
  int bcm;           // whole-step kernel: PSA_BC_PERIODIC, PSA_BC_NEUMANN or
                     //   PSA_BC_DIRICHLET (ghost cells filled by the kernel)
  T bc_l, bc_r;      // Neumann: the gradients of the left and right sides
  const T *bc_vals;  // Dirichlet: the 2 g ghost values (left, then right)
};

// Value of cell gi of a stage input with the non-periodic ghost fill of the
// whole-step kernel, from the interior cells of u (the state indexed by
// global cell): zero outside [0, m) (the 'same' convolutions' padding),
// the fill of scalar.py:441-511 in the ghost cells (the arithmetic of
// k_fill_neumann / k_fill_dirichlet), u itself inside.
template <typename T>
__device__ __forceinline__ T bc_cell(const StageArgs<T> &p, const T *u, int64_t gi) {
  const int64_t m = p.m;
  const int g = p.g;
  if (gi < 0 || gi >= m) return T(0);
  if (gi < g) {
    if (p.bcm == PSA_BC_DIRICHLET) return p.bc_vals[gi];
    const int k = g - 1 - static_cast<int>(gi);
    return u[g + k] - (T(2 * k + 1) * p.dx) * p.bc_l;
  }
  if (gi >= m - g) {
    if (p.bcm == PSA_BC_DIRICHLET) return p.bc_vals[g + (gi - (m - g))];
    const int k = static_cast<int>(m - 1 - gi);
    return u[m - 2 * g + k] + (T(2 * k + 1 - 2 * g) * p.dx) * p.bc_r;
  }
  return u[gi];
}

template <typename T, int VEL>
__device__ __forceinline__ T velocity_at(const T *s_v, int j, T vconst) {
  if constexpr (VEL == 2) {
    return s_v[j];
  } else {
    return vconst;
  }
}

// 16-byte vectors of the dtype
template <typename T>
struct Vec16;
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2;
};
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4;
};

template <typename T>
__device__ __forceinline__ typename Vec16<T>::type load16(const T *p) {
  return *reinterpret_cast<const typename Vec16<T>::type *>(p);
}
template <typename T>
__device__ __forceinline__ void store16(T *p, typename Vec16<T>::type v) {
  *reinterpret_cast<typename Vec16<T>::type *>(p) = v;
}
__device__ __forceinline__ unsigned long long vmax_bits(double2 v,
                                                        unsigned long long m) {
  const unsigned long long a = abs_bits(v.x), b = abs_bits(v.y);
  m = a > m ? a : m;
  return b > m ? b : m;
}
__device__ __forceinline__ unsigned long long vmax_bits(float4 v,
                                                        unsigned long long m) {
  const unsigned long long a = abs_bits(v.x), b = abs_bits(v.y);
  const unsigned long long c = abs_bits(v.z), d = abs_bits(v.w);
  m = a > m ? a : m;
  m = b > m ? b : m;
  m = c > m ? c : m;
  return d > m ? d : m;
}

// COUNT 16-byte chunks moved between global memory and LDS by BLOCK
// threads: every load of a thread is issued before its first LDS write, so a
// wave keeps all of its loads in flight at once.
template <typename T, int BLOCK, int COUNT>
struct Chunks {
  using V = typename Vec16<T>::type;
  static constexpr int VN = Vec16<T>::n;
  static constexpr int CH = (COUNT + BLOCK - 1) / BLOCK;
  V r[CH];

  __device__ __forceinline__ void load(const T *src, int tid) {
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int c = tid + k * BLOCK;
      if (k < CH - 1 || COUNT % BLOCK == 0 || c < COUNT) r[k] = load16(src + c * VN);
    }
  }
  __device__ __forceinline__ void store(T *dst, int tid) const {
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int c = tid + k * BLOCK;
      if (k < CH - 1 || COUNT % BLOCK == 0 || c < COUNT) store16(dst + c * VN, r[k]);
    }
  }
};


// Workgroup barrier for LDS hand-offs. With ASYNC, only LDS operations are
// waited for (the workgroup-scope fence of __syncthreads would also wait for
// every outstanding global load, including direct-to-LDS copies of the next
// tile that are meant to stay in flight).
template <bool ASYNC>
__device__ __forceinline__ void tile_sync() {
  if constexpr (ASYNC) {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  } else {
    __syncthreads();
  }
}

// Waits until every global memory operation of this wave has completed.
__device__ __forceinline__ void wait_vm_all() {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), other counters untouched
}

// Waits until at most N (< 16) global memory operations of this wave are
// outstanding: with the N most recent ones being stores, every older load
// (or direct-to-LDS copy) has completed.
template <int N>
__device__ __forceinline__ void wait_vm_but() {
  static_assert(N >= 0 && N < 16, "vmcnt range");
  __builtin_amdgcn_s_waitcnt(0x0F70 | N);
}

// Direct-to-LDS copy of COUNT 16-byte chunks from src to dst (both 16-byte
// aligned): a wave copies 1 KB per instruction, lane l the l-th chunk of its
// round, rounds dealt to the waves in turn; the copies complete
// asynchronously (wait_vm_all, then a barrier, before reading dst).
template <int BLOCK, int COUNT>
__device__ __forceinline__ void dma_chunks(const void *src, void *dst) {
  constexpr int WAVES = BLOCK / 64;
  constexpr int ROUNDS = (COUNT + 63) / 64;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const char *s = static_cast<const char *>(src);
  char *d = static_cast<char *>(dst);
#pragma unroll
  for (int r = wave; r < ROUNDS; r += WAVES) {
    const int c = r * 64 + lane;
    if (COUNT % 64 == 0 || c < COUNT)
      __builtin_amdgcn_global_load_lds(s + 16 * c, d + 1024 * r, 16, 0, 0);
  }
}

// Entries of the per-wave queue of learned samples (3 words each).
template <int CPT>
constexpr int learned_queue_words() {
  return 64 * 2 * CPT * 3;
}

// Learned reconstructions of a thread's CPT cells (wave-collective). Cells in
// the learned regime put the inputs of their two samples (right stencil,
// mirrored stencil) into the wave's LDS queue q; the wave evaluates the
// queue 64 samples at a time (and only the chains that hold samples; the
// network is fetched from src only if the queue is not empty), the weights
// come back through the same slots. Shock and smooth cells keep the values
// already in ul, ur (WENO-JS and ideal weights).
template <typename T, int CPT>
__device__ __forceinline__ void learned_cells(const MlpSource &src, const T *w,
                                              const int *regime, T *ul, T *ur,
                                              unsigned int *q) {
  const int lane = threadIdx.x & 63;
  unsigned int *qw = q + (threadIdx.x >> 6) * learned_queue_words<CPT>();
  int nmine = 0;
#pragma unroll
  for (int r = 0; r < CPT; ++r) nmine += regime[r] == 2 ? 2 : 0;
  // most waves of a smooth solution hold no learned cell (wave-uniform exit)
  if (__ballot(nmine != 0) == 0) return;
  int incl = nmine;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  const int ntot = __shfl(incl, 63, 64);
  const int first = incl - nmine;
  if (ntot > 0) {
    int slot = first;
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
      if (regime[r] == 2) {
        float x[MLP_IN], xm[MLP_IN];
        mlp_features(w[r], w[r + 1], w[r + 2], w[r + 3], w[r + 4], x);
        xm[0] = x[4];
        xm[1] = x[3];
        xm[2] = x[2];
        xm[3] = x[1];
        xm[4] = x[0];
        const mlp_u32x4_t a = mlp_pack_inputs(x), b = mlp_pack_inputs(xm);
        qw[3 * slot + 0] = a[0];
        qw[3 * slot + 1] = a[1];
        qw[3 * slot + 2] = a[2];
        qw[3 * slot + 3] = b[0];
        qw[3 * slot + 4] = b[1];
        qw[3 * slot + 5] = b[2];
        slot += 2;
      }
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    // the network's operand registers live only while the queue is evaluated
    MlpWave W;
    mlp_fetch(W, src);
    for (int r0 = 0; r0 < ntot; r0 += 64) {  // wave-uniform
      const int e = r0 + lane;
      mlp_u32x4_t own = {0u, 0u, 0u, 0u};
      if (e < ntot) {
        own[0] = qw[3 * e + 0];
        own[1] = qw[3 * e + 1];
        own[2] = qw[3 * e + 2];
      }
      const int left = ntot - r0;
      const int nch = left >= 64 ? 4 : (left + 15) >> 4;
      float z[MLP_OUT], wt[MLP_OUT];
      mlp_logits_packed(W, own, nch, z);
      mlp_softmax(z, wt);
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      if (e < ntot) {
        qw[3 * e + 0] = __float_as_uint(wt[0]);
        qw[3 * e + 1] = __float_as_uint(wt[1]);
        qw[3 * e + 2] = __float_as_uint(wt[2]);
      }
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  }
  if (ntot == 0) return;
  int slot = first;
#pragma unroll
  for (int r = 0; r < CPT; ++r) {
    if (regime[r] != 2) continue;
    T hr0, hr1, hr2, hl0, hl1, hl2;
    weno_uhat_diff(w[r + 2], w[r + 1] - w[r], w[r + 2] - w[r + 1], w[r + 3] - w[r + 2],
                   w[r + 4] - w[r + 3], hr0, hr1, hr2, hl0, hl1, hl2);
    const float wr[MLP_OUT] = {__uint_as_float(qw[3 * slot + 0]),
                               __uint_as_float(qw[3 * slot + 1]),
                               __uint_as_float(qw[3 * slot + 2])};
    const float wl[MLP_OUT] = {__uint_as_float(qw[3 * slot + 3]),
                               __uint_as_float(qw[3 * slot + 4]),
                               __uint_as_float(qw[3 * slot + 5])};
    ur[r] = mlp_combine(wr, hr0, hr1, hr2);
    ul[r] = mlp_combine(wl, hl0, hl1, hl2);
    slot += 2;
  }
}

// Left and right states of CPT consecutive cells from w = the cells
// -2 .. CPT + 1 around them (reconstruction.py:358-377 with WENOJS53). The
// mirrored left reconstruction uses the same three indicators with beta0 and
// beta2 swapped. Everything is built from the first differences of w, shared
// between neighbouring cells: the second differences of the indicators
// (t00 = D(i) - D(i-1)), t01 = t00 + 2 D(i), t11 = D(i+1) + D(i),
// t21 = t20 - 2 D(i+1) (weno.py:219-226 rearranged), and the candidates
// (weno_uhat_diff; in FP64 the middle candidates and the third differences
// of weno_diff_sums). LEARNED: the weights of cells outside the shock regime
// come from the network or the ideal weights (wave-collective, learned_cells).
template <typename T, int CPT, bool LEARNED>
__device__ __forceinline__ void reconstruct_cells(const StageArgs<T> &p,
                                                  const T *w, T *ul, T *ur,
                                                  const MlpSource &src,
                                                  unsigned int *q) {
  using C = Weno5Coeffs<T>;
  // D[k] = w[k + 1] - w[k]: D[k] is u(c) - u(c - 1) for cell c = k - 1
  T D[CPT + 3];
#pragma unroll
  for (int k = 0; k < CPT + 3; ++k) D[k] = w[k + 1] - w[k];
  // second differences at cells -1 .. CPT and 13/3 times their squares (the
  // indicators are carried as 4 beta_k with 4 eps, Weno5Coeffs::a0x4)
  T d2[CPT + 2], E[CPT + 2];
#pragma unroll
  for (int k = 0; k < CPT + 2; ++k) {
    d2[k] = D[k + 1] - D[k];
    E[k] = C::a0x4 * (d2[k] * d2[k]);
  }
  const T eps4 = T(4) * p.eps;

  // the third-difference form (weno_diff_sums, weno_diff_pair_f32) with
  // 13/3 d2^2 + 4 eps formed once per cell (each enters one indicator of
  // three cells). FP64: the shared reciprocal for every cell without a
  // branch; a cell whose denominator is not a positive normal number is
  // redone in the reference form from its stencil after one wave-uniform
  // test (its result depends on its own values only, not on the wave it
  // falls in). LEARNED: the same for the cells in the shock regime, whose
  // regime comes from the indicators 4 beta_k = t^2 + 13/3 d2^2 of
  // learned_right16.
  T Ee[CPT + 2], t3[CPT + 1], t2[CPT + 1];
#pragma unroll
  for (int k = 0; k < CPT + 2; ++k) Ee[k] = fma(C::a0x4, d2[k] * d2[k], eps4);
#pragma unroll
  for (int k = 0; k < CPT + 1; ++k) {
    // third difference of cells k - 1, k (d2 of cell k minus d2 of cell k - 1)
    const T t = d2[k + 1] - d2[k];
    t3[k] = t * T(1.0 / 3.0);
    t2[k] = T(0.5) * t;
  }
  int regime[CPT];
  unsigned int redo = 0;
#pragma unroll
  for (int r = 0; r < CPT; ++r) {
    const T t01 = fma(T(2), D[r + 1], d2[r]);
    const T t11 = D[r + 2] + D[r + 1];
    const T t21 = fma(T(-2), D[r + 2], d2[r + 2]);
    regime[r] = 0;
    if constexpr (LEARNED)
      regime[r] = mlp_regime(fma(t01, t01, E[r]), fma(t11, t11, E[r + 1]),
                             fma(t21, t21, E[r + 2]), eps4, p.tau, p.sigma);
    if (regime[r] == 0) {
      const T e0 = fma(t01, t01, Ee[r]), e1 = fma(t11, t11, Ee[r + 1]),
              e2 = fma(t21, t21, Ee[r + 2]);
      const T hr1 = fma(t11 + D[r + 2], T(1.0 / 6.0), w[r + 2]);
      const T hl1 = fma(t11 + D[r + 1], T(-1.0 / 6.0), w[r + 2]);
      if constexpr (std::is_same<T, double>::value) {
        const WenoDiffSums s =
            weno_diff_sums(e0, e1, e2, t3[r], t2[r], t3[r + 1], t2[r + 1]);
        redo |= weno_diff_fast(s) ? 0u : (1u << r);
        weno_diff_divide(s, hr1, hl1, ul[r], ur[r]);
      } else {
        weno_diff_pair_f32(e0, e1, e2, t3[r], t2[r], t3[r + 1], t2[r + 1], hr1, hl1,
                           ul[r], ur[r]);
      }
    } else {
      T hr0, hr1, hr2, hl0, hl1, hl2;
      weno_uhat_diff(w[r + 2], D[r], D[r + 1], D[r + 2], D[r + 3], hr0, hr1, hr2, hl0,
                     hl1, hl2);
      ur[r] = mlp_combine_ideal(hr0, hr1, hr2);
      ul[r] = mlp_combine_ideal(hl0, hl1, hl2);
    }
  }
  if constexpr (std::is_same<T, double>::value) {
    // (marked unlikely: the compiler then lays the redo code out after the
    // kernel's hot path, which keeps that path compact in the instruction
    // cache; 2.5% on the FP64 whole-step kernel)
    if (__builtin_expect(__any(redo != 0), 0)) {
#pragma unroll
      for (int r = 0; r < CPT; ++r) {
        if (redo & (1u << r)) {
          T b0, b1, b2, hr0, hr1, hr2, hl0, hl1, hl2;
          weno_beta(w[r], w[r + 1], w[r + 2], w[r + 3], w[r + 4], b0, b1, b2);
          weno_uhat(w[r], w[r + 1], w[r + 2], w[r + 3], w[r + 4], hr0, hr1, hr2);
          weno_uhat(w[r + 4], w[r + 3], w[r + 2], w[r + 1], w[r], hl0, hl1, hl2);
          const T e0 = p.eps + b0, e1 = p.eps + b1, e2 = p.eps + b2;
          weno_pair_ref_q(hr0, hr1, hr2, hl0, hl1, hl2, e0 * e0, e1 * e1, e2 * e2,
                          ul[r], ur[r]);
        }
      }
    }
  }
  if constexpr (LEARNED) learned_cells<T, CPT>(src, w, regime, ul, ur, q);
}

// Third-order reconstructions (WENO-JS3 or ES-WENO3 weights, weno3_pair) of
// CPT cells from the layout of reconstruct_cells (w holds cells -2 ..
// CPT + 1; cells -1 .. CPT are used). EDGE (the operator over the whole
// array): ES-WENO's tau is zero for the first and last cell of the array,
// gfirst being the global index of cell 0.
template <typename T, int CPT, bool EDGE>
__device__ __forceinline__ void reconstruct_cells3(const StageArgs<T> &p,
                                                   const T *w, T *ul, T *ur,
                                                   int64_t gfirst, T *om = nullptr) {
  // D[k] = u(k) - u(k - 1)
  T D[CPT + 1];
#pragma unroll
  for (int k = 0; k <= CPT; ++k) D[k] = w[k + 2] - w[k + 1];
  const bool es = p.es != 0;
#pragma unroll
  for (int r = 0; r < CPT; ++r) {
    bool tz = false;
    if constexpr (EDGE) {
      const int64_t gm = gfirst + r;
      tz = gm == 0 || gm == p.m - 1;
    }
    weno3_pair(w[r + 2], D[r], D[r + 1], p.eps, es, tz, ul[r], ur[r],
               om != nullptr ? &om[r] : nullptr);
  }
}

// MUSCL states (muscl_pair) of CPT cells from the layout of
// reconstruct_cells; EDGE: the array's first and last cell are padded as the
// reference pads them.
template <typename T, int CPT, bool EDGE>
__device__ __forceinline__ void reconstruct_cells_muscl(const StageArgs<T> &p,
                                                        const T *w, T *ul,
                                                        T *ur, int64_t gfirst) {
  T D[CPT + 1];
#pragma unroll
  for (int k = 0; k <= CPT; ++k) D[k] = w[k + 2] - w[k + 1];
#pragma unroll
  for (int r = 0; r < CPT; ++r) {
    bool lo = false, hi = false;
    if constexpr (EDGE) {
      const int64_t gm = gfirst + r;
      lo = gm == 0;
      hi = gm == p.m - 1;
    }
    muscl_pair(w[r + 2], D[r], D[r + 1], p.limiter, p.lim_theta, p.lim_atol, lo,
               hi, ul[r], ur[r]);
  }
}

// Slope-limited MUSCL states (muscls_pair) of CPT cells from the layout of
// reconstruct_cells; EDGE: du = 0 for the array's first and last cell.
template <typename T, int CPT, bool EDGE>
__device__ __forceinline__ void reconstruct_cells_muscls(const StageArgs<T> &p,
                                                         const T *w, T *ul,
                                                         T *ur, int64_t gfirst) {
#pragma unroll
  for (int r = 0; r < CPT; ++r) {
    bool edge = false;
    if constexpr (EDGE) {
      const int64_t gm = gfirst + r;
      edge = gm == 0 || gm == p.m - 1;
    }
    muscls_pair(w[r + 2], w[r + 2] - w[r + 1], w[r + 3] - w[r + 2],
                w[r + 3] + w[r + 1], p.limiter, p.dx, edge, ul[r], ur[r]);
  }
}

// The reconstruction of a fused kernel: REC 0 WENO-JS5 (or learned weights),
// REC 1 third order (reconstruct_cells3), REC 2 MUSCL, REC 3 slope-limited
// MUSCL.
template <typename T, int CPT, bool LEARNED, int REC, bool EDGE = false>
__device__ __forceinline__ void reconstruct_any(const StageArgs<T> &p,
                                                const T *w, T *ul, T *ur,
                                                const MlpSource &src,
                                                unsigned int *q,
                                                int64_t gfirst, T *om = nullptr) {
  if constexpr (REC == 1) {
    reconstruct_cells3<T, CPT, EDGE>(p, w, ul, ur, gfirst, om);
  } else if constexpr (REC == 2) {
    reconstruct_cells_muscl<T, CPT, EDGE>(p, w, ul, ur, gfirst);
  } else if constexpr (REC == 3) {
    reconstruct_cells_muscls<T, CPT, EDGE>(p, w, ul, ur, gfirst);
  } else {
    reconstruct_cells<T, CPT, LEARNED>(p, w, ul, ur, src, q);
  }
}

// Face flux of the stage kernels for Rusanov, Lax-Friedrichs (global speed
// vconst) with nu = 1 and Engquist-Osher (all three doubled, see
// stage_lambda), Godunov, or advection with a constant velocity.
template <typename T, int FLUX>
__device__ __forceinline__ T stage_flux(T urL, T ulR, T uL, T uR, T vconst) {
  if constexpr (FLUX == PSA_FLUX_RUSANOV) {
    return flux_rusanov_nu1_x2(urL, ulR, uL, uR);
  } else if constexpr (FLUX == PSA_FLUX_LAX_FRIEDRICHS) {
    return flux_lax_friedrichs_nu1_x2(urL, ulR, vconst);
  } else if constexpr (FLUX == PSA_FLUX_ENGQUIST_OSHER) {
    return flux_engquist_osher_x2(urL, ulR);
  } else if constexpr (FLUX == PSA_FLUX_BURGERS_UPWIND) {
    return flux_burgers_upwind(urL, ulR);
  } else {
    // a constant velocity reconstructs to itself up to rounding, so the
    // upwind switch only sees its sign
    return flux_advection_upwind(urL, ulR, vconst, vconst);
  }
}

// u + dt L(u) = u - lam (F(i+1/2) - F(i-1/2)) with lam = dt / dx, times the
// constant velocity for advection (L = -a dF / dx), times 1/2 for the
// doubled fluxes of stage_flux.
template <typename T, int FLUX>
__device__ __forceinline__ T stage_lambda(const StageArgs<T> &p, T dt) {
  const T lam = dt * p.inv_dx;
  if constexpr (FLUX == PSA_FLUX_RUSANOV || FLUX == PSA_FLUX_LAX_FRIEDRICHS ||
                FLUX == PSA_FLUX_ENGQUIST_OSHER) {
    return T(0.5) * lam;
  } else if constexpr (FLUX == PSA_FLUX_ADVECTION_UPWIND) {
    return lam * p.vconst;
  } else {
    return lam;
  }
}

// The SSP-RK3 stage combinations of timestepping.py:312-320 from
// y = uk + dt L(uk) = uk - lam dF and un = u^n.
template <typename T, int STAGE>
__device__ __forceinline__ T stage_value(T uk, T un, T lam, T dF) {
  const T y = fma(-lam, dF, uk);
  if constexpr (STAGE == 1) {
    return y;
  } else if constexpr (STAGE == 2) {
    return T(3.0 / 4.0) * un + T(1.0 / 4.0) * y;
  } else {
    return T(1.0 / 3.0) * un + T(2.0 / 3.0) * y;
  }
}

// LDS of one stage tile.
template <typename T, int VEL, int BLOCK, int CPT>
struct StageSmem {
  static constexpr int TILE = BLOCK * CPT;
  static constexpr int HALO = 4;
  static constexpr int NV = VEL == 2 ? TILE + 2 * HALO : 1;
  static constexpr int NVX = VEL == 2 ? BLOCK + 1 : 1;
  __align__(16) T u[TILE + 2 * HALO];
  __align__(16) T v[NV];
  __align__(16) T io[TILE];
  T urL[BLOCK + 1];
  T ulF[BLOCK + 1];
  T edge_val[2];   // learned right state of cell -1, left state of cell TILE
  int edge_js[2];  // 1 where those cells keep WENO-JS
  T arL[NVX];
  T alF[NVX];
  unsigned long long red[BLOCK / 64];
};

// One tile of one stage: cells [S, S + BLOCK * CPT) of the arrays of p.
//
// VEL: 0 Burgers (no velocity), 1 constant velocity, 2 velocity array.
// LEARNED: the reconstruction takes the learned weights of W (Burgers only,
//        every wave of the workgroup must run the tile).
// PRELOADED: a tile on the fast path is already in sm.u (and u^n in sm.io)
//        and visible to the workgroup; ASYNC: barriers wait for LDS only.
// wl: the network's operand registers in LDS (persistent kernels), or null
//        (then they are loaded from p.mlp where the network runs).
// STAGE: 0 operator (all m cells, faces 0 and m padded with zero flux),
//        1..3 SSP-RK3 stages (interior cells, ghost images rewritten).
//
// The workgroup stages the tile plus a 4-cell halo in LDS, every thread
// reconstructs CPT consecutive cells (both sides from one set of smoothness
// indicators), neighbouring threads exchange one right and one left state
// through LDS, each thread evaluates its CPT + 1 faces, applies the
// conservative difference and the stage combination, and the results leave
// through LDS as coalesced stores. For STAGE 3 (and stages 1 and 2 of
// Lax-Friedrichs), vmax accumulates the bit pattern of max |u| of the cells
// written. Ends with the LDS free for reuse only after the caller's next
// barrier.
template <typename T, int FLUX, int VEL, int STAGE, int BLOCK, int CPT,
          bool LEARNED = false, bool PRELOADED = false, bool ASYNC = false,
          int REC = 0>
__device__ __forceinline__ void stage_tile(const StageArgs<T> &p, int64_t S,
                                           T dt,
                                           StageSmem<T, VEL, BLOCK, CPT> &sm,
                                           unsigned long long &vmax,
                                           unsigned int *q = nullptr,
                                           const MlpWaveLds *wl = nullptr) {
  static_assert(!LEARNED || VEL == 0, "learned weights are for Burgers");
  static_assert(REC == 0 || !LEARNED,
                "learned weights replace the WENO-JS5 weights only");
  constexpr bool EDGE = REC != 0 && STAGE == 0;
  constexpr int TILE = BLOCK * CPT;
  constexpr int HALO = 4;
  constexpr int VN = Vec16<T>::n;
  // max|u| of the output: the next step's dt (stage 3) and the next stage's
  // Lax-Friedrichs speed (stages 1 and 2)
  constexpr bool kReducesOutput =
      STAGE == 3 || (STAGE > 0 && FLUX == PSA_FLUX_LAX_FRIEDRICHS);
  const int tid = threadIdx.x;
  const int64_t m = p.m;
  const int64_t n = p.n;
  const int g = p.g;

  // A tile whose cells and halo are interior needs neither the ghost
  // mapping nor bounds checks, and (STAGE > 0) writes no ghost images.
  const bool fast = p.aligned != 0 && STAGE > 0 && S - HALO >= g &&
                    S + TILE + HALO <= m - g;

  // the constant advection velocity, or the Lax-Friedrichs flux speed
  T fspeed = p.vconst;
  auto gather_speed = [&]() {
    if constexpr (FLUX == PSA_FLUX_LAX_FRIEDRICHS) {
      if (p.speed_slots != nullptr) fspeed = from_bits<T>(gather_slots(p.speed_slots));
    }
  };

  // ---- stage the tile and its halo
  if (fast && PRELOADED) {
    // staged by the caller
    gather_speed();
  } else if (fast) {
    // all loads of the tile, the velocity and u^n first, then the LDS writes
    Chunks<T, BLOCK, (TILE + 2 * HALO) / VN> cu;
    Chunks<T, BLOCK, VEL == 2 ? (TILE + 2 * HALO) / VN : 1> cv;
    Chunks<T, BLOCK, STAGE >= 2 ? TILE / VN : 1> cn;
    cu.load(p.uk + S - HALO, tid);
    if constexpr (VEL == 2) cv.load(p.vel + S - HALO, tid);
    if constexpr (STAGE >= 2) cn.load(p.un + S, tid);
    gather_speed();
    cu.store(sm.u, tid);
    if constexpr (VEL == 2) cv.store(sm.v, tid);
    if constexpr (STAGE >= 2) cn.store(sm.io, tid);
  } else {
    gather_speed();
    if (STAGE == 0 && p.nowrap) {
      for (int j = tid; j < TILE + 2 * HALO; j += BLOCK) {
        sm.u[j] = load_raw(p.uk, S - HALO + j, m);
      }
    } else {
      for (int j = tid; j < TILE + 2 * HALO; j += BLOCK) {
        sm.u[j] = load_filled(p.uk, S - HALO + j, m, n, g);
      }
    }
    if constexpr (VEL == 2) {
      for (int j = tid; j < TILE + 2 * HALO; j += BLOCK) {
        sm.v[j] = load_raw(p.vel, S - HALO + j, m);
      }
    }
    if constexpr (STAGE >= 2) {
      for (int j = tid; j < TILE; j += BLOCK) {
        const int64_t gm = S + j;
        sm.io[j] = gm < m ? p.un[gm] : T(0);
      }
    }
  }
  tile_sync<ASYNC>();

  // ---- reconstruct CPT consecutive cells
  const int l0 = tid * CPT;
  T w[CPT + 4];  // cells l0 - 2 .. l0 + CPT + 1
#pragma unroll
  for (int k = 0; k < CPT + 4; ++k) w[k] = sm.u[l0 + 2 + k];
  T ul[CPT], ur[CPT];
  const MlpSource src{p.mlp, wl};
  // ES-WENO: the reconstruction also gives each cell's first weight
  static_assert(FLUX != PSA_FLUX_ESWENO || REC == 1,
                "the ES-WENO flux takes its weights from the ES-WENO3 reconstruction");
  T om[FLUX == PSA_FLUX_ESWENO ? CPT : 1];
  reconstruct_any<T, CPT, LEARNED, REC, EDGE>(p, w, ul, ur, src, q, S + l0,
                                              FLUX == PSA_FLUX_ESWENO ? om : nullptr);

  T al[VEL == 2 ? CPT : 1], ar[VEL == 2 ? CPT : 1];
  if constexpr (VEL == 2 && REC != 0) {
    // the velocity is reconstructed the same way (advection/schemes.py:108)
    reconstruct_any<T, CPT, false, REC, EDGE>(p, &sm.v[l0 + 2], al, ar, src,
                                              nullptr, S + l0);
  } else if constexpr (VEL == 2) {
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
      const int b = l0 + HALO + r;
      weno_pair(sm.v[b - 2], sm.v[b - 1], sm.v[b], sm.v[b + 1], sm.v[b + 2],
                p.eps, al[r], ar[r]);
    }
  }

  // ---- exchange edge states with the neighbouring threads
  sm.urL[tid + 1] = ur[CPT - 1];
  sm.ulF[tid] = ul[0];
  if constexpr (VEL == 2) {
    sm.arL[tid + 1] = ar[CPT - 1];
    sm.alF[tid] = al[0];
  }
  if constexpr (LEARNED) {
    // wave 0 (a wave-uniform branch) runs one 16-sample chain: lane 0 the
    // right state of local cell -1, lane 1 the left state of local cell TILE
    // from its mirrored stencil (lanes 2..15 repeat lane 0); the WENO-JS
    // states of those cells come from the same code as in the WENO-JS
    // kernels below, and the face loop picks one or the other
    if ((tid >> 6) == 0) {
      const bool hi = tid == 1;
      const int c = hi ? TILE + HALO : HALO - 1;
      const int d = hi ? -1 : 1;
      bool js;
      const T r = learned_right16(src, sm.u[c - 2 * d], sm.u[c - d], sm.u[c],
                                  sm.u[c + d], sm.u[c + 2 * d], p.eps, p.tau,
                                  p.sigma, js);
      if (tid < 2) {
        sm.edge_val[tid] = r;
        sm.edge_js[tid] = js ? 1 : 0;
      }
    }
  }
  if (tid == 0) {
    // right state of local cell -1 (the arithmetic of every other cell)
    T el, er;
    reconstruct_any<T, 1, false, REC, EDGE>(p, &sm.u[1], &el, &er, src, nullptr,
                                            S - 1);
    sm.urL[0] = er;
    if constexpr (VEL == 2 && REC != 0) {
      T al1, ar1;
      reconstruct_any<T, 1, false, REC, EDGE>(p, &sm.v[1], &al1, &ar1, src,
                                              nullptr, S - 1);
      sm.arL[0] = ar1;
    } else if constexpr (VEL == 2) {
      sm.arL[0] = weno_right(sm.v[1], sm.v[2], sm.v[3], sm.v[4], sm.v[5], p.eps);
    }
  }
  if (tid == BLOCK - 1) {
    // left state of local cell TILE
    constexpr int c = TILE + HALO;
    T el, er;
    reconstruct_any<T, 1, false, REC, EDGE>(p, &sm.u[c - 2], &el, &er, src,
                                            nullptr, S + TILE);
    sm.ulF[BLOCK] = el;
    if constexpr (VEL == 2 && REC != 0) {
      T al1, ar1;
      reconstruct_any<T, 1, false, REC, EDGE>(p, &sm.v[c - 2], &al1, &ar1, src,
                                              nullptr, S + TILE);
      sm.alF[BLOCK] = al1;
    } else if constexpr (VEL == 2) {
      sm.alF[BLOCK] = weno_right(sm.v[c + 2], sm.v[c + 1], sm.v[c],
                                 sm.v[c - 1], sm.v[c - 2], p.eps);
    }
  }
  tile_sync<ASYNC>();

  // ---- faces l0 - 1/2 .. l0 + CPT - 1/2
  // ES-WENO: weight 0 of the cells w[1] .. w[CPT + 2] (cell l0 - 1 + r),
  // each shared by two faces
  T esw[FLUX == PSA_FLUX_ESWENO ? CPT + 2 : 1];
  if constexpr (FLUX == PSA_FLUX_ESWENO) {
    // the thread's own cells from the reconstruction, the two neighbours'
    // with the same arithmetic (es_omega0_pair), so a face's two threads
    // agree bit for bit
    bool end0 = false, end1 = false;
    if constexpr (EDGE) {
      const int64_t g0 = S + l0 - 1, g1 = S + l0 + CPT;
      end0 = g0 == 0 || g0 == m - 1;
      end1 = g1 == 0 || g1 == m - 1;
    }
    esw[0] = es_omega0_pair(w[1] - w[0], w[2] - w[1], p.eps, end0);
#pragma unroll
    for (int r = 0; r < CPT; ++r) esw[r + 1] = om[r];
    esw[CPT + 1] = es_omega0_pair(w[CPT + 2] - w[CPT + 1], w[CPT + 3] - w[CPT + 2], p.eps,
                                  end1);
  }
  T F[CPT + 1];
  const bool has_nu = (FLUX == PSA_FLUX_RUSANOV ||
                       FLUX == PSA_FLUX_LAX_FRIEDRICHS) && p.nu != nullptr;
  // the stage kernels' flux forms (stage_flux) where they apply
  const bool plain = STAGE > 0 && VEL != 2 && !has_nu && FLUX != PSA_FLUX_ESWENO;
#pragma unroll
  for (int k = 0; k <= CPT; ++k) {
    T urL = k == 0 ? sm.urL[tid] : ur[k - 1];
    T ulR = k == CPT ? sm.ulF[tid + 1] : ul[k];
    if constexpr (LEARNED) {
      if (k == 0 && tid == 0 && !sm.edge_js[0]) urL = sm.edge_val[0];
      if (k == CPT && tid == BLOCK - 1 && !sm.edge_js[1]) ulR = sm.edge_val[1];
    }
    T Fk;
    if (plain) {
      Fk = stage_flux<T, FLUX>(urL, ulR, w[k + 1], w[k + 2], fspeed);
    } else if constexpr (FLUX == PSA_FLUX_RUSANOV) {
      if (has_nu) {
        const int64_t gm = S + l0 + k;  // face between cells gm - 1 and gm
        const T nuk = (gm >= 1 && gm <= m - 1) ? p.nu[gm - 1] : T(1);
        Fk = flux_rusanov(urL, ulR, w[k + 1], w[k + 2], nuk);
      } else {
        Fk = flux_rusanov_nu1(urL, ulR, w[k + 1], w[k + 2]);
      }
    } else if constexpr (FLUX == PSA_FLUX_LAX_FRIEDRICHS) {
      T nuk = T(1);
      if (has_nu) {
        const int64_t gm = S + l0 + k;
        nuk = (gm >= 1 && gm <= m - 1) ? p.nu[gm - 1] : T(1);
      }
      Fk = flux_lax_friedrichs(urL, ulR, fspeed, nuk);
    } else if constexpr (FLUX == PSA_FLUX_ENGQUIST_OSHER) {
      Fk = flux_engquist_osher(urL, ulR);
    } else if constexpr (FLUX == PSA_FLUX_BURGERS_UPWIND) {
      Fk = flux_burgers_upwind(urL, ulR);
    } else if constexpr (FLUX == PSA_FLUX_ESWENO) {
      // the face between cells w[k + 1] and w[k + 2]
      Fk = flux_burgers_upwind(urL, ulR) +
           es_dissipation_w(esw[k], esw[k + 1], w[k + 1], w[k + 2], p.es_delta);
    } else {
      T arL, alR;
      if constexpr (VEL == 2) {
        arL = k == 0 ? sm.arL[tid] : ar[k - 1];
        alR = k == CPT ? sm.alF[tid + 1] : al[k];
      } else {
        arL = p.vconst;
        alR = p.vconst;
      }
      Fk = flux_advection_upwind(urL, ulR, arL, alR);
    }
    if constexpr (STAGE == 0) {
      const int64_t gm = S + l0 + k;
      if (gm <= 0 || gm >= m) Fk = T(0);
    }
    F[k] = Fk;
  }

  // ---- conservative difference and stage combination
  T lam = T(0);
  if constexpr (STAGE > 0) lam = plain ? stage_lambda<T, FLUX>(p, dt) : dt * p.inv_dx;
#pragma unroll
  for (int r = 0; r < CPT; ++r) {
    const T dF = F[r + 1] - F[r];
    const T uk = w[r + 2];
    T val;
    if constexpr (STAGE == 0) {
      if constexpr (FLUX == PSA_FLUX_ADVECTION_UPWIND) {
        val = (-velocity_at<T, VEL>(sm.v, l0 + HALO + r, p.vconst) * dF) / p.dx;
      } else {
        val = -dF / p.dx;
      }
    } else {
      T lr = lam;
      if constexpr (FLUX == PSA_FLUX_ADVECTION_UPWIND && VEL == 2)
        lr = lam * velocity_at<T, VEL>(sm.v, l0 + HALO + r, p.vconst);
      const T un = STAGE >= 2 ? sm.io[l0 + r] : T(0);
      val = stage_value<T, STAGE == 0 ? 1 : STAGE>(uk, un, lr, dF);
    }
    sm.io[l0 + r] = val;
  }
  tile_sync<ASYNC>();

  // ---- coalesced stores (+ ghost images, + max|u| of the new state)
  if (fast) {
    Chunks<T, BLOCK, TILE / VN> co;
#pragma unroll
    for (int k = 0; k < co.CH; ++k) {
      const int c = tid + k * BLOCK;
      if (k < co.CH - 1 || (TILE / VN) % BLOCK == 0 || c < TILE / VN)
        co.r[k] = load16(sm.io + c * VN);
    }
    co.store(p.out + S, tid);
    if constexpr (kReducesOutput) {
#pragma unroll
      for (int k = 0; k < co.CH; ++k) {
        const int c = tid + k * BLOCK;
        if (k < co.CH - 1 || (TILE / VN) % BLOCK == 0 || c < TILE / VN)
          vmax = vmax_bits(co.r[k], vmax);
      }
    }
  } else {
    for (int j = tid; j < TILE; j += BLOCK) {
      const int64_t gm = S + j;
      if constexpr (STAGE == 0) {
        if (gm < m) p.out[gm] = sm.io[j];
      } else {
        if (gm >= g && gm < m - g) {
          const T v = sm.io[j];
          p.out[gm] = v;
          if (gm < 2 * g) p.out[gm + n] = v;
          if (gm >= m - 2 * g) p.out[gm - n] = v;
          if constexpr (kReducesOutput) {
            const unsigned long long b = abs_bits(v);
            vmax = b > vmax ? b : vmax;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// whole-step tile

// ES: the ES-WENO flux, whose threads also exchange the first weight of
// their edge cells (omL, omR).
template <typename T, int BLOCK, int CPT, bool ES = false>
struct StepSmem {
  static constexpr int C = BLOCK * CPT;  // cells computed per stage
  static constexpr int L = C + 8;        // cells held per buffer
  static constexpr int TO = C - 16;      // output cells per tile
  __align__(16) T x0[L];                 // u^n, then k2
  __align__(16) T x1[L];                 // k1
  T urL[BLOCK + 1];
  T ulF[BLOCK + 1];
  T omL[ES ? BLOCK + 1 : 1];
  T omR[ES ? BLOCK + 1 : 1];
  unsigned long long red[BLOCK / 64];
};

// One stage of the tile from the LDS buffer X into val (the thread's CPT
// cells): the reconstruction, the face fluxes and the stage combination of
// stage_tile (Rusanov with nu = 1, Godunov, or advection with a constant
// velocity). un holds the thread's cells of u^n.
template <typename T, int FLUX, int VEL, int BLOCK, int CPT, int STAGE,
          bool LEARNED, int REC = 0>
__device__ __forceinline__ void step_stage(const StageArgs<T> &p, const T *X,
                                           StepSmem<T, BLOCK, CPT,
                                                    FLUX == PSA_FLUX_ESWENO> &sm,
                                           T dt,
                                           const T *un, T *val,
                                           const MlpSource &src,
                                           unsigned int *q) {
  const int tid = threadIdx.x;
  const int j0 = 4 + CPT * tid;
  T w[CPT + 4];  // cells j0 - 2 .. j0 + CPT + 1
#pragma unroll
  for (int k = 0; k < CPT + 4; ++k) w[k] = X[j0 - 2 + k];
  T ul[CPT], ur[CPT];
  static_assert(FLUX != PSA_FLUX_ESWENO || REC == 1,
                "the ES-WENO flux takes its weights from the ES-WENO3 reconstruction");
  T om[FLUX == PSA_FLUX_ESWENO ? CPT : 1];
  reconstruct_any<T, CPT, LEARNED, REC>(p, w, ul, ur, src, q, 0,
                                        FLUX == PSA_FLUX_ESWENO ? om : nullptr);

  sm.urL[tid + 1] = ur[CPT - 1];
  sm.ulF[tid] = ul[0];
  if constexpr (FLUX == PSA_FLUX_ESWENO) {
    sm.omL[tid + 1] = om[CPT - 1];
    sm.omR[tid] = om[0];
  }
  __syncthreads();

  // ES-WENO: the first weight of cells j0 - 1 .. j0 + CPT, the outer two from
  // the neighbouring threads
  T esw[FLUX == PSA_FLUX_ESWENO ? CPT + 2 : 1];
  if constexpr (FLUX == PSA_FLUX_ESWENO) {
    esw[0] = sm.omL[tid];
#pragma unroll
    for (int r = 0; r < CPT; ++r) esw[r + 1] = om[r];
    esw[CPT + 1] = sm.omR[tid + 1];
  }
  T F[CPT + 1];
#pragma unroll
  for (int k = 0; k <= CPT; ++k) {
    const T urL = k == 0 ? sm.urL[tid] : ur[k - 1];
    const T ulR = k == CPT ? sm.ulF[tid + 1] : ul[k];
    if constexpr (FLUX == PSA_FLUX_ESWENO) {
      F[k] = flux_burgers_upwind(urL, ulR) +
             es_dissipation_w(esw[k], esw[k + 1], w[k + 1], w[k + 2], p.es_delta);
    } else {
      F[k] = stage_flux<T, FLUX>(urL, ulR, w[k + 1], w[k + 2], p.vconst);
    }
  }
  const T lam = stage_lambda<T, FLUX>(p, dt);
#pragma unroll
  for (int r = 0; r < CPT; ++r)
    val[r] = stage_value<T, STAGE>(w[r + 2], un[r], lam, F[r + 1] - F[r]);
}


// Stores a thread's CPT values of a whole-step tile (cells j0 = 4 + CPT t ..,
// relative to base) that fall in the output cells [12, C - 4), and with VMAX
// accumulates the bit pattern of max|u| of the interior ones.
template <typename T, int BLOCK, int CPT, bool VMAX>
__device__ __forceinline__ void step_store(const StageArgs<T> &p, int64_t base,
                                           bool fast, const T *val,
                                           T *__restrict__ dst,
                                           unsigned long long &vmax) {
  constexpr int C = BLOCK * CPT;
  constexpr int VN = Vec16<T>::n;
  const int j0 = 4 + CPT * static_cast<int>(threadIdx.x);
  const int64_t a0 = base + j0;
  if (fast && j0 >= 12 && j0 + CPT <= C - 4) {
    if constexpr (CPT % VN == 0) {
#pragma unroll
      for (int r = 0; r < CPT; r += VN) {
        typename Vec16<T>::type v;
        if constexpr (VN == 2) {
          v = {val[r], val[r + 1]};
        } else {
          v = {val[r], val[r + 1], val[r + 2], val[r + 3]};
        }
        store16(dst + a0 + r, v);
      }
    } else {
#pragma unroll
      for (int r = 0; r < CPT; ++r) dst[a0 + r] = val[r];
    }
    if constexpr (VMAX) {
      // the maximum in the dtype (NaN is tracked apart: fmax skips it)
      T mx = T(0);
      bool nan = false;
#pragma unroll
      for (int r = 0; r < CPT; ++r) {
        mx = fmax(mx, fabs(val[r]));
        nan |= isnan(val[r]);
      }
      const unsigned long long b = nan ? nan_bits<T>() : abs_bits(mx);
      vmax = b > vmax ? b : vmax;
    }
  } else {
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
      const int j = j0 + r;
      const int64_t a = a0 + r;
      if (j >= 12 && j < C - 4 && a >= 0 && a < p.m) {
        dst[a] = val[r];
        if constexpr (VMAX) {
          if (a >= p.g && a < p.m - p.g) {
            const unsigned long long b = abs_bits(val[r]);
            vmax = b > vmax ? b : vmax;
          }
        }
      }
    }
  }
}

// The ghost cells of a non-periodic boundary in the LDS stage buffer X of a
// whole-step tile (X[j] is cell base + j) filled from the buffer's interior
// cells (bc_cell): thread k < g the left cell k, thread g + k the right cell
// m - g + k, where the tile holds them.
template <typename T, int BLOCK, int CPT>
__device__ __forceinline__ void step_refill(const StageArgs<T> &p, int64_t base, T *X) {
  constexpr int L = StepSmem<T, BLOCK, CPT>::L;
  const int tid = threadIdx.x;
  const int g = p.g;
  if (tid >= 2 * g) return;
  const int64_t gi = tid < g ? tid : p.m - g + (tid - g);
  const int64_t j = gi - base;
  if (j >= 0 && j < L) X[j] = bc_cell(p, X - base, gi);
}

// One whole SSP-RK3 step (timestepping.py:312-320) of the output tile
// [A, A + TO) from uin = u^n into uout = u^(n+1) (distinct arrays of p.m
// cells): the workgroup holds the cells base .. base + C + 8 of u^n,
// base = A - 12 (16-byte loads for interior tiles, the periodic image
// (a - g) mod n otherwise). Thread t computes the cells j0 = 4 + CPT t ..
// j0 + CPT - 1 (relative to base) in every stage, so each stage computes C
// cells: stage 1 is exact on [5, C + 3) (the two outermost faces have no
// neighbouring thread), stage 2 on [8, C), stage 3 on [11, C - 3), which
// contains the output cells [12, C - 4). The halo cells are recomputed by the
// neighbouring tiles (16 of every C cells), with the arithmetic of the stage
// kernels, so the results are the same bit for bit. Output cells in the ghost
// layers are written as the periodic images they are. vmax accumulates the
// bit pattern of max|u| of the interior cells written. With p.k1out, p.k2out
// the stage states k1 and k2 of the output cells are written as well (they
// are exact there). Ends with the LDS in use (a barrier before the next
// tile).
//
// Ctl supplies the step: ctl.issue() runs once the tile's loads are in
// flight, ctl.go() and ctl.dt() are read after the workgroup's first barrier
// (a step not taken returns false with nothing computed or stored).
template <typename T, int FLUX, int VEL, int BLOCK, int CPT, bool LEARNED, int REC,
          class Ctl>
__device__ __forceinline__ bool step_tile_ctl(const StageArgs<T> &p, int64_t A,
                                              const Ctl &ctl,
                                              const T *__restrict__ uin,
                                              T *__restrict__ uout,
                                              StepSmem<T, BLOCK, CPT,
                                                       FLUX == PSA_FLUX_ESWENO> &sm,
                                              unsigned long long &vmax,
                                              unsigned int *q,
                                              const MlpWaveLds *wl) {
  static_assert(VEL != 2, "no velocity array in the whole-step tile");
  using SM = StepSmem<T, BLOCK, CPT, FLUX == PSA_FLUX_ESWENO>;
  constexpr int C = SM::C;
  constexpr int L = SM::L;
  constexpr int VN = Vec16<T>::n;
  static_assert(C % 16 == 0, "tile geometry");
  const int tid = threadIdx.x;
  const int64_t m = p.m;
  const int64_t n = p.n;
  const int g = p.g;
  const int64_t base = A - 12;

  // ---- u^n with a 12-cell margin
  const bool fast = p.aligned != 0 && base >= g && base + L <= m - g;
  if (fast) {
    Chunks<T, BLOCK, L / VN> cu;
    cu.load(uin + base, tid);
    ctl.issue();
    cu.store(sm.x0, tid);
  } else if (p.bcm == PSA_BC_PERIODIC) {
    ctl.issue();
    for (int j = tid; j < L; j += BLOCK) {
      int64_t i = (base + j - g) % n;
      if (i < 0) i += n;
      sm.x0[j] = uin[g + i];
    }
  } else {
    ctl.issue();
    for (int j = tid; j < L; j += BLOCK) sm.x0[j] = bc_cell(p, uin, base + j);
  }
  // the k1 entries stage 2 reads outside the cells stage 1 writes, and the
  // exchange entries without a neighbouring thread (their results are
  // outside the exact ranges; finite values keep them on the fast paths)
  if (tid < 8) sm.x1[tid < 4 ? tid : C + tid] = T(0);
  if (tid == 0) {
    sm.urL[0] = T(0);
    sm.ulF[BLOCK] = T(0);
    if constexpr (FLUX == PSA_FLUX_ESWENO) {
      sm.omL[0] = T(0);
      sm.omR[BLOCK] = T(0);
    }
  }
  __syncthreads();
  if (!ctl.go()) return false;
  const T dt = ctl.dt();

  const int j0 = 4 + CPT * tid;
  T un[CPT], val[CPT];
#pragma unroll
  for (int r = 0; r < CPT; ++r) un[r] = sm.x0[j0 + r];
  const MlpSource src{p.mlp, wl};

  // a tile that holds ghost cells of a non-periodic boundary refills them
  // from each stage's output before the next stage reads it (the interior
  // cells never read the cells outside [0, m), whose stage values are
  // meaningless)
  const bool refill = p.bcm != PSA_BC_PERIODIC && (base < g || base + L > m - g);
  unsigned long long unused = 0;
  step_stage<T, FLUX, VEL, BLOCK, CPT, 1, LEARNED, REC>(p, sm.x0, sm, dt, un, val, src, q);
#pragma unroll
  for (int r = 0; r < CPT; ++r) sm.x1[j0 + r] = val[r];
  if (p.k1out != nullptr) step_store<T, BLOCK, CPT, false>(p, base, fast, val, p.k1out, unused);
  __syncthreads();
  if (refill) {
    step_refill<T, BLOCK, CPT>(p, base, sm.x1);
    __syncthreads();
  }
  // (every read of x0 ended before the exchange barrier of stage 1)
  step_stage<T, FLUX, VEL, BLOCK, CPT, 2, LEARNED, REC>(p, sm.x1, sm, dt, un, val, src, q);
#pragma unroll
  for (int r = 0; r < CPT; ++r) sm.x0[j0 + r] = val[r];
  if (p.k2out != nullptr) step_store<T, BLOCK, CPT, false>(p, base, fast, val, p.k2out, unused);
  __syncthreads();
  if (refill) {
    step_refill<T, BLOCK, CPT>(p, base, sm.x0);
    __syncthreads();
  }
  step_stage<T, FLUX, VEL, BLOCK, CPT, 3, LEARNED, REC>(p, sm.x0, sm, dt, un, val, src, q);

  step_store<T, BLOCK, CPT, true>(p, base, fast, val, uout, vmax);
  return true;
}

// A step of a given dt.
template <typename T>
struct FixedStep {
  T v;
  __device__ __forceinline__ void issue() const {}
  __device__ __forceinline__ bool go() const { return true; }
  __device__ __forceinline__ T dt() const { return v; }
};

template <typename T, int FLUX, int VEL, int BLOCK, int CPT, bool LEARNED,
          int REC = 0>
__device__ __forceinline__ void step_tile(const StageArgs<T> &p, int64_t A, T dt,
                                          const T *__restrict__ uin,
                                          T *__restrict__ uout,
                                          StepSmem<T, BLOCK, CPT,
                                                   FLUX == PSA_FLUX_ESWENO> &sm,
                                          unsigned long long &vmax,
                                          unsigned int *q,
                                          const MlpWaveLds *wl) {
  step_tile_ctl<T, FLUX, VEL, BLOCK, CPT, LEARNED, REC>(p, A, FixedStep<T>{dt}, uin, uout,
                                                        sm, vmax, q, wl);
}

}  // namespace psa
