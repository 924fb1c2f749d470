// Forward WENO-JS5 / Rusanov / Godunov / SSP-RK3 kernels for gfx950 and the
// C ABI declared in include/pyshocks_amd.h.
//
// Layout: every state is a contiguous array of m = n + 2 g cells (the
// reference's grid.x layout, grid.py:131-179). The fused stage kernel tiles
// the cell index space from 0 in tiles of BLOCK * CPT cells; each workgroup
// stages its tile plus a 4-cell halo per side in LDS from coalesced loads,
// every thread reconstructs CPT consecutive cells (both sides from one set
// of smoothness indicators), neighbouring threads exchange one right and one
// left state through LDS, each thread evaluates its CPT + 1 faces, applies
// the conservative difference and the SSP-RK3 stage combination, and the
// results leave through LDS as coalesced stores.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>

#include "../../include/pyshocks_amd.h"
#include "psa_internal.hpp"
#include "stage_device.hpp"
#include "weno5_device.hpp"

namespace psa {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int check_launch(const char *what) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    return set_error(PSA_ERR_HIP, std::string(what) + ": " +
                                      hipGetErrorString(err));
  }
  return PSA_OK;
}

// ---------------------------------------------------------------------------
// standalone weno.py kernels (reference expression order)

template <typename T>
__global__ void k_weno_fields(const T *__restrict__ u, int64_t m, T eps,
                              T *__restrict__ beta, T *__restrict__ uhat,
                              T *__restrict__ omega) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T um2 = load_raw(u, i - 2, m), um1 = load_raw(u, i - 1, m);
  const T u0 = u[i];
  const T up1 = load_raw(u, i + 1, m), up2 = load_raw(u, i + 2, m);

  T b0, b1, b2;
  weno_beta(um2, um1, u0, up1, up2, b0, b1, b2);
  if (beta != nullptr) {
    beta[i] = b0;
    beta[m + i] = b1;
    beta[2 * m + i] = b2;
  }
  if (uhat != nullptr) {
    T h0, h1, h2;
    weno_uhat(um2, um1, u0, up1, up2, h0, h1, h2);
    uhat[i] = h0;
    uhat[m + i] = h1;
    uhat[2 * m + i] = h2;
  }
  if (omega != nullptr) {
    using C = Weno5Coeffs<T>;
    const T e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
    const T a0 = C::d0 / (e0 * e0);
    const T a1 = C::d1 / (e1 * e1);
    const T a2 = C::d2 / (e2 * e2);
    const T s = (a0 + a1) + a2;
    omega[i] = a0 / s;
    omega[m + i] = a1 / s;
    omega[2 * m + i] = a2 / s;
  }
}

// Both states of the cells of f as given ('same' convolutions, zero padding):
// REC = PSA_REC_WENOJS53, _WENOJS32 or _ESWENO32.
template <typename T, int REC>
__global__ void k_weno_reconstruct(const T *__restrict__ f, int64_t m, T eps,
                                   T *__restrict__ ul, T *__restrict__ ur) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T um1 = load_raw(f, i - 1, m);
  const T u0 = f[i];
  const T up1 = load_raw(f, i + 1, m);
  if constexpr (REC == PSA_REC_WENOJS53) {
    const T um2 = load_raw(f, i - 2, m), up2 = load_raw(f, i + 2, m);
    ur[i] = weno_side_ref(um2, um1, u0, up1, up2, eps);
    ul[i] = weno_side_ref(up2, up1, u0, um1, um2, eps);
  } else {
    constexpr bool es = REC == PSA_REC_ESWENO32;
    const T tau = (es && i > 0 && i < m - 1) ? weno3_tau_ref(um1, u0, up1) : T(0);
    ur[i] = weno3_side_ref(um1, u0, up1, eps, es, tau);
    ul[i] = weno3_side_ref(up1, u0, um1, eps, es, tau);
  }
}

// reconstruction.py:279-296 slope-limited MUSCL on f as given.
template <typename T>
__global__ void k_muscls_reconstruct(const T *__restrict__ f, int64_t m,
                                     LimParams<T> lim, T *__restrict__ ul,
                                     T *__restrict__ ur) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T um1 = load_raw(f, i - 1, m), u0 = f[i], up1 = load_raw(f, i + 1, m);
  T l, r;
  muscls_pair(u0, u0 - um1, up1 - u0, up1 + um1, lim.lm, lim.dx,
              i == 0 || i == m - 1, l, r);
  ul[i] = l;
  ur[i] = r;
}

// reconstruction.py:205-252 MUSCL on f as given (muscl_pair with the ends of
// the array padded as the reference pads them).
template <typename T>
__global__ void k_muscl_reconstruct(const T *__restrict__ f, int64_t m,
                                    LimParams<T> lim, T *__restrict__ ul,
                                    T *__restrict__ ur) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T um1 = load_raw(f, i - 1, m), u0 = f[i], up1 = load_raw(f, i + 1, m);
  T l, r;
  muscl_pair(u0, u0 - um1, up1 - u0, lim.lm, lim.theta, lim.atol, i == 0,
             i == m - 1, l, r);
  ul[i] = l;
  ur[i] = r;
}

template <typename T, int K>
__global__ void k_esweno_dissipation(const T *__restrict__ u, T *__restrict__ fnum,
                                     int64_t m, T eps, T delta) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x + 1;
  if (j >= m) return;
  // cells j - 2 .. j + 1
  const T v0 = load_raw(u, j - 2, m), v1 = u[j - 1], v2 = u[j], v3 = load_raw(u, j + 1, m);
  fnum[j] = fnum[j] + es_dissipation<T, K>(v0, v1, v2, v3, eps, delta, j - 1 == 0,
                                           j == m - 1);
}

template <typename T>
__global__ void k_flux_divergence(const T *__restrict__ fnum, const T *__restrict__ vel,
                                  T *__restrict__ out, int64_t m, T dx) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T dF = fnum[i + 1] - fnum[i];
  out[i] = vel != nullptr ? (-vel[i] * dF) / dx : -dF / dx;
}

// scalar.py:441-457: ghost cells of a side from the given values (interior
// copied when out != u).
template <typename T>
__global__ void k_fill_dirichlet(const T *__restrict__ u, T *__restrict__ out,
                                 int64_t m, int g, const T *__restrict__ left,
                                 const T *__restrict__ right) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  if (i < g && left != nullptr) {
    out[i] = left[i];
  } else if (i >= m - g && right != nullptr) {
    out[i] = right[i - (m - g)];
  } else if (out != u) {
    out[i] = u[i];
  }
}

// scalar.py:477-511 on a uniform grid: ghost g-1-k <- u[g+k] - (2k+1) dx gl,
// ghost m-1-k <- u[m-2g+k] + (2k+1-2g) dx gr (sides: 1 left, 2 right).
template <typename T>
__global__ void k_fill_neumann(const T *__restrict__ u, T *__restrict__ out,
                               int64_t m, int g, T dx, T gl, T gr, int sides) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  if (i < g && (sides & 1)) {
    const int k = g - 1 - static_cast<int>(i);
    out[i] = u[g + k] - (T(2 * k + 1) * dx) * gl;
  } else if (i >= m - g && (sides & 2)) {
    const int k = static_cast<int>(m - 1 - i);
    out[i] = u[m - 2 * g + k] + (T(2 * k + 1 - 2 * g) * dx) * gr;
  } else if (out != u) {
    out[i] = u[i];
  }
}

template <typename T>
__global__ void k_apply_boundary(const T *__restrict__ u, T *__restrict__ out,
                                 int64_t m, int64_t n, int g) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  // in place is safe: ghosts read interior cells, interior cells read
  // themselves
  if (i < g || i >= m - g) {
    out[i] = load_filled(u, i, m, n, g);
  } else if (out != u) {
    out[i] = u[i];
  }
}

// numerical_flux on u as given: one thread per face j in [0, m], face j lies
// between cells j - 1 and j; fnum[0] = fnum[m] = 0 (jnp.pad(fnum, 1)).
// The states either side of face j (between cells j - 1 and j) from the
// cells v = j - 3 .. j + 2 of an array of m cells as given: ur of cell j - 1
// and ul of cell j, by the reconstruction REC.
template <typename T, int REC>
__device__ __forceinline__ void face_states(const T *v, int64_t j, int64_t m,
                                            T eps, T &urL, T &ulR,
                                            const LimParams<T> &lim) {
  if constexpr (REC == PSA_REC_MUSCL) {
    // cells j - 2 .. j + 1 are v[1] .. v[4]; cell j - 1 may be the first
    // of the array, cell j the last
    T ul, ur;
    muscl_pair(v[2], v[2] - v[1], v[3] - v[2], lim.lm, lim.theta, lim.atol,
               j - 1 == 0, false, ul, urL);
    muscl_pair(v[3], v[3] - v[2], v[4] - v[3], lim.lm, lim.theta, lim.atol, false,
               j == m - 1, ulR, ur);
  } else if constexpr (REC == PSA_REC_MUSCLS) {
    T ul, ur;
    muscls_pair(v[2], v[2] - v[1], v[3] - v[2], v[3] + v[1], lim.lm, lim.dx,
                j - 1 == 0, ul, urL);
    muscls_pair(v[3], v[3] - v[2], v[4] - v[3], v[4] + v[2], lim.lm, lim.dx,
                j == m - 1, ulR, ur);
  } else if constexpr (REC == PSA_REC_WENOJS53) {
    urL = weno_side_ref(v[0], v[1], v[2], v[3], v[4], eps);
    ulR = weno_side_ref(v[5], v[4], v[3], v[2], v[1], eps);
  } else {
    // cells j - 2 .. j + 1 are v[1] .. v[4]; tau is zero at the array ends
    constexpr bool es = REC == PSA_REC_ESWENO32;
    const T tL = (es && j - 1 > 0) ? weno3_tau_ref(v[1], v[2], v[3]) : T(0);
    const T tR = (es && j < m - 1) ? weno3_tau_ref(v[2], v[3], v[4]) : T(0);
    urL = weno3_side_ref(v[1], v[2], v[3], eps, es, tL);
    ulR = weno3_side_ref(v[4], v[3], v[2], eps, es, tR);
  }
}

template <typename T, int FLUX, int REC = PSA_REC_WENOJS53>
__global__ void k_numerical_flux(const T *__restrict__ u,
                                 const T *__restrict__ vel,
                                 const T *__restrict__ nu, int64_t m, T eps,
                                 T *__restrict__ fnum, LimParams<T> lim) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j > m) return;
  if (j == 0 || j == m) {
    fnum[j] = T(0);
    return;
  }
  // cells j - 3 .. j + 2
  T v[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = load_raw(u, j - 3 + k, m);
  // ur of cell j - 1 and ul of cell j
  T urL, ulR;
  face_states<T, REC>(v, j, m, eps, urL, ulR, lim);
  T F;
  if constexpr (FLUX == PSA_FLUX_RUSANOV) {
    const T nuj = nu != nullptr ? nu[j - 1] : T(1);
    F = flux_rusanov(urL, ulR, v[2], v[3], nuj);
  } else if constexpr (FLUX == PSA_FLUX_LAX_FRIEDRICHS) {
    // vel is the device scalar max|a|
    const T nuj = nu != nullptr ? nu[j - 1] : T(1);
    F = flux_lax_friedrichs(urL, ulR, *vel, nuj);
  } else if constexpr (FLUX == PSA_FLUX_ENGQUIST_OSHER) {
    F = flux_engquist_osher(urL, ulR);
  } else if constexpr (FLUX == PSA_FLUX_BURGERS_UPWIND) {
    F = flux_burgers_upwind(urL, ulR);
  } else {
    T a[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] = load_raw(vel, j - 3 + k, m);
    T arL, alR;
    face_states<T, REC>(a, j, m, eps, arL, alR, lim);
    F = flux_advection_upwind(urL, ulR, arL, alR);
  }
  fnum[j] = F;
}

// advection/schemes.py:135-186 at face j (cells j - 2 .. j + 1).
template <typename T>
__global__ void k_esweno_advection_flux(const T *__restrict__ u, const T *__restrict__ vel,
                                        const T *__restrict__ amax, int64_t m, T eps,
                                        T delta, T *__restrict__ fnum) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j > m) return;
  if (j == 0 || j == m) {
    fnum[j] = T(0);
    return;
  }
  T v[6], f[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    v[k] = load_raw(u, j - 3 + k, m);
    f[k] = load_raw(vel, j - 3 + k, m) * v[k];
  }
  const LimParams<T> none{};
  T urL, ulR, frL, flR;
  face_states<T, PSA_REC_ESWENO32>(v, j, m, eps, urL, ulR, none);
  face_states<T, PSA_REC_ESWENO32>(f, j, m, eps, frL, flR, none);
  const T a = *amax;
  const T F = T(0.5) * (flR + frL) - T(0.5) * a * (ulR - urL);
  fnum[j] = F + es_dissipation<T, 1>(v[1], v[2], v[3], v[4], eps, delta, j - 1 == 0,
                                     j == m - 1);
}

template <typename T>
__global__ void k_fill_zero_bits(unsigned long long *bits) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *bits = 0ull;
}

template <typename T>
__global__ void k_store_zero(T *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = T(0);
}

// max |u[g:m-g]| by atomicMax on the bit pattern of the (non-negative)
// result; out points to a scalar of the dtype (zeroed beforehand).
template <typename T, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_max_abs(const T *__restrict__ u,
                                                   int64_t m, int g,
                                                   T *__restrict__ out,
                                                   unsigned long long *bits) {
  __shared__ unsigned long long s_red[BLOCK / 64];
  unsigned long long v = 0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * BLOCK;
  for (int64_t i = g + static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x;
       i < m - g; i += stride) {
    const unsigned long long b = abs_bits(u[i]);
    v = b > v ? b : v;
  }
  v = block_max_bits<BLOCK>(v, s_red);
  if (threadIdx.x == 0) {
    if (bits != nullptr) {
      atomicMax(bits, v);
    } else if constexpr (sizeof(T) == 8) {
      atomicMax(reinterpret_cast<unsigned long long *>(out), v);
    } else {
      atomicMax(reinterpret_cast<unsigned int *>(out),
                static_cast<unsigned int>(v));
    }
  }
}

// ---------------------------------------------------------------------------
// device-side step control (timestepping.py:89-152)

template <typename T>
__global__ void k_control_init(psa_control *ctrl, double t, double tfinal,
                               double theta, double cfl_num, double fixed_dt,
                               long long maxit, int dt_mode,
                               int speed_from_state, double speed_const) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ctrl->t = t;
  ctrl->dt = 0.0;
  ctrl->tfinal = tfinal;
  ctrl->theta = theta;
  ctrl->cfl_num = cfl_num;
  ctrl->fixed_dt = fixed_dt;
  ctrl->m = 0;
  ctrl->maxit = maxit;
  ctrl->umax_bits = abs_bits(static_cast<T>(speed_const));
  for (int k = 0; k < PSA_UMAX_SLOTS; ++k) {
    ctrl->umax_slot[k][0] = 0ull;
    ctrl->speed_slot[0][k][0] = 0ull;
    ctrl->speed_slot[1][k][0] = 0ull;
    for (int b = 0; b < 3; ++b) ctrl->step_slot[b][k][0] = 0ull;
  }
  ctrl->umax_slot[0][0] = ctrl->umax_bits;
  for (int b = 0; b < 2; ++b) {
    ctrl->seq_t[b] = t;
    ctrl->seq_m[b] = 0;
    ctrl->seq_status[b] = PSA_OK;
  }
  ctrl->active = 0;
  ctrl->status = PSA_OK;
  ctrl->dt_mode = dt_mode;
  ctrl->speed_from_state = speed_from_state;
}

// Decision of stage 1, evaluated identically by every workgroup.
__device__ __forceinline__ unsigned long long gather_umax(
    const psa_control *ctrl) {
  return gather_slots(ctrl->umax_slot);
}

// Whether the step from a state at time t0 after m0 steps with status
// status0 is taken, and its dt (timestepping.py:89-152).
template <typename T>
__device__ __forceinline__ bool control_decide_at(const psa_control *ctrl, double t0,
                                                  long long m0, int status0,
                                                  unsigned long long umax_bits,
                                                  T &dt, int &status) {
  status = status0;
  if (status != PSA_OK) return false;
  const T t = static_cast<T>(t0);
  const T tfinal = static_cast<T>(ctrl->tfinal);
  if (t >= tfinal) return false;
  if (ctrl->maxit >= 0 && m0 >= ctrl->maxit) return false;

  if (ctrl->dt_mode == PSA_DT_FIXED) {
    dt = static_cast<T>(ctrl->fixed_dt);
  } else {
    const T smax = from_bits<T>(umax_bits);
    dt = static_cast<T>(ctrl->theta) * (static_cast<T>(ctrl->cfl_num) / smax);
  }
  if (!isinf(tfinal)) {
    const T dt_min = tfinal - t;
    dt = (dt < dt_min ? dt : dt_min) + static_cast<T>(1.0e-15);
  }
  if (!isfinite(dt)) {
    status = PSA_ERR_NONFINITE_DT;
    return false;
  }
  return true;
}

template <typename T>
__device__ __forceinline__ bool control_decide(const psa_control *ctrl,
                                               unsigned long long umax_bits,
                                               T &dt, int &status) {
  return control_decide_at<T>(ctrl, ctrl->t, ctrl->m, ctrl->status, umax_bits, dt,
                              status);
}

template <typename T>
__global__ void k_copy_if_inactive(const T *__restrict__ src,
                                   T *__restrict__ dst, int64_t m,
                                   const psa_control *ctrl) {
  if (ctrl->active != 0) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// fused stage kernel: one workgroup per tile

// Step control at the start of a stage launch (timestepping.py:89-152 on
// the device): stage 1 decides whether the step is taken and publishes dt,
// stage 2 advances t and m and clears the max|u| slots, stage 3 reads dt.
// Lax-Friedrichs (scalar.py:276-277): speed is max|u| of the stage input,
// from the max|u| slots (u^n) or speed_slot[0] (k1) or speed_slot[1] (k2),
// or from p.speed without a control block; stage 1 clears speed_slot[1] for
// stage 2 and stage 3 clears speed_slot[0] for the next step's stage 1
// (each after its last reader is done). DEFER: stages 2 and 3 leave the
// speed slots to stage_tile (p.speed_slots), which gathers them once its
// tile loads are in flight.
// Returns false when the step is not taken (the workgroup then exits).
template <typename T, int STAGE, int FLUX, bool DEFER = false>
__device__ __forceinline__ bool stage_control(const StageArgs<T> &p, T &dt,
                                              T &speed) {
  constexpr bool LF = FLUX == PSA_FLUX_LAX_FRIEDRICHS;
  const int tid = threadIdx.x;
  dt = T(0);
  speed = T(0);
  if constexpr (STAGE > 0) {
    if (p.ctrl != nullptr) {
      if constexpr (STAGE == 1) {
        int status;
        const unsigned long long umax = gather_umax(p.ctrl);
        const bool active = control_decide<T>(p.ctrl, umax, dt, status);
        if (blockIdx.x == 0 && tid == 0) {
          p.ctrl->active = active ? 1 : 0;
          p.ctrl->status = status;
          p.ctrl->umax_bits = umax;
          if (active) p.ctrl->dt = static_cast<double>(dt);
        }
        if (!active) return false;
        if constexpr (LF) {
          speed = from_bits<T>(umax);
          if (blockIdx.x == 0 && tid < PSA_UMAX_SLOTS)
            p.ctrl->speed_slot[1][tid][0] = 0ull;
        }
      } else {
        if (p.ctrl->active == 0) return false;
        dt = static_cast<T>(p.ctrl->dt);
        if constexpr (LF && !DEFER)
          speed = from_bits<T>(gather_slots(p.ctrl->speed_slot[STAGE - 2]));
        if constexpr (STAGE == 2) {
          if (blockIdx.x == 0 && tid == 0) {
            p.ctrl->t = static_cast<double>(static_cast<T>(p.ctrl->t) + dt);
            p.ctrl->m = p.ctrl->m + 1;
          }
          // stage 1 has read the slots; clear them for stage 3
          if (p.ctrl->speed_from_state && blockIdx.x == 0 &&
              tid < PSA_UMAX_SLOTS) {
            p.ctrl->umax_slot[tid][0] = 0ull;
          }
        } else if constexpr (LF) {
          // stage 2 has read speed_slot[0]
          if (blockIdx.x == 0 && tid < PSA_UMAX_SLOTS)
            p.ctrl->speed_slot[0][tid][0] = 0ull;
        }
      }
    } else {
      dt = *p.dt_ptr;
      if constexpr (LF) speed = *p.speed;
    }
  } else if constexpr (LF) {
    speed = *p.speed;
  }

  return true;
}

// The workgroup's part of the max|u| reduction of a stage launch's output
// (after its last tile): into the max|u| slots for the next step's dt
// (stage 3), into speed_slot[STAGE - 1] for the Lax-Friedrichs speed of the
// next stage (stages 1 and 2).
template <typename T, int STAGE, int FLUX, int BLOCK>
__device__ __forceinline__ void stage_reduce(const StageArgs<T> &p,
                                             unsigned long long vmax,
                                             unsigned long long *red) {
  constexpr bool lf = FLUX == PSA_FLUX_LAX_FRIEDRICHS && (STAGE == 1 || STAGE == 2);
  if constexpr (STAGE == 3 || lf) {
    if (p.ctrl != nullptr && p.ctrl->speed_from_state) {
      vmax = block_max_bits<BLOCK>(vmax, red);
      if (threadIdx.x == 0) {
        const int k = blockIdx.x % PSA_UMAX_SLOTS;
        if constexpr (lf) {
          atomicMax(&p.ctrl->speed_slot[STAGE - 1][k][0], vmax);
        } else {
          atomicMax(&p.ctrl->umax_slot[k][0], vmax);
        }
      }
    }
  }
}

template <typename T, int FLUX, int VEL, int STAGE, int BLOCK, int CPT,
          bool LEARNED, int REC = 0>
__global__ __launch_bounds__(BLOCK) void k_stage(StageArgs<T> p) {
  constexpr int TILE = BLOCK * CPT;
  __shared__ StageSmem<T, VEL, BLOCK, CPT> sm;
  __shared__ unsigned int q[LEARNED ? (BLOCK / 64) * learned_queue_words<CPT>() : 1];
  const int tid = threadIdx.x;
  const int64_t ntiles = (p.m + TILE - 1) / TILE;

  T dt, speed;
  if (!stage_control<T, STAGE, FLUX, true>(p, dt, speed)) return;
  StageArgs<T> pl = p;
  if constexpr (FLUX == PSA_FLUX_LAX_FRIEDRICHS) {
    pl.vconst = speed;
    if (STAGE >= 2 && p.ctrl != nullptr)
      pl.speed_slots = p.ctrl->speed_slot[STAGE >= 2 ? STAGE - 2 : 0];
  }

  unsigned long long vmax = 0;
  const int64_t tile = blockIdx.x;  // the grid has one workgroup per tile
  if (tile < ntiles) {
    stage_tile<T, FLUX, VEL, STAGE, BLOCK, CPT, LEARNED, false, false, REC>(
        pl, tile * TILE, dt, sm, vmax, q);
  }
  (void)tid;
  stage_reduce<T, STAGE, FLUX, BLOCK>(p, vmax, sm.red);
}

// Pipelined stage kernel: a persistent grid (the workgroups the device holds
// at once) walks the tiles with a stride of the grid size. While a
// workgroup computes one tile from one LDS buffer, the next tile (and its
// u^n) is copied straight from HBM into the other buffer, so every workgroup
// keeps its next loads in flight through its compute phase. Tiles at the
// domain ends take the ordinary path (ghost mapping, bounds checks). The
// per-tile arithmetic is that of k_stage, so results are bit-identical.
template <typename T, int FLUX, int VEL, int STAGE, int BLOCK, int CPT>
__global__ __launch_bounds__(BLOCK) void k_stage_pipe(StageArgs<T> p) {
  static_assert(VEL != 2 && STAGE > 0, "pipelined stages 1-3 without a velocity array");
  constexpr int TILE = BLOCK * CPT;
  constexpr int HALO = 4;
  constexpr int VN = Vec16<T>::n;
  __shared__ StageSmem<T, VEL, BLOCK, CPT> smA;
  __shared__ StageSmem<T, VEL, BLOCK, CPT> smB;
  const int tid = threadIdx.x;
  const int64_t ntiles = (p.m + TILE - 1) / TILE;

  T dt, speed;
  if (!stage_control<T, STAGE, FLUX>(p, dt, speed)) return;
  StageArgs<T> pl = p;
  if constexpr (FLUX == PSA_FLUX_LAX_FRIEDRICHS) pl.vconst = speed;

  auto fast_tile = [&](int64_t t) {
    const int64_t S = t * TILE;
    return p.aligned != 0 && S - HALO >= p.g && S + TILE + HALO <= p.m - p.g;
  };
  auto issue = [&](int64_t t, StageSmem<T, VEL, BLOCK, CPT> &sm) {
    if (t < ntiles && fast_tile(t)) {
      const int64_t S = t * TILE;
      dma_chunks<BLOCK, (TILE + 2 * HALO) / VN>(p.uk + S - HALO, sm.u);
      if constexpr (STAGE >= 2) dma_chunks<BLOCK, TILE / VN>(p.un + S, sm.io);
    }
  };
  unsigned long long vmax = 0;
  // one tile: wait until every wave is done with nxt (its last reads were
  // the previous tile's stores), start the copies of the following tile into
  // it, compute this tile from cur, then wait for the copies
  auto step = [&](int64_t t, StageSmem<T, VEL, BLOCK, CPT> &cur,
                  StageSmem<T, VEL, BLOCK, CPT> &nxt) {
    tile_sync<true>();
    issue(t + gridDim.x, nxt);
    stage_tile<T, FLUX, VEL, STAGE, BLOCK, CPT, false, true, true>(pl, t * TILE, dt,
                                                                 cur, vmax);
    // the copies into nxt are older than this tile's stores: on the fast path
    // (a known number of stores per wave) only the stores may stay in flight
    if (fast_tile(t)) {
      wait_vm_but<Chunks<T, BLOCK, TILE / VN>::CH>();
    } else {
      wait_vm_all();
    }
    tile_sync<true>();
  };

  int64_t t = blockIdx.x;
  issue(t, smA);
  wait_vm_all();
  tile_sync<true>();
  while (t < ntiles) {
    step(t, smA, smB);
    t += gridDim.x;
    if (t >= ntiles) break;
    step(t, smB, smA);
    t += gridDim.x;
  }
  (void)tid;
  stage_reduce<T, STAGE, FLUX, BLOCK>(p, vmax, smA.red);
}

// ---------------------------------------------------------------------------
// whole-step kernel (step_tile): the three SSP-RK3 stages of a tile in LDS

// Before the first whole-step launch of a psa_ssprk33_steps call (one wave):
// the state's t, m and status into entry 0 of the sequence record, max|u|
// of the state (or max|a| of an advection velocity array: the word-0 slots,
// as stage 1 reads them) into umax_bits, which the launches read for a
// speed that does not change, and into slot 0 of step_slot[2], which launch
// 0 reads otherwise; step_slot[0], which launch 0 reduces into, cleared.
__global__ void k_step_prologue(psa_control *ctrl) {
  const int lane = threadIdx.x;
  const unsigned long long umax = gather_umax(ctrl);
  ctrl->step_slot[2][lane][0] = lane == 0 ? umax : 0ull;
  ctrl->step_slot[0][lane][0] = 0ull;
  if (lane == 0) {
    ctrl->umax_bits = umax;
    ctrl->seq_t[0] = ctrl->t;
    ctrl->seq_m[0] = ctrl->m;
    ctrl->seq_status[0] = ctrl->status;
  }
}

// After the last whole-step launch s of the call (one wave): max|u| of the
// final state (the partials of step_slot[s % 3]) into word 0 of slot 0 for
// the stage kernels and the next call (the other word-0 entries zero).
__global__ void k_step_epilogue(psa_control *ctrl, int last_set) {
  const int lane = threadIdx.x;
  if (ctrl->speed_from_state == 0) return;
  const unsigned long long umax = wave_max_bits(ctrl->step_slot[last_set][lane][0]);
  ctrl->umax_slot[lane][0] = lane == 0 ? umax : 0ull;
}

// The step control of whole-step launch s = seq (mod 6) of a sequence
// (timestepping.py:89-152): wave 0 of every workgroup makes the same decision
// (control_decide_at) from the sequence record seq_*[s % 2] and max|u| of
// u^n, the maximum of the partials the previous launch left in
// step_slot[(s + 2) % 3], while the tile's loads are in flight, and leaves
// it in LDS for the workgroup. Wave 0 of workgroup 0 records the state after
// the step in seq_*[(s + 1) % 2] and in the fields the host and the stage
// kernels read (t, m, dt, active, status, umax_bits), and clears
// step_slot[(s + 1) % 3] for the next launch; a step not taken carries max|u|
// over into step_slot[s % 3].
template <typename T>
struct SeqStep {
  psa_control *ctrl;
  int seq;
  T *s_dt;
  int *s_go;

  __device__ __forceinline__ void issue() const {
    const int tid = threadIdx.x;
    if (tid >= 64) return;  // wave 0
    const int par = seq & 1;
    const bool from_state = ctrl->speed_from_state != 0;
    const unsigned long long umax =
        from_state ? wave_max_u64(ctrl->step_slot[(seq + 2) % 3][tid][0]) : ctrl->umax_bits;
    const double t0 = ctrl->seq_t[par];
    const long long m0 = ctrl->seq_m[par];
    T dt = T(0);
    int status;
    const bool active =
        control_decide_at<T>(ctrl, t0, m0, ctrl->seq_status[par], umax, dt, status);
    if (tid == 0) {
      *s_dt = dt;
      *s_go = active ? 1 : 0;
    }
    if (blockIdx.x != 0) return;
    ctrl->step_slot[(seq + 1) % 3][tid][0] = 0ull;
    if (tid == 0) {
      const double t1 = active ? static_cast<double>(static_cast<T>(t0) + dt) : t0;
      const long long m1 = active ? m0 + 1 : m0;
      ctrl->seq_t[par ^ 1] = t1;
      ctrl->seq_m[par ^ 1] = m1;
      ctrl->seq_status[par ^ 1] = status;
      ctrl->t = t1;
      ctrl->m = m1;
      ctrl->status = status;
      ctrl->active = active ? 1 : 0;
      ctrl->umax_bits = umax;
      if (active) ctrl->dt = static_cast<double>(dt);
      // the state does not change: its max|u| carries over to the next launch
      if (!active && from_state) atomicMax(&ctrl->step_slot[seq % 3][0][0], umax);
    }
  }
  __device__ __forceinline__ bool go() const { return *s_go != 0; }
  __device__ __forceinline__ T dt() const { return *s_dt; }
};

// One SSP-RK3 step (timestepping.py:312-320) of u^n = p.uk into
// u^(n+1) = p.out (distinct arrays) with the step control of the control
// block, launch s = p.step_seq (mod 6) of a sequence (SeqStep): a step not
// taken copies the state to p.out; the workgroups reduce max|u_int| of the
// new state into step_slot[s % 3]. No launch between the steps: each one starts
// from what the previous one left.
// Without a control block the step is taken with the device scalar p.dt_ptr
// (and p.k1out, p.k2out may receive the stage states).
// Waves per SIMD the compiler keeps the registers of the FP64 WENO-JS5
// whole-step kernels within (5: at most 96 VGPRs, which the LDS of a
// 256-thread tile allows as well); the other variants take what they need.
#ifndef PSA_STEP_F64_WAVES
#define PSA_STEP_F64_WAVES 5
#endif
template <typename T, bool LEARNED, int REC>
constexpr int step_min_waves() {
  return std::is_same<T, double>::value && !LEARNED && REC == 0 ? PSA_STEP_F64_WAVES
                                                                 : 1;
}

template <typename T, int FLUX, int VEL, int BLOCK, int CPT, bool LEARNED,
          int REC = 0>
__global__ __attribute__((amdgpu_flat_work_group_size(1, BLOCK),
                          amdgpu_waves_per_eu((step_min_waves<T, LEARNED, REC>()))))
void k_step(StageArgs<T> p) {
  static_assert(VEL != 2, "no velocity array in the whole-step kernel");
  using SM = StepSmem<T, BLOCK, CPT, FLUX == PSA_FLUX_ESWENO>;
  __shared__ SM sm;
  __shared__ unsigned int q[LEARNED ? (BLOCK / 64) * learned_queue_words<CPT>() : 1];
  const int tid = threadIdx.x;
  psa_control *ctrl = p.ctrl;
  const int64_t A = static_cast<int64_t>(blockIdx.x) * SM::TO;

  if (ctrl == nullptr) {
    // a fixed step of the device scalar dt (the adjoint's forward passes);
    // in a batch, instance blockIdx.y with its own dt
    const int64_t off = static_cast<int64_t>(blockIdx.y) * p.bstride;
    StageArgs<T> pb = p;
    if (p.k1out != nullptr) pb.k1out = p.k1out + off;
    if (p.k2out != nullptr) pb.k2out = p.k2out + off;
    unsigned long long vmax = 0;
    step_tile<T, FLUX, VEL, BLOCK, CPT, LEARNED, REC>(pb, A, p.dt_ptr[blockIdx.y],
                                                      p.uk + off, p.out + off, sm, vmax,
                                                      q, nullptr);
    return;
  }

  __shared__ T s_dt;
  __shared__ int s_go;
  const SeqStep<T> ctl{ctrl, p.step_seq, &s_dt, &s_go};
  unsigned long long vmax = 0;
  if (!step_tile_ctl<T, FLUX, VEL, BLOCK, CPT, LEARNED, REC>(p, A, ctl, p.uk, p.out, sm,
                                                             vmax, q, nullptr)) {
    // the step is not taken: the state moves to the output unchanged
    for (int j = tid; j < SM::TO; j += BLOCK) {
      const int64_t a = A + j;
      if (a < p.m) p.out[a] = p.uk[a];
    }
    return;
  }
  // ---- partial max|u| of the new state
  if (ctrl->speed_from_state != 0) {
    vmax = block_max_bits<BLOCK>(vmax, sm.red);
    if (tid == 0)
      atomicMax(&ctrl->step_slot[p.step_seq % 3][blockIdx.x % PSA_UMAX_SLOTS][0], vmax);
  }
}

template <typename T>
__global__ void k_copy(const T *__restrict__ src, T *__restrict__ dst,
                       int64_t m) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// ring schedule: a periodic state of n <= PSA_RING_MAX_CELLS cells held in
// LDS by one workgroup (thread i owns cell i) for all the steps of a call

constexpr int kRingMax = PSA_RING_MAX_CELLS;

// cell i of a stage output into a ring buffer R (R[k] is cell k - 2 mod n)
template <typename T>
__device__ __forceinline__ void ring_put(T *R, int n, int i, T v) {
  R[i + 2] = v;
  if (i < 2) R[n + 2 + i] = v;
  if (i >= n - 2) R[i - n + 2] = v;
}

// One stage of cell i from ring X into ring Y, with the arithmetic of
// step_stage (one cell per thread): reconstruct_any on the same five-cell
// stencil, the same face fluxes and stage combination, the neighbours' edge
// states (and ES-WENO weights) through LDS with the periodic wrap. Ends with
// Y written and visible to the workgroup.
template <typename T, int FLUX, int REC, int STAGE>
__device__ __forceinline__ T ring_stage(const StageArgs<T> &p, const T *X, T *Y, T lam,
                                        T un, int n, bool live, T *sUR, T *sUL, T *sOM) {
  constexpr bool ES = FLUX == PSA_FLUX_ESWENO;
  const int i = threadIdx.x;
  T w[5];
  T ul[1], ur[1], om[1];
  if (live) {
#pragma unroll
    for (int k = 0; k < 5; ++k) w[k] = X[i + k];  // cells i - 2 .. i + 2
    const MlpSource src{nullptr, nullptr};
    reconstruct_any<T, 1, false, REC>(p, w, ul, ur, src, nullptr, 0, ES ? om : nullptr);
    sUR[i] = ur[0];
    sUL[i] = ul[0];
    if constexpr (ES) sOM[i] = om[0];
  }
  __syncthreads();
  T val = T(0);
  if (live) {
    const int il = i == 0 ? n - 1 : i - 1;
    const int ir = i == n - 1 ? 0 : i + 1;
    T F0, F1;
    if constexpr (ES) {
      F0 = flux_burgers_upwind(sUR[il], ul[0]) +
           es_dissipation_w(sOM[il], om[0], w[1], w[2], p.es_delta);
      F1 = flux_burgers_upwind(ur[0], sUL[ir]) +
           es_dissipation_w(om[0], sOM[ir], w[2], w[3], p.es_delta);
    } else {
      F0 = stage_flux<T, FLUX>(sUR[il], ul[0], w[1], w[2], p.vconst);
      F1 = stage_flux<T, FLUX>(ur[0], sUL[ir], w[2], w[3], p.vconst);
    }
    val = stage_value<T, STAGE>(w[2], un, lam, F1 - F0);
    ring_put(Y, n, i, val);
  }
  __syncthreads();
  return val;
}

// max|u| bit pattern over the live cells (the NaN pattern if any is NaN), in
// every thread; s_red holds one entry per wave
template <typename T>
__device__ __forceinline__ unsigned long long ring_max_bits(T v, bool live,
                                                            unsigned long long *s_red) {
  unsigned long long b = 0;
  if (live) b = isnan(v) ? nan_bits<T>() : abs_bits(v);
  b = wave_max_u64(b);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = b;
  __syncthreads();
  const int nw = static_cast<int>(blockDim.x >> 6);
  b = s_red[0];
  for (int k = 1; k < nw; ++k) b = s_red[k] > b ? s_red[k] : b;
  return b;
}

// Up to nsteps steps of the control block's sequence (the decisions of the
// whole-step launches, control_decide_at, from the state's own max|u|), one
// workgroup of blockDim.x >= n threads, the state in LDS throughout; on exit
// u_out holds the state (ghosts filled periodically after a step, or u_in
// as given when no step was taken) and the control block what the
// whole-step schedule leaves (t, m, status, active, dt, umax_bits and the
// final state's max|u| in word 0 of umax_slot).
template <typename T, int FLUX, int REC>
__global__ __launch_bounds__(kRingMax) void k_ring(StageArgs<T> p, long long nsteps) {
  constexpr bool ES = FLUX == PSA_FLUX_ESWENO;
  __shared__ T ringA[kRingMax + 4];
  __shared__ T ringB[kRingMax + 4];
  __shared__ T sUR[kRingMax];
  __shared__ T sUL[kRingMax];
  __shared__ T sOM[ES ? kRingMax : 1];
  __shared__ unsigned long long s_red[kRingMax / 64];
  psa_control *ctrl = p.ctrl;
  const int i = threadIdx.x;
  const int n = static_cast<int>(p.n);
  const int g = p.g;
  const bool live = i < n;

  T u = live ? p.uk[g + i] : T(0);
  if (live) ring_put(ringA, n, i, u);
  T *cur = ringA;
  T *oth = ringB;

  const bool from_state = ctrl->speed_from_state != 0;
  unsigned long long umax = gather_umax(ctrl);
  double t = ctrl->t;
  long long mstep = ctrl->m;
  int status = ctrl->status;
  bool active = false;
  bool taken = false;
  T dt_last = T(0);
  unsigned long long umax_used = umax;  // max|u| of the last decision
  __syncthreads();
  for (long long s = 0; s < nsteps; ++s) {
    T dt = T(0);
    int st;
    umax_used = umax;
    active = control_decide_at<T>(ctrl, t, mstep, status, umax, dt, st);
    status = st;
    if (!active) break;
    const T lam = stage_lambda<T, FLUX>(p, dt);
    ring_stage<T, FLUX, REC, 1>(p, cur, oth, lam, u, n, live, sUR, sUL, sOM);
    ring_stage<T, FLUX, REC, 2>(p, oth, cur, lam, u, n, live, sUR, sUL, sOM);
    u = ring_stage<T, FLUX, REC, 3>(p, cur, oth, lam, u, n, live, sUR, sUL, sOM);
    T *tmp = cur;
    cur = oth;
    oth = tmp;
    t = static_cast<double>(static_cast<T>(t) + dt);
    mstep += 1;
    dt_last = dt;
    taken = true;
    if (from_state) umax = ring_max_bits(u, live, s_red);
  }

  if (taken) {
    if (live) {
      p.out[g + i] = u;
      if (i >= n - g) p.out[i - (n - g)] = u;  // left ghosts: cells n - g ..
      if (i < g) p.out[n + g + i] = u;          // right ghosts: cells 0 .. g - 1
    }
  } else {
    for (int j = i; j < n + 2 * g; j += static_cast<int>(blockDim.x)) p.out[j] = p.uk[j];
  }
  if (i < 64) {
    if (from_state) ctrl->umax_slot[i][0] = i == 0 ? umax : 0ull;
    if (i == 0) {
      ctrl->t = t;
      ctrl->m = mstep;
      ctrl->status = status;
      ctrl->active = active ? 1 : 0;
      ctrl->umax_bits = umax_used;
      if (taken) ctrl->dt = static_cast<double>(dt_last);
      ctrl->seq_t[0] = t;
      ctrl->seq_m[0] = mstep;
      ctrl->seq_status[0] = status;
    }
  }
}

}  // namespace psa

// ===========================================================================
// host side

using namespace psa;

namespace {

constexpr int kThreads = 256;

inline unsigned int blocks_for(int64_t work, int per_block) {
  return static_cast<unsigned int>((work + per_block - 1) / per_block);
}

int check_dtype(int dtype) {
  if (dtype != PSA_FLOAT64 && dtype != PSA_FLOAT32) {
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "unsupported dtype " + std::to_string(dtype));
  }
  return PSA_OK;
}

int check_grid(int64_t m, int g, int gmin = 3) {
  if (g < gmin) {
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "the reconstruction needs at least " + std::to_string(gmin) +
                         " ghost cells (nghosts >= stencil_width), got " +
                         std::to_string(g));
  }
  if (m - 2 * static_cast<int64_t>(g) < g) {
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "periodic ghost fill needs at least nghosts interior "
                     "cells, got m = " +
                         std::to_string(m));
  }
  return PSA_OK;
}

template <typename T>
int launch_max_abs(const T *u, int64_t m, int g, T *out,
                   unsigned long long *bits, hipStream_t stream) {
  constexpr int B = 256;
  const int64_t work = m - 2 * static_cast<int64_t>(g);
  unsigned int blocks = blocks_for(work, B * 8);
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL((k_max_abs<T, B>), dim3(blocks), dim3(B), 0, stream, u,
                     m, g, out, bits);
  return check_launch("max_abs");
}

}  // namespace

// --- stage dispatch ---------------------------------------------------------

namespace psa {

struct TileConfig {
  int block;
  int cpt;
  bool pipe;  // stages 1-3: the persistent pipelined kernel
  bool step;  // steps: the whole-step kernel where it applies
};

inline bool aligned16(const void *p) {
  return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

// Workgroups the device holds at once for a kernel (CUs x occupancy),
// queried once per kernel and device (the kernel is the template argument,
// so every kernel has a cache of its own).
template <auto Kernel>
unsigned int resident_blocks(int block) {
  constexpr int kMaxDevices = 64;
  static unsigned int cache[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 0;
  if (cache[dev] == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(Kernel), block, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess ||
        per_cu <= 0 || cus <= 0) {
      (void)hipGetLastError();
      return 0;
    }
    cache[dev] = static_cast<unsigned int>(per_cu) * static_cast<unsigned int>(cus);
  }
  return cache[dev];
}

template <typename T, int FLUX, int VEL, int STAGE, int BLOCK, int CPT,
          bool LEARNED = false, int REC = 0>
int launch_stage_cfg(const StageArgs<T> &args, hipStream_t stream) {
  constexpr int TILE = BLOCK * CPT;
  StageArgs<T> a = args;
  a.aligned = aligned16(a.uk) && aligned16(a.out) &&
              (a.un == nullptr || aligned16(a.un)) &&
              (a.vel == nullptr || aligned16(a.vel));
  const unsigned int blocks = blocks_for(a.m, TILE);
  hipLaunchKernelGGL((k_stage<T, FLUX, VEL, STAGE, BLOCK, CPT, LEARNED, REC>),
                     dim3(blocks), dim3(BLOCK), 0, stream, a);
  return check_launch("ssprk33 stage");
}

template <typename T, int FLUX, int VEL, int STAGE, int BLOCK, int CPT>
int launch_stage_pipe_cfg(const StageArgs<T> &args, hipStream_t stream) {
  constexpr int TILE = BLOCK * CPT;
  StageArgs<T> a = args;
  a.aligned = aligned16(a.uk) && aligned16(a.out) &&
              (a.un == nullptr || aligned16(a.un));
  const unsigned int cap = resident_blocks<k_stage_pipe<T, FLUX, VEL, STAGE, BLOCK, CPT>>(BLOCK);
  if (cap == 0) return launch_stage_cfg<T, FLUX, VEL, STAGE, BLOCK, CPT>(args, stream);
  unsigned int blocks = blocks_for(a.m, TILE);
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL((k_stage_pipe<T, FLUX, VEL, STAGE, BLOCK, CPT>), dim3(blocks),
                     dim3(BLOCK), 0, stream, a);
  return check_launch("ssprk33 stage (pipelined)");
}

template <typename T, int FLUX, int VEL, int STAGE>
int launch_stage_tiles(const StageArgs<T> &a, TileConfig cfg,
                       hipStream_t stream) {
  if constexpr (STAGE > 0 && VEL != 2) {
    if (cfg.pipe) {
      if (cfg.block == 256 && cfg.cpt == 4)
        return launch_stage_pipe_cfg<T, FLUX, VEL, STAGE, 256, 4>(a, stream);
      if (cfg.block == 256 && cfg.cpt == 2)
        return launch_stage_pipe_cfg<T, FLUX, VEL, STAGE, 256, 2>(a, stream);
    }
  }
  if (cfg.block == 256 && cfg.cpt == 1)
    return launch_stage_cfg<T, FLUX, VEL, STAGE, 256, 1>(a, stream);
  if (cfg.block == 256 && cfg.cpt == 2)
    return launch_stage_cfg<T, FLUX, VEL, STAGE, 256, 2>(a, stream);
  if (cfg.block == 256 && cfg.cpt == 4)
    return launch_stage_cfg<T, FLUX, VEL, STAGE, 256, 4>(a, stream);
  if (cfg.block == 128 && cfg.cpt == 4)
    return launch_stage_cfg<T, FLUX, VEL, STAGE, 128, 4>(a, stream);
  if (cfg.block == 512 && cfg.cpt == 2)
    return launch_stage_cfg<T, FLUX, VEL, STAGE, 512, 2>(a, stream);
  return set_error(PSA_ERR_INVALID_ARGUMENT,
                   "unsupported tile configuration block=" +
                       std::to_string(cfg.block) +
                       " cells_per_thread=" + std::to_string(cfg.cpt));
}

// Whole-step kernel: Burgers (WENO-JS or learned weights) and advection
// with a constant velocity.
template <typename T, int FLUX, int VEL, int BLOCK, int CPT, bool LEARNED,
          int REC = 0>
int launch_step_cfg(const StageArgs<T> &args, hipStream_t stream) {
  constexpr int TO = StepSmem<T, BLOCK, CPT>::TO;
  StageArgs<T> a = args;
  a.aligned = aligned16(a.uk) && aligned16(a.out) &&
              (a.bstride * static_cast<int64_t>(sizeof(T))) % 16 == 0;
  const unsigned int blocks = blocks_for(a.m, TO);
  const unsigned int batch = a.batch > 1 ? static_cast<unsigned int>(a.batch) : 1u;
  hipLaunchKernelGGL((k_step<T, FLUX, VEL, BLOCK, CPT, LEARNED, REC>), dim3(blocks, batch),
                     dim3(BLOCK), 0, stream, a);
  return check_launch("ssprk33 step");
}

template <typename T, int FLUX, int VEL, bool LEARNED, int REC = 0>
int launch_step_tiles(const StageArgs<T> &a, TileConfig cfg,
                      hipStream_t stream) {
  if (cfg.block == 256 && cfg.cpt == 4)
    return launch_step_cfg<T, FLUX, VEL, 256, 4, LEARNED, REC>(a, stream);
  if (cfg.block == 256 && cfg.cpt == 2)
    return launch_step_cfg<T, FLUX, VEL, 256, 2, LEARNED, REC>(a, stream);
  if constexpr (REC == 0) {
    // one cell per thread (WENO-JS5, learned weights or not): the shortest
    // launch for the smaller grids, whose steps are latency
    if (cfg.block == 256 && cfg.cpt == 1)
      return launch_step_cfg<T, FLUX, VEL, 256, 1, LEARNED, 0>(a, stream);
  }
  return set_error(PSA_ERR_INVALID_ARGUMENT,
                   "unsupported tile configuration for the whole-step kernel block=" +
                       std::to_string(cfg.block) +
                       " cells_per_thread=" + std::to_string(cfg.cpt));
}

// Whether the whole-step kernel runs the steps of this descriptor.
inline bool step_applies(const psa_solver_desc *d, TileConfig cfg, int vel) {
  if (!cfg.step || vel == 2 || d->nu != nullptr) return false;
  // the stages of Lax-Friedrichs need the global max|u| of the stage before
  if (d->flux == PSA_FLUX_LAX_FRIEDRICHS) return false;
  if (d->flux == PSA_FLUX_ADVECTION_UPWIND && d->mlp != nullptr) return false;
  if (cfg.block == 256 && cfg.cpt == 1)
    return d->rec == PSA_REC_WENOJS53 && d->flux != PSA_FLUX_ESWENO;
  return cfg.block == 256 && (cfg.cpt == 4 || cfg.cpt == 2);
}

template <typename T, int REC>
int launch_step_rec(const psa_solver_desc *d, const StageArgs<T> &a, TileConfig cfg,
                    hipStream_t stream) {
  if (d->flux == PSA_FLUX_RUSANOV)
    return launch_step_tiles<T, PSA_FLUX_RUSANOV, 0, false, REC>(a, cfg, stream);
  if (d->flux == PSA_FLUX_BURGERS_UPWIND)
    return launch_step_tiles<T, PSA_FLUX_BURGERS_UPWIND, 0, false, REC>(a, cfg, stream);
  if (d->flux == PSA_FLUX_ENGQUIST_OSHER)
    return launch_step_tiles<T, PSA_FLUX_ENGQUIST_OSHER, 0, false, REC>(a, cfg, stream);
  if constexpr (REC == 1) {
    if (d->flux == PSA_FLUX_ESWENO)
      return launch_step_tiles<T, PSA_FLUX_ESWENO, 0, false, REC>(a, cfg, stream);
  }
  return launch_step_tiles<T, PSA_FLUX_ADVECTION_UPWIND, 1, false, REC>(a, cfg, stream);
}

template <typename T>
int launch_step(const psa_solver_desc *d, int vel, const StageArgs<T> &a,
                TileConfig cfg, hipStream_t stream) {
  if (a.rec3) return launch_step_rec<T, 1>(d, a, cfg, stream);
  if (a.muscl) return launch_step_rec<T, 2>(d, a, cfg, stream);
  if (a.muscls) return launch_step_rec<T, 3>(d, a, cfg, stream);
  if (d->mlp != nullptr) {
    if (d->flux == PSA_FLUX_RUSANOV)
      return launch_step_tiles<T, PSA_FLUX_RUSANOV, 0, true>(a, cfg, stream);
    if (d->flux == PSA_FLUX_ENGQUIST_OSHER)
      return launch_step_tiles<T, PSA_FLUX_ENGQUIST_OSHER, 0, true>(a, cfg, stream);
    return launch_step_tiles<T, PSA_FLUX_BURGERS_UPWIND, 0, true>(a, cfg, stream);
  }
  if (d->flux == PSA_FLUX_RUSANOV)
    return launch_step_tiles<T, PSA_FLUX_RUSANOV, 0, false>(a, cfg, stream);
  if (d->flux == PSA_FLUX_BURGERS_UPWIND)
    return launch_step_tiles<T, PSA_FLUX_BURGERS_UPWIND, 0, false>(a, cfg, stream);
  if (d->flux == PSA_FLUX_ENGQUIST_OSHER)
    return launch_step_tiles<T, PSA_FLUX_ENGQUIST_OSHER, 0, false>(a, cfg, stream);
  (void)vel;
  return launch_step_tiles<T, PSA_FLUX_ADVECTION_UPWIND, 1, false>(a, cfg, stream);
}

// Whether the ring schedule runs the steps of this descriptor (with a
// control block): periodic, at most PSA_RING_MAX_CELLS cells, what the
// whole-step kernel takes without learned weights.
inline bool ring_applies(const psa_solver_desc *d, int vel) {
  if (d->schedule != PSA_SCHEDULE_DEFAULT && d->schedule != PSA_SCHEDULE_RING) return false;
  if (d->bc != PSA_BC_PERIODIC || d->mlp != nullptr || d->nu != nullptr || vel == 2)
    return false;
  if (d->flux == PSA_FLUX_LAX_FRIEDRICHS) return false;
  return d->n >= 8 && d->n <= PSA_RING_MAX_CELLS && d->g >= 2 && d->g <= d->n;
}

template <typename T, int REC>
int launch_ring_rec(const psa_solver_desc *d, const StageArgs<T> &a, long long nsteps,
                    hipStream_t stream) {
  const dim3 block(static_cast<unsigned int>((a.n + 63) / 64 * 64));
  if (d->flux == PSA_FLUX_RUSANOV) {
    hipLaunchKernelGGL((k_ring<T, PSA_FLUX_RUSANOV, REC>), dim3(1), block, 0, stream, a,
                       nsteps);
  } else if (d->flux == PSA_FLUX_BURGERS_UPWIND) {
    hipLaunchKernelGGL((k_ring<T, PSA_FLUX_BURGERS_UPWIND, REC>), dim3(1), block, 0, stream,
                       a, nsteps);
  } else if (d->flux == PSA_FLUX_ENGQUIST_OSHER) {
    hipLaunchKernelGGL((k_ring<T, PSA_FLUX_ENGQUIST_OSHER, REC>), dim3(1), block, 0, stream,
                       a, nsteps);
  } else if (d->flux == PSA_FLUX_ADVECTION_UPWIND) {
    hipLaunchKernelGGL((k_ring<T, PSA_FLUX_ADVECTION_UPWIND, REC>), dim3(1), block, 0,
                       stream, a, nsteps);
  } else if constexpr (REC == 1) {
    if (d->flux != PSA_FLUX_ESWENO)
      return set_error(PSA_ERR_NOT_IMPLEMENTED, "flux not on the ring schedule");
    hipLaunchKernelGGL((k_ring<T, PSA_FLUX_ESWENO, REC>), dim3(1), block, 0, stream, a,
                       nsteps);
  } else {
    return set_error(PSA_ERR_NOT_IMPLEMENTED, "flux not on the ring schedule");
  }
  return check_launch("ring steps");
}

// All nsteps steps of a call on the ring schedule (ring_applies).
template <typename T>
int launch_ring(const psa_solver_desc *d, const StageArgs<T> &a, long long nsteps,
                hipStream_t stream) {
  if (a.rec3) return launch_ring_rec<T, 1>(d, a, nsteps, stream);
  if (a.muscl) return launch_ring_rec<T, 2>(d, a, nsteps, stream);
  if (a.muscls) return launch_ring_rec<T, 3>(d, a, nsteps, stream);
  return launch_ring_rec<T, 0>(d, a, nsteps, stream);
}

// Learned weights: Burgers fluxes and three tile shapes.
template <typename T, int FLUX, int STAGE>
int launch_stage_learned(const StageArgs<T> &a, TileConfig cfg,
                         hipStream_t stream) {
  if (cfg.block == 256 && cfg.cpt == 4)
    return launch_stage_cfg<T, FLUX, 0, STAGE, 256, 4, true>(a, stream);
  if (cfg.block == 256 && cfg.cpt == 2)
    return launch_stage_cfg<T, FLUX, 0, STAGE, 256, 2, true>(a, stream);
  if (cfg.block == 256 && cfg.cpt == 1)
    return launch_stage_cfg<T, FLUX, 0, STAGE, 256, 1, true>(a, stream);
  return set_error(PSA_ERR_INVALID_ARGUMENT,
                   "unsupported tile configuration for learned weights block=" +
                       std::to_string(cfg.block) +
                       " cells_per_thread=" + std::to_string(cfg.cpt));
}

// The other reconstructions (REC 1 third order, REC 2 MUSCL): the stage
// kernels with tiles 256 x 4 or 256 x 2 (the pipelined schedule takes these
// too).
template <typename T, int FLUX, int VEL, int STAGE, int REC>
int launch_stage_tiles_rec(const StageArgs<T> &a, TileConfig cfg,
                           hipStream_t stream) {
  if (cfg.block == 256 && cfg.cpt == 4)
    return launch_stage_cfg<T, FLUX, VEL, STAGE, 256, 4, false, REC>(a, stream);
  if (cfg.block == 256 && cfg.cpt == 2)
    return launch_stage_cfg<T, FLUX, VEL, STAGE, 256, 2, false, REC>(a, stream);
  return set_error(PSA_ERR_INVALID_ARGUMENT,
                   "unsupported tile configuration for the third-order and "
                   "MUSCL reconstructions block=" + std::to_string(cfg.block) +
                       " cells_per_thread=" + std::to_string(cfg.cpt));
}

template <typename T, int STAGE, int REC>
int launch_stage_rec(int flux, int vel, const StageArgs<T> &a, TileConfig cfg,
                     hipStream_t stream) {
  if (flux == PSA_FLUX_RUSANOV)
    return launch_stage_tiles_rec<T, PSA_FLUX_RUSANOV, 0, STAGE, REC>(a, cfg, stream);
  if (flux == PSA_FLUX_BURGERS_UPWIND)
    return launch_stage_tiles_rec<T, PSA_FLUX_BURGERS_UPWIND, 0, STAGE, REC>(a, cfg,
                                                                            stream);
  if (flux == PSA_FLUX_ENGQUIST_OSHER)
    return launch_stage_tiles_rec<T, PSA_FLUX_ENGQUIST_OSHER, 0, STAGE, REC>(a, cfg,
                                                                            stream);
  if (flux == PSA_FLUX_LAX_FRIEDRICHS)
    return launch_stage_tiles_rec<T, PSA_FLUX_LAX_FRIEDRICHS, 0, STAGE, REC>(a, cfg,
                                                                            stream);
  if constexpr (REC == 1) {
    if (flux == PSA_FLUX_ESWENO)
      return launch_stage_tiles_rec<T, PSA_FLUX_ESWENO, 0, STAGE, REC>(a, cfg, stream);
  }
  if (flux == PSA_FLUX_ADVECTION_UPWIND && vel == 1)
    return launch_stage_tiles_rec<T, PSA_FLUX_ADVECTION_UPWIND, 1, STAGE, REC>(
        a, cfg, stream);
  return launch_stage_tiles_rec<T, PSA_FLUX_ADVECTION_UPWIND, 2, STAGE, REC>(a, cfg,
                                                                            stream);
}

template <typename T, int STAGE>
int launch_stage(int flux, int vel, const StageArgs<T> &a, TileConfig cfg,
                 hipStream_t stream) {
  if (a.rec3) return launch_stage_rec<T, STAGE, 1>(flux, vel, a, cfg, stream);
  if (a.muscl) return launch_stage_rec<T, STAGE, 2>(flux, vel, a, cfg, stream);
  if (a.muscls) return launch_stage_rec<T, STAGE, 3>(flux, vel, a, cfg, stream);
  if (a.mlp != nullptr) {
    if (flux == PSA_FLUX_RUSANOV)
      return launch_stage_learned<T, PSA_FLUX_RUSANOV, STAGE>(a, cfg, stream);
    if (flux == PSA_FLUX_BURGERS_UPWIND)
      return launch_stage_learned<T, PSA_FLUX_BURGERS_UPWIND, STAGE>(a, cfg,
                                                                    stream);
    if (flux == PSA_FLUX_ENGQUIST_OSHER)
      return launch_stage_learned<T, PSA_FLUX_ENGQUIST_OSHER, STAGE>(a, cfg, stream);
    if (flux == PSA_FLUX_LAX_FRIEDRICHS)
      return launch_stage_learned<T, PSA_FLUX_LAX_FRIEDRICHS, STAGE>(a, cfg, stream);
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "learned weights are implemented for Burgers only");
  }
  if (flux == PSA_FLUX_RUSANOV)
    return launch_stage_tiles<T, PSA_FLUX_RUSANOV, 0, STAGE>(a, cfg, stream);
  if (flux == PSA_FLUX_BURGERS_UPWIND)
    return launch_stage_tiles<T, PSA_FLUX_BURGERS_UPWIND, 0, STAGE>(a, cfg,
                                                                    stream);
  if (flux == PSA_FLUX_ENGQUIST_OSHER)
    return launch_stage_tiles<T, PSA_FLUX_ENGQUIST_OSHER, 0, STAGE>(a, cfg,
                                                                    stream);
  if (flux == PSA_FLUX_LAX_FRIEDRICHS)
    return launch_stage_tiles<T, PSA_FLUX_LAX_FRIEDRICHS, 0, STAGE>(a, cfg,
                                                                    stream);
  if (flux == PSA_FLUX_ADVECTION_UPWIND) {
    if (vel == 1)
      return launch_stage_tiles<T, PSA_FLUX_ADVECTION_UPWIND, 1, STAGE>(
          a, cfg, stream);
    return launch_stage_tiles<T, PSA_FLUX_ADVECTION_UPWIND, 2, STAGE>(a, cfg,
                                                                      stream);
  }
  return set_error(PSA_ERR_NOT_IMPLEMENTED,
                   "unsupported flux " + std::to_string(flux));
}

TileConfig default_tiles(int dtype) {
  (void)dtype;
  return TileConfig{256, 4, false, true};
}

int check_desc(const psa_solver_desc *d) {
  if (d == nullptr) return set_error(PSA_ERR_INVALID_ARGUMENT, "null desc");
  int st = check_dtype(d->dtype);
  if (st != PSA_OK) return st;
  if (d->n <= 0) return set_error(PSA_ERR_INVALID_ARGUMENT, "n must be > 0");
  if (d->rec < PSA_REC_WENOJS53 || d->rec > PSA_REC_MUSCLS)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "unsupported reconstruction " + std::to_string(d->rec));
  if (d->rec == PSA_REC_MUSCL) {
    st = check_limiter(d->limiter, d->limiter_theta, d->muscl_atol);
    if (st != PSA_OK) return st;
  }
  if (d->rec == PSA_REC_MUSCLS) {
    st = check_slope_limiter(d->limiter);
    if (st != PSA_OK) return st;
  }
  if (d->bc < PSA_BC_PERIODIC || d->bc > PSA_BC_DIRICHLET)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad bc " + std::to_string(d->bc));
  if (d->bc == PSA_BC_NEUMANN && !(std::isfinite(d->bc_left) && std::isfinite(d->bc_right)))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "the Neumann gradients must be finite");
  if (d->bc == PSA_BC_DIRICHLET && d->bc_values == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "a Dirichlet boundary needs its 2 g ghost values (bc_values)");
  st = check_grid(d->n + 2 * static_cast<int64_t>(d->g), d->g,
                  d->rec == PSA_REC_WENOJS53 ? 3 : 2);
  if (st != PSA_OK) return st;
  if (d->rec != PSA_REC_WENOJS53 && d->mlp != nullptr)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "learned weights replace the WENO-JS5 weights only");

  if (!(d->dx > 0.0) || !std::isfinite(d->dx))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "dx must be positive");
  if (!(d->eps >= 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "eps must be >= 0");
  if (d->flux == PSA_FLUX_ADVECTION_UPWIND && !d->velocity_is_const &&
      d->velocity == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "advection needs a velocity array or a constant");
  if (d->flux < 0 || d->flux > PSA_FLUX_ESWENO)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "unsupported flux " + std::to_string(d->flux));
  if (d->schedule < PSA_SCHEDULE_DEFAULT || d->schedule > PSA_SCHEDULE_RING)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "unknown schedule " + std::to_string(d->schedule));
  if (d->flux == PSA_FLUX_ESWENO &&
      (d->rec != PSA_REC_ESWENO32 || !(d->eps > 0.0) || !(d->es_delta >= 0.0)))
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "the ES-WENO flux needs the ES-WENO3 reconstruction, eps > 0 "
                     "and es_delta >= 0");
  if (d->mlp != nullptr &&
      (d->flux == PSA_FLUX_ADVECTION_UPWIND || d->flux == PSA_FLUX_ESWENO))
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "learned weights are implemented for the Burgers Rusanov, "
                     "Godunov, Lax-Friedrichs and Engquist-Osher fluxes");
  if (d->mlp != nullptr && (!(d->mlp_tau >= 0.0) || !(d->mlp_sigma >= 0.0)))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "mlp_tau and mlp_sigma must be >= 0");
  return PSA_OK;
}

TileConfig tiles_of(const psa_solver_desc *d) {
  TileConfig c = default_tiles(d->dtype);
  // below ~2^20 cells a grid of 1024-cell tiles leaves most of the device
  // idle: half-size tiles double the workgroups in flight
  if (d->n < (int64_t(1) << 20)) c.cpt = 2;
  if (d->block > 0) c.block = d->block;
  if (d->cells_per_thread > 0) c.cpt = d->cells_per_thread;
  if (d->schedule == PSA_SCHEDULE_TILE) c.step = false;
  if (d->schedule == PSA_SCHEDULE_PIPELINE) {
    c.pipe = true;
    c.step = false;
  }
  // below 2^17 cells, one cell per thread where the whole-step kernel takes
  // it (WENO-JS5, learned weights or not): the steps there are latency, and
  // the shorter per-thread chain makes them 8-15% shorter
  if (d->n < (int64_t(1) << 17) && d->block <= 0 && d->cells_per_thread <= 0) {
    TileConfig one = c;
    one.cpt = 1;
    const int vel =
        d->flux != PSA_FLUX_ADVECTION_UPWIND ? 0 : (d->velocity_is_const ? 1 : 2);
    if (step_applies(d, one, vel)) c.cpt = 1;
  }
  return c;
}

template <typename T>
StageArgs<T> make_args(const psa_solver_desc *d) {
  StageArgs<T> a{};
  a.m = d->n + 2 * static_cast<int64_t>(d->g);
  a.n = d->n;
  a.g = d->g;
  a.dx = static_cast<T>(d->dx);
  a.inv_dx = static_cast<T>(1.0) / static_cast<T>(d->dx);
  a.eps = static_cast<T>(d->eps);
  a.vconst = static_cast<T>(d->velocity_const);
  a.vel = static_cast<const T *>(d->velocity);
  if (d->flux == PSA_FLUX_LAX_FRIEDRICHS) {
    a.vel = nullptr;
    a.speed = static_cast<const T *>(d->velocity);
  }
  a.nu = static_cast<const T *>(d->nu);
  a.aligned = 1;
  a.es = d->rec == PSA_REC_ESWENO32 ? 1 : 0;
  a.rec3 = (d->rec == PSA_REC_WENOJS32 || d->rec == PSA_REC_ESWENO32) ? 1 : 0;
  a.muscl = d->rec == PSA_REC_MUSCL ? 1 : 0;
  a.muscls = d->rec == PSA_REC_MUSCLS ? 1 : 0;
  a.es_delta = static_cast<T>(d->es_delta);
  a.nowrap = d->bc == PSA_BC_GIVEN ? 1 : 0;
  a.bcm = (d->bc == PSA_BC_NEUMANN || d->bc == PSA_BC_DIRICHLET) ? d->bc : PSA_BC_PERIODIC;
  a.bc_l = static_cast<T>(d->bc_left);
  a.bc_r = static_cast<T>(d->bc_right);
  a.bc_vals = static_cast<const T *>(d->bc_values);
  a.limiter = d->limiter;
  a.lim_theta = static_cast<T>(d->limiter_theta);
  a.lim_atol = static_cast<T>(d->muscl_atol);
  a.mlp = d->mlp;
  a.tau = static_cast<T>(d->mlp_tau);
  a.sigma = static_cast<T>(d->mlp_sigma);
  return a;
}

inline int vel_mode(const psa_solver_desc *d) {
  if (d->flux != PSA_FLUX_ADVECTION_UPWIND) return 0;
  return d->velocity_is_const ? 1 : 2;
}

template <typename T>
int run_steps(const psa_solver_desc *d, const void *u_in, void *u_out,
              void *k1, void *k2, void *control, const void *dt_dev,
              long long nsteps, hipStream_t stream) {
  StageArgs<T> a = make_args<T>(d);
  a.ctrl = static_cast<psa_control *>(control);
  a.dt_ptr = static_cast<const T *>(dt_dev);
  const TileConfig cfg = tiles_of(d);
  const int vel = vel_mode(d);
  const T *un = static_cast<const T *>(u_in);
  T *uo = static_cast<T *>(u_out);
  T *b1 = static_cast<T *>(k1);
  T *b2 = static_cast<T *>(k2);
  int st = PSA_OK;
  if (a.ctrl != nullptr && ring_applies(d, vel)) {
    if (nsteps <= 0) return PSA_OK;
    a.uk = un;
    a.out = uo;
    return launch_ring<T>(d, a, nsteps, stream);
  }
  if (a.ctrl != nullptr && step_applies(d, cfg, vel)) {
    // one launch per step; u^n and u^(n+1) are distinct arrays: the state
    // goes u_in -> k1 -> k2 -> k1 ... -> u_out (u_in -> k1 -> u_out, by a
    // copy, for one in-place step)
    if (nsteps <= 0) return PSA_OK;
    hipLaunchKernelGGL(k_step_prologue, dim3(1), dim3(64), 0, stream, a.ctrl);
    st = check_launch("step prologue");
    if (st != PSA_OK) return st;
    const bool copy_back = nsteps == 1 && un == uo;
    for (long long s = 0; s < nsteps; ++s) {
      a.uk = s == 0 ? un : ((s - 1) % 2 == 0 ? b1 : b2);
      a.out = (s == nsteps - 1 && !copy_back) ? uo : (s % 2 == 0 ? b1 : b2);
      a.un = a.uk;
      a.step_seq = static_cast<int>(s % 6);
      st = launch_step<T>(d, vel, a, cfg, stream);
      if (st != PSA_OK) return st;
    }
    hipLaunchKernelGGL(k_step_epilogue, dim3(1), dim3(64), 0, stream, a.ctrl,
                       static_cast<int>((nsteps - 1) % 3));
    st = check_launch("step epilogue");
    if (st != PSA_OK) return st;
    if (copy_back) {
      hipLaunchKernelGGL(k_copy<T>, dim3(blocks_for(a.m, 256)), dim3(256), 0,
                         stream, b1, uo, a.m);
      st = check_launch("copy");
      if (st != PSA_OK) return st;
    }
    if (a.bcm != PSA_BC_PERIODIC) {
      // the steps fill the ghost cells of every stage input from its interior;
      // the final state gets them too (the arithmetic of the fill kernels)
      const unsigned int blocks = blocks_for(a.m, 256);
      if (a.bcm == PSA_BC_NEUMANN) {
        hipLaunchKernelGGL(k_fill_neumann<T>, dim3(blocks), dim3(256), 0, stream, uo, uo,
                           a.m, a.g, a.dx, a.bc_l, a.bc_r, 3);
      } else {
        hipLaunchKernelGGL(k_fill_dirichlet<T>, dim3(blocks), dim3(256), 0, stream, uo,
                           uo, a.m, a.g, a.bc_vals, a.bc_vals + a.g);
      }
      st = check_launch("boundary fill");
      if (st != PSA_OK) return st;
    }
    return st;
  }
  for (long long s = 0; s < nsteps; ++s) {
    // after the first step the state lives in u_out
    const T *cur = s == 0 ? un : uo;
    a.un = cur;
    a.uk = cur;
    a.out = b1;
    st = launch_stage<T, 1>(d->flux, vel, a, cfg, stream);
    if (st != PSA_OK) return st;
    a.uk = b1;
    a.out = b2;
    st = launch_stage<T, 2>(d->flux, vel, a, cfg, stream);
    if (st != PSA_OK) return st;
    a.uk = b2;
    a.out = uo;
    st = launch_stage<T, 3>(d->flux, vel, a, cfg, stream);
    if (st != PSA_OK) return st;
    if (s == 0 && un != uo && a.ctrl != nullptr) {
      // the first step reads u_in and writes u_out; if it was not taken,
      // u_out must still receive the state
      const unsigned int blocks = blocks_for(a.m, 256);
      hipLaunchKernelGGL(k_copy_if_inactive<T>, dim3(blocks), dim3(256), 0,
                         stream, un, uo, a.m, a.ctrl);
      st = check_launch("copy_if_inactive");
      if (st != PSA_OK) return st;
    }
  }
  return PSA_OK;
}

// One fixed-dt whole step u_in -> u_out (distinct arrays) with the device
// scalar dt, writing the stage states into k1, k2 if they are not null, for
// the adjoint's forward and recompute passes. Returns
// PSA_ERR_NOT_IMPLEMENTED (and launches nothing) where the whole-step kernel
// does not apply; the caller then runs the stage kernels.
int step_fixed_dt(const psa_solver_desc *d, const void *u_in, void *u_out, void *k1,
                  void *k2, const void *dt, hipStream_t stream, long long batch,
                  int64_t ld) {
  TileConfig cfg = tiles_of(d);
  // a batch fills the device with its instances: there the one-cell tiles of
  // the small single grids give way to two cells per thread (the learned
  // step 6% faster in batches of 64, the plain one the same)
  if (batch > 1 && cfg.cpt == 1 && d->cells_per_thread <= 0) cfg.cpt = 2;
  const int vel = vel_mode(d);
  if (!step_applies(d, cfg, vel)) return PSA_ERR_NOT_IMPLEMENTED;
  if (batch < 1 || batch > 65535)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "batch must be in [1, 65535]");
  if (d->dtype == PSA_FLOAT64) {
    StageArgs<double> a = make_args<double>(d);
    a.uk = a.un = static_cast<const double *>(u_in);
    a.out = static_cast<double *>(u_out);
    a.k1out = static_cast<double *>(k1);
    a.k2out = static_cast<double *>(k2);
    a.dt_ptr = static_cast<const double *>(dt);
    a.bstride = ld;
    a.batch = static_cast<int>(batch);
    return launch_step<double>(d, vel, a, cfg, stream);
  }
  StageArgs<float> a = make_args<float>(d);
  a.uk = a.un = static_cast<const float *>(u_in);
  a.out = static_cast<float *>(u_out);
  a.k1out = static_cast<float *>(k1);
  a.k2out = static_cast<float *>(k2);
  a.dt_ptr = static_cast<const float *>(dt);
  a.bstride = ld;
  a.batch = static_cast<int>(batch);
  return launch_step<float>(d, vel, a, cfg, stream);
}

}  // namespace psa

namespace {

int weno_fields(int dtype, const void *u, void *beta, void *uhat,
                       void *omega, int64_t m, double eps, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  if (u == nullptr || m <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null input or m <= 0");
  auto s = static_cast<hipStream_t>(stream);
  const unsigned int blocks = blocks_for(m, kThreads);
  if (dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_weno_fields<double>, dim3(blocks), dim3(kThreads), 0,
                       s, static_cast<const double *>(u), m, eps,
                       static_cast<double *>(beta), static_cast<double *>(uhat),
                       static_cast<double *>(omega));
  } else {
    hipLaunchKernelGGL(k_weno_fields<float>, dim3(blocks), dim3(kThreads), 0,
                       s, static_cast<const float *>(u), m,
                       static_cast<float>(eps), static_cast<float *>(beta),
                       static_cast<float *>(uhat), static_cast<float *>(omega));
  }
  return check_launch("weno_fields");
}

template <typename T, int REC>
void launch_numerical_flux(int flux, const T *uu, const T *vv, const T *nn, T *ff,
                           int64_t m, T e, hipStream_t s,
                           LimParams<T> lim = LimParams<T>{}) {
  const unsigned int blocks = blocks_for(m + 1, kThreads);
  if (flux == PSA_FLUX_RUSANOV) {
    hipLaunchKernelGGL((k_numerical_flux<T, PSA_FLUX_RUSANOV, REC>), dim3(blocks),
                       dim3(kThreads), 0, s, uu, vv, nn, m, e, ff, lim);
  } else if (flux == PSA_FLUX_BURGERS_UPWIND) {
    hipLaunchKernelGGL((k_numerical_flux<T, PSA_FLUX_BURGERS_UPWIND, REC>),
                       dim3(blocks), dim3(kThreads), 0, s, uu, vv, nn, m, e, ff, lim);
  } else if (flux == PSA_FLUX_ENGQUIST_OSHER) {
    hipLaunchKernelGGL((k_numerical_flux<T, PSA_FLUX_ENGQUIST_OSHER, REC>),
                       dim3(blocks), dim3(kThreads), 0, s, uu, vv, nn, m, e, ff, lim);
  } else if (flux == PSA_FLUX_LAX_FRIEDRICHS) {
    hipLaunchKernelGGL((k_numerical_flux<T, PSA_FLUX_LAX_FRIEDRICHS, REC>),
                       dim3(blocks), dim3(kThreads), 0, s, uu, vv, nn, m, e, ff, lim);
  } else {
    hipLaunchKernelGGL((k_numerical_flux<T, PSA_FLUX_ADVECTION_UPWIND, REC>),
                       dim3(blocks), dim3(kThreads), 0, s, uu, vv, nn, m, e, ff, lim);
  }
}

template <typename T>
int numerical_flux_t(int flux, int rec, const void *u, const void *velocity,
                     const void *nu, void *fnum, int64_t m, double eps,
                     hipStream_t s) {
  const T *uu = static_cast<const T *>(u);
  const T *vv = static_cast<const T *>(velocity);
  const T *nn = static_cast<const T *>(nu);
  T *ff = static_cast<T *>(fnum);
  const T e = static_cast<T>(eps);
  if (rec == PSA_REC_WENOJS32) {
    launch_numerical_flux<T, PSA_REC_WENOJS32>(flux, uu, vv, nn, ff, m, e, s);
  } else if (rec == PSA_REC_ESWENO32) {
    launch_numerical_flux<T, PSA_REC_ESWENO32>(flux, uu, vv, nn, ff, m, e, s);
  } else {
    launch_numerical_flux<T, PSA_REC_WENOJS53>(flux, uu, vv, nn, ff, m, e, s);
  }
  return check_launch("numerical_flux");
}

}  // namespace

// ===========================================================================
// C ABI

extern "C" {

int psa_version(void) { return 1; }

int psa_control_bytes(void) { return static_cast<int>(sizeof(psa_control)); }

int psa_get_last_error(char *buf, size_t n) {
  if (buf == nullptr || n == 0) return PSA_ERR_INVALID_ARGUMENT;
  const std::string &e = psa::g_last_error;
  const size_t k = e.size() < n - 1 ? e.size() : n - 1;
  std::memcpy(buf, e.data(), k);
  buf[k] = '\0';
  return PSA_OK;
}

int psa_device_count(int *count) {
  if (count == nullptr) return set_error(PSA_ERR_INVALID_ARGUMENT, "null");
  hipError_t err = hipGetDeviceCount(count);
  if (err != hipSuccess) {
    *count = 0;
    return set_error(PSA_ERR_HIP, hipGetErrorString(err));
  }
  return PSA_OK;
}

int psa_weno_smoothness(int dtype, const void *u, void *beta, int64_t m,
                        void *stream) {
  if (beta == nullptr) return set_error(PSA_ERR_INVALID_ARGUMENT, "null beta");
  return weno_fields(dtype, u, beta, nullptr, nullptr, m, 0.0, stream);
}

int psa_weno_interp(int dtype, const void *u, void *uhat, int64_t m,
                    void *stream) {
  if (uhat == nullptr) return set_error(PSA_ERR_INVALID_ARGUMENT, "null uhat");
  return weno_fields(dtype, u, nullptr, uhat, nullptr, m, 0.0, stream);
}

int psa_weno_js_weights(int dtype, const void *u, void *omega, int64_t m,
                        double eps, void *stream) {
  if (omega == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null omega");
  return weno_fields(dtype, u, nullptr, nullptr, omega, m, eps, stream);
}

int psa_reconstruct(int dtype, int rec, const void *f, void *ul, void *ur,
                    int64_t m, double eps, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  if (rec < PSA_REC_WENOJS53 || rec > PSA_REC_ESWENO32)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "unsupported reconstruction " + std::to_string(rec));
  if (f == nullptr || ul == nullptr || ur == nullptr || m <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer or m <= 0");
  auto s = static_cast<hipStream_t>(stream);
  const unsigned int blocks = blocks_for(m, kThreads);
  auto launch = [&](auto tag, auto rtag) {
    using T = decltype(tag);
    constexpr int R = decltype(rtag)::value;
    hipLaunchKernelGGL((k_weno_reconstruct<T, R>), dim3(blocks), dim3(kThreads), 0,
                       s, static_cast<const T *>(f), m, static_cast<T>(eps),
                       static_cast<T *>(ul), static_cast<T *>(ur));
  };
  using R0 = std::integral_constant<int, PSA_REC_WENOJS53>;
  using R1 = std::integral_constant<int, PSA_REC_WENOJS32>;
  using R2 = std::integral_constant<int, PSA_REC_ESWENO32>;
  if (dtype == PSA_FLOAT64) {
    if (rec == PSA_REC_WENOJS53) launch(0.0, R0{});
    if (rec == PSA_REC_WENOJS32) launch(0.0, R1{});
    if (rec == PSA_REC_ESWENO32) launch(0.0, R2{});
  } else {
    if (rec == PSA_REC_WENOJS53) launch(0.0f, R0{});
    if (rec == PSA_REC_WENOJS32) launch(0.0f, R1{});
    if (rec == PSA_REC_ESWENO32) launch(0.0f, R2{});
  }
  return check_launch("reconstruct");
}

int psa_weno_reconstruct(int dtype, const void *f, void *ul, void *ur,
                         int64_t m, double eps, void *stream) {
  return psa_reconstruct(dtype, PSA_REC_WENOJS53, f, ul, ur, m, eps, stream);
}

int psa_muscl_reconstruct(int dtype, int limiter, double theta, double atol,
                          const void *f, void *ul, void *ur, int64_t m,
                          void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  st = check_limiter(limiter, theta, atol);
  if (st != PSA_OK) return st;
  if (f == nullptr || ul == nullptr || ur == nullptr || m <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer or m <= 0");
  auto s = static_cast<hipStream_t>(stream);
  const unsigned int blocks = blocks_for(m, kThreads);
  if (dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_muscl_reconstruct<double>, dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const double *>(f), m,
                       LimParams<double>{limiter, theta, atol, 0.0},
                       static_cast<double *>(ul), static_cast<double *>(ur));
  } else {
    hipLaunchKernelGGL(k_muscl_reconstruct<float>, dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const float *>(f), m,
                       LimParams<float>{limiter, static_cast<float>(theta),
                                        static_cast<float>(atol), 0.0f},
                       static_cast<float *>(ul), static_cast<float *>(ur));
  }
  return check_launch("muscl_reconstruct");
}

int psa_muscls_reconstruct(int dtype, int limiter, const void *f, void *ul,
                           void *ur, int64_t m, double dx, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  st = check_slope_limiter(limiter);
  if (st != PSA_OK) return st;
  if (f == nullptr || ul == nullptr || ur == nullptr || m <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer or m <= 0");
  if (!(dx > 0.0)) return set_error(PSA_ERR_INVALID_ARGUMENT, "dx must be positive");
  auto s = static_cast<hipStream_t>(stream);
  const unsigned int blocks = blocks_for(m, kThreads);
  if (dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_muscls_reconstruct<double>, dim3(blocks), dim3(kThreads), 0,
                       s, static_cast<const double *>(f), m,
                       LimParams<double>{limiter, 0.0, 0.0, dx},
                       static_cast<double *>(ul), static_cast<double *>(ur));
  } else {
    hipLaunchKernelGGL(k_muscls_reconstruct<float>, dim3(blocks), dim3(kThreads), 0,
                       s, static_cast<const float *>(f), m,
                       LimParams<float>{limiter, 0.0f, 0.0f, static_cast<float>(dx)},
                       static_cast<float *>(ul), static_cast<float *>(ur));
  }
  return check_launch("muscls_reconstruct");
}

int psa_numerical_flux_desc(const psa_solver_desc *desc, const void *u,
                            const void *speed, void *fnum, void *stream) {
  int st = check_desc(desc);
  if (st != PSA_OK) return st;
  if (u == nullptr || fnum == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (desc->flux == PSA_FLUX_ADVECTION_UPWIND && desc->velocity == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "advection upwind flux needs a velocity array");
  if (desc->flux == PSA_FLUX_LAX_FRIEDRICHS && speed == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "the Lax-Friedrichs flux needs the device scalar max|a|");
  if (desc->mlp != nullptr)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "learned weights have psa_learned_reconstruct");
  if (desc->bc == PSA_BC_NEUMANN || desc->bc == PSA_BC_DIRICHLET)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the face flux takes the ghost cells as given (PSA_BC_GIVEN after a "
                     "fill)");
  auto s = static_cast<hipStream_t>(stream);
  const int64_t m = desc->n + 2 * static_cast<int64_t>(desc->g);
  const void *vel = desc->flux == PSA_FLUX_LAX_FRIEDRICHS ? speed : desc->velocity;
  if (desc->rec != PSA_REC_MUSCL && desc->rec != PSA_REC_MUSCLS)
    return psa_numerical_flux_rec(desc->dtype, desc->flux, desc->rec, u, vel,
                                  desc->nu, fnum, m, desc->eps, stream);
  auto launch = [&](auto tag) {
    using T = decltype(tag);
    const LimParams<T> lim{desc->limiter, static_cast<T>(desc->limiter_theta),
                           static_cast<T>(desc->muscl_atol), static_cast<T>(desc->dx)};
    const T *uu = static_cast<const T *>(u);
    const T *vv = static_cast<const T *>(vel);
    const T *nn = static_cast<const T *>(desc->nu);
    T *ff = static_cast<T *>(fnum);
    if (desc->rec == PSA_REC_MUSCL) {
      launch_numerical_flux<T, PSA_REC_MUSCL>(desc->flux, uu, vv, nn, ff, m,
                                              static_cast<T>(desc->eps), s, lim);
    } else {
      launch_numerical_flux<T, PSA_REC_MUSCLS>(desc->flux, uu, vv, nn, ff, m,
                                               static_cast<T>(desc->eps), s, lim);
    }
  };
  if (desc->dtype == PSA_FLOAT64) {
    launch(0.0);
  } else {
    launch(0.0f);
  }
  return check_launch("numerical_flux");
}

int psa_apply_boundary_periodic(int dtype, const void *u, void *out,
                                int64_t m, int g, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  if (g < 0 || m - 2 * static_cast<int64_t>(g) < g || m <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad grid size");
  if (u == nullptr || out == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  auto s = static_cast<hipStream_t>(stream);
  const int64_t n = m - 2 * static_cast<int64_t>(g);
  const unsigned int blocks = blocks_for(m, kThreads);
  if (dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_apply_boundary<double>, dim3(blocks), dim3(kThreads),
                       0, s, static_cast<const double *>(u),
                       static_cast<double *>(out), m, n, g);
  } else {
    hipLaunchKernelGGL(k_apply_boundary<float>, dim3(blocks), dim3(kThreads),
                       0, s, static_cast<const float *>(u),
                       static_cast<float *>(out), m, n, g);
  }
  return check_launch("apply_boundary");
}

int psa_esweno_dissipation(int dtype, const void *u, void *fnum, int64_t m,
                           double eps, double delta, int weight, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  if (u == nullptr || fnum == nullptr || m < 2)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer or m < 2");
  if (weight != 0 && weight != 1)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "weight must be 0 or 1");
  if (!(eps > 0.0) || !(delta >= 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "eps must be > 0 and delta >= 0");
  auto s = static_cast<hipStream_t>(stream);
  const unsigned int blocks = blocks_for(m - 1, kThreads);
  auto launch = [&](auto tag) {
    using T = decltype(tag);
    const T *uu = static_cast<const T *>(u);
    T *ff = static_cast<T *>(fnum);
    if (weight == 0) {
      hipLaunchKernelGGL((k_esweno_dissipation<T, 0>), dim3(blocks), dim3(kThreads), 0, s,
                         uu, ff, m, static_cast<T>(eps), static_cast<T>(delta));
    } else {
      hipLaunchKernelGGL((k_esweno_dissipation<T, 1>), dim3(blocks), dim3(kThreads), 0, s,
                         uu, ff, m, static_cast<T>(eps), static_cast<T>(delta));
    }
  };
  if (dtype == PSA_FLOAT64) {
    launch(0.0);
  } else {
    launch(0.0f);
  }
  return check_launch("esweno_dissipation");
}

int psa_esweno_advection_flux(int dtype, const void *u, const void *velocity,
                              const void *amax, void *fnum, int64_t m, double eps,
                              double delta, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  if (u == nullptr || velocity == nullptr || amax == nullptr || fnum == nullptr || m < 2)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer or m < 2");
  if (!(eps > 0.0) || !(delta >= 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "eps must be > 0 and delta >= 0");
  auto s = static_cast<hipStream_t>(stream);
  const unsigned int blocks = blocks_for(m + 1, kThreads);
  if (dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_esweno_advection_flux<double>, dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const double *>(u), static_cast<const double *>(velocity),
                       static_cast<const double *>(amax), m, eps, delta,
                       static_cast<double *>(fnum));
  } else {
    hipLaunchKernelGGL(k_esweno_advection_flux<float>, dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const float *>(u), static_cast<const float *>(velocity),
                       static_cast<const float *>(amax), m, static_cast<float>(eps),
                       static_cast<float>(delta), static_cast<float *>(fnum));
  }
  return check_launch("esweno_advection_flux");
}

int psa_flux_divergence(int dtype, const void *fnum, const void *velocity, void *out,
                        int64_t m, double dx, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  if (fnum == nullptr || out == nullptr || m <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer or m <= 0");
  if (!(dx > 0.0)) return set_error(PSA_ERR_INVALID_ARGUMENT, "dx must be positive");
  auto s = static_cast<hipStream_t>(stream);
  const unsigned int blocks = blocks_for(m, kThreads);
  if (dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_flux_divergence<double>, dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const double *>(fnum),
                       static_cast<const double *>(velocity), static_cast<double *>(out),
                       m, dx);
  } else {
    hipLaunchKernelGGL(k_flux_divergence<float>, dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const float *>(fnum), static_cast<const float *>(velocity),
                       static_cast<float *>(out), m, static_cast<float>(dx));
  }
  return check_launch("flux_divergence");
}

int psa_apply_boundary_dirichlet(int dtype, const void *u, void *out,
                                 int64_t m, int g, const void *left,
                                 const void *right, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  if (g < 0 || m - 2 * static_cast<int64_t>(g) < g || m <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad grid size");
  if (u == nullptr || out == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  auto s = static_cast<hipStream_t>(stream);
  const unsigned int blocks = blocks_for(m, kThreads);
  if (dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_fill_dirichlet<double>, dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const double *>(u), static_cast<double *>(out), m,
                       g, static_cast<const double *>(left),
                       static_cast<const double *>(right));
  } else {
    hipLaunchKernelGGL(k_fill_dirichlet<float>, dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const float *>(u), static_cast<float *>(out), m, g,
                       static_cast<const float *>(left),
                       static_cast<const float *>(right));
  }
  return check_launch("apply_boundary_dirichlet");
}

int psa_apply_boundary_neumann(int dtype, const void *u, void *out, int64_t m,
                               int g, double dx, double gl, double gr,
                               int sides, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  if (g < 0 || m - 2 * static_cast<int64_t>(g) < g || m <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad grid size");
  if (u == nullptr || out == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (sides < 0 || sides > 3)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "sides must be 0 .. 3");
  auto s = static_cast<hipStream_t>(stream);
  const unsigned int blocks = blocks_for(m, kThreads);
  if (dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_fill_neumann<double>, dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const double *>(u), static_cast<double *>(out), m,
                       g, dx, gl, gr, sides);
  } else {
    hipLaunchKernelGGL(k_fill_neumann<float>, dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const float *>(u), static_cast<float *>(out), m, g,
                       static_cast<float>(dx), static_cast<float>(gl),
                       static_cast<float>(gr), sides);
  }
  return check_launch("apply_boundary_neumann");
}

int psa_numerical_flux(int dtype, int flux, const void *u,
                       const void *velocity, const void *nu, void *fnum,
                       int64_t m, double eps, void *stream) {
  return psa_numerical_flux_rec(dtype, flux, PSA_REC_WENOJS53, u, velocity, nu,
                                fnum, m, eps, stream);
}

int psa_numerical_flux_rec(int dtype, int flux, int rec, const void *u,
                           const void *velocity, const void *nu, void *fnum,
                           int64_t m, double eps, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  if (rec < PSA_REC_WENOJS53 || rec > PSA_REC_ESWENO32)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "unsupported reconstruction " + std::to_string(rec));
  if (flux < 0 || flux > 4)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "unsupported flux " + std::to_string(flux));
  if (u == nullptr || fnum == nullptr || m < 2)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer or m < 2");
  if (flux == PSA_FLUX_ADVECTION_UPWIND && velocity == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "advection upwind flux needs a velocity array");
  if (flux == PSA_FLUX_LAX_FRIEDRICHS && velocity == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "the Lax-Friedrichs flux needs the device scalar max|a|");
  auto s = static_cast<hipStream_t>(stream);
  if (dtype == PSA_FLOAT64)
    return numerical_flux_t<double>(flux, rec, u, velocity, nu, fnum, m, eps, s);
  return numerical_flux_t<float>(flux, rec, u, velocity, nu, fnum, m, eps, s);
}

int psa_apply_operator(int dtype, int flux, const void *u,
                       const void *velocity, const void *nu, void *out,
                       int64_t m, int g, double dx, double eps, void *stream) {
  return psa_apply_operator_rec(dtype, flux, PSA_REC_WENOJS53, u, velocity, nu,
                                out, m, g, dx, eps, stream);
}

int psa_apply_operator_rec(int dtype, int flux, int rec, const void *u,
                           const void *velocity, const void *nu, void *out,
                           int64_t m, int g, double dx, double eps,
                           void *stream) {
  if (rec == PSA_REC_MUSCL || rec == PSA_REC_MUSCLS)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "MUSCL takes its limiter from a descriptor: use "
                     "psa_ssprk33_stage with stage 0");
  if (flux == PSA_FLUX_ESWENO)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "the ES-WENO flux takes es_delta from a descriptor: use "
                     "psa_ssprk33_stage with stage 0");
  psa_solver_desc d{};
  d.dtype = dtype;
  d.flux = flux;
  d.rec = rec;
  d.g = g;
  d.n = m - 2 * static_cast<int64_t>(g);
  d.dx = dx;
  d.eps = eps;
  d.velocity = velocity;
  d.velocity_is_const = 0;
  d.nu = nu;
  int st = check_desc(&d);
  if (st != PSA_OK) return st;
  if (u == nullptr || out == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (u == out)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "apply_operator cannot write into its input");
  if (flux == PSA_FLUX_LAX_FRIEDRICHS && velocity == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "the Lax-Friedrichs operator needs the device scalar max|u|");
  auto s = static_cast<hipStream_t>(stream);
  const TileConfig cfg = default_tiles(dtype);
  const int vel = vel_mode(&d);
  if (dtype == PSA_FLOAT64) {
    StageArgs<double> a = make_args<double>(&d);
    a.uk = static_cast<const double *>(u);
    a.out = static_cast<double *>(out);
    return launch_stage<double, 0>(flux, vel, a, cfg, s);
  }
  StageArgs<float> a = make_args<float>(&d);
  a.uk = static_cast<const float *>(u);
  a.out = static_cast<float *>(out);
  return launch_stage<float, 0>(flux, vel, a, cfg, s);
}

int psa_max_abs_interior(int dtype, const void *u, void *out, int64_t m,
                         int g, void *stream) {
  int st = check_dtype(dtype);
  if (st != PSA_OK) return st;
  if (u == nullptr || out == nullptr || g < 0 || m - 2 * (int64_t)g <= 0)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad arguments");
  auto s = static_cast<hipStream_t>(stream);
  if (dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_store_zero<double>, dim3(1), dim3(64), 0, s,
                       static_cast<double *>(out));
    st = check_launch("max_abs init");
    if (st != PSA_OK) return st;
    return launch_max_abs<double>(static_cast<const double *>(u), m, g,
                                  static_cast<double *>(out), nullptr, s);
  }
  hipLaunchKernelGGL(k_store_zero<float>, dim3(1), dim3(64), 0, s,
                     static_cast<float *>(out));
  st = check_launch("max_abs init");
  if (st != PSA_OK) return st;
  return launch_max_abs<float>(static_cast<const float *>(u), m, g,
                               static_cast<float *>(out), nullptr, s);
}

int psa_control_init(const psa_solver_desc *desc, const void *u,
                     void *control, double tstart, double tfinal,
                     long long maxit, int dt_mode, double theta,
                     double cfl_num, double fixed_dt, void *stream) {
  int st = check_desc(desc);
  if (st != PSA_OK) return st;
  if (control == nullptr || u == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (dt_mode != PSA_DT_CFL && dt_mode != PSA_DT_FIXED)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad dt mode");
  auto s = static_cast<hipStream_t>(stream);
  const bool advection = desc->flux == PSA_FLUX_ADVECTION_UPWIND;
  // Lax-Friedrichs reads max|u^n| as the flux speed of stage 1 also with a
  // fixed dt
  const bool lf = desc->flux == PSA_FLUX_LAX_FRIEDRICHS;
  const int speed_from_state =
      ((dt_mode == PSA_DT_CFL && !advection) || lf) ? 1 : 0;
  auto *ctrl = static_cast<psa_control *>(control);
  // a constant advection velocity without an array: max|a| = |a|
  const bool const_speed = advection && desc->velocity == nullptr;
  const double speed = const_speed ? desc->velocity_const : 0.0;
  if (desc->dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_control_init<double>, dim3(1), dim3(64), 0, s, ctrl,
                       tstart, tfinal, theta, cfl_num, fixed_dt, maxit,
                       dt_mode, speed_from_state, speed);
  } else {
    hipLaunchKernelGGL(k_control_init<float>, dim3(1), dim3(64), 0, s, ctrl,
                       tstart, tfinal, theta, cfl_num, fixed_dt, maxit,
                       dt_mode, speed_from_state, speed);
  }
  st = check_launch("control_init");
  if (st != PSA_OK) return st;
  if ((dt_mode != PSA_DT_CFL && !lf) || const_speed) return PSA_OK;

  const int64_t m = desc->n + 2 * static_cast<int64_t>(desc->g);
  const void *src = advection ? desc->velocity : u;
  if (desc->dtype == PSA_FLOAT64)
    return launch_max_abs<double>(static_cast<const double *>(src), m,
                                  desc->g, nullptr, &ctrl->umax_slot[0][0], s);
  return launch_max_abs<float>(static_cast<const float *>(src), m, desc->g,
                               nullptr, &ctrl->umax_slot[0][0], s);
}

int psa_ssprk33_steps(const psa_solver_desc *desc, const void *u_in,
                      void *u_out, void *k1, void *k2, void *control,
                      long long nsteps, void *stream) {
  int st = check_desc(desc);
  if (st != PSA_OK) return st;
  if (u_in == nullptr || u_out == nullptr || k1 == nullptr || k2 == nullptr ||
      control == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (k1 == k2 || k1 == u_in || k2 == u_in || k1 == u_out || k2 == u_out)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "stage buffers must be distinct from the state");
  if (desc->bc == PSA_BC_GIVEN)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "PSA_BC_GIVEN is for the operator (stage 0); the steps fill "
                     "periodic, Neumann or Dirichlet ghost cells themselves");
  if ((desc->bc == PSA_BC_NEUMANN || desc->bc == PSA_BC_DIRICHLET) &&
      !step_applies(desc, tiles_of(desc), vel_mode(desc)))
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "Neumann and Dirichlet fills run in the whole-step kernel only "
                     "(schedule default or step, no velocity array or Rusanov nu, not "
                     "Lax-Friedrichs, tiles 256 x 4 or 256 x 2)");
  if (desc->schedule == PSA_SCHEDULE_RING && !ring_applies(desc, vel_mode(desc)))
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the ring schedule takes periodic states of 8 to "
                     "PSA_RING_MAX_CELLS cells, no learned weights, velocity array, "
                     "Rusanov nu or Lax-Friedrichs");
  if (nsteps < 0) return set_error(PSA_ERR_INVALID_ARGUMENT, "nsteps < 0");
  if (nsteps == 0) return PSA_OK;
  auto s = static_cast<hipStream_t>(stream);
  if (desc->dtype == PSA_FLOAT64)
    return run_steps<double>(desc, u_in, u_out, k1, k2, control, nullptr,
                             nsteps, s);
  return run_steps<float>(desc, u_in, u_out, k1, k2, control, nullptr, nsteps,
                          s);
}

int psa_ssprk33_stage(const psa_solver_desc *desc, int stage, const void *un,
                      const void *uk, void *out, void *control,
                      const void *dt, void *stream) {
  int st = check_desc(desc);
  if (st != PSA_OK) return st;
  if (stage < 0 || stage > 3)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "stage must be 0 (operator), 1, 2 or 3");
  if (uk == nullptr || out == nullptr || (stage > 1 && un == nullptr))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null state pointer");
  if (stage > 0 && control == nullptr && dt == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "need a control block or dt");
  if ((stage > 0 && desc->bc != PSA_BC_PERIODIC) || desc->bc == PSA_BC_NEUMANN ||
      desc->bc == PSA_BC_DIRICHLET)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the stage kernels are periodic (the operator also takes "
                     "PSA_BC_GIVEN); Neumann and Dirichlet fills run in "
                     "psa_ssprk33_steps' whole-step kernel");
  if (out == uk)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "a stage cannot write into its stencil input");
  if (desc->flux == PSA_FLUX_LAX_FRIEDRICHS &&
      (stage == 0 || control == nullptr) && desc->velocity == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "Lax-Friedrichs without a control block needs the device "
                     "scalar max|uk| in desc->velocity");
  auto s = static_cast<hipStream_t>(stream);
  const TileConfig cfg = tiles_of(desc);
  const int vel = vel_mode(desc);
  if (desc->dtype == PSA_FLOAT64) {
    StageArgs<double> a = make_args<double>(desc);
    a.un = static_cast<const double *>(stage == 1 ? uk : un);
    a.uk = static_cast<const double *>(uk);
    a.out = static_cast<double *>(out);
    a.ctrl = static_cast<psa_control *>(control);
    a.dt_ptr = static_cast<const double *>(dt);
    if (stage == 0) {
      a.ctrl = nullptr;
      return launch_stage<double, 0>(desc->flux, vel, a, cfg, s);
    }
    if (stage == 1) return launch_stage<double, 1>(desc->flux, vel, a, cfg, s);
    if (stage == 2) return launch_stage<double, 2>(desc->flux, vel, a, cfg, s);
    return launch_stage<double, 3>(desc->flux, vel, a, cfg, s);
  }
  StageArgs<float> a = make_args<float>(desc);
  a.un = static_cast<const float *>(stage == 1 ? uk : un);
  a.uk = static_cast<const float *>(uk);
  a.out = static_cast<float *>(out);
  a.ctrl = static_cast<psa_control *>(control);
  a.dt_ptr = static_cast<const float *>(dt);
  if (stage == 0) {
    a.ctrl = nullptr;
    return launch_stage<float, 0>(desc->flux, vel, a, cfg, s);
  }
  if (stage == 1) return launch_stage<float, 1>(desc->flux, vel, a, cfg, s);
  if (stage == 2) return launch_stage<float, 2>(desc->flux, vel, a, cfg, s);
  return launch_stage<float, 3>(desc->flux, vel, a, cfg, s);
}

int psa_ssprk33_advance(const psa_solver_desc *desc, const void *u,
                        void *out, void *k1, void *k2, const void *dt,
                        void *stream) {
  int st = check_desc(desc);
  if (st != PSA_OK) return st;
  if (u == nullptr || out == nullptr || k1 == nullptr || k2 == nullptr ||
      dt == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (k1 == k2 || k1 == u || k2 == u || k1 == out || k2 == out)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "stage buffers must be distinct from the state");
  if (desc->flux == PSA_FLUX_LAX_FRIEDRICHS)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "Lax-Friedrichs steps need a control block for the stage "
                     "maxima (psa_ssprk33_steps)");
  if (desc->bc != PSA_BC_PERIODIC)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the fused SSP-RK3 stages are periodic (bc = PSA_BC_PERIODIC)");
  auto s = static_cast<hipStream_t>(stream);
  if (desc->dtype == PSA_FLOAT64)
    return run_steps<double>(desc, u, out, k1, k2, nullptr, dt, 1, s);
  return run_steps<float>(desc, u, out, k1, k2, nullptr, dt, 1, s);
}

}  // extern "C"
