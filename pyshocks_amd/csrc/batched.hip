// Batched solver: many independent periodic Burgers instances of uneven size
// advanced to their final times in one persistent launch.
//
// Each workgroup takes the next instance from a queue (an atomic counter over
// a caller-provided order, longest first) and runs its whole time loop: the
// CFL time step of timestepping.py:89-152 with predict_timestep of
// burgers/schemes.py:42-49,121-127 (Rusanov with alpha = 1 or Godunov), and
// the three SSP-RK3 stages of timestepping.py:312-320, each a sweep over the
// instance's tiles with the same tile computation as the single-instance
// stage kernels, separated by workgroup barriers. max|u| of each new state is
// reduced inside the workgroup, so no step needs a grid-wide
// synchronisation. A workgroup exits when the queue is empty; an instance
// stops at its final time, at max_steps (status PSA_ERR_INVALID_ARGUMENT) or
// at a non-finite time step or one too small to advance t
// (PSA_ERR_NONFINITE_DT), so every workgroup terminates.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/pyshocks_amd.h"
#include "psa_internal.hpp"
#include "stage_device.hpp"

namespace psa {
namespace {

template <typename T>
struct BatchArgs {
  T *u;
  T *k1;
  T *k2;
  const long long *offsets;
  const int *ncells;
  const double *dx;
  const double *tfinal;
  const int *order;
  long long ninst;
  unsigned int *counter;
  double *t_out;
  long long *steps_out;
  int *status_out;
  long long max_steps;
  T eps;
  T theta;
  int g;
  const float *mlp;
  T tau;
};

template <int BLOCK>
__device__ __forceinline__ unsigned long long block_max_bcast(
    unsigned long long v, unsigned long long *s_red,
    unsigned long long *s_out) {
  v = block_max_bits<BLOCK>(v, s_red);
  if (threadIdx.x == 0) *s_out = v;
  __syncthreads();
  return *s_out;
}

template <typename T, int FLUX, int BLOCK, int CPT, bool LEARNED>
__global__ __launch_bounds__(BLOCK) void k_batch(BatchArgs<T> a) {
  constexpr int TILE = BLOCK * CPT;
  __shared__ StageSmem<T, 0, BLOCK, CPT> sm;
  __shared__ unsigned long long s_bcast;
  __shared__ long long s_next;
  const int tid = threadIdx.x;
  MlpWave W;
  if constexpr (LEARNED) mlp_load(W, a.mlp);

  for (;;) {
    __syncthreads();  // previous instance done with the LDS and s_next
    if (tid == 0) s_next = static_cast<long long>(atomicAdd(a.counter, 1u));
    __syncthreads();
    const long long k = s_next;
    if (k >= a.ninst) break;
    const int i = a.order[k];

    StageArgs<T> p{};
    p.n = a.ncells[i];
    p.g = a.g;
    p.m = p.n + 2 * static_cast<int64_t>(a.g);
    p.dx = static_cast<T>(a.dx[i]);
    p.inv_dx = T(1) / p.dx;
    p.eps = a.eps;
    p.mlp = a.mlp;
    p.tau = a.tau;
    p.aligned = 1;  // instance offsets are 16-byte aligned
    T *u = a.u + a.offsets[i];
    T *k1 = a.k1 + a.offsets[i];
    T *k2 = a.k2 + a.offsets[i];
    const int64_t ntiles = (p.m + TILE - 1) / TILE;

    // max |u_int| of the initial state
    unsigned long long vb = 0;
    for (int64_t j = a.g + tid; j < p.m - a.g; j += BLOCK) {
      const unsigned long long b = abs_bits(u[j]);
      vb = b > vb ? b : vb;
    }
    unsigned long long umax = block_max_bcast<BLOCK>(vb, sm.red, &s_bcast);

    const T tf = static_cast<T>(a.tfinal[i]);
    const T cfl = T(0.5) * p.dx;  // 0.5 dx^(2 - alpha), alpha = 1
    T t = T(0);
    long long steps = 0;
    int status = PSA_OK;
    for (;;) {
      if (t >= tf) break;
      if (steps >= a.max_steps) {
        status = PSA_ERR_INVALID_ARGUMENT;
        break;
      }
      T dt = a.theta * (cfl / from_bits<T>(umax));
      if (!isinf(tf)) {
        const T dt_min = tf - t;
        dt = (dt < dt_min ? dt : dt_min) + static_cast<T>(1.0e-15);
      }
      // a step that does not advance t (dt absorbed by t) would never end
      if (!isfinite(dt) || !(t + dt > t)) {
        status = PSA_ERR_NONFINITE_DT;
        break;
      }

      unsigned long long vdummy = 0;
      p.un = u;
      p.uk = u;
      p.out = k1;
      for (int64_t tile = 0; tile < ntiles; ++tile) {
        __syncthreads();
        stage_tile<T, FLUX, 0, 1, BLOCK, CPT, LEARNED>(p, tile * TILE, dt, sm,
                                                          vdummy, W);
      }
      __syncthreads();  // k1 complete
      p.uk = k1;
      p.out = k2;
      for (int64_t tile = 0; tile < ntiles; ++tile) {
        __syncthreads();
        stage_tile<T, FLUX, 0, 2, BLOCK, CPT, LEARNED>(p, tile * TILE, dt, sm,
                                                          vdummy, W);
      }
      __syncthreads();  // k2 complete
      p.uk = k2;
      p.out = u;
      unsigned long long vmax = 0;
      for (int64_t tile = 0; tile < ntiles; ++tile) {
        __syncthreads();
        stage_tile<T, FLUX, 0, 3, BLOCK, CPT, LEARNED>(p, tile * TILE, dt, sm,
                                                       vmax, W);
      }
      __syncthreads();  // u complete
      umax = block_max_bcast<BLOCK>(vmax, sm.red, &s_bcast);
      t = t + dt;
      steps += 1;
    }
    if (tid == 0) {
      a.t_out[i] = static_cast<double>(t);
      a.steps_out[i] = steps;
      a.status_out[i] = status;
    }
  }
}

__global__ void k_batch_reset(unsigned int *counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

// Workgroups the device holds at once for a kernel (CUs x occupancy).
template <typename K>
unsigned int resident(K kernel, int block) {
  int dev = 0, per_cu = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void *>(kernel), block, 0) !=
          hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess ||
      per_cu <= 0 || cus <= 0) {
    (void)hipGetLastError();
    return 256;
  }
  return static_cast<unsigned int>(per_cu) * static_cast<unsigned int>(cus);
}

template <typename T, int FLUX, int BLOCK, int CPT, bool LEARNED = false>
int launch_batch_cfg(const BatchArgs<T> &a, int grid, hipStream_t s) {
  unsigned int blocks =
      grid > 0 ? static_cast<unsigned int>(grid)
               : resident(k_batch<T, FLUX, BLOCK, CPT, LEARNED>, BLOCK);
  if (static_cast<long long>(blocks) > a.ninst)
    blocks = static_cast<unsigned int>(a.ninst);
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.counter);
  int st = check_launch("batch reset");
  if (st != PSA_OK) return st;
  hipLaunchKernelGGL((k_batch<T, FLUX, BLOCK, CPT, LEARNED>), dim3(blocks),
                     dim3(BLOCK), 0, s, a);
  return check_launch("batch solve");
}

template <typename T, int FLUX>
int launch_batch_tiles(const BatchArgs<T> &a, int block, int cpt, int grid,
                       hipStream_t s) {
  if (a.mlp != nullptr) {
    if (block == 256 && cpt == 4)
      return launch_batch_cfg<T, FLUX, 256, 4, true>(a, grid, s);
    if (block == 512 && cpt == 2)
      return launch_batch_cfg<T, FLUX, 512, 2, true>(a, grid, s);
    if (block == 1024 && cpt == 2)
      return launch_batch_cfg<T, FLUX, 1024, 2, true>(a, grid, s);
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "unsupported batch tile configuration for learned "
                     "weights block=" + std::to_string(block) +
                         " cells_per_thread=" + std::to_string(cpt));
  }
  if (block == 256 && cpt == 4) return launch_batch_cfg<T, FLUX, 256, 4>(a, grid, s);
  if (block == 256 && cpt == 2) return launch_batch_cfg<T, FLUX, 256, 2>(a, grid, s);
  if (block == 512 && cpt == 2) return launch_batch_cfg<T, FLUX, 512, 2>(a, grid, s);
  if (block == 512 && cpt == 4) return launch_batch_cfg<T, FLUX, 512, 4>(a, grid, s);
  if (block == 1024 && cpt == 2) return launch_batch_cfg<T, FLUX, 1024, 2>(a, grid, s);
  return set_error(PSA_ERR_INVALID_ARGUMENT,
                   "unsupported batch tile configuration block=" +
                       std::to_string(block) + " cells_per_thread=" +
                       std::to_string(cpt));
}

template <typename T>
int launch_batch(const psa_batch_desc *d, void *u, void *k1, void *k2,
                 unsigned int *counter, double *t_out, long long *steps_out,
                 int *status_out, hipStream_t s) {
  BatchArgs<T> a{};
  a.u = static_cast<T *>(u);
  a.k1 = static_cast<T *>(k1);
  a.k2 = static_cast<T *>(k2);
  a.offsets = d->offsets;
  a.ncells = d->ncells;
  a.dx = d->dx;
  a.tfinal = d->tfinal;
  a.order = d->order;
  a.ninst = d->ninst;
  a.counter = counter;
  a.t_out = t_out;
  a.steps_out = steps_out;
  a.status_out = status_out;
  a.max_steps = d->max_steps;
  a.eps = static_cast<T>(d->eps);
  a.theta = static_cast<T>(d->theta);
  a.g = d->g;
  a.mlp = d->mlp;
  a.tau = static_cast<T>(d->mlp_tau);
  const int block = d->block > 0 ? d->block : 256;
  const int cpt = d->cells_per_thread > 0 ? d->cells_per_thread : 4;
  if (d->flux == PSA_FLUX_RUSANOV)
    return launch_batch_tiles<T, PSA_FLUX_RUSANOV>(a, block, cpt, d->grid, s);
  return launch_batch_tiles<T, PSA_FLUX_BURGERS_UPWIND>(a, block, cpt, d->grid,
                                                        s);
}

}  // namespace
}  // namespace psa

using namespace psa;

extern "C" int psa_batch_solve(const psa_batch_desc *desc, void *u, void *k1,
                               void *k2, unsigned int *counter, double *t_out,
                               long long *steps_out, int *status_out,
                               void *stream) {
  if (desc == nullptr) return set_error(PSA_ERR_INVALID_ARGUMENT, "null desc");
  if (desc->dtype != PSA_FLOAT64 && desc->dtype != PSA_FLOAT32)
    return set_error(PSA_ERR_NOT_IMPLEMENTED, "unsupported dtype");
  if (desc->flux != PSA_FLUX_RUSANOV && desc->flux != PSA_FLUX_BURGERS_UPWIND)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the batched solver takes Burgers Rusanov or Godunov");
  if (desc->g < 3)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "needs at least 3 ghost cells");
  if (desc->ninst < 0 || desc->max_steps < 0 || !(desc->eps >= 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad batch parameters");
  if (desc->mlp != nullptr && !(desc->mlp_tau >= 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "mlp_tau must be >= 0");
  if (u == nullptr || k1 == nullptr || k2 == nullptr || counter == nullptr ||
      t_out == nullptr || steps_out == nullptr || status_out == nullptr ||
      desc->offsets == nullptr || desc->ncells == nullptr ||
      desc->dx == nullptr || desc->tfinal == nullptr || desc->order == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (desc->ninst == 0) return PSA_OK;
  auto s = static_cast<hipStream_t>(stream);
  if (desc->dtype == PSA_FLOAT64)
    return launch_batch<double>(desc, u, k1, k2, counter, t_out, steps_out,
                                status_out, s);
  return launch_batch<float>(desc, u, k1, k2, counter, t_out, steps_out,
                             status_out, s);
}
