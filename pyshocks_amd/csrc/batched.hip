// Batched solver: many independent periodic Burgers instances of uneven size
// advanced to their final times in one persistent launch.
//
// Work is a queue of slots (an atomic counter over a caller-provided list,
// longest instances first). A slot is one member of an instance's team: an
// instance of team size 1 is run by one workgroup alone, a large instance by
// a team of workgroups that split its tiles and meet at a global barrier
// after every stage. Every member runs the instance's whole time loop: the
// CFL time step of timestepping.py:89-152 with predict_timestep of
// burgers/schemes.py:42-49,121-127 (Rusanov with alpha = 1 or Godunov), and
// the three SSP-RK3 stages of timestepping.py:312-320, each a sweep over its
// tiles with the same tile computation as the single-instance stage kernels.
// max|u| of each new state is reduced in the workgroup and, for a team,
// through an atomic in the team's state block; all members take the same dt
// from the same reduced value, so they stop at the same step.
//
// Termination: the grid never exceeds the workgroups the device holds at
// once, and teams never exceed the grid, so every slot is eventually taken
// (slots are taken in order, so only the most recent team can be missing
// members, and the workgroups of all other teams finish and take them). A
// member that waits at a barrier longer than the wait limit marks the team
// aborted (status PSA_ERR_HIP) and every member leaves. An instance stops at
// its final time, at max_steps (status PSA_ERR_INVALID_ARGUMENT) or at a
// non-finite time step or one too small to advance t (PSA_ERR_NONFINITE_DT).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <type_traits>

#include "../../include/pyshocks_amd.h"
#include "psa_internal.hpp"
#include "stage_device.hpp"

namespace psa {
namespace {

// Synchronisation state of one team (PSA_TEAM_STATE_BYTES per instance).
struct TeamState {
  unsigned int arrive;  // barrier arrivals, monotone
  int abort;
  unsigned int pad0[2];
  // max|u| bits: stage sweeps alternate between the first two (cleared
  // after the first barrier of a step); whole-step sweeps rotate over all
  // three (one barrier per step)
  unsigned long long umax[3];
  // Lax-Friedrichs: max|k1| and max|k2| of the step, the flux speeds of
  // stages 2 and 3 (each cleared after the barrier that follows its reads)
  unsigned long long speed[2];
  unsigned long long pad1[1];
};
static_assert(sizeof(TeamState) == PSA_TEAM_STATE_BYTES, "team state size");

template <typename T>
struct BatchArgs {
  T *u;
  T *k1;
  T *k2;
  const long long *offsets;
  const int *ncells;
  const double *dx;
  const double *tfinal;
  const int *order;       // slot -> instance (nslots entries)
  const int *slot_rank;   // slot -> member rank, or null (all ranks 0)
  const int *team;        // instance -> team size, or null (all 1)
  TeamState *teams;       // instance -> team state (needed if team != null)
  long long nslots;
  unsigned int *counter;
  double *t_out;
  long long *steps_out;
  int *status_out;
  long long max_steps;
  long long wait_ticks;   // barrier wait limit in wall-clock ticks
  T eps;
  T theta;
  double theta64;
  int g;
  int time_f64;
  const float *mlp;
  T tau;
  T sigma;
  int rec;       // PSA_REC_* (ES-WENO3 and WENO-JS3 share the kernels)
  int limiter;   // PSA_LIMITER_* of the MUSCL reconstructions
  T lim_theta;
  T lim_atol;
  // host only: when set, the dispatch fills in the selected kernel's
  // attributes (and, if resident is set too, the workgroups the device holds
  // at once) instead of launching it
  hipFuncAttributes *query;
  unsigned int *resident;
};

// A value every lane of the wave holds (read from LDS after a broadcast)
// moved to scalar registers, so that what is derived from it (instance
// sizes, pointers, time control) stays out of the vector registers.
__device__ __forceinline__ int wave_uniform(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ unsigned long long wave_uniform(unsigned long long v) {
  const unsigned int lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned int>(v));
  const unsigned int hi =
      __builtin_amdgcn_readfirstlane(static_cast<unsigned int>(v >> 32));
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

template <int BLOCK>
__device__ __forceinline__ unsigned long long block_max_bcast(
    unsigned long long v, unsigned long long *s_red,
    unsigned long long *s_out) {
  v = block_max_bits<BLOCK>(v, s_red);
  if (threadIdx.x == 0) *s_out = v;
  __syncthreads();
  return *s_out;
}

// Barrier of the `size` members of a team: arrivals count up monotonically,
// so the k-th barrier (k = 1, 2, ...) waits for size * k arrivals. Writes of
// the workgroup before the barrier are released at device scope and the
// other members' writes acquired after it. Returns false if the team was
// aborted (by a member that waited too long).
__device__ __forceinline__ bool team_barrier(TeamState *ts, int size,
                                             unsigned int k, long long wait_ticks,
                                             int *s_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(&ts->arrive, 1u, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int target = static_cast<unsigned int>(size) * k;
    const long long start = wall_clock64();
    int ok = 1;
    for (;;) {
      const unsigned int got = __hip_atomic_load(&ts->arrive, __ATOMIC_ACQUIRE,
                                                 __HIP_MEMORY_SCOPE_AGENT);
      if (got >= target) break;
      if (__hip_atomic_load(&ts->abort, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT) != 0) {
        ok = 0;
        break;
      }
      if (wall_clock64() - start > wait_ticks) {
        __hip_atomic_store(&ts->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *s_flag = ok;
  }
  __syncthreads();
  return wave_uniform(*s_flag) != 0;
}

// WHOLE: one sweep per step with the three stages of each tile in LDS
// (step_tile), the state alternating between u and k1; otherwise one sweep
// per stage (stage_tile).
// REC: the reconstruction class of the fused kernels (0 WENO-JS5 or learned,
// 1 third order, 2 MUSCL, 3 slope-limited MUSCL).
template <typename T, int FLUX, int BLOCK, int CPT, bool LEARNED, bool WHOLE, int REC>
__global__ __launch_bounds__(BLOCK) void k_batch(BatchArgs<T> a) {
  using Smem = typename std::conditional<WHOLE,
                                         StepSmem<T, BLOCK, CPT, FLUX == PSA_FLUX_ESWENO>,
                                         StageSmem<T, 0, BLOCK, CPT>>::type;
  constexpr int TILE = WHOLE ? StepSmem<T, BLOCK, CPT>::TO : BLOCK * CPT;
  __shared__ Smem sm;
  __shared__ unsigned long long s_bcast;
  __shared__ long long s_next;
  __shared__ int s_flag;
  __shared__ unsigned int q[LEARNED ? (BLOCK / 64) * learned_queue_words<CPT>() : 1];
  const int tid = threadIdx.x;
  // a persistent workgroup keeps the network's operand registers in LDS
  __shared__ MlpLdsSlot<LEARNED> wl;
  const MlpWaveLds *wlp = nullptr;
  if constexpr (LEARNED) {
    if (tid < 64) {
      MlpWave W;
      mlp_load(W, a.mlp);
      mlp_store_lds(W, wl.w);
    }
    __syncthreads();
    wlp = &wl.w;
  }

  for (;;) {
    __syncthreads();  // previous slot done with the LDS and s_next
    if (tid == 0) s_next = static_cast<long long>(atomicAdd(a.counter, 1u));
    __syncthreads();
    const long long k = wave_uniform(static_cast<int>(s_next));
    if (k >= a.nslots) break;
    const int i = a.order[k];
    const int rank = a.slot_rank != nullptr ? a.slot_rank[k] : 0;
    const int size = a.team != nullptr ? a.team[i] : 1;
    TeamState *ts = size > 1 ? a.teams + i : nullptr;

    StageArgs<T> p{};
    p.n = a.ncells[i];
    p.g = a.g;
    p.m = p.n + 2 * static_cast<int64_t>(a.g);
    p.dx = static_cast<T>(a.dx[i]);
    p.inv_dx = T(1) / p.dx;
    p.eps = a.eps;
    p.mlp = a.mlp;
    p.tau = a.tau;
    p.sigma = a.sigma;
    p.es = a.rec == PSA_REC_ESWENO32 ? 1 : 0;
    p.limiter = a.limiter;
    p.lim_theta = a.lim_theta;
    p.lim_atol = a.lim_atol;
    p.aligned = 1;  // instance offsets are 16-byte aligned
    T *u = a.u + a.offsets[i];
    T *k1 = a.k1 + a.offsets[i];
    T *k2 = a.k2 + a.offsets[i];
    const int64_t ntiles = (p.m + TILE - 1) / TILE;
    // this member's tiles
    const int64_t tlo = ntiles * rank / size;
    const int64_t thi = ntiles * (rank + 1) / size;
    unsigned int nbar = 0;
    bool alive = true;

    // team reduction of max|u| bits into buffer b (a single workgroup keeps
    // the block maximum)
    // (b = 0 .. 2: umax[b]; 3, 4: speed[b - 3])
    auto team_max = [&](unsigned long long v, int b) -> unsigned long long {
      v = wave_uniform(block_max_bcast<BLOCK>(v, sm.red, &s_bcast));
      if (ts == nullptr) return v;
      unsigned long long *slot = b < 3 ? &ts->umax[b] : &ts->speed[b - 3];
      if (tid == 0 && v != 0)
        __hip_atomic_fetch_max(slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      alive = alive && team_barrier(ts, size, ++nbar, a.wait_ticks, &s_flag);
      if (tid == 0)
        s_bcast = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      return wave_uniform(s_bcast);
    };
    auto barrier = [&]() {
      if (ts == nullptr) {
        __syncthreads();
      } else {
        alive = alive && team_barrier(ts, size, ++nbar, a.wait_ticks, &s_flag);
      }
    };

    // max |u_int| of the initial state over this member's cells
    unsigned long long vb = 0;
    {
      int64_t lo = tlo * TILE, hi = thi * TILE;
      lo = lo > a.g ? lo : a.g;
      hi = hi < p.m - a.g ? hi : p.m - a.g;
      for (int64_t j = lo + tid; j < hi; j += BLOCK) {
        const unsigned long long b = abs_bits(u[j]);
        vb = b > vb ? b : vb;
      }
    }
    unsigned long long umax = team_max(vb, WHOLE ? 2 : 1);

    // time control: in the state dtype (timestepping.py:136-152), or with
    // time_f64 in FP64 (t and the unclipped dt; the stages use dt rounded to
    // the state dtype and t advances by exactly that value)
    const double tf64 = a.tfinal[i];
    const double cfl64 = 0.5 * a.dx[i];  // 0.5 dx^(2 - alpha), alpha = 1
    const T tf = static_cast<T>(tf64);
    const T cfl = T(0.5) * p.dx;
    T t = T(0);
    double t64 = 0.0;
    long long steps = 0;
    int status = alive ? PSA_OK : PSA_ERR_HIP;
    while (alive) {
      T dt;
      // time_f64: the clipped last step advances t by the FP64 step (t
      // reaches tfinal + 1e-15 as in timestepping.py:140-142; the rounding of
      // that step to the state dtype must not leave t just short of tfinal)
      double dt_t64 = 0.0;
      if (a.time_f64) {
        if (t64 >= tf64) break;
      } else {
        if (t >= tf) break;
      }
      if (steps >= a.max_steps) {
        status = PSA_ERR_INVALID_ARGUMENT;
        break;
      }
      if (a.time_f64) {
        double dt64 = a.theta64 * (cfl64 / static_cast<double>(from_bits<T>(umax)));
        bool clipped = false;
        if (!isinf(tf64)) {
          const double dt_min = tf64 - t64;
          clipped = !(dt64 < dt_min);
          dt64 = (dt64 < dt_min ? dt64 : dt_min) + 1.0e-15;
        }
        dt = static_cast<T>(dt64);
        dt_t64 = clipped ? dt64 : static_cast<double>(dt);
        if (!isfinite(dt) || !(dt > T(0)) || !(t64 + static_cast<double>(dt) > t64)) {
          status = PSA_ERR_NONFINITE_DT;
          break;
        }
      } else {
        dt = a.theta * (cfl / from_bits<T>(umax));
        if (!isinf(tf)) {
          const T dt_min = tf - t;
          dt = (dt < dt_min ? dt : dt_min) + static_cast<T>(1.0e-15);
        }
        // a step that does not advance t (dt absorbed by t) would never end
        if (!isfinite(dt) || !(t + dt > t)) {
          status = PSA_ERR_NONFINITE_DT;
          break;
        }
      }

      if constexpr (WHOLE) {
        // every member has read the max|u| of two steps ago (before the
        // barrier of the previous step): clear it for the next step
        if (ts != nullptr && rank == 0 && tid == 0)
          __hip_atomic_store(&ts->umax[(steps + 1) % 3], 0ull, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        const T *src = (steps & 1) ? k1 : u;
        T *dst = (steps & 1) ? u : k1;
        unsigned long long vmax = 0;
        for (int64_t tile = tlo; tile < thi; ++tile) {
          __syncthreads();
          step_tile<T, FLUX, 0, BLOCK, CPT, LEARNED, REC>(p, tile * TILE, dt, src, dst,
                                                          sm, vmax, q, wlp);
        }
        umax = team_max(vmax, static_cast<int>(steps % 3));  // u^(n+1) complete
      } else {
        // Lax-Friedrichs (scalar.py:276-277): the flux speed of each stage is
        // max|u| of its input over the instance (its ghosts are images of
        // interior cells): u^n's from the step control, k1's and k2's from
        // the maxima the stages before reduce over the team
        constexpr bool kLF = FLUX == PSA_FLUX_LAX_FRIEDRICHS;
        unsigned long long v1 = 0, v2 = 0;
        p.un = u;
        p.uk = u;
        p.out = k1;
        if constexpr (kLF) p.vconst = from_bits<T>(umax);
        for (int64_t tile = tlo; tile < thi; ++tile) {
          __syncthreads();
          stage_tile<T, FLUX, 0, 1, BLOCK, CPT, LEARNED, false, false, REC>(
              p, tile * TILE, dt, sm, v1, q, wlp);
        }
        if constexpr (kLF) {
          p.vconst = from_bits<T>(team_max(v1, 3));  // k1 complete
        } else {
          barrier();  // k1 complete
        }
        // every member has read the max|u| reduced two steps ago: clear that
        // buffer for the next step
        if (ts != nullptr && rank == 0 && tid == 0)
          __hip_atomic_store(&ts->umax[(steps + 1) & 1], 0ull, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        p.uk = k1;
        p.out = k2;
        for (int64_t tile = tlo; tile < thi; ++tile) {
          __syncthreads();
          stage_tile<T, FLUX, 0, 2, BLOCK, CPT, LEARNED, false, false, REC>(
              p, tile * TILE, dt, sm, v2, q, wlp);
        }
        if constexpr (kLF) {
          p.vconst = from_bits<T>(team_max(v2, 4));  // k2 complete
          // every member has read max|k1| (after the barrier before this
          // one); the next writes come after the step's last barrier
          if (ts != nullptr && rank == 0 && tid == 0)
            __hip_atomic_store(&ts->speed[0], 0ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
          barrier();  // k2 complete
        }
        p.uk = k2;
        p.out = u;
        unsigned long long vmax = 0;
        for (int64_t tile = tlo; tile < thi; ++tile) {
          __syncthreads();
          stage_tile<T, FLUX, 0, 3, BLOCK, CPT, LEARNED, false, false, REC>(
              p, tile * TILE, dt, sm, vmax, q, wlp);
        }
        umax = team_max(vmax, static_cast<int>(steps & 1));  // u complete
        // every member has read max|k2| (after the barrier before this one);
        // the next writes come after the next step's first barrier
        if constexpr (kLF) {
          if (ts != nullptr && rank == 0 && tid == 0)
            __hip_atomic_store(&ts->speed[1], 0ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      t = t + dt;
      t64 = t64 + dt_t64;
      steps += 1;
    }
    if (!alive) status = PSA_ERR_HIP;
    if constexpr (WHOLE) {
      if (steps & 1) {
        // the last state is in k1: this member's cells back to u
        int64_t lo = tlo * TILE, hi = thi * TILE;
        hi = hi < p.m ? hi : p.m;
        for (int64_t j = lo + tid; j < hi; j += BLOCK) u[j] = k1[j];
      }
    }
    if (tid == 0 && rank == 0) {
      a.t_out[i] = a.time_f64 ? t64 : static_cast<double>(t);
      a.steps_out[i] = steps;
      a.status_out[i] = status;
    }
  }
}

// Clears the slot counter and the team states (a kernel, so that the launch
// can be captured in a graph).
__global__ void k_batch_reset(unsigned int *counter, TeamState *teams,
                              long long nteams) {
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx == 0) *counter = 0u;
  for (long long j = idx; j < nteams; j += static_cast<long long>(gridDim.x) * blockDim.x) {
    teams[j].arrive = 0u;
    teams[j].abort = 0;
    teams[j].umax[0] = 0ull;
    teams[j].umax[1] = 0ull;
    teams[j].umax[2] = 0ull;
    teams[j].speed[0] = 0ull;
    teams[j].speed[1] = 0ull;
  }
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
    return 0;
  }
  return static_cast<unsigned int>(per_cu) * static_cast<unsigned int>(cus);
}

template <typename T, int FLUX, int BLOCK, int CPT, bool LEARNED, bool WHOLE, int REC = 0>
int launch_batch_cfg(const BatchArgs<T> &a, int grid, int max_team, long long nteams,
                     hipStream_t s) {
  if (a.query != nullptr) {
    const hipError_t err = hipFuncGetAttributes(
        a.query,
        reinterpret_cast<const void *>(k_batch<T, FLUX, BLOCK, CPT, LEARNED, WHOLE, REC>));
    if (err != hipSuccess)
      return set_error(PSA_ERR_HIP, std::string("hipFuncGetAttributes: ") +
                                        hipGetErrorString(err));
    if (a.resident != nullptr) {
      *a.resident = resident(k_batch<T, FLUX, BLOCK, CPT, LEARNED, WHOLE, REC>, BLOCK);
      if (*a.resident == 0)
        return set_error(PSA_ERR_HIP, "cannot query the resident workgroup count");
    }
    return PSA_OK;
  }
  const unsigned int cap =
      resident(k_batch<T, FLUX, BLOCK, CPT, LEARNED, WHOLE, REC>, BLOCK);
  if (cap == 0)
    return set_error(PSA_ERR_HIP, "cannot query the resident workgroup count");
  unsigned int blocks = grid > 0 ? static_cast<unsigned int>(grid) : cap;
  if (blocks > cap && max_team > 1)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "teams need a grid the device holds at once (<= " +
                         std::to_string(cap) + " workgroups)");
  if (static_cast<long long>(blocks) > a.nslots)
    blocks = static_cast<unsigned int>(a.nslots);
  if (blocks == 0) blocks = 1;
  if (max_team > static_cast<int>(blocks))
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "a team of " + std::to_string(max_team) +
                         " members does not fit the grid of " +
                         std::to_string(blocks) + " workgroups");
  const long long nt = a.teams != nullptr ? nteams : 0;
  unsigned int rb = static_cast<unsigned int>((nt + 255) / 256);
  if (rb < 1) rb = 1;
  if (rb > 1024) rb = 1024;
  hipLaunchKernelGGL(k_batch_reset, dim3(rb), dim3(256), 0, s, a.counter, a.teams, nt);
  int st = check_launch("batch reset");
  if (st != PSA_OK) return st;
  hipLaunchKernelGGL((k_batch<T, FLUX, BLOCK, CPT, LEARNED, WHOLE, REC>), dim3(blocks),
                     dim3(BLOCK), 0, s, a);
  return check_launch("batch solve");
}

// The other reconstructions take tiles 256 x 4 or 256 x 2, as the fused
// single-instance kernels.
template <typename T, int FLUX, bool WHOLE, int REC>
int launch_batch_rec(const BatchArgs<T> &a, int block, int cpt, int grid,
                     int max_team, long long nteams, hipStream_t s) {
  if (block == 256 && cpt == 4)
    return launch_batch_cfg<T, FLUX, 256, 4, false, WHOLE, REC>(a, grid, max_team, nteams, s);
  if (block == 256 && cpt == 2)
    return launch_batch_cfg<T, FLUX, 256, 2, false, WHOLE, REC>(a, grid, max_team, nteams, s);
  return set_error(PSA_ERR_INVALID_ARGUMENT,
                   "the third-order and MUSCL reconstructions take tiles 256 x 4 or "
                   "256 x 2, not block=" + std::to_string(block) +
                       " cells_per_thread=" + std::to_string(cpt));
}

template <typename T, int FLUX, bool WHOLE>
int launch_batch_tiles(const BatchArgs<T> &a, int block, int cpt, int grid,
                       int max_team, long long nteams, hipStream_t s) {
  if (a.rec == PSA_REC_WENOJS32 || a.rec == PSA_REC_ESWENO32)
    return launch_batch_rec<T, FLUX, WHOLE, 1>(a, block, cpt, grid, max_team, nteams, s);
  if (a.rec == PSA_REC_MUSCL)
    return launch_batch_rec<T, FLUX, WHOLE, 2>(a, block, cpt, grid, max_team, nteams, s);
  if (a.rec == PSA_REC_MUSCLS)
    return launch_batch_rec<T, FLUX, WHOLE, 3>(a, block, cpt, grid, max_team, nteams, s);
  if (a.mlp != nullptr) {
    if (block == 256 && cpt == 4)
      return launch_batch_cfg<T, FLUX, 256, 4, true, WHOLE>(a, grid, max_team, nteams, s);
    if (block == 512 && cpt == 2)
      return launch_batch_cfg<T, FLUX, 512, 2, true, WHOLE>(a, grid, max_team, nteams, s);
    if (block == 1024 && cpt == 2)
      return launch_batch_cfg<T, FLUX, 1024, 2, true, WHOLE>(a, grid, max_team, nteams, s);
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "unsupported batch tile configuration for learned "
                     "weights block=" + std::to_string(block) +
                         " cells_per_thread=" + std::to_string(cpt));
  }
  if (block == 256 && cpt == 4)
    return launch_batch_cfg<T, FLUX, 256, 4, false, WHOLE>(a, grid, max_team, nteams, s);
  if (block == 256 && cpt == 2)
    return launch_batch_cfg<T, FLUX, 256, 2, false, WHOLE>(a, grid, max_team, nteams, s);
  if (block == 512 && cpt == 2)
    return launch_batch_cfg<T, FLUX, 512, 2, false, WHOLE>(a, grid, max_team, nteams, s);
  if (block == 512 && cpt == 4)
    return launch_batch_cfg<T, FLUX, 512, 4, false, WHOLE>(a, grid, max_team, nteams, s);
  if (block == 1024 && cpt == 2)
    return launch_batch_cfg<T, FLUX, 1024, 2, false, WHOLE>(a, grid, max_team, nteams, s);
  return set_error(PSA_ERR_INVALID_ARGUMENT,
                   "unsupported batch tile configuration block=" +
                       std::to_string(block) + " cells_per_thread=" +
                       std::to_string(cpt));
}

// Lax-Friedrichs: stage sweeps only (every stage needs the maximum of its
// input over the instance), WENO-JS5 weights.
template <typename T>
int launch_batch_lf(const BatchArgs<T> &a, int block, int cpt, int grid, int max_team,
                    long long nteams, hipStream_t s) {
  constexpr int F = PSA_FLUX_LAX_FRIEDRICHS;
  if (block == 256 && cpt == 4)
    return launch_batch_cfg<T, F, 256, 4, false, false>(a, grid, max_team, nteams, s);
  if (block == 256 && cpt == 2)
    return launch_batch_cfg<T, F, 256, 2, false, false>(a, grid, max_team, nteams, s);
  if (block == 512 && cpt == 2)
    return launch_batch_cfg<T, F, 512, 2, false, false>(a, grid, max_team, nteams, s);
  if (block == 512 && cpt == 4)
    return launch_batch_cfg<T, F, 512, 4, false, false>(a, grid, max_team, nteams, s);
  if (block == 1024 && cpt == 2)
    return launch_batch_cfg<T, F, 1024, 2, false, false>(a, grid, max_team, nteams, s);
  return set_error(PSA_ERR_INVALID_ARGUMENT,
                   "unsupported batch tile configuration block=" + std::to_string(block) +
                       " cells_per_thread=" + std::to_string(cpt));
}

template <typename T, int FLUX>
int launch_batch_flux(const BatchArgs<T> &a, bool whole, int block, int cpt,
                      int grid, int max_team, long long nteams, hipStream_t s) {
  if (whole)
    return launch_batch_tiles<T, FLUX, true>(a, block, cpt, grid, max_team, nteams, s);
  return launch_batch_tiles<T, FLUX, false>(a, block, cpt, grid, max_team, nteams, s);
}

template <typename T>
int launch_batch(const psa_batch_desc *d, void *u, void *k1, void *k2,
                 unsigned int *counter, double *t_out, long long *steps_out,
                 int *status_out, hipStream_t s, hipFuncAttributes *query = nullptr,
                 unsigned int *resident = nullptr) {
  BatchArgs<T> a{};
  a.query = query;
  a.resident = resident;
  a.u = static_cast<T *>(u);
  a.k1 = static_cast<T *>(k1);
  a.k2 = static_cast<T *>(k2);
  a.offsets = d->offsets;
  a.ncells = d->ncells;
  a.dx = d->dx;
  a.tfinal = d->tfinal;
  const bool teams = d->team != nullptr;
  a.order = teams ? d->slot_inst : d->order;
  a.slot_rank = teams ? d->slot_rank : nullptr;
  a.team = teams ? d->team : nullptr;
  a.teams = teams ? static_cast<TeamState *>(d->team_state) : nullptr;
  a.nslots = teams ? d->nslots : d->ninst;
  a.counter = counter;
  a.t_out = t_out;
  a.steps_out = steps_out;
  a.status_out = status_out;
  a.max_steps = d->max_steps;
  a.eps = static_cast<T>(d->eps);
  a.theta = static_cast<T>(d->theta);
  a.theta64 = d->theta;
  a.time_f64 = d->time_f64;
  a.g = d->g;
  a.mlp = d->mlp;
  a.tau = static_cast<T>(d->mlp_tau);
  a.sigma = static_cast<T>(d->mlp_sigma);
  a.rec = d->rec;
  a.limiter = d->limiter;
  a.lim_theta = static_cast<T>(d->limiter_theta);
  a.lim_atol = static_cast<T>(d->muscl_atol);
  int dev = 0, rate_khz = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) !=
          hipSuccess ||
      rate_khz <= 0) {
    (void)hipGetLastError();
    rate_khz = 100000;  // 100 MHz
  }
  const double wait_s = d->team_wait_s > 0.0 ? d->team_wait_s : 60.0;
  a.wait_ticks = static_cast<long long>(wait_s * 1000.0 * rate_khz);
  const int block = d->block > 0 ? d->block : 256;
  const int cpt = d->cells_per_thread > 0 ? d->cells_per_thread : 4;
  const int max_team = teams ? d->max_team : 1;
  const bool whole = d->schedule != PSA_SCHEDULE_TILE;
  if (d->flux == PSA_FLUX_LAX_FRIEDRICHS)
    return launch_batch_lf<T>(a, block, cpt, d->grid, max_team, d->ninst, s);
  if (d->flux == PSA_FLUX_RUSANOV)
    return launch_batch_flux<T, PSA_FLUX_RUSANOV>(a, whole, block, cpt, d->grid,
                                                  max_team, d->ninst, s);
  if (d->flux == PSA_FLUX_ENGQUIST_OSHER)
    return launch_batch_flux<T, PSA_FLUX_ENGQUIST_OSHER>(a, whole, block, cpt,
                                                         d->grid, max_team,
                                                         d->ninst, s);
  return launch_batch_flux<T, PSA_FLUX_BURGERS_UPWIND>(a, whole, block, cpt, d->grid,
                                                       max_team, d->ninst, s);
}

}  // namespace
}  // namespace psa

using namespace psa;

namespace {

// The descriptor checks of psa_batch_solve that do not involve buffers.
int check_batch_desc(const psa_batch_desc *desc) {
  if (desc == nullptr) return set_error(PSA_ERR_INVALID_ARGUMENT, "null desc");
  if (desc->dtype != PSA_FLOAT64 && desc->dtype != PSA_FLOAT32)
    return set_error(PSA_ERR_NOT_IMPLEMENTED, "unsupported dtype");
  if (desc->flux != PSA_FLUX_RUSANOV && desc->flux != PSA_FLUX_BURGERS_UPWIND &&
      desc->flux != PSA_FLUX_ENGQUIST_OSHER && desc->flux != PSA_FLUX_LAX_FRIEDRICHS)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the batched solver takes Burgers Rusanov, Godunov, "
                     "Lax-Friedrichs or Engquist-Osher");
  if (desc->flux == PSA_FLUX_LAX_FRIEDRICHS &&
      (desc->rec != PSA_REC_WENOJS53 || desc->mlp != nullptr))
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the batched Lax-Friedrichs solver takes the WENO-JS5 weights");
  if (desc->rec < PSA_REC_WENOJS53 || desc->rec > PSA_REC_MUSCLS)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "unsupported reconstruction " + std::to_string(desc->rec));
  if (desc->rec != PSA_REC_WENOJS53 && desc->mlp != nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "learned weights replace the WENO-JS5 weights only");
  if (desc->rec == PSA_REC_MUSCL) {
    const int st = check_limiter(desc->limiter, desc->limiter_theta, desc->muscl_atol);
    if (st != PSA_OK) return st;
  }
  if (desc->rec == PSA_REC_MUSCLS) {
    const int st = check_slope_limiter(desc->limiter);
    if (st != PSA_OK) return st;
  }
  const int gmin = desc->rec == PSA_REC_WENOJS53 ? 3 : 2;
  if (desc->g < gmin)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "needs at least " + std::to_string(gmin) + " ghost cells");
  if (desc->ninst < 0 || desc->max_steps < 0 || !(desc->eps >= 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad batch parameters");
  if (desc->schedule != PSA_SCHEDULE_DEFAULT && desc->schedule != PSA_SCHEDULE_TILE &&
      desc->schedule != PSA_SCHEDULE_STEP)
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "the batched solver takes the step or tile schedule");
  if (desc->mlp != nullptr && (!(desc->mlp_tau >= 0.0) || !(desc->mlp_sigma >= 0.0)))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "mlp_tau and mlp_sigma must be >= 0");
  return PSA_OK;
}

}  // namespace

extern "C" int psa_batch_kernel_info(const psa_batch_desc *desc, int *vgprs,
                                     long long *private_bytes, long long *lds_bytes) {
  const int st = check_batch_desc(desc);
  if (st != PSA_OK) return st;
  if (vgprs == nullptr || private_bytes == nullptr || lds_bytes == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  hipFuncAttributes attr{};
  const int r = desc->dtype == PSA_FLOAT64
                    ? launch_batch<double>(desc, nullptr, nullptr, nullptr, nullptr,
                                           nullptr, nullptr, nullptr, nullptr, &attr)
                    : launch_batch<float>(desc, nullptr, nullptr, nullptr, nullptr,
                                          nullptr, nullptr, nullptr, nullptr, &attr);
  if (r != PSA_OK) return r;
  *vgprs = attr.numRegs;
  *private_bytes = static_cast<long long>(attr.localSizeBytes);
  *lds_bytes = static_cast<long long>(attr.sharedSizeBytes);
  return PSA_OK;
}

extern "C" int psa_batch_resident_workgroups(const psa_batch_desc *desc,
                                             int *workgroups) {
  const int st = check_batch_desc(desc);
  if (st != PSA_OK) return st;
  if (workgroups == nullptr) return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  hipFuncAttributes attr{};
  unsigned int cap = 0;
  const int r = desc->dtype == PSA_FLOAT64
                    ? launch_batch<double>(desc, nullptr, nullptr, nullptr, nullptr,
                                           nullptr, nullptr, nullptr, nullptr, &attr, &cap)
                    : launch_batch<float>(desc, nullptr, nullptr, nullptr, nullptr,
                                          nullptr, nullptr, nullptr, nullptr, &attr, &cap);
  if (r != PSA_OK) return r;
  *workgroups = static_cast<int>(cap);
  return PSA_OK;
}

extern "C" int psa_batch_solve(const psa_batch_desc *desc, void *u, void *k1,
                               void *k2, unsigned int *counter, double *t_out,
                               long long *steps_out, int *status_out,
                               void *stream) {
  const int chk = check_batch_desc(desc);
  if (chk != PSA_OK) return chk;
  if (u == nullptr || k1 == nullptr || k2 == nullptr || counter == nullptr ||
      t_out == nullptr || steps_out == nullptr || status_out == nullptr ||
      desc->offsets == nullptr || desc->ncells == nullptr ||
      desc->dx == nullptr || desc->tfinal == nullptr || desc->order == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (desc->team != nullptr &&
      (desc->slot_inst == nullptr || desc->slot_rank == nullptr ||
       desc->team_state == nullptr || desc->nslots < desc->ninst ||
       desc->max_team < 1))
    return set_error(PSA_ERR_INVALID_ARGUMENT,
                     "teams need slot_inst, slot_rank, team_state, nslots and "
                     "max_team");
  if (desc->ninst == 0) return PSA_OK;
  auto s = static_cast<hipStream_t>(stream);
  if (desc->dtype == PSA_FLOAT64)
    return launch_batch<double>(desc, u, k1, k2, counter, t_out, steps_out,
                                status_out, s);
  return launch_batch<float>(desc, u, k1, k2, counter, t_out, steps_out,
                             status_out, s);
}
