// Interior norms of grid.py:445-506 (_norm / norm) as deterministic two-pass
// reductions: a fixed grid of workgroups writes one partial per workgroup to
// a caller-provided workspace, and one workgroup combines the partials in a
// fixed order, so repeated calls return bit-identical results. Sums
// accumulate in FP64 for both dtypes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/pyshocks_amd.h"
#include "psa_internal.hpp"

namespace psa {
namespace {

constexpr int kRedBlock = 256;
constexpr int kRedBlocks = 512;

template <int P>
__device__ __forceinline__ double combine(double a, double b) {
  if constexpr (P == PSA_NORM_INF) {
    // NaN-propagating maximum (jnp.max semantics)
    if (a != a) return a;
    if (b != b) return b;
    return a > b ? a : b;
  } else {
    return a + b;
  }
}

template <typename T, int P>
__device__ __forceinline__ double term(const T *u, int64_t i, int64_t hi) {
  if constexpr (P == PSA_NORM_L1) {
    return fabs(static_cast<double>(u[i]));
  } else if constexpr (P == PSA_NORM_L2) {
    const double v = static_cast<double>(u[i]);
    return v * v;
  } else if constexpr (P == PSA_NORM_INF) {
    return fabs(static_cast<double>(u[i]));
  } else {
    // total variation of |u| over the slice (_norm takes |u| first,
    // grid.py:446): ||u[i + 1]| - |u[i]||, i + 1 < hi
    if (i + 1 >= hi) return 0.0;
    return fabs(fabs(static_cast<double>(u[i + 1])) -
                fabs(static_cast<double>(u[i])));
  }
}

template <int P>
__device__ double block_reduce(double v, double *s_red) {
  for (int off = 32; off > 0; off >>= 1) {
    v = combine<P>(v, __shfl_xor(v, off, 64));
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) s_red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kRedBlock / 64; ++w) v = combine<P>(v, s_red[w]);
  }
  return v;
}

template <typename T, int P>
__global__ __launch_bounds__(kRedBlock) void k_norm_partial(
    const T *__restrict__ u, int64_t lo, int64_t hi,
    double *__restrict__ partial) {
  __shared__ double s_red[kRedBlock / 64];
  double v = 0.0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kRedBlock;
  for (int64_t i = lo + static_cast<int64_t>(blockIdx.x) * kRedBlock +
                   threadIdx.x;
       i < hi; i += stride) {
    v = combine<P>(v, term<T, P>(u, i, hi));
  }
  v = block_reduce<P>(v, s_red);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

template <typename T, int P>
__global__ __launch_bounds__(kRedBlock) void k_norm_final(
    const double *__restrict__ partial, int nparts, double weight,
    T *__restrict__ out) {
  __shared__ double s_red[kRedBlock / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kRedBlock) {
    v = combine<P>(v, partial[i]);
  }
  v = block_reduce<P>(v, s_red);
  if (threadIdx.x == 0) {
    double r;
    if constexpr (P == PSA_NORM_L1) {
      r = v * weight;
    } else if constexpr (P == PSA_NORM_L2) {
      r = sqrt(v * weight);
    } else {
      r = v;
    }
    *out = static_cast<T>(r);
  }
}

template <typename T, int P>
int launch_norm(const void *u, int64_t lo, int64_t hi, double weight,
                void *out, void *workspace, hipStream_t s) {
  int64_t nparts = (hi - lo + kRedBlock - 1) / kRedBlock;
  if (nparts > kRedBlocks) nparts = kRedBlocks;
  if (nparts < 1) nparts = 1;
  auto *part = static_cast<double *>(workspace);
  hipLaunchKernelGGL((k_norm_partial<T, P>), dim3(nparts), dim3(kRedBlock), 0,
                     s, static_cast<const T *>(u), lo, hi, part);
  int st = check_launch("norm partial");
  if (st != PSA_OK) return st;
  hipLaunchKernelGGL((k_norm_final<T, P>), dim3(1), dim3(kRedBlock), 0, s,
                     part, static_cast<int>(nparts), weight,
                     static_cast<T *>(out));
  return check_launch("norm final");
}

template <typename T>
int dispatch_norm(int p, const void *u, int64_t lo, int64_t hi, double weight,
                  void *out, void *ws, hipStream_t s) {
  switch (p) {
    case PSA_NORM_L1:
      return launch_norm<T, PSA_NORM_L1>(u, lo, hi, weight, out, ws, s);
    case PSA_NORM_L2:
      return launch_norm<T, PSA_NORM_L2>(u, lo, hi, weight, out, ws, s);
    case PSA_NORM_INF:
      return launch_norm<T, PSA_NORM_INF>(u, lo, hi, weight, out, ws, s);
    case PSA_NORM_TVD:
      return launch_norm<T, PSA_NORM_TVD>(u, lo, hi, weight, out, ws, s);
    default:
      return set_error(PSA_ERR_NOT_IMPLEMENTED,
                       "unsupported norm order " + std::to_string(p));
  }
}

}  // namespace
}  // namespace psa

extern "C" {

int psa_norm_workspace_bytes(void) {
  return static_cast<int>(psa::kRedBlocks * sizeof(double));
}

int psa_interior_norm(int dtype, int p, const void *u, int64_t lo, int64_t hi,
                      double weight, void *out, void *workspace,
                      size_t workspace_bytes, void *stream) {
  using namespace psa;
  if (dtype != PSA_FLOAT64 && dtype != PSA_FLOAT32)
    return set_error(PSA_ERR_NOT_IMPLEMENTED, "unsupported dtype");
  if (u == nullptr || out == nullptr || workspace == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (lo < 0 || hi <= lo)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "empty slice");
  if (workspace_bytes < static_cast<size_t>(kRedBlocks) * sizeof(double))
    return set_error(PSA_ERR_WORKSPACE, "norm workspace too small");
  auto s = static_cast<hipStream_t>(stream);
  if (dtype == PSA_FLOAT64)
    return dispatch_norm<double>(p, u, lo, hi, weight, out, workspace, s);
  return dispatch_norm<float>(p, u, lo, hi, weight, out, workspace, s);
}

}  // extern "C"
