// Learned WENO weights: standalone inference, the training loss and its
// gradient, the Adam update, and the learned reconstruction of a whole array.
// The perceptron and its matrix-core layout are described in mlp_device.hpp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/pyshocks_amd.h"
#include "mlp_device.hpp"
#include "psa_internal.hpp"
#include "stage_device.hpp"
#include "weno5_device.hpp"

namespace psa {
namespace {

static_assert(MLP_NPARAM == PSA_MLP_NPARAM, "parameter count");

// ---------------------------------------------------------------------------
// inference on given inputs

__global__ __launch_bounds__(256) void k_mlp_weights(
    const float *__restrict__ P, const float *__restrict__ feat, long long n,
    float *__restrict__ omega, float *__restrict__ logits) {
  MlpWave W;
  mlp_load(W, P);
  const int lane = threadIdx.x & 63;
  const long long waves = static_cast<long long>(gridDim.x) * (blockDim.x >> 6);
  const long long wave = static_cast<long long>(blockIdx.x) * (blockDim.x >> 6) +
                         (threadIdx.x >> 6);
  // wave-uniform trip count: every lane of a wave runs the collective
  for (long long base = wave * 64; base < n; base += waves * 64) {
    const long long s = base + lane;
    float x[MLP_IN] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (s < n) {
#pragma unroll
      for (int k = 0; k < MLP_IN; ++k) x[k] = feat[s * MLP_IN + k];
    }
    float z[MLP_OUT], w[MLP_OUT];
    mlp_logits(W, x, z);
    mlp_softmax(z, w);
    if (s < n) {
#pragma unroll
      for (int k = 0; k < MLP_OUT; ++k) {
        omega[s * MLP_OUT + k] = w[k];
        if (logits != nullptr) logits[s * MLP_OUT + k] = z[k];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// training loss and gradient
//
// Sample record (PSA_MLP_RECORD floats): inputs x[5], normalised candidate
// errors e[3] (candidate minus exact face value, over the stencil scale),
// WENO-JS weights wjs[3], sample weight c. Loss of a sample:
//   r = sum_k omega_k e_k,  L = c (r^2 / (mean_k e_k^2 + delta)
//                               + lambda sum_k (omega_k - wjs_k)^2)
// The batch loss is the sum over samples divided by n.

constexpr int GRAD_BLOCK = 128;
constexpr int GRAD_WAVES = GRAD_BLOCK / 64;
constexpr int GRAD_PART = MLP_NPARAM + 1;  // gradient entries + loss
constexpr int GRAD_MAX_BLOCKS = 1024;
constexpr int HS = MLP_HID + 1;  // padded LDS rows
constexpr int HSB = MLP_HID + 2;

// The activations are BF16 values, stored as such.
struct GradSmem {
  float x[GRAD_BLOCK][MLP_IN];
  __bf16 h1[GRAD_BLOCK][HSB];
  __bf16 h2[GRAD_BLOCK][HSB];
  float dh1[GRAD_BLOCK][HS];
  float dh2[GRAD_BLOCK][HS];
  float dz[GRAD_BLOCK][MLP_OUT];
  float red[GRAD_WAVES];
};

__device__ __forceinline__ float bf16_value(float v) {
  return static_cast<float>(static_cast<__bf16>(v));
}

__global__ __launch_bounds__(GRAD_BLOCK) void k_mlp_grad(
    const float *__restrict__ P, const float *__restrict__ rec, long long n,
    float lambda_js, float delta, float *__restrict__ part) {
  __shared__ GradSmem sm;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int q = lane >> 4;
  const int col = lane & 15;

  MlpWave W;
  mlp_load(W, P);
  // transposed operands of the backward products
  bf16x8_t w3t[2], w2t[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int row = 16 * b + col;  // unit of the layer below
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float a3 = (q == 0 && j < MLP_OUT) ? P[MLP_W3 + j * MLP_HID + row] : 0.0f;
      w3t[b][j] = static_cast<__bf16>(a3);
      w2t[b][j] = static_cast<__bf16>(P[MLP_W2 + mlp_slot_unit(q, j) * MLP_HID + row]);
    }
  }

  // gradient entries owned by this thread
  const int n1 = tid & 31;
  const int grp = tid >> 5;  // 0..3
  float gW2[8], gW1[2] = {0.0f, 0.0f}, gW3 = 0.0f, gb = 0.0f, lsum = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) gW2[k] = 0.0f;

  const f32x4_t zero4 = {0.0f, 0.0f, 0.0f, 0.0f};
  for (long long base = static_cast<long long>(blockIdx.x) * GRAD_BLOCK; base < n;
       base += static_cast<long long>(gridDim.x) * GRAD_BLOCK) {
    const int ls = wave * 64 + lane;  // local sample of this lane
    const long long s = base + ls;
    float x[MLP_IN] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float e[3] = {0.0f, 0.0f, 0.0f}, wjs[3] = {0.0f, 0.0f, 0.0f}, c = 0.0f;
    if (s < n) {
      const float *r = rec + s * PSA_MLP_RECORD;
#pragma unroll
      for (int k = 0; k < MLP_IN; ++k) x[k] = r[k];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        e[k] = r[5 + k];
        wjs[k] = r[8 + k];
      }
      c = r[11];
    }
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) sm.x[ls][k] = bf16_value(x[k]);

    // ---- forward, keeping the BF16 activations in LDS
    float z[MLP_OUT] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) {
      const int src = 16 * ch + col;
      const int srow = wave * 64 + src;  // LDS row of this column's sample
      bf16x8_t B;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float xs = j < MLP_IN ? __shfl(x[j < MLP_IN ? j : 0], src, 64) : 0.0f;
        B[j] = static_cast<__bf16>(q == 0 ? xs : 0.0f);
      }
      f32x4_t h0 = mlp_mfma(W.w1[0], B, W.b1[0]);
      f32x4_t h1 = mlp_mfma(W.w1[1], B, W.b1[1]);
      B = mlp_relu_pack(h0, h1);
#pragma unroll
      for (int j = 0; j < 8; ++j) sm.h1[srow][mlp_slot_unit(q, j)] = B[j];
      h0 = mlp_mfma(W.w2[0], B, W.b2[0]);
      h1 = mlp_mfma(W.w2[1], B, W.b2[1]);
      B = mlp_relu_pack(h0, h1);
#pragma unroll
      for (int j = 0; j < 8; ++j) sm.h2[srow][mlp_slot_unit(q, j)] = B[j];
      const f32x4_t o = mlp_mfma(W.w3, B, W.b3);
      const float t0 = __shfl(o[0], col, 64);
      const float t1 = __shfl(o[1], col, 64);
      const float t2 = __shfl(o[2], col, 64);
      if (q == ch) {
        z[0] = t0;
        z[1] = t1;
        z[2] = t2;
      }
    }

    // ---- loss and d loss / d logits of this lane's sample
    float w[MLP_OUT];
    mlp_softmax(z, w);
    const float r = (w[0] * e[0] + w[1] * e[1]) + w[2] * e[2];
    const float den = ((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]) * (1.0f / 3.0f) + delta;
    const float inv_den = den > 0.0f ? 1.0f / den : 0.0f;
    const float dj0 = w[0] - wjs[0], dj1 = w[1] - wjs[1], dj2 = w[2] - wjs[2];
    lsum += c * (r * r * inv_den + lambda_js * ((dj0 * dj0 + dj1 * dj1) + dj2 * dj2));
    const float gr = 2.0f * c * r * inv_den;
    const float gj = 2.0f * c * lambda_js;
    const float g0 = gr * e[0] + gj * dj0;
    const float g1 = gr * e[1] + gj * dj1;
    const float g2 = gr * e[2] + gj * dj2;
    const float gw = (w[0] * g0 + w[1] * g1) + w[2] * g2;
    float dz[MLP_OUT] = {w[0] * (g0 - gw), w[1] * (g1 - gw), w[2] * (g2 - gw)};
#pragma unroll
    for (int k = 0; k < MLP_OUT; ++k) sm.dz[ls][k] = dz[k];

    // ---- backward through the hidden layers (each lane reads back only
    // the LDS entries it wrote itself)
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) {
      const int src = 16 * ch + col;
      const int srow = wave * 64 + src;
      bf16x8_t B;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float ds = j < MLP_OUT ? __shfl(dz[j < MLP_OUT ? j : 0], src, 64) : 0.0f;
        B[j] = static_cast<__bf16>(q == 0 ? ds : 0.0f);
      }
      f32x4_t d0 = mlp_mfma(w3t[0], B, zero4);
      f32x4_t d1 = mlp_mfma(w3t[1], B, zero4);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int u0 = 4 * q + v, u1 = 16 + 4 * q + v;
        d0[v] = static_cast<float>(sm.h2[srow][u0]) > 0.0f ? d0[v] : 0.0f;
        d1[v] = static_cast<float>(sm.h2[srow][u1]) > 0.0f ? d1[v] : 0.0f;
        sm.dh2[srow][u0] = d0[v];
        sm.dh2[srow][u1] = d1[v];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        B[j] = static_cast<__bf16>(d0[j]);
        B[4 + j] = static_cast<__bf16>(d1[j]);
      }
      d0 = mlp_mfma(w2t[0], B, zero4);
      d1 = mlp_mfma(w2t[1], B, zero4);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int u0 = 4 * q + v, u1 = 16 + 4 * q + v;
        sm.dh1[srow][u0] = static_cast<float>(sm.h1[srow][u0]) > 0.0f ? d0[v] : 0.0f;
        sm.dh1[srow][u1] = static_cast<float>(sm.h1[srow][u1]) > 0.0f ? d1[v] : 0.0f;
      }
    }
    __syncthreads();

    // ---- parameter gradients, FP32 sums over the block's samples
    for (int t = 0; t < GRAD_BLOCK; ++t) {
      const float h1v = static_cast<float>(sm.h1[t][n1]);
#pragma unroll
      for (int k = 0; k < 8; ++k) gW2[k] += sm.dh2[t][grp + 4 * k] * h1v;
      const float dh1v = sm.dh1[t][n1];
      gW1[0] += dh1v * sm.x[t][grp];
      if (grp == 0) gW1[1] += dh1v * sm.x[t][4];
      if (grp < 3) gW3 += sm.dz[t][grp] * static_cast<float>(sm.h2[t][n1]);
      if (tid < 32) {
        gb += dh1v;
      } else if (tid < 64) {
        gb += sm.dh2[t][n1];
      } else if (tid < 64 + MLP_OUT) {
        gb += sm.dz[t][tid - 64];
      }
    }
    __syncthreads();
  }

  // ---- partial sums of this block
  float *out = part + static_cast<long long>(blockIdx.x) * GRAD_PART;
#pragma unroll
  for (int k = 0; k < 8; ++k) out[MLP_W2 + (grp + 4 * k) * MLP_HID + n1] = gW2[k];
  out[MLP_W1 + n1 * MLP_IN + grp] = gW1[0];
  if (grp == 0) out[MLP_W1 + n1 * MLP_IN + 4] = gW1[1];
  if (grp < 3) out[MLP_W3 + grp * MLP_HID + n1] = gW3;
  if (tid < 32) {
    out[MLP_B1 + tid] = gb;
  } else if (tid < 64) {
    out[MLP_B2 + tid - 32] = gb;
  } else if (tid < 64 + MLP_OUT) {
    out[MLP_B3 + tid - 64] = gb;
  }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  if (lane == 0) sm.red[wave] = lsum;
  __syncthreads();
  if (tid == 0) {
    float l = 0.0f;
    for (int k = 0; k < GRAD_WAVES; ++k) l += sm.red[k];
    out[MLP_NPARAM] = l;
  }
}

// Sum of the block partials in block order, divided by n.
__global__ void k_mlp_grad_reduce(const float *__restrict__ part, int nblocks,
                                  double inv_n, float *__restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= GRAD_PART) return;
  double acc = 0.0;
  for (int b = 0; b < nblocks; ++b) acc += part[static_cast<long long>(b) * GRAD_PART + e];
  grad[e] = static_cast<float>(acc * inv_n);
}

__global__ void k_mlp_adam(float *__restrict__ p, const float *__restrict__ g,
                           float *__restrict__ m, float *__restrict__ v, float lr,
                           float beta1, float beta2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= MLP_NPARAM) return;
  const float ge = g[e];
  const float me = beta1 * m[e] + (1.0f - beta1) * ge;
  const float ve = beta2 * v[e] + (1.0f - beta2) * ge * ge;
  m[e] = me;
  v[e] = ve;
  p[e] -= lr * (me * c1) / (sqrtf(ve * c2) + eps);
}

// ---------------------------------------------------------------------------
// learned reconstruction of an array (the learned counterpart of
// reconstruction.py:358-377, same zero padding outside [0, m))

template <typename T>
__global__ __launch_bounds__(256) void k_learned_reconstruct(
    const float *__restrict__ P, T tau, T sigma, T eps, const T *__restrict__ u,
    T *__restrict__ ul, T *__restrict__ ur, int64_t m) {
  MlpWave W;
  mlp_load(W, P);
  const int lane = threadIdx.x & 63;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * (blockDim.x >> 6);
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) +
                       (threadIdx.x >> 6);
  for (int64_t base = wave * 64; base < m; base += waves * 64) {
    const int64_t i = base + lane;
    const T um2 = load_raw(u, i - 2, m), um1 = load_raw(u, i - 1, m),
            u0 = load_raw(u, i, m), up1 = load_raw(u, i + 1, m),
            up2 = load_raw(u, i + 2, m);
    T l, r;
    learned_cell(W, um2, um1, u0, up1, up2, eps, tau, sigma, l, r);
    if (i < m) {
      ul[i] = l;
      ur[i] = r;
    }
  }
}

unsigned int grid_for(long long items, int per_block, unsigned int cap) {
  long long b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned int>(b);
}

}  // namespace
}  // namespace psa

using namespace psa;

extern "C" int psa_mlp_weights(const float *params, const float *features,
                               long long n, float *omega, float *logits,
                               void *stream) {
  if (n < 0) return set_error(PSA_ERR_INVALID_ARGUMENT, "negative sample count");
  if (n == 0) return PSA_OK;
  if (params == nullptr || features == nullptr || omega == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  hipLaunchKernelGGL(k_mlp_weights, dim3(grid_for(n, 256, 4096)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), params, features, n, omega,
                     logits);
  return check_launch("mlp weights");
}

extern "C" int psa_mlp_grad_workspace_bytes(long long n, size_t *bytes) {
  if (n < 0 || bytes == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad arguments");
  const unsigned int nb = grid_for(n, GRAD_BLOCK, GRAD_MAX_BLOCKS);
  *bytes = static_cast<size_t>(nb) * GRAD_PART * sizeof(float);
  return PSA_OK;
}

extern "C" int psa_mlp_grad(const float *params, const float *samples, long long n,
                            double lambda_js, double delta, float *grad, void *work,
                            size_t work_bytes, void *stream) {
  if (n <= 0) return set_error(PSA_ERR_INVALID_ARGUMENT, "no samples");
  if (params == nullptr || samples == nullptr || grad == nullptr || work == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (!(lambda_js >= 0.0) || !(delta >= 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "lambda and delta must be >= 0");
  const unsigned int nb = grid_for(n, GRAD_BLOCK, GRAD_MAX_BLOCKS);
  if (work_bytes < static_cast<size_t>(nb) * GRAD_PART * sizeof(float))
    return set_error(PSA_ERR_WORKSPACE, "gradient workspace too small");
  auto s = static_cast<hipStream_t>(stream);
  float *part = static_cast<float *>(work);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(GRAD_BLOCK), 0, s, params, samples, n,
                     static_cast<float>(lambda_js), static_cast<float>(delta), part);
  int st = check_launch("mlp gradient");
  if (st != PSA_OK) return st;
  hipLaunchKernelGGL(k_mlp_grad_reduce, dim3((GRAD_PART + 255) / 256), dim3(256), 0, s,
                     part, static_cast<int>(nb), 1.0 / static_cast<double>(n), grad);
  return check_launch("mlp gradient reduce");
}

extern "C" int psa_mlp_adam(float *params, const float *grad, float *m, float *v,
                            long long step, double lr, double beta1, double beta2,
                            double eps, void *stream) {
  if (params == nullptr || grad == nullptr || m == nullptr || v == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (step < 1 || !(lr > 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) ||
      !(beta2 >= 0.0 && beta2 < 1.0) || !(eps > 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "bad Adam parameters");
  const double c1 = 1.0 / (1.0 - std::pow(beta1, static_cast<double>(step)));
  const double c2 = 1.0 / (1.0 - std::pow(beta2, static_cast<double>(step)));
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0,
                     static_cast<hipStream_t>(stream), params, grad, m, v,
                     static_cast<float>(lr), static_cast<float>(beta1),
                     static_cast<float>(beta2), static_cast<float>(eps),
                     static_cast<float>(c1), static_cast<float>(c2));
  return check_launch("mlp adam");
}

extern "C" int psa_learned_reconstruct(int dtype, const float *params, double tau,
                                       double sigma, const void *u, void *ul, void *ur,
                                       int64_t m, double eps, void *stream) {
  if (m < 0) return set_error(PSA_ERR_INVALID_ARGUMENT, "negative size");
  if (m == 0) return PSA_OK;
  if (params == nullptr || u == nullptr || ul == nullptr || ur == nullptr)
    return set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (!(tau >= 0.0) || !(sigma >= 0.0) || !(eps >= 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "tau, sigma and eps must be >= 0");
  auto s = static_cast<hipStream_t>(stream);
  const unsigned int nb = grid_for(m, 256, 4096);
  if (dtype == PSA_FLOAT64) {
    hipLaunchKernelGGL(k_learned_reconstruct<double>, dim3(nb), dim3(256), 0, s, params,
                       tau, sigma, eps, static_cast<const double *>(u),
                       static_cast<double *>(ul), static_cast<double *>(ur), m);
  } else if (dtype == PSA_FLOAT32) {
    hipLaunchKernelGGL(k_learned_reconstruct<float>, dim3(nb), dim3(256), 0, s, params,
                       static_cast<float>(tau), static_cast<float>(sigma),
                       static_cast<float>(eps),
                       static_cast<const float *>(u), static_cast<float *>(ul),
                       static_cast<float *>(ur), m);
  } else {
    return set_error(PSA_ERR_NOT_IMPLEMENTED, "unsupported dtype");
  }
  return check_launch("learned reconstruct");
}
