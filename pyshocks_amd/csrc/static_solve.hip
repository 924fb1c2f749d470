// psa_ssprk33_solve_static: the step loop of psa_control_init +
// psa_ssprk33_steps with its scratch memory (the control block and the two
// stage buffers) in a static device array instead of caller buffers, for
// callers that cannot allocate.
//
// Host code never uses the array's name as an address: in host code it is a
// stand-in, not a device pointer. Each device's copy is resolved once with
// hipGetSymbolAddress (with that device current) and that pointer is what the
// kernels receive.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/pyshocks_amd.h"
#include "psa_internal.hpp"

namespace {

constexpr size_t kControlBytes = 30976;  // sizeof(psa_control) rounded to 256
constexpr size_t kStageBytes = static_cast<size_t>(PSA_STATIC_MAX_CELLS) * 8;
constexpr size_t kScratchBytes = kControlBytes + 2 * kStageBytes;
static_assert(sizeof(psa_control) <= kControlBytes, "control block layout");
static_assert(kScratchBytes == PSA_STATIC_SCRATCH_BYTES, "scratch size");

}  // namespace

// control block | stage buffer 1 | stage buffer 2 (each 256-byte aligned)
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's copy of scratch.
int scratch_of_current_device(void **out) {
  int dev = -1;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess)
    return psa::set_error(PSA_ERR_HIP, std::string("hipGetDevice: ") +
                                           hipGetErrorString(err));
  if (dev < 0 || dev >= kMaxDevices)
    return psa::set_error(PSA_ERR_INVALID_ARGUMENT, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || p == nullptr)
      return psa::set_error(PSA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                             hipGetErrorString(err));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return PSA_OK;
}

__global__ void k_static_result(const psa_control *ctrl, psa_static_result *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  out->t = ctrl->t;
  out->dt = ctrl->dt;
  out->m = ctrl->m;
  out->status = ctrl->status;
  out->pad_ = 0;
}

}  // namespace

extern "C" int psa_ssprk33_solve_static(const psa_solver_desc *desc,
                                        const void *u_in, void *u_out,
                                        double tstart, double tfinal,
                                        long long maxit, int dt_mode,
                                        double theta, double cfl_num,
                                        double fixed_dt, long long nsteps,
                                        void *result, void *stream) {
  if (desc == nullptr) return psa::set_error(PSA_ERR_INVALID_ARGUMENT, "null descriptor");
  if (desc->n < 1 || desc->g < 0)
    return psa::set_error(PSA_ERR_INVALID_ARGUMENT, "bad grid size");
  const int64_t m = desc->n + 2 * static_cast<int64_t>(desc->g);
  if (m > PSA_STATIC_MAX_CELLS)
    return psa::set_error(PSA_ERR_WORKSPACE,
                          "psa_ssprk33_solve_static: m = " + std::to_string(m) +
                              " cells exceeds PSA_STATIC_MAX_CELLS = " +
                              std::to_string(PSA_STATIC_MAX_CELLS));
  if (u_in == nullptr || u_out == nullptr)
    return psa::set_error(PSA_ERR_INVALID_ARGUMENT, "null pointer");
  if (nsteps < 0) return psa::set_error(PSA_ERR_INVALID_ARGUMENT, "nsteps < 0");

  void *base = nullptr;
  int st = scratch_of_current_device(&base);
  if (st != PSA_OK) return st;
  auto *bytes = static_cast<unsigned char *>(base);
  void *control = bytes;
  void *k1 = bytes + kControlBytes;
  void *k2 = bytes + kControlBytes + kStageBytes;

  // the control block starts from the init kernel (no memset), so the
  // previous call's contents do not matter
  st = psa_control_init(desc, u_in, control, tstart, tfinal, maxit, dt_mode,
                        theta, cfl_num, fixed_dt, stream);
  if (st != PSA_OK) return st;
  if (nsteps > 0) {
    st = psa_ssprk33_steps(desc, u_in, u_out, k1, k2, control, nsteps, stream);
    if (st != PSA_OK) return st;
  } else if (u_out != u_in) {
    const size_t esize = desc->dtype == PSA_FLOAT64 ? 8 : 4;
    hipError_t err = hipMemcpyAsync(u_out, u_in, static_cast<size_t>(m) * esize,
                                    hipMemcpyDeviceToDevice,
                                    static_cast<hipStream_t>(stream));
    if (err != hipSuccess)
      return psa::set_error(PSA_ERR_HIP, std::string("copy: ") + hipGetErrorString(err));
  }
  if (result != nullptr) {
    hipLaunchKernelGGL(k_static_result, dim3(1), dim3(64), 0,
                       static_cast<hipStream_t>(stream),
                       static_cast<const psa_control *>(control),
                       static_cast<psa_static_result *>(result));
    st = psa::check_launch("static result");
  }
  return st;
}
