// Host-side helpers shared by the translation units of libpyshocks_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/pyshocks_amd.h"

namespace psa {

// Records msg as the thread-local last error and returns code.
int set_error(int code, const std::string &msg);

// Maps a pending HIP launch error to PSA_ERR_HIP.
int check_launch(const char *what);

// One fixed-dt whole step (forward.hip): u_in -> u_out, the stage states
// into k1, k2 if not null; PSA_ERR_NOT_IMPLEMENTED where the whole-step kernel
// does not apply (nothing launched). A batch of instances takes one launch:
// instance b at offset b * ld (elements) of every array, dt[b] its step.
int step_fixed_dt(const psa_solver_desc *d, const void *u_in, void *u_out, void *k1,
                  void *k2, const void *dt, hipStream_t stream, long long batch = 1,
                  int64_t ld = 0);

// Limiter of the flux-limited MUSCL: a known PSA_LIMITER_*, minmod's theta
// in [1, 2] (limiters.py:166-198), a positive slope-ratio cut-off.
inline int check_limiter(int lm, double theta, double atol) {
  if (lm < PSA_LIMITER_MINMOD || lm > PSA_LIMITER_NONE)
    return set_error(PSA_ERR_NOT_IMPLEMENTED, "unsupported limiter " + std::to_string(lm));
  if (lm == PSA_LIMITER_MINMOD && !(theta >= 1.0 && theta <= 2.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "'theta' must be in [1, 2]");
  if (!(atol > 0.0))
    return set_error(PSA_ERR_INVALID_ARGUMENT, "the MUSCL slope cut-off atol must be > 0");
  return PSA_OK;
}

// Limiter of the slope-limited MUSCL: every limiter but Koren, which has no
// slope form (limiters.py:381-399).
inline int check_slope_limiter(int lm) {
  if (lm == PSA_LIMITER_KOREN)
    return set_error(PSA_ERR_NOT_IMPLEMENTED,
                     "the Koren limiter has no slope form (limiters.py:381-389)");
  if (lm < PSA_LIMITER_MINMOD || lm > PSA_LIMITER_NONE)
    return set_error(PSA_ERR_NOT_IMPLEMENTED, "unsupported limiter " + std::to_string(lm));
  return PSA_OK;
}

}  // namespace psa
