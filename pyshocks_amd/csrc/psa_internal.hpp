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
// does not apply (nothing launched).
int step_fixed_dt(const psa_solver_desc *d, const void *u_in, void *u_out, void *k1,
                  void *k2, const void *dt, hipStream_t stream);

}  // namespace psa
