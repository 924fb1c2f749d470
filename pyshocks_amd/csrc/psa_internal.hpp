// Host-side helpers shared by the translation units of libpyshocks_amd.
#pragma once

#include <string>

namespace psa {

// Records msg as the thread-local last error and returns code.
int set_error(int code, const std::string &msg);

// Maps a pending HIP launch error to PSA_ERR_HIP.
int check_launch(const char *what);

}  // namespace psa
