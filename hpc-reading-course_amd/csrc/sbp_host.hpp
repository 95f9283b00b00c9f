// sbp_host.hpp -- error reporting shared by the host translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/sbp_hip.h"

namespace sbp {
namespace host {
// Records msg as the thread's last error (sbp_last_error) and returns code.
int fail(int code, const std::string &msg);
}  // namespace host
}  // namespace sbp

#define SBP_HIP_OR_FAIL(expr)                                                             \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ::sbp::host::fail(e_ == hipErrorOutOfMemory ? SBP_ENOMEM : SBP_EHIP,         \
                               std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)
