// lbm_internal.hpp -- error plumbing shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/lbm_capi.h"

namespace lbm {

void set_error(const char* fmt, ...);
void clear_error();

}  // namespace lbm

#define LBM_HIP_CHECK(call)                                                                  \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      ::lbm::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,      \
                       __LINE__);                                                            \
      return e_ == hipErrorOutOfMemory ? LBM_ERR_OOM : LBM_ERR_HIP;                          \
    }                                                                                        \
  } while (0)

#define LBM_REQUIRE(cond, status, ...)  \
  do {                                  \
    if (!(cond)) {                      \
      ::lbm::set_error(__VA_ARGS__);    \
      return status;                    \
    }                                   \
  } while (0)

// Launch-error check after a kernel launch (launch configuration errors only;
// execution errors surface at the next synchronising call).
#define LBM_LAUNCH_CHECK() LBM_HIP_CHECK(hipGetLastError())
