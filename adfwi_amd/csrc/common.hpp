// Shared host/device helpers of libadfwi_amd.so: error reporting and argument checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "adfwi_amd.h"

namespace adfwi {

// Thread-local text of the last failure (adfwi_last_error).
void set_error(const char *fmt, ...);
const char *last_error();

// Returns ADFWI_E_FAULT_INJECTED when the environment variable
// ADFWI_FAULT_INJECT names `entry` (or "all"); used by host-side failure tests.
int fault_injected(const char *entry);

#define ADFWI_CHECK_ARG(cond, ...)                 \
  do {                                             \
    if (!(cond)) {                                 \
      ::adfwi::set_error(__VA_ARGS__);             \
      return ADFWI_E_ARG;                          \
    }                                              \
  } while (0)

#define ADFWI_CHECK_HIP(expr)                                                   \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ::adfwi::set_error("%s failed: %s", #expr, hipGetErrorString(e_));        \
      return ADFWI_E_LAUNCH;                                                    \
    }                                                                           \
  } while (0)

inline int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("launch of %s failed: %s", what, hipGetErrorString(e));
    return ADFWI_E_LAUNCH;
  }
  return ADFWI_OK;
}

}  // namespace adfwi
