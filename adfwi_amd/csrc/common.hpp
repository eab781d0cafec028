// Shared host/device helpers of libadfwi_amd.so: error reporting, argument checks, and the
// receiver row mask of the step kernels.
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

// Rows base .. base+127 that hold receivers, as a wavefront-uniform bit mask built once per
// wavefront from the map's row counts: a row without receivers then costs one scalar test, and
// a lane's receiver index is read only on rows that have some.
struct RowMask {
  unsigned long long lo, hi;
  int base;
  __device__ bool has(int row) const {
    const int i = row - base;
    if (i < 0 || i >= 128) return false;
    return i < 64 ? ((lo >> i) & 1ull) != 0 : ((hi >> (i - 64)) & 1ull) != 0;
  }
};
__device__ __forceinline__ RowMask receiver_rows(const int *map, int Z, int base, int lane) {
  const int r0 = base + lane, r1 = base + 64 + lane;
  const bool a = map && r0 >= 0 && r0 < Z && map[r0] > 0;
  const bool b = map && r1 >= 0 && r1 < Z && map[r1] > 0;
  return RowMask{(unsigned long long)__ballot(a), (unsigned long long)__ballot(b), base};
}

}  // namespace adfwi
