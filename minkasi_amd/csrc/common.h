// Shared helpers for the gfx950 kernels behind libmkgpu.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

#include "../../include/mkgpu.h"

namespace mk {

// Thread-local last error message, returned by mk_last_error().
void set_error(const char* fmt, ...);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Checks the launch that was just queued.
int check_launch(const char* what);

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Grid size for a grid-stride loop over n items: enough workgroups to cover
// every CU several times (256 CUs), capped so small inputs launch few groups.
inline unsigned grid_for(int64_t n, int block, int64_t max_groups = 256 * 16) {
  int64_t g = cdiv(n, block);
  if (g > max_groups) g = max_groups;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// Counts lanes with bad == true into *nbad (when nbad is not null): one
// global atomic per wave that has any such lane.
__device__ inline void count_bad(bool bad, int64_t* nbad) {
  unsigned long long m = __ballot(bad);
  if (nbad && m && (threadIdx.x & 63) == (__ffsll((long long)m) - 1))
    atomicAdd((unsigned long long*)nbad, (unsigned long long)__popcll(m));
}

}  // namespace mk

#define MK_REQUIRE(cond, code, ...)      \
  do {                                   \
    if (!(cond)) {                       \
      mk::set_error(__VA_ARGS__);        \
      return (code);                     \
    }                                    \
  } while (0)

#define MK_HIP(call)                                                     \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      mk::set_error("%s failed: %s", #call, hipGetErrorString(e_));      \
      return MK_ERR_HIP;                                                 \
    }                                                                    \
  } while (0)
