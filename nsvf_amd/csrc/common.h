// Shared helpers for the gfx950 kernels and their C-ABI launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/nsvf_amd.h"

namespace nsvf {

// Sentinel depth of padded hit slots and samples (fairnr/clib/__init__.py:36).
constexpr float kMaxDepth = 10000.0f;

// Records the error text for nsvf_last_error() and returns `code`.
int set_error(int code, const char* fmt, ...);
void clear_error();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Checks the launch that was just enqueued.
inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(NSVF_ELAUNCH, "%s: %s", what, hipGetErrorString(e));
  return NSVF_OK;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Makes this wavefront's earlier LDS writes visible to all its lanes (a
// barrier for one wave; waves of a block do not wait for each other).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Slab test of one ray against one axis-aligned box (fairnr/clib/src/
// intersect_gpu.cu:73-122): the same comparisons and ternaries, so an axis
// whose bounds come out NaN (direction 0 and origin on the slab plane) fails
// every comparison and is skipped.  `inv` holds IEEE reciprocals of the
// direction, computed once per ray.  Built with -ffp-contract=off, so every
// product and difference below is rounded on its own.
__device__ __forceinline__ bool ray_aabb(const float o[3], const float inv[3],
                                         float cx, float cy, float cz, float h,
                                         float* out_lo, float* out_hi) {
  float f_low = 0.0f;
  float f_high = 100000.0f;
  const float c[3] = {cx, cy, cz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float lo = (c[a] - h - o[a]) * inv[a];
    float hi = (c[a] + h - o[a]) * inv[a];
    if (hi < lo) {
      float t = lo;
      lo = hi;
      hi = t;
    }
    if (hi < f_low) return false;
    if (lo > f_high) return false;
    f_low = (lo > f_low) ? lo : f_low;
    f_high = (hi < f_high) ? hi : f_high;
    if (f_low > f_high) return false;
  }
  *out_lo = f_low;
  *out_hi = f_high;
  return true;
}

}  // namespace nsvf
