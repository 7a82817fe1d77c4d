// Per-ray sampling procedures shared by the op-level kernels and the fused
// render pipeline.  Each procedure walks one ray sequentially (one lane per
// ray) and hands every sample to a `Sink`, so the same arithmetic feeds a
// padded [R,S] row, a compacted sample stream, or a counter.
#pragma once

#include "common.h"

namespace nsvf {

// Hands samples to a padded output row; slots >= S are dropped and counted.
struct RowSink {
  int32_t* idx;
  float* depth;
  float* dists;
  int S;
  int dropped;
  __device__ __forceinline__ void operator()(int s, int32_t id, float dist, float mid) {
    if (s < S) {
      idx[s] = id;
      dists[s] = dist;
      depth[s] = mid;
    } else {
      ++dropped;
    }
  }
  // Marks slots [s, S) as padding: idx -1, depth 0, dists 0.
  __device__ __forceinline__ void pad_from(int s) {
    for (int t = s; t < S; ++t) {
      idx[t] = -1;
      depth[t] = 0.0f;
      dists[t] = 0.0f;
    }
  }
};

// Counts samples only.
struct CountSink {
  __device__ __forceinline__ void operator()(int, int32_t, float, float) {}
};

// Inverse-CDF sampling of one ray's bins (fairnr/clib/src/sample_gpu.cu:
// 139-200).  `Bins` provides, for bin k of this ray: id(k), lo(k), hi(k),
// prob(k); plus next_first() (the slot read one past the last bin: the
// first id of the next ray in the batch buffer, or -1 past its end) and
// ray0_id(k) (slot k of the first ray of the batch row, which the
// reference's trailing-loop stop test reads).  `Noise` provides at(k).
//
// Arithmetic: 1/steps in double rounded to float (sample_gpu.cu:147);
// z = fma(u, hi-lo, lo) (the contraction of sample_gpu.cu:180); everything
// else separately rounded.  Returns the number of samples produced (capped
// near `limit`, the sink's capacity).
template <class Bins, class Noise, class Sink>
__device__ __forceinline__ int icdf_ray(const Bins& bins, int P, float steps,
                                        float fixed_step_size, const Noise& noise,
                                        int compat, int limit, Sink& sink) {
  int curr_bin = 0, s = 0;
  float curr_min_depth = bins.lo(0);
  float curr_max_depth = bins.hi(0);
  float curr_min_cdf = 0.0f;
  float curr_max_cdf = bins.prob(0);
  float step_size = float(1.0 / double(steps));
  float z_low = curr_min_depth;
  const int total_steps = int(ceilf(steps));
  bool done = false;
  if (fixed_step_size > 0.0f) step_size = fixed_step_size;

  for (int curr_step = 0; curr_step < total_steps; ++curr_step) {
    // Past `limit` every further sample would be dropped, so the ray stops:
    // this bounds the loop on any input (steps may be huge or NaN-derived)
    // without changing what is written.
    if (s >= limit) return s;
    const float curr_cdf = (float(curr_step) + noise.at(curr_step)) * step_size;
    while (curr_cdf > curr_max_cdf) {
      sink(s, bins.id(curr_bin), curr_max_depth - z_low, (curr_max_depth + z_low) * 0.5f);
      ++curr_bin;
      ++s;
      if (curr_bin >= P || bins.id(curr_bin) == -1) {
        done = true;
        break;
      }
      curr_min_depth = bins.lo(curr_bin);
      curr_max_depth = bins.hi(curr_bin);
      curr_min_cdf = curr_max_cdf;
      curr_max_cdf = curr_max_cdf + bins.prob(curr_bin);
      z_low = curr_min_depth;
    }
    if (done) break;
    const float u = (curr_cdf - curr_min_cdf) / (curr_max_cdf - curr_min_cdf);
    const float z = __fmaf_rn(u, curr_max_depth - curr_min_depth, curr_min_depth);
    sink(s, bins.id(curr_bin), z - z_low, (z + z_low) * 0.5f);
    z_low = z;
    ++s;
  }
  if (done && !compat) return s;
  // Trailing loop.  The reference's condition `(z_low < max) && (~done)` is
  // true whenever z_low < max (`~` of a bool is never 0), so after `done` it
  // writes one more tail whose id is read at slot curr_bin: a pad (-1), or,
  // when all P slots were bins, the next ray's first id.
  while (z_low < curr_max_depth) {
    const int32_t id = (curr_bin < P) ? bins.id(curr_bin) : bins.next_first();
    sink(s, id, curr_max_depth - z_low, (curr_max_depth + z_low) * 0.5f);
    ++curr_bin;
    ++s;
    if (curr_bin >= P) break;
    const int32_t probe = compat ? bins.ray0_id(curr_bin) : bins.id(curr_bin);
    if (probe == -1) break;
    curr_min_depth = bins.lo(curr_bin);
    curr_max_depth = bins.hi(curr_bin);
    z_low = curr_min_depth;
  }
  return s;
}

struct ConstNoise {
  float v;
  __device__ __forceinline__ float at(int) const { return v; }
};

// Training-mode noise (clib/__init__.py:252-256 draws noise.uniform_()
// clamped to [0.001, 0.999] when the sampler is not deterministic): a
// counter-based uniform value per (seed, ray row, step) from the PCG hash
// (Jarzynski and Olano, "Hash functions for GPU rendering", JCGT 2020), so a
// kernel needs no random state and any implementation can reproduce the
// draws: key = hash(row + hash(seed)), u = (hash(step + key) >> 8) / 2^24,
// clamped.
__host__ __device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
  const uint32_t state = v * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

struct HashNoise {
  uint32_t key;  // pcg_hash(row + pcg_hash(seed))
  __device__ __forceinline__ static HashNoise of(uint32_t seed, uint32_t row) {
    return HashNoise{pcg_hash(row + pcg_hash(seed))};
  }
  __device__ __forceinline__ float at(int k) const {
    const float u = float(pcg_hash(uint32_t(k) + key) >> 8) * (1.0f / 16777216.0f);
    return fminf(fmaxf(u, 0.001f), 0.999f);
  }
};

}  // namespace nsvf
