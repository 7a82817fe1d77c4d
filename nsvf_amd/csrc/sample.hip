// Ray sampling kernels for gfx950: inverse-CDF and uniform, one lane per ray.
//
// Op-level kernels keep the `_ext` layout ([B,R,P] bins in, [B,R,S] padded
// samples out; fairnr/clib/src/sample.cpp:23-95).  The row kernel serves the
// Python-level `inverse_cdf_sampling` (fairnr/clib/__init__.py:231-300) in one
// launch: it takes the [N,P] rows directly and derives, for each ray, the
// neighbours the reference's G=200 regrouping and 800-column chunking give it
// (which the trailing-loop quirks read), instead of materialising the
// regrouped chunks.
#include "sample_core.h"

namespace nsvf {
namespace {

constexpr int kSampleBlock = 256;

// Bins of ray j of an op-level [B,R,P] batch buffer.
struct OpBins {
  const int32_t* idx;   // this ray's row
  const float* lo_;
  const float* hi_;
  const float* pr;
  const int32_t* row0;  // first ray of the batch row
  int32_t next;         // slot P of this row in the flat buffer
  __device__ __forceinline__ int32_t id(int k) const { return idx[k]; }
  __device__ __forceinline__ float lo(int k) const { return lo_[k]; }
  __device__ __forceinline__ float hi(int k) const { return hi_[k]; }
  __device__ __forceinline__ float prob(int k) const { return pr[k]; }
  __device__ __forceinline__ int32_t next_first() const { return next; }
  __device__ __forceinline__ int32_t ray0_id(int k) const { return row0[k]; }
};

// Noise read from the flat batch buffer (a read past the row continues into
// the next ray's noise, as the reference's pointer arithmetic does; past the
// end of the buffer it reads 0.5).
struct FlatNoise {
  const float* row;
  int64_t avail;  // readable elements from `row` to the end of the buffer
  __device__ __forceinline__ float at(int k) const { return k < avail ? row[k] : 0.5f; }
};

__global__ __launch_bounds__(kSampleBlock) void icdf_op_kernel(
    const int32_t* __restrict__ pts_idx, const float* __restrict__ min_depth,
    const float* __restrict__ max_depth, const float* __restrict__ noise,
    const float* __restrict__ probs, const float* __restrict__ steps, int64_t BR,
    int R, int P, int S, float fixed_step_size, int compat,
    int32_t* __restrict__ out_idx, float* __restrict__ out_depth,
    float* __restrict__ out_dists, unsigned long long* __restrict__ flags) {
  const int64_t r = int64_t(blockIdx.x) * kSampleBlock + threadIdx.x;
  if (r >= BR) return;
  const int64_t b = r / R;
  const int64_t H = r * P;
  OpBins bins{pts_idx + H, min_depth + H, max_depth + H, probs + H, pts_idx + b * R * P,
              (r + 1 < BR) ? pts_idx[H + P] : -1};
  FlatNoise nz{noise + r * S, (BR - r) * S};
  RowSink sink{out_idx + r * S, out_depth + r * S, out_dists + r * S, S, 0};
  const int s = icdf_ray(bins, P, steps[r], fixed_step_size, nz, compat, S, sink);
  sink.pad_from(min(s, S));
  if (flags) {
    if (sink.dropped) atomicAdd(&flags[1], (unsigned long long)sink.dropped);
    const int total = int(ceilf(steps[r]));
    if (int64_t(total) > (BR - r) * S) atomicAdd(&flags[2], 1ull);
  }
}

// Bins of hit ray n of [N,P] rows under the reference's regrouping.
__device__ __forceinline__ void grouped_neighbours(int64_t n, int64_t N, int G, int chunk,
                                                   int64_t* ray0, int64_t* next) {
  const int64_t Kc = (N + G - 1) / G;  // columns per group row (H / G)
  const int64_t g = n / Kc, c = n - g * Kc;
  const int64_t c0 = (c / chunk) * chunk;
  const int64_t W = min(int64_t(chunk), Kc - c0);
  int64_t n0 = g * Kc + c0;
  *ray0 = n0 < N ? n0 : 0;  // rows past N are copies of row 0
  int64_t nn;
  if (c - c0 < W - 1)
    nn = n + 1;
  else if (g < G - 1)
    nn = (g + 1) * Kc + c0;
  else
    nn = -1;  // one past the chunk buffer
  *next = (nn >= N) ? 0 : nn;
}

__global__ __launch_bounds__(kSampleBlock) void icdf_rows_kernel(
    const int32_t* __restrict__ pts_idx, const float* __restrict__ min_depth,
    const float* __restrict__ max_depth, const float* __restrict__ noise,
    const float* __restrict__ probs, const float* __restrict__ steps, int64_t N,
    int P, int S, int G, int chunk, float fixed_step_size, int compat,
    int32_t* __restrict__ out_idx, float* __restrict__ out_depth,
    float* __restrict__ out_dists, unsigned long long* __restrict__ flags) {
  const int64_t n = int64_t(blockIdx.x) * kSampleBlock + threadIdx.x;
  if (n >= N) return;
  int64_t ray0, next;
  grouped_neighbours(n, N, G, chunk, &ray0, &next);
  const int64_t H = n * P;
  OpBins bins{pts_idx + H, min_depth + H, max_depth + H, probs + H, pts_idx + ray0 * P,
              next < 0 ? -1 : pts_idx[next * P]};
  RowSink sink{out_idx + n * S, out_depth + n * S, out_dists + n * S, S, 0};
  int s;
  if (noise) {
    FlatNoise nz{noise + n * S, S};
    s = icdf_ray(bins, P, steps[n], fixed_step_size, nz, compat, S, sink);
  } else {
    ConstNoise nz{0.5f};
    s = icdf_ray(bins, P, steps[n], fixed_step_size, nz, compat, S, sink);
  }
  sink.pad_from(min(s, S));
  if (flags && sink.dropped) atomicAdd(&flags[1], (unsigned long long)sink.dropped);
}

// Uniform sampling of one ray (fairnr/clib/src/sample_gpu.cu:42-105) into a
// padded row; the row is first filled as the reference's host wrapper
// allocates it (idx -1, depth 0, dists 0), since the phases read it back.
__global__ __launch_bounds__(kSampleBlock) void uniform_op_kernel(
    const int32_t* __restrict__ pts_idx_all, const float* __restrict__ min_depth_all,
    const float* __restrict__ max_depth_all, const float* __restrict__ noise_all,
    int64_t BR, int P, int S, float step_size, int32_t* __restrict__ out_idx,
    float* __restrict__ out_depth, float* __restrict__ out_dists,
    unsigned long long* __restrict__ flags) {
  const int64_t r = int64_t(blockIdx.x) * kSampleBlock + threadIdx.x;
  if (r >= BR) return;
  const int64_t H = r * P, K = r * S;
  const int32_t* pts_idx = pts_idx_all + H;
  const float* min_depth = min_depth_all + H;
  const float* max_depth = max_depth_all + H;
  const float* noise = noise_all + K;
  int32_t* sidx = out_idx + K;
  float* sdep = out_depth + K;
  float* sdis = out_dists + K;
  for (int t = 0; t < S; ++t) {
    sidx[t] = -1;
    sdep[t] = 0.0f;
    sdis[t] = 0.0f;
  }
  int oob = 0, dropped = 0, stuck = 0;
  int s = 0, ucur = 0, umin = 0, umax = 0;
  float last_min_depth, last_max_depth, curr_depth = 0.0f;
  while (true) {
    if (umax == P || ucur == S || pts_idx[umax] == -1) break;
    // Past S every write is dropped and nothing later reads this phase's
    // counters, so stopping here changes no output and bounds the loop.
    if (s >= S) break;
    last_min_depth = (umin < P) ? min_depth[umin] : 10000.0f;
    last_max_depth = (umax < P) ? max_depth[umax] : 10000.0f;
    if (ucur < S) curr_depth = __fmaf_rn(float(ucur) + noise[ucur], step_size, min_depth[0]);
    if (last_max_depth <= curr_depth && last_max_depth <= last_min_depth) {
      if (s < S) {
        sdep[s] = last_max_depth;
        sidx[s] = pts_idx[umax];
      } else {
        ++dropped;
      }
      umax++;
      s++;
      continue;
    }
    if (curr_depth <= last_min_depth && curr_depth <= last_max_depth) {
      // umin == 0 only when a noise value is exactly 0: the reference then
      // reads the previous ray's last slot (-1 before the buffer start).
      int32_t id;
      if (umin > 0) {
        id = pts_idx[umin - 1];
      } else if (r > 0) {
        id = pts_idx_all[H - 1];
      } else {
        id = -1;
        ++oob;
      }
      if (s < S) {
        sdep[s] = curr_depth;
        sidx[s] = id;
      } else {
        ++dropped;
      }
      ucur++;
      s++;
      continue;
    }
    if (last_min_depth <= curr_depth && last_min_depth <= last_max_depth) {
      if (s < S) {
        sdep[s] = last_min_depth;
        sidx[s] = pts_idx[umin];
      } else {
        ++dropped;
      }
      umin++;
      s++;
      continue;
    }
    ++stuck;  // NaN depths: no branch applies (the reference would spin)
    break;
  }
  int step = 0;
  umin = 0;
  umax = 0;
  for (ucur = 0; ucur < S - 1; ucur++) {
    if (sidx[ucur + 1] == -1) break;
    const float l_depth = sdep[ucur];
    const float r_depth = sdep[ucur + 1];
    sdep[ucur] = (l_depth + r_depth) * 0.5f;
    sdis[ucur] = r_depth - l_depth;
    if (umin < P && sdep[ucur] >= min_depth[umin] && pts_idx[umin] > -1) umin++;
    if (umax < P && sdep[ucur] >= max_depth[umax] && pts_idx[umax] > -1) umax++;
    if (umax == P || pts_idx[umax] == -1) break;
    if (umin - 1 == umax && sdis[ucur] > 0) {
      sdep[step] = sdep[ucur];
      sdis[step] = sdis[ucur];
      sidx[step] = sidx[ucur];
      step++;
    }
  }
  for (int t = step; t < S; t++) sidx[t] = -1;
  if (flags) {
    if (oob) atomicAdd(&flags[0], (unsigned long long)oob);
    if (dropped) atomicAdd(&flags[1], (unsigned long long)dropped);
    if (stuck) atomicAdd(&flags[2], (unsigned long long)stuck);
  }
}

}  // namespace
}  // namespace nsvf

using namespace nsvf;

extern "C" {

int nsvf_inverse_cdf_sampling(const int32_t* pts_idx, const float* min_depth,
                              const float* max_depth, const float* noise,
                              const float* probs, const float* steps, int B,
                              int R, int P, int S, float fixed_step_size,
                              int compat, int32_t* out_idx, float* out_depth,
                              float* out_dists, int64_t* flags, void* stream) {
  clear_error();
  if (B < 0 || R < 0 || P < 0 || S < 0)
    return set_error(NSVF_EINVAL, "inverse_cdf_sampling: negative size");
  const int64_t BR = int64_t(B) * R;
  if (BR == 0 || S == 0) return NSVF_OK;
  if (P == 0) return set_error(NSVF_EINVAL, "inverse_cdf_sampling: no bins (P = 0)");
  if (!pts_idx || !min_depth || !max_depth || !noise || !probs || !steps || !out_idx ||
      !out_depth || !out_dists)
    return set_error(NSVF_EINVAL, "inverse_cdf_sampling: null pointer");
  hipLaunchKernelGGL(icdf_op_kernel, dim3(unsigned(ceil_div(BR, kSampleBlock))), dim3(kSampleBlock),
                     0, as_stream(stream), pts_idx, min_depth, max_depth, noise, probs, steps, BR, R,
                     P, S, fixed_step_size, compat, out_idx, out_depth, out_dists,
                     reinterpret_cast<unsigned long long*>(flags));
  return check_launch("inverse_cdf_sampling");
}

int nsvf_inverse_cdf_sampling_rows(const int32_t* pts_idx, const float* min_depth,
                                   const float* max_depth, const float* noise,
                                   const float* probs, const float* steps, int N,
                                   int P, int S, int group, int chunk,
                                   float fixed_step_size, int compat,
                                   int32_t* out_idx, float* out_depth,
                                   float* out_dists, int64_t* flags, void* stream) {
  clear_error();
  if (N < 0 || P < 0 || S < 0 || group <= 0 || chunk <= 0)
    return set_error(NSVF_EINVAL, "inverse_cdf_sampling_rows: bad size");
  if (N == 0 || S == 0) return NSVF_OK;
  if (P == 0) return set_error(NSVF_EINVAL, "inverse_cdf_sampling_rows: no bins (P = 0)");
  if (!pts_idx || !min_depth || !max_depth || !probs || !steps || !out_idx || !out_depth ||
      !out_dists)
    return set_error(NSVF_EINVAL, "inverse_cdf_sampling_rows: null pointer");
  hipLaunchKernelGGL(icdf_rows_kernel, dim3(unsigned(ceil_div(N, kSampleBlock))),
                     dim3(kSampleBlock), 0, as_stream(stream), pts_idx, min_depth, max_depth,
                     noise, probs, steps, int64_t(N), P, S, group, chunk, fixed_step_size, compat,
                     out_idx, out_depth, out_dists, reinterpret_cast<unsigned long long*>(flags));
  return check_launch("inverse_cdf_sampling_rows");
}

int nsvf_uniform_ray_sampling(const int32_t* pts_idx, const float* min_depth,
                              const float* max_depth, const float* noise,
                              int B, int R, int P, int S, float step_size,
                              int32_t* out_idx, float* out_depth,
                              float* out_dists, int64_t* flags, void* stream) {
  clear_error();
  if (B < 0 || R < 0 || P < 0 || S < 0)
    return set_error(NSVF_EINVAL, "uniform_ray_sampling: negative size");
  const int64_t BR = int64_t(B) * R;
  if (BR == 0 || S == 0) return NSVF_OK;
  if (P == 0) return set_error(NSVF_EINVAL, "uniform_ray_sampling: no bins (P = 0)");
  if (!pts_idx || !min_depth || !max_depth || !noise || !out_idx || !out_depth || !out_dists)
    return set_error(NSVF_EINVAL, "uniform_ray_sampling: null pointer");
  hipLaunchKernelGGL(uniform_op_kernel, dim3(unsigned(ceil_div(BR, kSampleBlock))),
                     dim3(kSampleBlock), 0, as_stream(stream), pts_idx, min_depth, max_depth, noise,
                     BR, P, S, step_size, out_idx, out_depth, out_dists,
                     reinterpret_cast<unsigned long long*>(flags));
  return check_launch("uniform_ray_sampling");
}

}  // extern "C"
