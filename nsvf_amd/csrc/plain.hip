// Comparator kernels: the reference algorithm in its naive form on MI355X, for
// the plain-FP32 baseline the fast path is measured against.
//
// They keep the reference's launch geometry (fairnr/clib/include/
// cuda_utils.h:18-24: one workgroup per batch row, the largest power of two
// <= rows (capped at 512) threads, each thread looping over rays with a
// workgroup-size stride), one thread per ray, no LDS: the SVO traversal keeps
// a 256-entry stack per thread in private (scratch) memory and the brute-force
// test recomputes the direction reciprocals per voxel, as intersect_gpu.cu
// does.  Results equal the fast kernels'.
#include "sample_core.h"

namespace nsvf {
namespace {

int opt_threads(int work) {
  int p = 1;
  while (p * 2 <= work && p * 2 <= 512) p *= 2;
  return p;
}

__global__ void plain_aabb_kernel(int n, int m, float half_voxel, int n_max,
                                  const float* __restrict__ ray_start,
                                  const float* __restrict__ ray_dir,
                                  const float* __restrict__ points, int32_t* __restrict__ idx,
                                  float* __restrict__ min_depth, float* __restrict__ max_depth) {
  const int b = blockIdx.x;
  points += int64_t(b) * n * 3;
  ray_start += int64_t(b) * m * 3;
  ray_dir += int64_t(b) * m * 3;
  idx += int64_t(b) * m * n_max;
  min_depth += int64_t(b) * m * n_max;
  max_depth += int64_t(b) * m * n_max;
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    for (int l = 0; l < n_max; ++l) idx[int64_t(j) * n_max + l] = -1;
    for (int k = 0, cnt = 0; k < n && cnt < n_max; ++k) {
      const float o[3] = {ray_start[j * 3], ray_start[j * 3 + 1], ray_start[j * 3 + 2]};
      const float inv[3] = {1.0f / ray_dir[j * 3], 1.0f / ray_dir[j * 3 + 1],
                            1.0f / ray_dir[j * 3 + 2]};
      float lo, hi;
      if (ray_aabb(o, inv, points[k * 3], points[k * 3 + 1], points[k * 3 + 2], half_voxel, &lo,
                   &hi)) {
        idx[int64_t(j) * n_max + cnt] = k;
        min_depth[int64_t(j) * n_max + cnt] = lo;
        max_depth[int64_t(j) * n_max + cnt] = hi;
        ++cnt;
      }
    }
  }
}

__global__ void plain_svo_kernel(int n, int m, float half_voxel, int n_max,
                                 const float* __restrict__ ray_start,
                                 const float* __restrict__ ray_dir,
                                 const float* __restrict__ points,
                                 const int32_t* __restrict__ children,
                                 int32_t* __restrict__ idx, float* __restrict__ min_depth,
                                 float* __restrict__ max_depth) {
  const int b = blockIdx.x;
  points += int64_t(b) * n * 3;
  children += int64_t(b) * n * 9;
  ray_start += int64_t(b) * m * 3;
  ray_dir += int64_t(b) * m * 3;
  idx += int64_t(b) * m * n_max;
  min_depth += int64_t(b) * m * n_max;
  max_depth += int64_t(b) * m * n_max;
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    for (int l = 0; l < n_max; ++l) idx[int64_t(j) * n_max + l] = -1;
    int stack[256];
    int ptr = 0, cnt = 0;
    stack[0] = n - 1;
    while (ptr > -1 && cnt < n_max) {
      const int k = stack[ptr--];
      const float o[3] = {ray_start[j * 3], ray_start[j * 3 + 1], ray_start[j * 3 + 2]};
      const float inv[3] = {1.0f / ray_dir[j * 3], 1.0f / ray_dir[j * 3 + 1],
                            1.0f / ray_dir[j * 3 + 2]};
      const int sc = children[int64_t(k) * 9 + 8];
      float lo, hi;
      if (!ray_aabb(o, inv, points[int64_t(k) * 3], points[int64_t(k) * 3 + 1],
                    points[int64_t(k) * 3 + 2], half_voxel * float(sc), &lo, &hi))
        continue;
      if (sc == 1) {
        idx[int64_t(j) * n_max + cnt] = k;
        min_depth[int64_t(j) * n_max + cnt] = lo;
        max_depth[int64_t(j) * n_max + cnt] = hi;
        ++cnt;
        continue;
      }
      // a full stack stops the ray (the reference asserts ptr < 256)
      for (int u = 0; u < 8 && ptr < 254; ++u) {
        const int c = children[int64_t(k) * 9 + u];
        if (c > -1 && c < n) stack[++ptr] = c;
      }
    }
  }
}

struct PlainBins {
  const int32_t* idx;
  const float* lo_;
  const float* hi_;
  const float* pr;
  const int32_t* row0;
  int32_t next;
  __device__ __forceinline__ int32_t id(int k) const { return idx[k]; }
  __device__ __forceinline__ float lo(int k) const { return lo_[k]; }
  __device__ __forceinline__ float hi(int k) const { return hi_[k]; }
  __device__ __forceinline__ float prob(int k) const { return pr[k]; }
  __device__ __forceinline__ int32_t next_first() const { return next; }
  __device__ __forceinline__ int32_t ray0_id(int k) const { return row0[k]; }
};

struct PlainNoise {
  const float* row;
  int64_t avail;
  __device__ __forceinline__ float at(int k) const { return k < avail ? row[k] : 0.5f; }
};

__global__ void plain_icdf_kernel(int B, int R, int P, int S, float fixed_step_size, int compat,
                                  const int32_t* __restrict__ pts_idx,
                                  const float* __restrict__ min_depth,
                                  const float* __restrict__ max_depth,
                                  const float* __restrict__ noise, const float* __restrict__ probs,
                                  const float* __restrict__ steps, int32_t* __restrict__ out_idx,
                                  float* __restrict__ out_depth, float* __restrict__ out_dists) {
  const int b = blockIdx.x;
  const int64_t BR = int64_t(B) * R;
  for (int j = threadIdx.x; j < R; j += blockDim.x) {
    const int64_t r = int64_t(b) * R + j;
    const int64_t H = r * P;
    PlainBins bins{pts_idx + H, min_depth + H, max_depth + H, probs + H,
                   pts_idx + int64_t(b) * R * P, (r + 1 < BR) ? pts_idx[H + P] : -1};
    PlainNoise nz{noise + r * S, (BR - r) * S};
    RowSink sink{out_idx + r * S, out_depth + r * S, out_dists + r * S, S, 0};
    icdf_ray(bins, P, steps[r], fixed_step_size, nz, compat, S, sink);
  }
}

}  // namespace
}  // namespace nsvf

using namespace nsvf;

extern "C" {

int nsvf_plain_aabb_intersect(const float* ray_start, const float* ray_dir, const float* points,
                              int B, int M, int N, float voxelsize, int n_max, int32_t* idx,
                              float* min_depth, float* max_depth, void* stream) {
  clear_error();
  if (B <= 0 || M <= 0 || n_max <= 0) return NSVF_OK;
  hipLaunchKernelGGL(plain_aabb_kernel, dim3(B), dim3(opt_threads(M)), 0, as_stream(stream), N, M,
                     voxelsize * 0.5f, n_max, ray_start, ray_dir, points, idx, min_depth, max_depth);
  return check_launch("plain_aabb_intersect");
}

int nsvf_plain_svo_intersect(const float* ray_start, const float* ray_dir, const float* points,
                             const int32_t* children, int B, int M, int T, float voxelsize,
                             int n_max, int32_t* idx, float* min_depth, float* max_depth,
                             void* stream) {
  clear_error();
  if (B <= 0 || M <= 0 || n_max <= 0) return NSVF_OK;
  if (T <= 0) return set_error(NSVF_EINVAL, "plain_svo_intersect: empty tree");
  hipLaunchKernelGGL(plain_svo_kernel, dim3(B), dim3(opt_threads(M)), 0, as_stream(stream), T, M,
                     voxelsize * 0.5f, n_max, ray_start, ray_dir, points, children, idx,
                     min_depth, max_depth);
  return check_launch("plain_svo_intersect");
}

int nsvf_plain_inverse_cdf_sampling(const int32_t* pts_idx, const float* min_depth,
                                    const float* max_depth, const float* noise,
                                    const float* probs, const float* steps, int B, int R, int P,
                                    int S, float fixed_step_size, int compat, int32_t* out_idx,
                                    float* out_depth, float* out_dists, void* stream) {
  clear_error();
  if (B <= 0 || R <= 0 || S <= 0) return NSVF_OK;
  if (P <= 0) return set_error(NSVF_EINVAL, "plain_inverse_cdf_sampling: no bins");
  hipLaunchKernelGGL(plain_icdf_kernel, dim3(B), dim3(opt_threads(R)), 0, as_stream(stream), B, R,
                     P, S, fixed_step_size, compat, pts_idx, min_depth, max_depth, noise, probs, steps,
                     out_idx, out_depth, out_dists);
  return check_launch("plain_inverse_cdf_sampling");
}

}  // extern "C"
