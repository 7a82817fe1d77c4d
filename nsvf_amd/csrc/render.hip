// Fused render pipeline kernels (one view or a batch of rays, scene resident
// on the device).  Stages, each one kernel:
//
//   hits      octree traversal + depth sort + segment-length sum per ray
//             (encoder.py:498-535 with nsvf.py:65-74 folded in)
//   compact   ray ranks -> list of hit rays (fairnr_model.py:157-162)
//   sample    inverse-CDF sampling of a ray's bins into a compacted sample
//             stream; counting pass + writing pass around a device scan
//             (encoder.py:538-556, clib/__init__.py:231-300); coarse bins are
//             the sorted hits, fine bins are the coarse samples
//             (nerf.py:64-79)
//   composite front-to-back compositing of one ray's samples by one
//             wavefront with a prefix scan (renderer.py:193-218)
//   finish    background fill of every ray of the frame (nsvf.py:89-101)
//
// Sample streams are compacted (CSR by hit ray): per-ray offset and count,
// per-sample voxel id, midpoint depth, interval length and ray id.  No
// padded [R, S] sample tensor exists on this path; padding only appears in
// the op-level kernels that keep the reference's layout.
#include "render.h"
#include "sample_core.h"

namespace nsvf {
namespace {

constexpr int kBlock = 256;
constexpr int kHitsBlock = 128;
constexpr int kMaxLevels = 30;

// ---------------------------------------------------------------------------
// hits

// Traverses the octree for one ray, appending leaf hits to `row`.  Child
// slots are visited in the order (i ^ mask), i = 0..7, with mask the ray's
// direction-sign octant (front to back), or 7..0 (the reference's order).
// Returns the hit count (<= n_max).
template <bool kFrontToBack>
__device__ int traverse(const float o[3], const float inv[3], int mask,
                        const float* __restrict__ cen, const int32_t* __restrict__ ch, int T,
                        float half_voxel, int n_max, HitRec* __restrict__ row, int32_t* stack) {
  int cnt = 0, level = -1;
  const int root = T - 1;
  {
    const int sc = ch[int64_t(root) * 9 + 8];
    float lo, hi;
    if (n_max > 0 && ray_aabb(o, inv, cen[int64_t(root) * 3], cen[int64_t(root) * 3 + 1],
                              cen[int64_t(root) * 3 + 2], half_voxel * float(sc), &lo, &hi)) {
      if (sc == 1) {
        row[0] = HitRec{lo, hi, root, 0};
        return 1;
      }
      level = 0;
      stack[0] = (root << 4) | 8;
    }
  }
  while (level >= 0 && cnt < n_max) {
    const int32_t e = stack[level * kHitsBlock];
    const int node = e >> 4;
    const int left = e & 15;
    if (left == 0) {
      --level;
      continue;
    }
    stack[level * kHitsBlock] = (node << 4) | (left - 1);
    const int slot = kFrontToBack ? ((8 - left) ^ mask) : (left - 1);
    const int c = ch[int64_t(node) * 9 + slot];
    if (c < 0 || c >= T) continue;
    const int sc = ch[int64_t(c) * 9 + 8];
    float lo, hi;
    if (!ray_aabb(o, inv, cen[int64_t(c) * 3], cen[int64_t(c) * 3 + 1], cen[int64_t(c) * 3 + 2],
                  half_voxel * float(sc), &lo, &hi))
      continue;
    if (sc == 1) {
      row[cnt++] = HitRec{lo, hi, c, 0};
      continue;
    }
    const int ps = ch[int64_t(node) * 9 + 8];
    if (sc < ps && level + 1 < kMaxLevels) {
      ++level;
      stack[level * kHitsBlock] = (c << 4) | 8;
    }
  }
  return cnt;
}

// Per ray: hits sorted by entry depth, count, and the left-to-right sum of
// the segment lengths.  The front-to-back traversal yields the hits already
// sorted (leaves are disjoint voxels and a ray crosses octants in increasing
// (slot ^ mask) order), so it is the common path.  A ray whose hits reach
// n_max (the kept subset then depends on traversal order), or whose entry
// depths are not strictly increasing (equal entries: the tie order must be
// the reference's), is redone in the reference's order and stably
// insertion-sorted, which is exactly the reference pipeline's result.
__global__ __launch_bounds__(kHitsBlock) void hits_kernel(
    const float* __restrict__ ray_start, const float* __restrict__ ray_dir, int64_t R,
    const float* __restrict__ cen, const int32_t* __restrict__ ch, int T, float half_voxel,
    int n_max, HitRec* __restrict__ hits, int32_t* __restrict__ count, float* __restrict__ tot_out,
    int32_t* __restrict__ hit_flag) {
  __shared__ int32_t stack[kMaxLevels * kHitsBlock];
  const int64_t r = int64_t(blockIdx.x) * kHitsBlock + threadIdx.x;
  if (r >= R) return;
  float o[3], inv[3];
  int mask = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    o[a] = ray_start[r * 3 + a];
    const float d = ray_dir[r * 3 + a];
    inv[a] = 1.0f / d;
    mask |= (d < 0.0f ? 1 : 0) << a;
  }
  HitRec* row = hits + r * n_max;
  int32_t* st = stack + threadIdx.x;
  int cnt = traverse<true>(o, inv, mask, cen, ch, T, half_voxel, n_max, row, st);
  bool ok = cnt < n_max;
  for (int k = 1; ok && k < cnt; ++k) ok = row[k - 1].lo < row[k].lo;
  if (!ok) {
    cnt = traverse<false>(o, inv, mask, cen, ch, T, half_voxel, n_max, row, st);
    for (int k = 1; k < cnt; ++k) {  // stable insertion sort by entry depth
      const HitRec h = row[k];
      int j = k - 1;
      while (j >= 0 && row[j].lo > h.lo) {
        row[j + 1] = row[j];
        --j;
      }
      row[j + 1] = h;
    }
  }
  float tot = 0.0f;
  for (int k = 0; k < cnt; ++k) tot = tot + (row[k].hi - row[k].lo);
  count[r] = cnt;
  tot_out[r] = tot;
  hit_flag[r] = cnt > 0 ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void compact_kernel(const int32_t* __restrict__ flag,
                                                         const int32_t* __restrict__ rank,
                                                         int64_t R, int32_t* __restrict__ list) {
  const int64_t r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (r < R && flag[r]) list[rank[r]] = int32_t(r);
}

// ---------------------------------------------------------------------------
// sample

// Bins of a hit ray taken from its sorted hit row (coarse pass).  Slots past
// the hit count read as the reference's padding (-1, MAX_DEPTH, prob 0);
// probs are seg/sum computed on the fly, as nsvf.py:70 computes them.
struct HitBins {
  const HitRec* row;
  int cnt;
  float tot;
  const HitRec* row0;
  int cnt0;
  int32_t next;
  __device__ __forceinline__ int32_t id(int k) const { return k < cnt ? row[k].idx : -1; }
  __device__ __forceinline__ float lo(int k) const { return k < cnt ? row[k].lo : kMaxDepth; }
  __device__ __forceinline__ float hi(int k) const { return k < cnt ? row[k].hi : kMaxDepth; }
  __device__ __forceinline__ float prob(int k) const {
    return k < cnt ? (row[k].hi - row[k].lo) / tot : 0.0f;
  }
  __device__ __forceinline__ int32_t next_first() const { return next; }
  __device__ __forceinline__ int32_t ray0_id(int k) const { return k < cnt0 ? row0[k].idx : -1; }
};

// Bins of a hit ray taken from its coarse samples (fine pass, nerf.py:64-79):
// [depth - dist/2, depth + dist/2], probs (p + 1e-5) / sum over all P columns.
struct SampleBins {
  const int32_t* idx;
  const float* depth;
  const float* dist;
  const float* prob_c;
  int cnt;
  float den;  // column sum of (p + 1e-5)
  const int32_t* idx0;
  int cnt0;
  int32_t next;
  __device__ __forceinline__ int32_t id(int k) const { return k < cnt ? idx[k] : -1; }
  __device__ __forceinline__ float lo(int k) const {
    return k < cnt ? depth[k] - dist[k] * 0.5f : kMaxDepth;
  }
  __device__ __forceinline__ float hi(int k) const {
    return k < cnt ? depth[k] + dist[k] * 0.5f : kMaxDepth;
  }
  __device__ __forceinline__ float prob(int k) const {
    return (k < cnt ? prob_c[k] + 1e-5f : 1e-5f) / den;
  }
  __device__ __forceinline__ int32_t next_first() const { return next; }
  __device__ __forceinline__ int32_t ray0_id(int k) const { return k < cnt0 ? idx0[k] : -1; }
};

// Counts the samples with a voxel id (pads are dropped from the stream).
struct ValidCountSink {
  int n = 0;
  __device__ __forceinline__ void operator()(int, int32_t id, float, float) {
    if (id != -1) ++n;
  }
};

// Writes valid samples to the stream (dists clamped at 0: encoder.py:547).
struct StreamSink {
  int32_t* idx;
  float* depth;
  float* dist;
  int32_t* ray;
  int32_t ray_id;
  int n = 0;
  __device__ __forceinline__ void operator()(int, int32_t id, float d, float mid) {
    if (id == -1) return;
    idx[n] = id;
    depth[n] = mid;
    dist[n] = d > 0.0f ? d : 0.0f;
    ray[n] = ray_id;
    ++n;
  }
};

__device__ __forceinline__ void grouped_neighbours(int64_t n, int64_t N, int64_t* ray0,
                                                   int64_t* next) {
  constexpr int G = 200, chunk = 800;  // clib/__init__.py:234,259
  const int64_t Kc = (N + G - 1) / G;
  const int64_t g = n / Kc, c = n - g * Kc;
  const int64_t c0 = (c / chunk) * chunk;
  const int64_t W = min(int64_t(chunk), Kc - c0);
  const int64_t n0 = g * Kc + c0;
  *ray0 = n0 < N ? n0 : 0;
  int64_t nn;
  if (c - c0 < W - 1)
    nn = n + 1;
  else if (g < G - 1)
    nn = (g + 1) * Kc + c0;
  else
    nn = -1;
  *next = nn >= N ? 0 : nn;
}

template <bool kWrite>
__global__ __launch_bounds__(kBlock) void coarse_sample_kernel(
    const int32_t* __restrict__ hit_list, const int32_t* __restrict__ n_hit_dev,
    const HitRec* __restrict__ hits, const int32_t* __restrict__ count,
    const float* __restrict__ tot, int n_max, float step_size, int compat,
    int32_t* __restrict__ counts_out, int32_t* __restrict__ max_count,
    const int32_t* __restrict__ offsets, SampleStream out, int64_t n_hit_max) {
  const int64_t n = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t N = *n_hit_dev;
  if (n >= N) {
    // counts past the hit rays are zero, so a scan over n_hit_max entries
    // gives the right offsets and total
    if (!kWrite && n < n_hit_max) counts_out[n] = 0;
    return;
  }
  const int32_t r = hit_list[n];
  int64_t n0, nn;
  grouped_neighbours(n, N, &n0, &nn);
  const int32_t r0 = hit_list[n0];
  HitBins bins{hits + int64_t(r) * n_max, count[r], tot[r], hits + int64_t(r0) * n_max, count[r0],
               nn < 0 ? -1 : hits[int64_t(hit_list[nn]) * n_max].idx};
  const float steps = tot[r] / step_size;
  ConstNoise noise{0.5f};
  const int limit = int(ceilf(steps)) + n_max + 2;
  if (!kWrite) {
    ValidCountSink sink;
    icdf_ray(bins, n_max, steps, -1.0f, noise, compat, limit, sink);
    counts_out[n] = sink.n;
    atomicMax(max_count, sink.n);
  } else {
    const int64_t off = offsets[n];
    StreamSink sink{out.idx + off, out.depth + off, out.dist + off, out.ray + off, r};
    icdf_ray(bins, n_max, steps, -1.0f, noise, compat, limit, sink);
  }
}

template <bool kWrite>
__global__ __launch_bounds__(kBlock) void fine_sample_kernel(
    const int32_t* __restrict__ hit_list, const int32_t* __restrict__ n_hit_dev,
    const int32_t* __restrict__ c_off, const int32_t* __restrict__ c_cnt,
    const int32_t* __restrict__ c_max_dev, SampleStream coarse, const float* __restrict__ c_prob,
    float fine_steps, int compat, int32_t* __restrict__ counts_out, int32_t* __restrict__ max_count,
    const int32_t* __restrict__ offsets, SampleStream out, int64_t n_hit_max) {
  const int64_t n = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t N = *n_hit_dev;
  if (n >= N) {
    if (!kWrite && n < n_hit_max) counts_out[n] = 0;
    return;
  }
  const int P = *c_max_dev;  // bin columns = longest coarse row
  if (P <= 0) {
    if (!kWrite) counts_out[n] = 0;
    return;
  }
  int64_t n0, nn;
  grouped_neighbours(n, N, &n0, &nn);
  const int64_t off = c_off[n];
  const int cnt = c_cnt[n];
  // Column sum of (p + 1e-5) over all P columns, left to right.
  float den = 0.0f;
  for (int k = 0; k < P; ++k) den = den + ((k < cnt ? c_prob[off + k] : 0.0f) + 1e-5f);
  SampleBins bins{coarse.idx + off, coarse.depth + off, coarse.dist + off, c_prob + off, cnt, den,
                  coarse.idx + c_off[n0], c_cnt[n0],
                  nn < 0 ? -1 : (c_cnt[nn] > 0 ? coarse.idx[c_off[nn]] : -1)};
  ConstNoise noise{0.5f};
  const int limit = int(ceilf(fine_steps)) + P + 2;
  const int32_t r = hit_list[n];
  if (!kWrite) {
    ValidCountSink sink;
    icdf_ray(bins, P, fine_steps, -1.0f, noise, compat, limit, sink);
    counts_out[n] = sink.n;
    atomicMax(max_count, sink.n);
  } else {
    const int64_t o = offsets[n];
    StreamSink sink{out.idx + o, out.depth + o, out.dist + o, out.ray + o, r};
    icdf_ray(bins, P, fine_steps, -1.0f, noise, compat, limit, sink);
  }
}

// ---------------------------------------------------------------------------
// composite

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// One wavefront per hit ray, 64 samples per step.  Free energy
// fe = relu(sigma) * dist * 7 (renderer.py:119-121); transmittance before a
// sample is exp(-(sum of the earlier fe)), an exclusive prefix sum across the
// wave plus the carry of earlier steps (renderer.py:195-197).
__global__ __launch_bounds__(kBlock) void composite_kernel(
    const int32_t* __restrict__ hit_list, const int32_t* __restrict__ n_hit_dev,
    const int32_t* __restrict__ off_arr, const int32_t* __restrict__ cnt_arr,
    const float* __restrict__ depth, const float* __restrict__ dist,
    const float4* __restrict__ field, float* __restrict__ prob_out, float* __restrict__ ray_rgb,
    float* __restrict__ ray_depth, float* __restrict__ ray_missed) {
  const int64_t n = int64_t(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (n >= *n_hit_dev) return;
  const int64_t off = off_arr[n];
  const int cnt = cnt_arr[n];
  float carry = 0.0f;
  float acc_p = 0.0f, acc_z = 0.0f, acc_r = 0.0f, acc_g = 0.0f, acc_b = 0.0f;
  for (int base = 0; base < cnt; base += 64) {
    const int k = base + lane;
    float fe = 0.0f;
    float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
    float z = 0.0f;
    if (k < cnt) {
      f = field[off + k];
      const float sg = f.x > 0.0f ? f.x : 0.0f;
      fe = (sg * dist[off + k]) * 7.0f;
      z = depth[off + k];
    }
    // inclusive scan of fe over the wave
    float incl = fe;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const float u = __shfl_up(incl, d, 64);
      if (lane >= d) incl += u;
    }
    const float excl = carry + (incl - fe);
    const float a = 1.0f - expf(-fe);
    const float tr = expf(-excl);
    const float p = k < cnt ? a * tr : 0.0f;
    if (k < cnt && prob_out) prob_out[off + k] = p;
    acc_p += p;
    acc_z += p * z;
    acc_r += p * f.y;
    acc_g += p * f.z;
    acc_b += p * f.w;
    carry += __shfl(incl, 63, 64);
  }
  acc_p = wave_sum(acc_p);
  acc_z = wave_sum(acc_z);
  acc_r = wave_sum(acc_r);
  acc_g = wave_sum(acc_g);
  acc_b = wave_sum(acc_b);
  if (lane == 0) {
    ray_rgb[n * 3 + 0] = acc_r;
    ray_rgb[n * 3 + 1] = acc_g;
    ray_rgb[n * 3 + 2] = acc_b;
    ray_depth[n] = acc_z;
    ray_missed[n] = 1.0f - acc_p;
  }
}

// Frame outputs for every ray: the hit rays' composite plus the background
// weighted by `missed`; rays that hit nothing are pure background.
__global__ __launch_bounds__(kBlock) void finish_kernel(
    int64_t R, const int32_t* __restrict__ hit_flag, const int32_t* __restrict__ rank,
    const float* __restrict__ ray_rgb, const float* __restrict__ ray_depth,
    const float* __restrict__ ray_missed, float bg_r, float bg_g, float bg_b, float bg_depth,
    float* __restrict__ rgb, float* __restrict__ depth, float* __restrict__ missed) {
  const int64_t r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (r >= R) return;
  float cr = 0.f, cg = 0.f, cb = 0.f, d = 0.f, m = 1.f;
  if (hit_flag[r]) {
    const int64_t n = rank[r];
    cr = ray_rgb[n * 3];
    cg = ray_rgb[n * 3 + 1];
    cb = ray_rgb[n * 3 + 2];
    d = ray_depth[n];
    m = ray_missed[n];
  }
  rgb[r * 3 + 0] = cr + m * bg_r;
  rgb[r * 3 + 1] = cg + m * bg_g;
  rgb[r * 3 + 2] = cb + m * bg_b;
  depth[r] = d + m * bg_depth;
  missed[r] = m;
}

}  // namespace
}  // namespace nsvf

using namespace nsvf;

extern "C" {

int nsvf_render_hits(const float* ray_start, const float* ray_dir, int64_t R,
                     const float* centers, const int32_t* children, int T, float voxelsize,
                     int n_max, void* hits, int32_t* count, float* tot, int32_t* hit_flag,
                     void* stream) {
  clear_error();
  if (R < 0 || T <= 0 || n_max <= 0) return set_error(NSVF_EINVAL, "render_hits: bad size");
  if (R == 0) return NSVF_OK;
  if (!ray_start || !ray_dir || !centers || !children || !hits || !count || !tot || !hit_flag)
    return set_error(NSVF_EINVAL, "render_hits: null pointer");
  if (T >= (1 << 27)) return set_error(NSVF_EINVAL, "render_hits: more than 2^27 nodes");
  hipLaunchKernelGGL(hits_kernel, dim3(unsigned(ceil_div(R, kHitsBlock))), dim3(kHitsBlock), 0,
                     as_stream(stream), ray_start, ray_dir, R, centers, children, T,
                     voxelsize * 0.5f, n_max, static_cast<HitRec*>(hits), count, tot, hit_flag);
  return check_launch("render_hits");
}

int nsvf_render_compact(const int32_t* hit_flag, const int32_t* rank, int64_t R, int32_t* list,
                        void* stream) {
  clear_error();
  if (R <= 0) return NSVF_OK;
  hipLaunchKernelGGL(compact_kernel, dim3(unsigned(ceil_div(R, kBlock))), dim3(kBlock), 0,
                     as_stream(stream), hit_flag, rank, R, list);
  return check_launch("render_compact");
}

int nsvf_render_sample_coarse(int write, const int32_t* hit_list, const int32_t* n_hit_dev,
                              int64_t n_hit_max, const void* hits, const int32_t* count,
                              const float* tot, int n_max, float step_size, int compat,
                              int32_t* counts_out, int32_t* max_count, const int32_t* offsets,
                              int32_t* s_idx, float* s_depth, float* s_dist, int32_t* s_ray,
                              void* stream) {
  clear_error();
  if (n_hit_max <= 0) return NSVF_OK;
  const dim3 grid(unsigned(ceil_div(n_hit_max, kBlock)));
  SampleStream out{s_idx, s_depth, s_dist, s_ray};
  const HitRec* h = static_cast<const HitRec*>(hits);
  if (write)
    hipLaunchKernelGGL(coarse_sample_kernel<true>, grid, dim3(kBlock), 0, as_stream(stream),
                       hit_list, n_hit_dev, h, count, tot, n_max, step_size, compat, counts_out,
                       max_count, offsets, out, n_hit_max);
  else
    hipLaunchKernelGGL(coarse_sample_kernel<false>, grid, dim3(kBlock), 0, as_stream(stream),
                       hit_list, n_hit_dev, h, count, tot, n_max, step_size, compat, counts_out,
                       max_count, offsets, out, n_hit_max);
  return check_launch("render_sample_coarse");
}

int nsvf_render_sample_fine(int write, const int32_t* hit_list, const int32_t* n_hit_dev,
                            int64_t n_hit_max, const int32_t* c_off, const int32_t* c_cnt,
                            const int32_t* c_max, const int32_t* c_idx, const float* c_depth,
                            const float* c_dist, const float* c_prob, float fine_steps, int compat,
                            int32_t* counts_out, int32_t* max_count, const int32_t* offsets,
                            int32_t* s_idx, float* s_depth, float* s_dist, int32_t* s_ray,
                            void* stream) {
  clear_error();
  if (n_hit_max <= 0) return NSVF_OK;
  const dim3 grid(unsigned(ceil_div(n_hit_max, kBlock)));
  SampleStream coarse{const_cast<int32_t*>(c_idx), const_cast<float*>(c_depth),
                      const_cast<float*>(c_dist), nullptr};
  SampleStream out{s_idx, s_depth, s_dist, s_ray};
  if (write)
    hipLaunchKernelGGL(fine_sample_kernel<true>, grid, dim3(kBlock), 0, as_stream(stream),
                       hit_list, n_hit_dev, c_off, c_cnt, c_max, coarse, c_prob, fine_steps, compat,
                       counts_out, max_count, offsets, out, n_hit_max);
  else
    hipLaunchKernelGGL(fine_sample_kernel<false>, grid, dim3(kBlock), 0, as_stream(stream),
                       hit_list, n_hit_dev, c_off, c_cnt, c_max, coarse, c_prob, fine_steps, compat,
                       counts_out, max_count, offsets, out, n_hit_max);
  return check_launch("render_sample_fine");
}

int nsvf_render_composite(const int32_t* hit_list, const int32_t* n_hit_dev, int64_t n_hit_max,
                          const int32_t* off, const int32_t* cnt, const float* depth,
                          const float* dist, const float* field, float* prob_out,
                          float* ray_rgb, float* ray_depth, float* ray_missed, void* stream) {
  clear_error();
  if (n_hit_max <= 0) return NSVF_OK;
  const int per_block = kBlock / 64;
  hipLaunchKernelGGL(composite_kernel, dim3(unsigned(ceil_div(n_hit_max, per_block))),
                     dim3(kBlock), 0, as_stream(stream), hit_list, n_hit_dev, off, cnt, depth, dist,
                     reinterpret_cast<const float4*>(field), prob_out, ray_rgb, ray_depth,
                     ray_missed);
  return check_launch("render_composite");
}

int nsvf_render_finish(int64_t R, const int32_t* hit_flag, const int32_t* rank,
                       const float* ray_rgb, const float* ray_depth, const float* ray_missed,
                       float bg_r, float bg_g, float bg_b, float bg_depth, float* rgb,
                       float* depth, float* missed, void* stream) {
  clear_error();
  if (R <= 0) return NSVF_OK;
  hipLaunchKernelGGL(finish_kernel, dim3(unsigned(ceil_div(R, kBlock))), dim3(kBlock), 0,
                     as_stream(stream), R, hit_flag, rank, ray_rgb, ray_depth, ray_missed, bg_r,
                     bg_g, bg_b, bg_depth, rgb, depth, missed);
  return check_launch("render_finish");
}

}  // extern "C"
