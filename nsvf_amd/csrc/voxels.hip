// Voxel splitting (fairnr/data/geometry.py:250-274 via encoder.py:656-676):
// every voxel becomes 8 children of half the size, the corner-key set is
// rebuilt from the children's corners, and each new corner's embedding is the
// trilinear interpolation of its parent's 8 corner embeddings.
//
// Integer work, done as in the reference but on the device:
//   coords = round((p - min p) / q)      q = vs / 4, per axis (fp32, ties even)
//   child  = coords + o_ch, corner = child + o_c,  o in {-1, +1}^3 (4x+2y+z)
//   keys   = unique(sorted) of the 64 corners of each voxel's children
//   feats  = each child corner's index in keys (the unique inverse)
// The 64N corners are packed into uint64 keys ((x+4) << 42 | (y+4) << 21 |
// z+4, which orders like the reference's lexicographic unique over rows),
// stable-radix-sorted with their positions, and the first position of each
// run of equal keys gives the key's id (a scan of run heads) and its parent
// voxel -- the lowest index among the voxels sharing the corner.  The
// reference takes the parent from a scatter_ whose winner among duplicates
// the device chooses (geometry.py:260-261); any parent gives the same value
// up to rounding, and this one is deterministic.
// Values: p = (key - coords[parent]) * 0.25 + 0.5 per axis (fp32), weights
// (wx * wy) * wz with w = p or 1 - p by corner bit, sum over the parent's
// corners in corner order (geometry.py:195-200, 264-271).
#include <hipcub/hipcub.hpp>

#include "common.h"

namespace nsvf {
namespace {

constexpr int kBlock = 256;
constexpr int64_t kCoordLimit = (int64_t(1) << 21) - 8;

inline int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

struct SplitWorkspace {
  uint64_t* keys_in;   // [64N]
  uint64_t* keys_out;
  int32_t* pos_in;     // [64N]
  int32_t* pos_out;
  int32_t* head;       // [64N] run heads of the sorted keys
  int32_t* rank;       // [64N] exclusive scan of head
  int32_t* scan_ws;
  uint64_t* ukeys;     // [64N] (first K' used) unique keys
  int32_t* parent;     // [64N] (first K' used) parent voxel of each key
  float* bounds;       // [8]: min xyz, pad, max xyz, pad
  int32_t* count;      // [2]: K', pad
  void* sort_tmp;
  size_t sort_tmp_bytes;
  int64_t total;
};

int split_layout(int64_t N, hipStream_t stream, char* base, SplitWorkspace* w) {
  const int64_t n = 64 * N;
  size_t tb = 0;
  const hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      nullptr, tb, static_cast<const uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
      static_cast<const int32_t*>(nullptr), static_cast<int32_t*>(nullptr), int(n), 0, 63, stream);
  if (e != hipSuccess)
    return set_error(NSVF_ELAUNCH, "split: radix sort size query: %s", hipGetErrorString(e));
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    char* p = base ? base + o : nullptr;
    o += align256(bytes);
    return p;
  };
  w->keys_in = reinterpret_cast<uint64_t*>(take(n * 8));
  w->keys_out = reinterpret_cast<uint64_t*>(take(n * 8));
  w->pos_in = reinterpret_cast<int32_t*>(take(n * 4));
  w->pos_out = reinterpret_cast<int32_t*>(take(n * 4));
  w->head = reinterpret_cast<int32_t*>(take(n * 4));
  w->rank = reinterpret_cast<int32_t*>(take(n * 4));
  w->scan_ws = reinterpret_cast<int32_t*>(take(nsvf_scan_workspace_ints(n) * 4));
  w->ukeys = reinterpret_cast<uint64_t*>(take(n * 8));
  w->parent = reinterpret_cast<int32_t*>(take(n * 4));
  w->bounds = reinterpret_cast<float*>(take(32));
  w->count = reinterpret_cast<int32_t*>(take(8));
  w->sort_tmp_bytes = tb;
  w->sort_tmp = take(int64_t(tb));
  w->total = o;
  return NSVF_OK;
}

// Per-axis min and max of the points (one workgroup; min/max are exact in
// any order).
__global__ __launch_bounds__(1024) void bounds_kernel(const float* __restrict__ p, int64_t N,
                                                      float* __restrict__ bounds) {
  __shared__ float red[6][1024];
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = threadIdx.x; i < N; i += 1024)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      mn[a] = fminf(mn[a], p[i * 3 + a]);
      mx[a] = fmaxf(mx[a], p[i * 3 + a]);
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    red[a][threadIdx.x] = mn[a];
    red[3 + a][threadIdx.x] = mx[a];
  }
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
        red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      bounds[a] = red[a][0];
      bounds[4 + a] = red[3 + a][0];
    }
}

// discretize_points (geometry.py:241-247) of one axis at quarter-voxel units.
__device__ __forceinline__ int coord_of(float p, float mn, float q) {
  return int(rintf((p - mn) / q));
}

// Thread per (voxel, child, corner): the packed corner key; corner 0's thread
// also writes the child centre p + o_ch * q (geometry.py:252).
__global__ __launch_bounds__(kBlock) void split_keys_kernel(
    const float* __restrict__ points, int64_t N, float q, const float* __restrict__ bounds,
    uint64_t* __restrict__ keys, int32_t* __restrict__ pos, float* __restrict__ new_points) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= 64 * N) return;
  const int64_t n = i >> 6;
  const int ch = int(i >> 3) & 7, c = int(i) & 7;
  uint64_t key = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int bit = 2 - a;
    const int och = ((ch >> bit) & 1) ? 1 : -1;
    const int oc = ((c >> bit) & 1) ? 1 : -1;
    const float p = points[n * 3 + a];
    const int k = coord_of(p, bounds[a], q) + och + oc;
    key = (key << 21) | uint64_t(k + 4);
    if (c == 0) new_points[((n * 8 + ch) * 3) + a] = p + float(och) * q;
  }
  keys[i] = key;
  pos[i] = int32_t(i);
}

__global__ __launch_bounds__(kBlock) void run_heads_kernel(const uint64_t* __restrict__ keys,
                                                           int64_t n, int32_t* __restrict__ head) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// Sorted position i -> key id u: feats of the original position, and the
// key and parent voxel at the run's first position.
__global__ __launch_bounds__(kBlock) void split_ids_kernel(
    const uint64_t* __restrict__ keys, const int32_t* __restrict__ pos,
    const int32_t* __restrict__ head, const int32_t* __restrict__ rank, int64_t n,
    int32_t* __restrict__ new_feats, uint64_t* __restrict__ ukeys,
    int32_t* __restrict__ parent) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t u = rank[i] + head[i] - 1;
  new_feats[pos[i]] = u;
  if (head[i]) {
    ukeys[u] = keys[i];
    parent[u] = pos[i] >> 6;
  }
}

// Thread per (key, feature): the interpolated embedding; feature 0's thread
// also writes the key's integer coordinates.
__global__ __launch_bounds__(kBlock) void split_values_kernel(
    const float* __restrict__ points, const int32_t* __restrict__ feats,
    const float* __restrict__ values, int D, float q, const float* __restrict__ bounds,
    const uint64_t* __restrict__ ukeys, const int32_t* __restrict__ parent, int64_t K,
    int64_t* __restrict__ new_keys, float* __restrict__ new_values) {
  const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= K * D) return;
  const int64_t u = t / D;
  const int d = int(t - u * D);
  const uint64_t key = ukeys[u];
  const int64_t n = parent[u];
  float p[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int k = int((key >> (21 * (2 - a))) & ((uint64_t(1) << 21) - 1)) - 4;
    const int c0 = coord_of(points[n * 3 + a], bounds[a], q);
    p[a] = float(k - c0) * 0.25f + 0.5f;
    if (d == 0) new_keys[u * 3 + a] = k;
  }
  const float r[3] = {1.0f - p[0], 1.0f - p[1], 1.0f - p[2]};
  float acc = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float w = (((c & 4) ? p[0] : r[0]) * ((c & 2) ? p[1] : r[1])) * ((c & 1) ? p[2] : r[2]);
    acc = acc + w * values[int64_t(feats[n * 8 + c]) * D + d];
  }
  new_values[t] = acc;
}

}  // namespace
}  // namespace nsvf

using namespace nsvf;

extern "C" {

int64_t nsvf_split_workspace_bytes(int64_t N, void* stream) {
  clear_error();
  if (N <= 0 || 64 * N >= (int64_t(1) << 31)) {
    set_error(NSVF_EINVAL, "split_workspace_bytes: bad voxel count");
    return -1;
  }
  SplitWorkspace w;
  if (split_layout(N, as_stream(stream), nullptr, &w)) return -1;
  return w.total;
}

int nsvf_split_voxels(const float* points, int64_t N, float voxel_size, void* workspace,
                      int64_t workspace_bytes, float* new_points, int32_t* new_feats,
                      int32_t* num_keys, void* stream) {
  clear_error();
  if (N <= 0 || 64 * N >= (int64_t(1) << 31))
    return set_error(NSVF_EINVAL, "split_voxels: bad voxel count");
  if (!points || !workspace || !new_points || !new_feats || !num_keys)
    return set_error(NSVF_EINVAL, "split_voxels: null pointer");
  hipStream_t st = as_stream(stream);
  SplitWorkspace w;
  int rc = split_layout(N, st, static_cast<char*>(workspace), &w);
  if (rc) return rc;
  if (workspace_bytes < w.total)
    return set_error(NSVF_EINVAL, "split_voxels: workspace too small");
  const int64_t n = 64 * N;
  // geometry.py:251: quarter = (vs / 2) * .5
  const float q = (voxel_size * 0.5f) * 0.5f;
  hipLaunchKernelGGL(bounds_kernel, dim3(1), dim3(1024), 0, st, points, N, w.bounds);
  // the extent must fit the 21-bit packed coordinates
  float b[8];
  if (hipMemcpyAsync(b, w.bounds, sizeof(b), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_error(NSVF_ELAUNCH, "split_voxels: bounds: %s", hipGetErrorString(hipGetLastError()));
  for (int a = 0; a < 3; ++a)
    if (!(double(b[4 + a] - b[a]) / q < double(kCoordLimit - 8)))
      return set_error(NSVF_EINVAL, "split_voxels: extent beyond %lld quarter voxels",
                       (long long)(kCoordLimit - 8));
  hipLaunchKernelGGL(split_keys_kernel, dim3(unsigned(ceil_div(n, kBlock))), dim3(kBlock), 0, st,
                     points, N, q, w.bounds, w.keys_in, w.pos_in, new_points);
  size_t tb = w.sort_tmp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, tb, w.keys_in, w.keys_out,
                                                    w.pos_in, w.pos_out, int(n), 0, 63, st);
  if (e != hipSuccess)
    return set_error(NSVF_ELAUNCH, "split_voxels: sort: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(run_heads_kernel, dim3(unsigned(ceil_div(n, kBlock))), dim3(kBlock), 0, st,
                     w.keys_out, n, w.head);
  rc = nsvf_scan_exclusive_i32(w.head, w.rank, n, w.scan_ws, w.count, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(split_ids_kernel, dim3(unsigned(ceil_div(n, kBlock))), dim3(kBlock), 0, st,
                     w.keys_out, w.pos_out, w.head, w.rank, n, new_feats, w.ukeys, w.parent);
  if (hipMemcpyAsync(num_keys, w.count, sizeof(int32_t), hipMemcpyDeviceToDevice, st) !=
      hipSuccess)
    return set_error(NSVF_ELAUNCH, "split_voxels: count copy");
  return check_launch("split_voxels");
}

int nsvf_split_values(const float* points, const int32_t* feats, const float* values,
                      int64_t N, int D, float voxel_size, const void* workspace,
                      int64_t workspace_bytes, int64_t K_new, int64_t* new_keys,
                      float* new_values, void* stream) {
  clear_error();
  if (N <= 0 || 64 * N >= (int64_t(1) << 31) || D <= 0 || K_new < 0 || K_new > 64 * N)
    return set_error(NSVF_EINVAL, "split_values: bad size");
  if (K_new == 0) return NSVF_OK;
  if (!points || !feats || !values || !workspace || !new_keys || !new_values)
    return set_error(NSVF_EINVAL, "split_values: null pointer");
  hipStream_t st = as_stream(stream);
  SplitWorkspace w;
  int rc = split_layout(N, st, static_cast<char*>(const_cast<void*>(workspace)), &w);
  if (rc) return rc;
  if (workspace_bytes < w.total)
    return set_error(NSVF_EINVAL, "split_values: workspace too small");
  const float q = (voxel_size * 0.5f) * 0.5f;
  hipLaunchKernelGGL(split_values_kernel, dim3(unsigned(ceil_div(K_new * D, kBlock))),
                     dim3(kBlock), 0, st, points, feats, values, D, q, w.bounds, w.ukeys, w.parent,
                     K_new, new_keys, new_values);
  return check_launch("split_values");
}

}  // extern "C"
