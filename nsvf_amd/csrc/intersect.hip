// Ray / sparse-voxel intersection kernels for gfx950.
//
// aabb: brute force over every voxel, the voxel centres staged through LDS in
//       tiles that all 256 lanes of a workgroup test against (one lane per
//       ray; LDS reads are wave-wide broadcasts).  The reciprocal of the ray
//       direction is computed once per ray.
// svo:  depth-first traversal of the loose octree, one lane per ray.  The
//       traversal keeps one (node, next child slot) entry per tree level in
//       LDS, so the per-lane state is `depth` words instead of the
//       reference's 256-entry stack in scratch memory; the visit order is the
//       reference's (children popped from slot 7 down to slot 0).
//
// Semantics follow fairnr/clib/src/intersect_gpu.cu:73-237 and
// intersect.cpp:49-112: first n_max hits per ray, idx -1 padded, depths 0
// padded (the kernels write the pads, so outputs need no pre-fill).
#include "common.h"

namespace nsvf {
namespace {

constexpr int kAabbBlock = 256;
constexpr int kAabbTile = 1024;

__global__ __launch_bounds__(kAabbBlock) void aabb_intersect_kernel(
    const float* __restrict__ ray_start, const float* __restrict__ ray_dir,
    const float* __restrict__ points, int M, int N, int64_t points_bstride,
    int blocks_per_batch, float half_voxel, int n_max,
    int32_t* __restrict__ idx, float* __restrict__ min_depth,
    float* __restrict__ max_depth) {
  __shared__ float tile[3][kAabbTile];
  const int b = blockIdx.x / blocks_per_batch;
  const int j = (blockIdx.x % blocks_per_batch) * kAabbBlock + threadIdx.x;
  const bool active = j < M;
  const int64_t r = int64_t(b) * M + j;
  const float* pts = points + b * points_bstride;

  float o[3] = {0.f, 0.f, 0.f}, inv[3] = {0.f, 0.f, 0.f};
  if (active) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      o[a] = ray_start[r * 3 + a];
      inv[a] = 1.0f / ray_dir[r * 3 + a];
    }
  }
  int32_t* oi = idx + r * n_max;
  float* olo = min_depth + r * n_max;
  float* ohi = max_depth + r * n_max;
  int cnt = 0;
  bool done = !active;

  for (int base = 0; base < N; base += kAabbTile) {
    const int n = min(kAabbTile, N - base);
    __syncthreads();
    for (int t = threadIdx.x; t < n * 3; t += kAabbBlock) {
      const int k = t / 3, a = t - k * 3;
      tile[a][k] = pts[int64_t(base) * 3 + t];
    }
    __syncthreads();
    if (!done) {
      for (int k = 0; k < n; ++k) {
        float lo, hi;
        if (ray_aabb(o, inv, tile[0][k], tile[1][k], tile[2][k], half_voxel, &lo, &hi)) {
          oi[cnt] = base + k;
          olo[cnt] = lo;
          ohi[cnt] = hi;
          if (++cnt == n_max) {
            done = true;
            break;
          }
        }
      }
    }
    if (__syncthreads_and(done ? 1 : 0)) break;
  }
  if (active) {
    for (int l = cnt; l < n_max; ++l) {
      oi[l] = -1;
      olo[l] = 0.0f;
      ohi[l] = 0.0f;
    }
  }
}

constexpr int kSvoBlock = 128;
constexpr int kSvoMaxLevels = 30;

// Depth-first traversal, one lane per ray.  stack[level] packs the node id
// (high bits) and the number of child slots still to visit (low 4 bits);
// slots are visited from 7 down to 0, the order in which the reference pops
// them after pushing 0..7.
//
// Guards that keep every lane finite on any input: child ids outside [0,T)
// are skipped, a child is descended into only if its scale is smaller than
// its parent's (so a malformed tree cannot cycle), and at most
// kSvoMaxLevels internal levels are held (a valid tree over int32
// coordinates has at most 31; deeper subtrees are not descended).
__global__ __launch_bounds__(kSvoBlock) void svo_intersect_kernel(
    const float* __restrict__ ray_start, const float* __restrict__ ray_dir,
    const float* __restrict__ centers, const int32_t* __restrict__ children,
    int64_t B_M, int M, int T, int64_t centers_bstride, int64_t children_bstride,
    float half_voxel, int n_max, int32_t* __restrict__ idx,
    float* __restrict__ min_depth, float* __restrict__ max_depth) {
  __shared__ int32_t stack[kSvoMaxLevels][kSvoBlock];
  const int64_t r = int64_t(blockIdx.x) * kSvoBlock + threadIdx.x;
  if (r >= B_M) return;
  const int b = int(r / M);
  const float* cen = centers + b * centers_bstride;
  const int32_t* ch = children + b * children_bstride;
  int32_t* st = &stack[0][threadIdx.x];

  float o[3], inv[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    o[a] = ray_start[r * 3 + a];
    inv[a] = 1.0f / ray_dir[r * 3 + a];
  }
  int32_t* oi = idx + r * n_max;
  float* olo = min_depth + r * n_max;
  float* ohi = max_depth + r * n_max;
  int cnt = 0;
  int level = -1;
  {
    const int root = T - 1;
    const int sc = ch[int64_t(root) * 9 + 8];
    float lo, hi;
    if (n_max > 0 &&
        ray_aabb(o, inv, cen[int64_t(root) * 3], cen[int64_t(root) * 3 + 1],
                 cen[int64_t(root) * 3 + 2], half_voxel * float(sc), &lo, &hi)) {
      if (sc == 1) {
        oi[0] = root;
        olo[0] = lo;
        ohi[0] = hi;
        cnt = 1;
      } else {
        level = 0;
        st[0] = (root << 4) | 8;
      }
    }
  }
  while (level >= 0 && cnt < n_max) {
    const int32_t e = st[level * kSvoBlock];
    const int node = e >> 4;
    const int left = e & 15;
    if (left == 0) {
      --level;
      continue;
    }
    st[level * kSvoBlock] = (node << 4) | (left - 1);
    const int c = ch[int64_t(node) * 9 + (left - 1)];
    if (c < 0 || c >= T) continue;
    const int sc = ch[int64_t(c) * 9 + 8];
    float lo, hi;
    if (!ray_aabb(o, inv, cen[int64_t(c) * 3], cen[int64_t(c) * 3 + 1],
                  cen[int64_t(c) * 3 + 2], half_voxel * float(sc), &lo, &hi))
      continue;
    if (sc == 1) {
      oi[cnt] = c;
      olo[cnt] = lo;
      ohi[cnt] = hi;
      ++cnt;
      continue;
    }
    const int ps = ch[int64_t(node) * 9 + 8];
    if (sc < ps && level + 1 < kSvoMaxLevels) {
      ++level;
      st[level * kSvoBlock] = (c << 4) | 8;
    }
  }
  for (int l = cnt; l < n_max; ++l) {
    oi[l] = -1;
    olo[l] = 0.0f;
    ohi[l] = 0.0f;
  }
}

}  // namespace
}  // namespace nsvf

using namespace nsvf;

extern "C" {

int nsvf_aabb_intersect(const float* ray_start, const float* ray_dir,
                        const float* points, int B, int M, int N,
                        int points_shared, float voxelsize, int n_max,
                        int32_t* idx, float* min_depth, float* max_depth,
                        void* stream) {
  clear_error();
  if (B < 0 || M < 0 || N < 0 || n_max < 0)
    return set_error(NSVF_EINVAL, "aabb_intersect: negative size");
  if (int64_t(B) * M == 0 || n_max == 0) return NSVF_OK;
  if (!ray_start || !ray_dir || !idx || !min_depth || !max_depth || (N > 0 && !points))
    return set_error(NSVF_EINVAL, "aabb_intersect: null pointer");
  const int bpb = int(ceil_div(M, kAabbBlock));
  const int64_t blocks = int64_t(bpb) * B;
  if (blocks > INT32_MAX) return set_error(NSVF_EINVAL, "aabb_intersect: too many rays");
  const int64_t pstride = points_shared ? 0 : int64_t(N) * 3;
  hipLaunchKernelGGL(aabb_intersect_kernel, dim3(unsigned(blocks)), dim3(kAabbBlock), 0,
                     as_stream(stream), ray_start, ray_dir, points, M, N, pstride, bpb,
                     voxelsize * 0.5f, n_max, idx, min_depth, max_depth);
  return check_launch("aabb_intersect");
}

int nsvf_svo_intersect(const float* ray_start, const float* ray_dir,
                       const float* points, const int32_t* children, int B,
                       int M, int T, int tree_shared, float voxelsize,
                       int n_max, int32_t* idx, float* min_depth,
                       float* max_depth, void* stream) {
  clear_error();
  if (B < 0 || M < 0 || T < 0 || n_max < 0)
    return set_error(NSVF_EINVAL, "svo_intersect: negative size");
  if (int64_t(B) * M == 0 || n_max == 0) return NSVF_OK;
  if (T == 0) return set_error(NSVF_EINVAL, "svo_intersect: empty tree");
  if (T >= (1 << 27)) return set_error(NSVF_EINVAL, "svo_intersect: more than 2^27 nodes");
  if (!ray_start || !ray_dir || !points || !children || !idx || !min_depth || !max_depth)
    return set_error(NSVF_EINVAL, "svo_intersect: null pointer");
  const int64_t BM = int64_t(B) * M;
  const int64_t blocks = ceil_div(BM, kSvoBlock);
  if (blocks > INT32_MAX) return set_error(NSVF_EINVAL, "svo_intersect: too many rays");
  const int64_t cstride = tree_shared ? 0 : int64_t(T) * 3;
  const int64_t hstride = tree_shared ? 0 : int64_t(T) * 9;
  hipLaunchKernelGGL(svo_intersect_kernel, dim3(unsigned(blocks)), dim3(kSvoBlock), 0,
                     as_stream(stream), ray_start, ray_dir, points, children, BM, M, T, cstride,
                     hstride, voxelsize * 0.5f, n_max, idx, min_depth, max_depth);
  return check_launch("svo_intersect");
}

}  // extern "C"
