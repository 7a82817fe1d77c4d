// Fused render pipeline kernels (one view or a batch of rays, scene resident
// on the device).  Stages, each one kernel:
//
//   hits      octree traversal + depth sort + segment-length sum per ray
//             (encoder.py:498-535 with nsvf.py:65-74 folded in)
//   compact   ray ranks -> list of hit rays (fairnr_model.py:157-162)
//   sample    inverse-CDF sampling of a ray's bins into a sample stream
//             (encoder.py:538-556, clib/__init__.py:231-300); coarse bins are
//             the sorted hits, fine bins are the coarse samples
//             (nerf.py:64-79).  Each ray owns a slot range sized by an upper
//             bound on its sample count, known before sampling (ceil(steps) +
//             bins + 2), so one pass writes the samples and unused slots are
//             marked with voxel id -1; no counting pass, no compaction.
//   composite front-to-back compositing of one ray's samples by one
//             wavefront with a prefix scan (renderer.py:193-218)
//   finish    background fill of every ray of the frame (nsvf.py:89-101)
//
// Sample streams are per-ray slot ranges (offset from a device scan of the
// bounds, count of written samples) of 16-byte records {voxel id, midpoint
// depth, interval length, ray id}.  No padded [R, S] sample tensor exists on
// this path; padding only appears in the op-level kernels that keep the
// reference's layout.
#include <algorithm>

#include "render.h"
#include "sample_core.h"
#include "sample_wave.h"

namespace nsvf {
namespace {

constexpr int kBlock = 256;
constexpr int kHitsBlock = 128;
constexpr int kMaxLevels = 24;

// ---------------------------------------------------------------------------
// hits

// Slab test against a box given by its corners; with mn = fp32(c - h) and
// mx = fp32(c + h) this is the reference's test (intersect_gpu.cu:93-94
// evaluates (c - h) - o and (c + h) - o) bit for bit.  The reference returns
// early (per axis: hi < f_low, lo > f_high, then f_low > f_high); here every
// axis runs and only the final f_low > f_high decides, which is the same
// test: f_low only grows and f_high only shrinks, so any early exit leaves
// f_low > f_high at the end, and without one the running values are the
// reference's.  The swap and the updates are compare-and-selects, as the
// reference's ternaries, so an axis whose bounds come out NaN (direction 0
// and origin on the slab plane) fails every comparison and changes nothing
// (a min / max pair would not: it drops the NaN operand).  No branches: the
// traversal runs one test per child and lane without divergence.
__device__ __forceinline__ bool ray_box(const float o[3], const float inv[3], const float4& mn,
                                        const float4& mx, float* out_lo, float* out_hi) {
  float f_low = 0.0f;
  float f_high = 100000.0f;
  const float a0[3] = {mn.x, mn.y, mn.z};
  const float a1[3] = {mx.x, mx.y, mx.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float l0 = (a0[a] - o[a]) * inv[a];
    const float h0 = (a1[a] - o[a]) * inv[a];
    const bool sw = h0 < l0;
    const float lo = sw ? h0 : l0;
    const float hi = sw ? l0 : h0;
    f_low = (lo > f_low) ? lo : f_low;
    f_high = (hi < f_high) ? hi : f_high;
  }
  *out_lo = f_low;
  *out_hi = f_high;
  return !(f_low > f_high);
}

// A ray's hit row as the traversal fills it, with what the kernel needs
// afterwards kept on the fly: whether entry depths strictly increase, and
// the left-to-right sum of the segment lengths (in append order, which is
// the row order whenever the row is kept as is).
struct HitRow {
  HitRec* row;
  int cnt;
  float prev_lo;
  float tot;
  bool increasing;
  __device__ __forceinline__ void push(float lo, float hi, int32_t id) {
    row[cnt] = HitRec{lo, hi, id, 0};
    increasing = increasing && (cnt == 0 || prev_lo < lo);
    prev_lo = lo;
    tot = tot + (hi - lo);
    ++cnt;
  }
};

// Wave-union traversal of the octree packed as node records: record k (16
// float4) holds, for child slot s, piece 2s = (box min xyz, scale bits) and
// piece 2s+1 = (box max xyz, child id bits), id -1 for an empty slot; record
// T holds the root in slot 0.  Leaf boxes are the voxel boxes; an internal
// node's box is the exact bounding box of its leaves' boxes (computed with
// the same fp32 values), so it prunes every subtree the ray misses while
// keeping every leaf the ray hits -- the slab test is monotone under box
// containment -- and the hit set equals the reference's, whose loose boxes
// (twice the octant) only prune less.
//
// The lanes of `group` (rays sharing the child order dmask) walk ONE node
// stack: a node is expanded when any of them hits it, so its record is
// wave-uniform -- one coalesced 256-byte load into this level's LDS slot,
// read back by broadcast -- and each lane only runs its own slab tests.  A
// lane's hits come out in the same order as from a walk of its own -- depth
// first, children in order slot = i ^ dmask -- because the union walk visits
// a superset of its nodes in that order and a lane that misses a node misses
// all of its descendants (containment again).  Expanding a node tests all
// its children at once; if every child hit by some lane is a leaf, each lane
// emits its hits right away in order, otherwise the union hit mask goes on
// the stack and the children are visited from it in order (a leaf among
// them is re-tested when visited).  The stack of masks lives in one register
// across the lanes (level l in lane l), the records of the open nodes in
// LDS.
__device__ __forceinline__ unsigned expand_union(int node, int psc, const float o[3],
                                                 const float inv[3], int dmask, bool act,
                                                 const float4* __restrict__ recs, int T,
                                                 int n_max, float4* lrec, HitRow& out) {
  const int lane = threadIdx.x & 63;
  if (lane < 16) lrec[lane] = recs[int64_t(node) * 16 + lane];
  wave_sync();
  const bool live = act && out.cnt < n_max;
  unsigned U = 0, leafs = 0, mine = 0;
  float los[8], his[8];
  int ids[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int s = i ^ dmask;
    const float4 mn = lrec[2 * s], mx = lrec[2 * s + 1];
    const int id = __builtin_amdgcn_readfirstlane(__float_as_int(mx.w));
    const int sc = __builtin_amdgcn_readfirstlane(__float_as_int(mn.w));
    ids[i] = id;
    los[i] = 0.0f;
    his[i] = 0.0f;
    if (unsigned(id) < unsigned(T) && sc < psc) {
      // every lane runs the (branch-free) test; a lane's bounds are read
      // only where its bit in `mine` is set
      const bool h = ray_box(o, inv, mn, mx, &los[i], &his[i]) & live;
      mine |= unsigned(h) << i;
      if (__ballot(h)) {
        U |= 1u << i;
        if (sc == 1) leafs |= 1u << i;
      }
    }
  }
  if ((U & ~leafs) == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (((mine >> i) & 1u) && out.cnt < n_max) out.push(los[i], his[i], ids[i]);
    return 0u;
  }
  return U;
}

// Appends the hits of the lanes with `act` set, all of which share dmask.
// lrec: this wave's LDS, kMaxLevels records.
__device__ void traverse_union(const float o[3], const float inv[3], int dmask, bool act,
                               const float4* __restrict__ recs, int T, int n_max, float4* lrec,
                               HitRow& out) {
  const bool here0 = (threadIdx.x & 63) == 0;
  int st_mask = 0;  // level l's remaining children in lane l
  int level = 0;
  {
    // record T: the root in slot 0 under a parent of unbounded scale
    const unsigned m = expand_union(T, 0x7fffffff, o, inv, 0, act, recs, T, n_max, lrec, out);
    st_mask = here0 ? int(m) : st_mask;
  }
  while (level >= 0) {
    if (!__ballot(act && out.cnt < n_max)) break;
    const unsigned m = unsigned(__builtin_amdgcn_readlane(st_mask, level));
    if (m == 0) {
      --level;
      continue;
    }
    const int i = __builtin_ctz(m);
    const bool here = int(threadIdx.x & 63) == level;
    st_mask = here ? int(m & (m - 1)) : st_mask;
    // record T's only child is the root, visited in slot 0 whatever dmask
    const int s = level == 0 ? i : (i ^ dmask);
    const float4* rec = lrec + level * 16;
    const float4 mn = rec[2 * s], mx = rec[2 * s + 1];
    const int id = __builtin_amdgcn_readfirstlane(__float_as_int(mx.w));
    const int sc = __builtin_amdgcn_readfirstlane(__float_as_int(mn.w));
    if (sc == 1) {
      float lo = 0.0f, hi = 0.0f;
      if (act && out.cnt < n_max && ray_box(o, inv, mn, mx, &lo, &hi)) out.push(lo, hi, id);
      continue;
    }
    if (level + 1 < kMaxLevels) {
      ++level;
      const unsigned mm =
          expand_union(id, sc, o, inv, dmask, act, recs, T, n_max, lrec + level * 16, out);
      const bool there = int(threadIdx.x & 63) == level;
      st_mask = there ? int(mm) : st_mask;
    }
  }
}

// The ray of pixel (row i, column j) of a camera (geometry.py:76-89,116-122,
// data_utils.py:57-65): uv = (j * su, i * sv) with su = float(W / w), sv =
// float(H / h); x = (u - cx) / fx, y = (v - cy) / fy (times z = 1); rt =
// [R | t] [x, y, 1, 1]; d = rt - t; dir = d / (|d| + 1e-8).  Every product
// and sum rounded on its own, sums left to right.
__device__ __forceinline__ void camera_ray(const nsvf_camera& cam, float su, float sv, int i,
                                           int j, float o[3], float d[3]) {
  const float u = float(j) * su, v = float(i) * sv;
  const float x = (u - cam.cx) / cam.fx, y = (v - cam.cy) / cam.fy;
  float q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float* m = cam.inv_rt + 4 * a;
    const float rt = ((m[0] * x + m[1] * y) + m[2]) + m[3];
    o[a] = m[3];
    q[a] = rt - m[3];
  }
  const float n = sqrtf((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + 1e-8f;
#pragma unroll
  for (int a = 0; a < 3; ++a) d[a] = q[a] / n;
}

__global__ __launch_bounds__(kBlock) void ray_gen_kernel(nsvf_camera cam, float su, float sv,
                                                         float* __restrict__ ray_start,
                                                         float* __restrict__ ray_dir) {
  const int64_t r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (r >= int64_t(cam.h) * cam.w) return;
  float o[3], d[3];
  camera_ray(cam, su, sv, int(r / cam.w), int(r % cam.w), o, d);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    ray_start[r * 3 + a] = o[a];
    ray_dir[r * 3 + a] = d[a];
  }
}

// Per ray: hits sorted by entry depth, count, and the left-to-right sum of
// the segment lengths.  The front-to-back traversal yields the hits already
// sorted (leaves are disjoint voxels and a ray crosses octants in increasing
// (slot ^ mask) order), so it is the common path.  A ray whose hits reach
// n_max (the kept subset then depends on traversal order), or whose entry
// depths are not strictly increasing (equal entries: the tie order must be
// the reference's), is traversed again in the reference's order -- child
// slots 7..0 (intersect_gpu.cu:201-230), which is slot = i ^ 7 -- and stably
// insertion-sorted, which is exactly the reference pipeline's result: the
// exact boxes prune only subtrees without a hit, so both traversals emit the
// same leaves in the same order and stop at the same n_max-th hit.
template <bool kCamera>
__global__ __launch_bounds__(kHitsBlock) void hits_kernel(
    const float* __restrict__ ray_start, const float* __restrict__ ray_dir, int64_t R, int width,
    nsvf_camera cam, float su, float sv, float* __restrict__ gen_start,
    float* __restrict__ gen_dir, const float4* __restrict__ recs, int T, int n_max,
    float step_size,
    HitRec* __restrict__ hits, int32_t* __restrict__ count, float* __restrict__ tot_out,
    int32_t* __restrict__ hit_flag, int32_t* __restrict__ bound,
    const int32_t* __restrict__ order) {
  __shared__ float4 lrec_all[kHitsBlock / 64][kMaxLevels * 16];
  float4* lrec = lrec_all[threadIdx.x >> 6];
  // the lane's ray: consecutive rays; for rays that are the rows of an
  // image `width` pixels wide, one 8x8 pixel tile per wave (neighbouring
  // pixels share most of their octree paths, which keeps the union small);
  // with `order`, ray order[t] (a caller's grouping of coherent rays;
  // entries outside [0, R) are skipped); lanes outside R or the image take
  // part in the wave's votes only
  const int64_t t = int64_t(blockIdx.x) * kHitsBlock + threadIdx.x;
  int64_t r = t;
  bool valid = t < R;
  if (order) {
    r = valid ? order[t] : 0;
    valid = valid && r >= 0 && r < R;
    r = valid ? r : 0;
  } else if (width > 0) {
    const int64_t tiles_x = (width + 7) / 8, k = t >> 6;
    const int q = int(t & 63);
    const int64_t y = (k / tiles_x) * 8 + (q >> 3), x = (k % tiles_x) * 8 + (q & 7);
    valid = x < width && y < R / width;
    r = y * width + x;
  }
  float o[3] = {0.0f, 0.0f, 0.0f}, inv[3] = {1.0f, 1.0f, 1.0f};
  int mask = 0;
  if (valid) {
    float dir[3];
    if constexpr (kCamera) {
      // the ray of pixel r of the camera, also stored for the later stages
      camera_ray(cam, su, sv, int(r / width), int(r % width), o, dir);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        gen_start[r * 3 + a] = o[a];
        gen_dir[r * 3 + a] = dir[a];
      }
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        o[a] = ray_start[r * 3 + a];
        dir[a] = ray_dir[r * 3 + a];
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      inv[a] = 1.0f / dir[a];
      mask |= (dir[a] < 0.0f ? 1 : 0) << a;
    }
  }
  HitRec* row = hits + (valid ? r : 0) * n_max;
  HitRow hr{row, 0, 0.0f, 0.0f, true};
  // one union walk per child order present in the wave
  uint64_t pending = __ballot(valid);
  while (pending) {
    const int d = __builtin_amdgcn_readlane(mask, first_lane(pending));
    const bool grp = valid && mask == d;
    pending &= ~__ballot(grp);
    traverse_union(o, inv, d, grp, recs, T, n_max, lrec, hr);
  }
  int cnt = hr.cnt;
  float tot = hr.tot;
  const bool redo = valid && !(cnt < n_max && hr.increasing);
  if (__ballot(redo)) {
    HitRow ref{row, 0, 0.0f, 0.0f, true};
    traverse_union(o, inv, 7, redo, recs, T, n_max, lrec, ref);
    if (redo) {
      cnt = ref.cnt;
      for (int k = 1; k < cnt; ++k) {  // stable insertion sort by entry depth
        const HitRec h = row[k];
        int j = k - 1;
        while (j >= 0 && row[j].lo > h.lo) {
          row[j + 1] = row[j];
          --j;
        }
        row[j + 1] = h;
      }
      tot = 0.0f;
      for (int k = 0; k < cnt; ++k) tot = tot + (row[k].hi - row[k].lo);
    }
  }
  if (!valid) return;
  count[r] = cnt;
  tot_out[r] = tot;
  hit_flag[r] = cnt > 0 ? 1 : 0;
  // at most ceil(steps) + cnt samples carry a voxel id (one per CDF step and
  // one tail per bin; the compat duplicate replaces a step sample)
  bound[r] = cnt > 0 ? int(ceilf(tot / step_size)) + cnt + 2 : 0;
}

// Sort key of a ray for the traversal's grouping (nsvf_ray_order_keys):
// direction octant (bits 42-44), origin cell on a 16^3 grid over the box
// [bounds[0..2], bounds[3..5]] (bits 30-41), Morton code of the direction
// quantised to 10 bits per axis (bits 0-29).
__device__ __forceinline__ int64_t spread3(int64_t q) {
  q &= 0x3FF;
  q = (q | (q << 16)) & 0x030000FF;
  q = (q | (q << 8)) & 0x0300F00F;
  q = (q | (q << 4)) & 0x030C30C3;
  q = (q | (q << 2)) & 0x09249249;
  return q;
}

__global__ __launch_bounds__(kBlock) void ray_key_kernel(const float* __restrict__ ray_start,
                                                         const float* __restrict__ ray_dir,
                                                         int64_t R,
                                                         const float* __restrict__ bounds,
                                                         int64_t* __restrict__ keys) {
  const int64_t r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (r >= R) return;
  int64_t oct = 0, cell = 0, morton = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float d = ray_dir[r * 3 + a], o = ray_start[r * 3 + a];
    oct |= int64_t(d < 0.0f) << a;
    const float span = fmaxf(bounds[3 + a] - bounds[a], 1e-30f);
    const int c = min(max(int((o - bounds[a]) / span * 15.999f), 0), 15);
    cell |= int64_t(c) << (4 * (2 - a));
    const float dc = fminf(fmaxf(d, -1.0f), 1.0f);
    const int q = min(max(__float2int_rn((dc + 1.0f) * 511.5f), 0), 1023);
    morton |= spread3(q) << a;
  }
  keys[r] = (oct << 42) | (cell << 30) | morton;
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
// One 64-sample pass of a ray's compositing weights (renderer.py:119-121,195-197):
// lane k of the pass holds fe_k = 7 max(sigma_k, 0) dist_k (0 past the
// row); p_k = (1 - exp(-fe_k)) exp(-(carry + fe_0 + ... + fe_{k-1})) by an
// inclusive wave scan, and carry grows by the pass's sum.  The coarse
// composite and the fine sampler (which composites the coarse row itself)
// both use it, so their weights are the same bits.
__device__ __forceinline__ float composite_pass(float fe, bool in, float& carry) {
  const int lane = threadIdx.x & 63;
  float incl = fe;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float u = __shfl_up(incl, d, 64);
    if (lane >= d) incl += u;
  }
  const float excl = carry + (incl - fe);
  const float a = 1.0f - expf(-fe);
  const float tr = expf(-excl);
  carry += __shfl(incl, 63, 64);
  return in ? a * tr : 0.0f;
}

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
  __device__ __forceinline__ float staged_prob(int k) const { return prob(k); }
  __device__ __forceinline__ void before_seq() const {}
  __device__ __forceinline__ int32_t next_first() const { return next; }
  __device__ __forceinline__ int32_t ray0_id(int k) const { return k < cnt0 ? row0[k].idx : -1; }
};

// Bins of a hit ray taken from its coarse samples (fine pass, nerf.py:64-79):
// [depth - dist/2, depth + dist/2], probs (p + 1e-5) / sum over all P columns.
constexpr int kRegPasses = 2;  // 64-bin passes of coarse weights a fine-sampler lane holds

// kReg: the fine sampler composited the coarse row itself; prob_c holds all
// the weights, each stored by the lane of its bin (bin 64 q + lane), and the
// first kRegPasses * 64 are in registers too (lane's pr[q]).  A staged bin
// is the lane's own: from pr, or read back from prob_c; lane 0's sequential
// walk reads them after before_seq().
template <bool kReg>
struct SampleBinsT {
  const SampleRec* rec;
  const float* prob_c;
  int cnt;
  float den;  // column sum of (p + 1e-5)
  const SampleRec* rec0;
  int cnt0;
  int32_t next;
  float pr[kRegPasses];
  __device__ __forceinline__ int32_t id(int k) const { return k < cnt ? rec[k].idx : -1; }
  __device__ __forceinline__ float lo(int k) const {
    return k < cnt ? rec[k].depth - rec[k].dist * 0.5f : kMaxDepth;
  }
  __device__ __forceinline__ float hi(int k) const {
    return k < cnt ? rec[k].depth + rec[k].dist * 0.5f : kMaxDepth;
  }
  __device__ __forceinline__ float prob(int k) const {
    return (k < cnt ? prob_c[k] + 1e-5f : 1e-5f) / den;
  }
  // prob(k) of the lane's own bin k = 64 q + lane of a staging pass
  __device__ __forceinline__ float staged_prob(int k) const {
    if constexpr (kReg) {
      const int q = k >> 6;
      if (q >= kRegPasses) return prob(k);
      float v = pr[0];
#pragma unroll
      for (int j = 1; j < kRegPasses; ++j) v = q == j ? pr[j] : v;
      return (k < cnt ? v + 1e-5f : 1e-5f) / den;
    } else {
      return prob(k);
    }
  }
  __device__ __forceinline__ void before_seq() const {
    if constexpr (kReg) {
      // the weights were stored by every lane of the wave (one work-group)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
  }
  __device__ __forceinline__ float depth(int k) const { return k < cnt ? rec[k].depth : 0.0f; }
  __device__ __forceinline__ int32_t next_first() const { return next; }
  __device__ __forceinline__ int32_t ray0_id(int k) const { return k < cnt0 ? rec0[k].idx : -1; }
};
using SampleBins = SampleBinsT<false>;

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

// Waves per sampler block and the grid-stride loop over hit rays.
constexpr int kSampleWaves = kBlock / 64;
// Bins a sampler wave stages in LDS (rows with more take the sequential
// walk).  Coarse rows have at most n_max bins (60 by default); fine rows as
// many as the ray's coarse samples.  The LDS per block (16 bytes per bin and
// wave) and the registers set the occupancy: the coarse kernel runs 8 waves
// per SIMD (128 bins, 63 VGPRs), the fine kernel 6 (384 bins; with fewer
// bins and registers for 7 or 8 it spills and runs slower).
constexpr int kCoarseBins = 128, kFineBins = 384;

// Ray segments of a batch of several views rendered together: segment s
// holds the frame rays [bounds[s], bounds[s + 1]).  The hierarchical pass
// normalises a ray's bins over the columns of the longest coarse row of its
// own view (nerf.py:64-79 pads every row to the longest one of the batch it
// is rendered in), so each view of a batch gets the result it gets alone.
constexpr int kMaxSegments = 16;

__device__ __forceinline__ int segment_of(const int32_t* __restrict__ bounds, int n_seg,
                                          int32_t r) {
  int s = 0;
  while (s + 1 < n_seg && r >= bounds[s + 1]) ++s;
  return s;
}

// Coarse pass: one wavefront per hit ray n (ray r = hit_list[n]) writes the
// ray's samples into its slot range [c_off[r], c_off[r] + bound[r]); count,
// fine-pass bound and the longest row -- of the frame, and with segments
// (n_seg > 0) of each segment -- are recorded for the fine pass.  kNoise:
// training-mode steps, u = HashNoise(seed, r) (else 0.5).
template <bool kNoise>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void coarse_sample_kernel(
    const int32_t* __restrict__ hit_list, const int32_t* __restrict__ n_hit_dev,
    const HitRec* __restrict__ hits, const int32_t* __restrict__ count,
    const float* __restrict__ tot, const int32_t* __restrict__ bound,
    const int32_t* __restrict__ c_off, int n_max, float step_size, int compat, float fine_steps,
    SampleRec* __restrict__ out, int32_t* __restrict__ c_cnt, int32_t* __restrict__ f_bound,
    int32_t* __restrict__ max_count, const int32_t* __restrict__ seg_bounds, int n_seg,
    int32_t* __restrict__ seg_max, uint32_t seed) {
  __shared__ float lds_c[kSampleWaves][kCoarseBins], lds_lo[kSampleWaves][kCoarseBins],
      lds_hi[kSampleWaves][kCoarseBins];
  __shared__ int32_t lds_id[kSampleWaves][kCoarseBins];
  __shared__ int block_max;
  __shared__ int block_seg_max[kMaxSegments];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t N = *n_hit_dev;
  if (threadIdx.x == 0) block_max = 0;
  if (threadIdx.x < kMaxSegments) block_seg_max[threadIdx.x] = 0;
  __syncthreads();
  int wave_max = 0;  // longest row of this wave's rays (one atomic per block at the end)
  for (int64_t n = int64_t(blockIdx.x) * kSampleWaves + wave; n < N;
       n += int64_t(gridDim.x) * kSampleWaves) {
    const int32_t r = hit_list[n];
    int64_t n0, nn;
    grouped_neighbours(n, N, &n0, &nn);
    const int32_t r0 = hit_list[n0];
    const int cnt = count[r];
    HitBins bins{hits + int64_t(r) * n_max, cnt, tot[r], hits + int64_t(r0) * n_max, count[r0],
                 nn < 0 ? -1 : hits[int64_t(hit_list[nn]) * n_max].idx};
    const float steps = tot[r] / step_size;
    const WaveBinsLds lds{lds_c[wave], lds_lo[wave], lds_hi[wave], lds_id[wave], kCoarseBins};
    int m;
    if constexpr (kNoise)
      m = icdf_wave(bins, n_max, cnt, steps, compat, bound[r], r, out + c_off[r], lds,
                    NoSlotHook{}, HashNoise::of(seed, uint32_t(r)));
    else
      m = icdf_wave(bins, n_max, cnt, steps, compat, bound[r], r, out + c_off[r], lds);
    if (lane == 0) {
      c_cnt[r] = m;
      f_bound[r] = int(ceilf(fine_steps)) + m + 2;
      if (n_seg > 0 && m > 0) atomicMax(&block_seg_max[segment_of(seg_bounds, n_seg, r)], m);
    }
    wave_max = m > wave_max ? m : wave_max;
  }
  if (lane == 0) atomicMax(&block_max, wave_max);
  __syncthreads();
  if (threadIdx.x == 0 && block_max > 0) atomicMax(max_count, block_max);
  if (int(threadIdx.x) < n_seg && block_seg_max[threadIdx.x] > 0)
    atomicMax(&seg_max[threadIdx.x], block_seg_max[threadIdx.x]);
}

// Column sum of (p + 1e-5) over the P bin columns of a fine-pass ray
// (nerf.py:64-79 normalises over all columns, pads included), in a fixed
// order: lane j sums columns j, j + 64, ... in turn, then an xor butterfly
// over the lanes (d = 1, 2, ..., 32).  Every lane ends with the same value.
__device__ __forceinline__ float fine_den(const float* prob, int cnt, int P) {
  const int lane = threadIdx.x & 63;
  float part = 0.0f;
  for (int k = lane; k < P; k += 64) part = part + ((k < cnt ? prob[k] : 0.0f) + 1e-5f);
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) part = part + __shfl_xor(part, d, 64);
  return part;
}

// fine_den with the ray's first kRegPasses * 64 weights in registers (column
// 64 q + lane in pr[q]); later columns from prob, stored by this same lane.
__device__ __forceinline__ float fine_den_reg(const float (&pr)[kRegPasses], const float* prob,
                                              int cnt, int P) {
  const int lane = threadIdx.x & 63;
  float part = 0.0f;
  for (int k = lane; k < P; k += 64) {
    const int q = k >> 6;
    float v = 0.0f;
    if (k < cnt) {
      if (q < kRegPasses) {
        v = pr[0];
#pragma unroll
        for (int j = 1; j < kRegPasses; ++j) v = q == j ? pr[j] : v;
      } else {
        v = prob[k];
      }
    }
    part = part + (v + 1e-5f);
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) part = part + __shfl_xor(part, d, 64);
  return part;
}

// Field reuse in the fine pass.  The fine pass evaluates the same field as
// the coarse pass (nerf.py:58: field_fine only with --use-fine-model), and
// the field of a sample depends only on its ray, voxel id and depth.  A
// coarse bin that no fine step splits comes back as a fine sample of the
// same voxel whose midpoint (lo + hi) / 2 of [depth - dist/2, depth +
// dist/2] is nearly always the coarse depth again; where it is bit for bit,
// the sample's field value is the coarse one.  The sampler only marks the
// slots: src[slot] = k reads bin k's coarse value, -1 = evaluate, -2 =
// unused (stores only: its serial walk never waits on an extra load).
struct FieldReuse {
  static constexpr bool kBins = true;
  int16_t* src;  // the ray's fine slots, or null for no reuse
  __device__ __forceinline__ void put(int pos, int bound, int k) {
    if (src && pos < bound) src[pos] = int16_t(k >= 0 ? k : -1);
  }
  __device__ __forceinline__ void seq(int n) {
    if (src)
      for (int k = threadIdx.x & 63; k < n; k += 64) src[k] = -1;
  }
  __device__ __forceinline__ void unused(int pos) {
    if (src) src[pos] = -2;
  }
};

// Fine pass: one wavefront per hit ray; bins are the ray's coarse samples and
// the *c_max columns of the longest coarse row (nerf.py:64-79) -- with
// segments (n_seg > 0), the seg_max[s] columns of the longest row of the
// ray's segment s.  With src
// (optional), the slots are marked for field reuse as FieldReuse describes.
// kComposite: the wave first composites the ray's coarse row from c_field
// (composite_pass, as the coarse composite does) and stores the weights to
// c_prob, keeping the staged bins' weights in registers; otherwise c_prob
// holds them.  kNoise: training-mode steps, u = HashNoise(seed, r) (else 0.5).
template <bool kComposite, bool kNoise>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void fine_sample_kernel(
    const int32_t* __restrict__ hit_list, const int32_t* __restrict__ n_hit_dev,
    const int32_t* __restrict__ c_off, const int32_t* __restrict__ c_cnt,
    const int32_t* __restrict__ c_max_dev, const SampleRec* __restrict__ coarse,
    const float4* __restrict__ c_field, float* __restrict__ c_prob,
    const int32_t* __restrict__ f_bound,
    const int32_t* __restrict__ f_off, float fine_steps, int compat, SampleRec* __restrict__ out,
    int32_t* __restrict__ f_cnt, int16_t* __restrict__ src, const int32_t* __restrict__ seg_bounds,
    int n_seg, const int32_t* __restrict__ seg_max, uint32_t seed) {
  __shared__ float lds_c[kSampleWaves][kFineBins], lds_lo[kSampleWaves][kFineBins],
      lds_hi[kSampleWaves][kFineBins];
  __shared__ int32_t lds_id[kSampleWaves][kFineBins];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t N = *n_hit_dev;
  const int P_frame = *c_max_dev;  // bin columns = longest coarse row
  for (int64_t n = int64_t(blockIdx.x) * kSampleWaves + wave; n < N;
       n += int64_t(gridDim.x) * kSampleWaves) {
    const int32_t r = hit_list[n];
    const int P = n_seg > 0 ? seg_max[segment_of(seg_bounds, n_seg, r)] : P_frame;
    const int b = f_bound[r];
    SampleRec* dst = out + f_off[r];
    int m = 0;
    FieldReuse reuse{src ? src + f_off[r] : nullptr};
    if (P > 0) {
      int64_t n0, nn;
      grouped_neighbours(n, N, &n0, &nn);
      const int32_t r0 = hit_list[n0];
      const int64_t off = c_off[r];
      const int cnt = c_cnt[r];
      int32_t next = -1;
      if (nn >= 0) {
        const int32_t rn = hit_list[nn];
        next = c_cnt[rn] > 0 ? coarse[c_off[rn]].idx : -1;
      }
      const WaveBinsLds lds{lds_c[wave], lds_lo[wave], lds_hi[wave], lds_id[wave], kFineBins};
      if constexpr (kComposite) {
        SampleBinsT<true> bins{coarse + off, c_prob + off, cnt, 0.0f, coarse + c_off[r0],
                               c_cnt[r0], next};
#pragma unroll
        for (int q = 0; q < kRegPasses; ++q) bins.pr[q] = 0.0f;
        float carry = 0.0f;
        for (int base = 0; base < cnt; base += 64) {
          const int k = base + lane;
          float fe = 0.0f;
          if (k < cnt) {
            const float sg = c_field[off + k].x;
            fe = ((sg > 0.0f ? sg : 0.0f) * coarse[off + k].dist) * 7.0f;
          }
          const float p = composite_pass(fe, k < cnt, carry);
          if (k < cnt) c_prob[off + k] = p;
          const int q = base >> 6;
#pragma unroll
          for (int j = 0; j < kRegPasses; ++j) bins.pr[j] = q == j ? p : bins.pr[j];
        }
        bins.den = fine_den_reg(bins.pr, c_prob + off, cnt, P);
        if constexpr (kNoise)
          m = icdf_wave(bins, P, cnt, fine_steps, compat, b, r, dst, lds, reuse,
                        HashNoise::of(seed, uint32_t(r)));
        else
          m = icdf_wave(bins, P, cnt, fine_steps, compat, b, r, dst, lds, reuse);
      } else {
        const float den = fine_den(c_prob + off, cnt, P);
        SampleBins bins{coarse + off, c_prob + off, cnt, den, coarse + c_off[r0], c_cnt[r0],
                        next};
        if constexpr (kNoise)
          m = icdf_wave(bins, P, cnt, fine_steps, compat, b, r, dst, lds, reuse,
                        HashNoise::of(seed, uint32_t(r)));
        else
          m = icdf_wave(bins, P, cnt, fine_steps, compat, b, r, dst, lds, reuse);
      }
    } else {
      for (int k = lane; k < b; k += 64) {
        dst[k] = SampleRec{-1, 0.0f, 0.0f, r};
        reuse.unused(k);
      }
    }
    if (lane == 0) f_cnt[r] = m;
  }
}

// The field's evaluation list of a marked fine stream: the slots s < *n_dev
// with src[s] == -1, in slot order.  A stream compaction over all slots in
// two kernels around a scan of per-block counts: a block of 4 waves takes
// 4096 consecutive slots, each wave 16 groups of 64 (one slot per lane, so
// every load and store of a group is one coalesced access), and a group's
// list entries come from a ballot.
constexpr int kMarkGroups = 16, kMarkPerBlock = kBlock * kMarkGroups;

// The wave's ballots of eval slots, group by group; returns their count.
__device__ __forceinline__ int mark_ballots(const int16_t* __restrict__ src, int64_t s0,
                                            int64_t n, uint64_t (&mask)[kMarkGroups]) {
  const int lane = threadIdx.x & 63;
  int16_t v[kMarkGroups];
#pragma unroll
  for (int q = 0; q < kMarkGroups; ++q) {
    const int64_t s = s0 + 64 * q + lane;
    v[q] = s < n ? src[s] : int16_t(-2);
  }
  int c = 0;
#pragma unroll
  for (int q = 0; q < kMarkGroups; ++q) {
    mask[q] = __ballot(v[q] == -1);
    c += __popcll(mask[q]);
  }
  return c;
}

__global__ __launch_bounds__(kBlock) void mark_count_kernel(const int16_t* __restrict__ src,
                                                            const int32_t* __restrict__ n_dev,
                                                            int64_t n_max,
                                                            int32_t* __restrict__ block_cnt) {
  __shared__ int wcnt[kBlock / 64];
  const int64_t n = min(int64_t(*n_dev), n_max);
  const int wave = threadIdx.x >> 6;
  const int64_t s0 = int64_t(blockIdx.x) * kMarkPerBlock + int64_t(wave) * 64 * kMarkGroups;
  uint64_t mask[kMarkGroups];
  const int c = mark_ballots(src, s0, n, mask);
  if ((threadIdx.x & 63) == 0) wcnt[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += wcnt[w];
    block_cnt[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void mark_compact_kernel(const int16_t* __restrict__ src,
                                                              const int32_t* __restrict__ n_dev,
                                                              int64_t n_max,
                                                              const int32_t* __restrict__ block_off,
                                                              int32_t* __restrict__ list) {
  __shared__ int wcnt[kBlock / 64];
  const int64_t n = min(int64_t(*n_dev), n_max);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t s0 = int64_t(blockIdx.x) * kMarkPerBlock + int64_t(wave) * 64 * kMarkGroups;
  uint64_t mask[kMarkGroups];
  const int c = mark_ballots(src, s0, n, mask);
  if (lane == 0) wcnt[wave] = c;
  __syncthreads();
  int pos = block_off[blockIdx.x];
  for (int w = 0; w < wave; ++w) pos += wcnt[w];
  const uint64_t below = (1ull << lane) - 1;
#pragma unroll
  for (int q = 0; q < kMarkGroups; ++q) {
    if ((mask[q] >> lane) & 1ull)
      list[pos + __popcll(mask[q] & below)] = int32_t(s0 + 64 * q + lane);
    pos += __popcll(mask[q]);
  }
}

// The reused fine field values written out (f_field[slot] = the coarse
// value it reads), for callers that keep the fine field; one wavefront per
// hit ray.
__global__ __launch_bounds__(kBlock) void reuse_fill_kernel(
    const int32_t* __restrict__ hit_list, const int32_t* __restrict__ n_hit_dev,
    const int32_t* __restrict__ c_off, const int32_t* __restrict__ f_off,
    const int32_t* __restrict__ f_cnt, const int16_t* __restrict__ src,
    const float4* __restrict__ c_field, float4* __restrict__ f_field) {
  const int64_t n = int64_t(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (n >= *n_hit_dev) return;
  const int32_t r = hit_list[n];
  const int64_t coff = c_off[r], off = f_off[r];
  const int m = f_cnt[r];
  for (int k = lane; k < m; k += 64) {
    const int sk = src[off + k];
    if (sk >= 0) f_field[off + k] = c_field[coff + sk];
  }
}

// ---------------------------------------------------------------------------
// composite

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Per-voxel maximum over rays of a ray's summed sample probabilities
// (encoder.py:594-603 track_voxel_probs: scatter_add of p over each ray's
// samples into voxel columns, then max over rays, then a running max; -1
// samples ignored).  One wavefront per hit ray.  A ray crosses a convex
// voxel once and its samples are depth-ordered, so a voxel's samples are
// one contiguous run: lane j sums run j left to right.  If an id comes back
// later in the ray (the compat-mode sample with the next ray's id), the
// lane of its first run sums every sample of that id to the end of the ray,
// again left to right.  Sums are >= 0, so an integer max on their bits
// orders them.
constexpr int kTrackRuns = 256;

__global__ __launch_bounds__(kBlock) void track_probs_kernel(
    const int32_t* __restrict__ hit_list, const int32_t* __restrict__ n_hit_dev,
    const int32_t* __restrict__ off_arr, const int32_t* __restrict__ cnt_arr,
    const SampleRec* __restrict__ samples, const float* __restrict__ probs,
    float* __restrict__ vmax, int64_t N) {
  __shared__ int32_t run_start[kSampleWaves][kTrackRuns + 1];
  __shared__ int32_t run_id[kSampleWaves][kTrackRuns];
  __shared__ int32_t run_dup[kSampleWaves][kTrackRuns];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t NH = *n_hit_dev;
  for (int64_t n = int64_t(blockIdx.x) * kSampleWaves + wave; n < NH;
       n += int64_t(gridDim.x) * kSampleWaves) {
    const int32_t r = hit_list[n];
    const SampleRec* rec = samples + off_arr[r];
    const float* pr = probs + off_arr[r];
    const int c = __builtin_amdgcn_readfirstlane(cnt_arr[r]);
    // run heads
    int nruns = 0;
    int32_t prev = INT32_MIN;  // no sample has this id
    for (int k0 = 0; k0 < c; k0 += 64) {
      const int k = k0 + lane;
      const int32_t id = k < c ? rec[k].idx : INT32_MIN;
      int32_t idp = __shfl_up(id, 1, 64);
      if (lane == 0) idp = prev;
      const bool head = k < c && id != idp;
      const uint64_t hm = __ballot(head);
      const int pos = nruns + __popcll(hm & ((1ull << lane) - 1));
      if (head && pos < kTrackRuns) {
        run_start[wave][pos] = k;
        run_id[wave][pos] = id;
      }
      nruns += __popcll(hm);
      prev = __shfl(id, 63, 64);
    }
    if (nruns > kTrackRuns) {
      // more runs than LDS holds (not met on rendered rays): lane 0 does the
      // whole ray, first occurrence by first occurrence
      if (lane == 0) {
        for (int s0 = 0; s0 < c; ++s0) {
          const int32_t id = rec[s0].idx;
          if (id < 0 || id >= N) continue;
          bool seen = false;
          for (int t = 0; t < s0 && !seen; ++t) seen = rec[t].idx == id;
          if (seen) continue;
          float acc = 0.0f;
          for (int s = s0; s < c; ++s)
            if (rec[s].idx == id) acc = acc + pr[s];
          atomicMax(reinterpret_cast<int*>(vmax + id), __float_as_int(acc));
        }
      }
      wave_sync();
      continue;
    }
    if (lane == 0) run_start[wave][nruns] = c;
    wave_sync();
    // runs whose id an earlier run already has
    bool any_dup = false;
    for (int j0 = 0; j0 < nruns; j0 += 64) {
      const int j = j0 + lane;
      bool dup = false;
      if (j < nruns) {
        const int32_t idj = run_id[wave][j];
        for (int i = 0; i < j && !dup; ++i) dup = run_id[wave][i] == idj;
        run_dup[wave][j] = dup ? 1 : 0;
      }
      any_dup = any_dup || __ballot(dup) != 0;
    }
    wave_sync();
    for (int j0 = 0; j0 < nruns; j0 += 64) {
      const int j = j0 + lane;
      if (j < nruns && !run_dup[wave][j]) {
        const int32_t id = run_id[wave][j];
        if (id >= 0 && id < N) {
          float acc = 0.0f;
          const int e = any_dup ? c : run_start[wave][j + 1];
          for (int s = run_start[wave][j]; s < e; ++s)
            if (!any_dup || rec[s].idx == id) acc = acc + pr[s];
          atomicMax(reinterpret_cast<int*>(vmax + id), __float_as_int(acc));
        }
      }
    }
    wave_sync();  // the next ray's run lists must not overtake this ray's reads
  }
}

// One wavefront per hit ray, 64 samples per step.  Free energy
// fe = relu(sigma) * dist * 7 (renderer.py:119-121); transmittance before a
// sample is exp(-(sum of the earlier fe)), an exclusive prefix sum across the
// wave plus the carry of earlier steps (renderer.py:195-197).
template <bool kReuse>
__global__ __launch_bounds__(kBlock) void composite_kernel(
    const int32_t* __restrict__ hit_list, const int32_t* __restrict__ n_hit_dev,
    const int32_t* __restrict__ off_arr, const int32_t* __restrict__ cnt_arr,
    const SampleRec* __restrict__ samples, const float4* __restrict__ field,
    float* __restrict__ prob_out, float* __restrict__ ray_rgb, float* __restrict__ ray_depth,
    float* __restrict__ ray_missed, const int16_t* __restrict__ src,
    const int32_t* __restrict__ src_off, const float4* __restrict__ src_field) {
  const int64_t n = int64_t(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (n >= *n_hit_dev) return;
  const int32_t r = hit_list[n];
  const int64_t off = off_arr[r];
  const int cnt = cnt_arr[r];
  const int64_t soff = kReuse ? src_off[r] : 0;
  // the marks of the ray's first 4 passes, loaded together up front so a
  // pass's field load does not wait on its own mark load
  constexpr int kPre = 4;
  int16_t pre[kPre];
#pragma unroll
  for (int q = 0; q < kPre; ++q) {
    const int k = 64 * q + lane;
    pre[q] = (kReuse && k < cnt) ? src[off + k] : int16_t(-1);
  }
  float carry = 0.0f;
  float acc_p = 0.0f, acc_z = 0.0f, acc_r = 0.0f, acc_g = 0.0f, acc_b = 0.0f;
  for (int base = 0; base < cnt; base += 64) {
    const int k = base + lane;
    float fe = 0.0f;
    float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
    float z = 0.0f;
    if (k < cnt) {
      const int q = base >> 6;
      int sk = -1;
      if constexpr (kReuse) {
        sk = pre[0];
#pragma unroll
        for (int j = 1; j < kPre; ++j) sk = q == j ? pre[j] : sk;
        if (q >= kPre) sk = src[off + k];
      }
      // one load from whichever array holds the value
      const float4* fp = sk >= 0 ? src_field + soff + sk : field + off + k;
      f = *fp;
      const SampleRec sr = samples[off + k];
      const float sg = f.x > 0.0f ? f.x : 0.0f;
      fe = (sg * sr.dist) * 7.0f;
      z = sr.depth;
    }
    const float p = composite_pass(fe, k < cnt, carry);
    if (k < cnt && prob_out) prob_out[off + k] = p;
    acc_p += p;
    acc_z += p * z;
    acc_r += p * f.y;
    acc_g += p * f.z;
    acc_b += p * f.w;
  }
  acc_p = wave_sum(acc_p);
  acc_z = wave_sum(acc_z);
  acc_r = wave_sum(acc_r);
  acc_g = wave_sum(acc_g);
  acc_b = wave_sum(acc_b);
  if (lane == 0) {
    ray_rgb[int64_t(r) * 3 + 0] = acc_r;
    ray_rgb[int64_t(r) * 3 + 1] = acc_g;
    ray_rgb[int64_t(r) * 3 + 2] = acc_b;
    ray_depth[r] = acc_z;
    ray_missed[r] = 1.0f - acc_p;
  }
}

// Gradient of composite_kernel's outputs with respect to the field values of
// each sample (the autograd of renderer.py:119-121,195-218).  With
// g_p(k) = <g_rgb, c_k> + g_depth z_k - g_missed and u_k = g_p(k) p_k:
//   d/dfe_k    = g_p(k) T_k exp(-fe_k) - sum_{j>k} u_j
//   d/dsigma_k = d/dfe_k * 7 dist_k  if sigma_k > 0, else 0
//   d/dc_k     = p_k g_rgb
// One wavefront per hit ray.  Pass 1 runs the forward scan again and parks
// each sample's transmittance exponent in grad_field[k].x; pass 2 walks the
// ray back to front, so the suffix sums of u are accumulated directly (no
// difference of two large prefix sums), and overwrites each record with its
// gradient.  T_k and p_k are recomputed with the forward's exact operations.
__global__ __launch_bounds__(kBlock) void composite_backward_kernel(
    const int32_t* __restrict__ hit_list, const int32_t* __restrict__ n_hit_dev,
    const int32_t* __restrict__ off_arr, const int32_t* __restrict__ cnt_arr,
    const SampleRec* __restrict__ samples, const float4* __restrict__ field,
    const float* __restrict__ g_rgb, const float* __restrict__ g_depth,
    const float* __restrict__ g_missed, float4* __restrict__ grad_field) {
  const int64_t n = int64_t(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (n >= *n_hit_dev) return;
  const int32_t r = hit_list[n];
  const int64_t off = off_arr[r];
  const int cnt = cnt_arr[r];
  float carry = 0.0f;
  for (int base = 0; base < cnt; base += 64) {
    const int k = base + lane;
    float fe = 0.0f;
    if (k < cnt) {
      const float sg = field[off + k].x > 0.0f ? field[off + k].x : 0.0f;
      fe = (sg * samples[off + k].dist) * 7.0f;
    }
    float incl = fe;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const float u = __shfl_up(incl, d, 64);
      if (lane >= d) incl += u;
    }
    if (k < cnt) grad_field[off + k].x = carry + (incl - fe);
    carry += __shfl(incl, 63, 64);
  }
  const float gr = g_rgb[int64_t(r) * 3 + 0], gg = g_rgb[int64_t(r) * 3 + 1],
              gb = g_rgb[int64_t(r) * 3 + 2];
  const float gd = g_depth[r], gm = g_missed[r];
  float suffix = 0.0f;  // sum of u over the samples after the current chunk
  for (int base = (cnt - 1) & ~63; base >= 0; base -= 64) {
    const int k = base + lane;
    float u = 0.0f, gp = 0.0f, p = 0.0f, tnext = 0.0f, sig = 0.0f, dist = 0.0f;
    if (k < cnt) {
      const float4 f = field[off + k];
      const SampleRec sr = samples[off + k];
      const float excl = grad_field[off + k].x;
      sig = f.x;
      dist = sr.dist;
      const float sg = f.x > 0.0f ? f.x : 0.0f;
      const float fe = (sg * sr.dist) * 7.0f;
      const float e = expf(-fe);
      const float tr = expf(-excl);
      p = (1.0f - e) * tr;
      tnext = tr * e;
      gp = gr * f.y + gg * f.z + gb * f.w + gd * sr.depth - gm;
      u = gp * p;
    }
    // inclusive suffix scan of u within the wave (lanes >= this one)
    float incl = u;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const float v = __shfl_down(incl, d, 64);
      if (lane + d < 64) incl += v;
    }
    if (k < cnt) {
      const float later = suffix + (incl - u);
      const float dfe = gp * tnext - later;
      const float ds = sig > 0.0f ? (dfe * dist) * 7.0f : 0.0f;
      grad_field[off + k] = make_float4(ds, p * gr, p * gg, p * gb);
    }
    suffix += __shfl(incl, 0, 64);
  }
}

// Frame outputs for every ray: the hit rays' composite plus the background
// weighted by `missed`; rays that hit nothing are pure background.
__global__ __launch_bounds__(kBlock) void finish_kernel(
    int64_t R, const int32_t* __restrict__ hit_flag, const float* __restrict__ ray_rgb,
    const float* __restrict__ ray_depth, const float* __restrict__ ray_missed, float bg_r,
    float bg_g, float bg_b, float bg_depth, float* __restrict__ rgb, float* __restrict__ depth,
    float* __restrict__ missed) {
  const int64_t r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (r >= R) return;
  float cr = 0.f, cg = 0.f, cb = 0.f, d = 0.f, m = 1.f;
  if (hit_flag[r]) {
    cr = ray_rgb[r * 3];
    cg = ray_rgb[r * 3 + 1];
    cb = ray_rgb[r * 3 + 2];
    d = ray_depth[r];
    m = ray_missed[r];
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

// The traversal keeps one LDS record per level: level 0 holds the root, a
// node of depth d (children[8] = 1 << (d + 1)) is expanded at level
// depth - d + 1, and depth-0 nodes push their leaves without expanding
// them, so a tree of depth `depth` needs depth + 2 levels.
static int check_tree_depth(int depth, const char* what) {
  if (depth < 0) return set_error(NSVF_EINVAL, "%s: tree depth %d", what, depth);
  if (depth + 2 > kMaxLevels)
    return set_error(NSVF_ESTACK, "%s: octree depth %d needs %d traversal levels (at most %d)",
                     what, depth, depth + 2, kMaxLevels);
  return NSVF_OK;
}

int nsvf_render_hits(const float* ray_start, const float* ray_dir, int64_t R, int width,
                     const void* records, int T, int depth, int n_max, float step_size,
                     void* hits, int32_t* count, float* tot, int32_t* hit_flag, int32_t* bound,
                     void* stream) {
  clear_error();
  if (R < 0 || T <= 0 || n_max <= 0) return set_error(NSVF_EINVAL, "render_hits: bad size");
  if (const int rc = check_tree_depth(depth, "render_hits")) return rc;
  if (width < 0 || (width > 0 && R % width != 0))
    return set_error(NSVF_EINVAL, "render_hits: %lld rays are not rows of width %d",
                     static_cast<long long>(R), width);
  if (R == 0) return NSVF_OK;
  if (!ray_start || !ray_dir || !records || !hits || !count || !tot || !hit_flag || !bound)
    return set_error(NSVF_EINVAL, "render_hits: null pointer");
  if (T >= (1 << 27)) return set_error(NSVF_EINVAL, "render_hits: more than 2^27 nodes");
  const int64_t threads =
      width > 0 ? ceil_div(R / width, 8) * ceil_div(width, 8) * 64 : R;
  hipLaunchKernelGGL(hits_kernel<false>, dim3(unsigned(ceil_div(threads, kHitsBlock))),
                     dim3(kHitsBlock), 0, as_stream(stream), ray_start, ray_dir, R, width,
                     nsvf_camera{}, 0.0f, 0.0f, nullptr, nullptr,
                     static_cast<const float4*>(records), T, n_max, step_size,
                     static_cast<HitRec*>(hits), count, tot, hit_flag, bound, nullptr);
  return check_launch("render_hits");
}

int nsvf_ray_order_keys(const float* ray_start, const float* ray_dir, int64_t R,
                        const float* bounds, int64_t* keys, void* stream) {
  clear_error();
  if (R < 0) return set_error(NSVF_EINVAL, "ray_order_keys: bad size");
  if (R == 0) return NSVF_OK;
  if (!ray_start || !ray_dir || !bounds || !keys)
    return set_error(NSVF_EINVAL, "ray_order_keys: null pointer");
  hipLaunchKernelGGL(ray_key_kernel, dim3(unsigned(ceil_div(R, kBlock))), dim3(kBlock), 0,
                     as_stream(stream), ray_start, ray_dir, R, bounds, keys);
  return check_launch("ray_order_keys");
}

int nsvf_render_hits_ordered(const float* ray_start, const float* ray_dir, int64_t R,
                             const int32_t* order, const void* records, int T, int depth,
                             int n_max, float step_size, void* hits, int32_t* count, float* tot,
                             int32_t* hit_flag, int32_t* bound, void* stream) {
  clear_error();
  if (R < 0 || T <= 0 || n_max <= 0)
    return set_error(NSVF_EINVAL, "render_hits_ordered: bad size");
  if (const int rc = check_tree_depth(depth, "render_hits_ordered")) return rc;
  if (R == 0) return NSVF_OK;
  if (!ray_start || !ray_dir || !order || !records || !hits || !count || !tot || !hit_flag ||
      !bound)
    return set_error(NSVF_EINVAL, "render_hits_ordered: null pointer");
  if (T >= (1 << 27))
    return set_error(NSVF_EINVAL, "render_hits_ordered: more than 2^27 nodes");
  hipLaunchKernelGGL(hits_kernel<false>, dim3(unsigned(ceil_div(R, kHitsBlock))),
                     dim3(kHitsBlock), 0, as_stream(stream), ray_start, ray_dir, R, 0,
                     nsvf_camera{}, 0.0f, 0.0f, nullptr, nullptr,
                     static_cast<const float4*>(records), T, n_max, step_size,
                     static_cast<HitRec*>(hits), count, tot, hit_flag, bound, order);
  return check_launch("render_hits_ordered");
}

static int check_camera(const nsvf_camera* cam, const char* what) {
  if (!cam) return set_error(NSVF_EINVAL, "%s: null camera", what);
  if (cam->h <= 0 || cam->w <= 0 || cam->H <= 0 || cam->W <= 0)
    return set_error(NSVF_EINVAL, "%s: bad image size %dx%d of %dx%d", what, cam->h, cam->w,
                     cam->H, cam->W);
  if (int64_t(cam->h) * cam->w >= (int64_t(1) << 31))
    return set_error(NSVF_EINVAL, "%s: more than 2^31 rays", what);
  return NSVF_OK;
}

int nsvf_generate_rays(const nsvf_camera* cam, float* ray_start, float* ray_dir, void* stream) {
  clear_error();
  if (const int rc = check_camera(cam, "generate_rays")) return rc;
  if (!ray_start || !ray_dir) return set_error(NSVF_EINVAL, "generate_rays: null pointer");
  // float(W / w): the quotient in double, rounded to float (data_utils.py:63-64)
  const float su = float(double(cam->W) / cam->w), sv = float(double(cam->H) / cam->h);
  const int64_t R = int64_t(cam->h) * cam->w;
  hipLaunchKernelGGL(ray_gen_kernel, dim3(unsigned(ceil_div(R, kBlock))), dim3(kBlock), 0,
                     as_stream(stream), *cam, su, sv, ray_start, ray_dir);
  return check_launch("generate_rays");
}

int nsvf_render_hits_camera(const nsvf_camera* cam, const void* records, int T, int depth,
                            int n_max, float step_size, float* ray_start, float* ray_dir,
                            void* hits, int32_t* count, float* tot, int32_t* hit_flag,
                            int32_t* bound, void* stream) {
  clear_error();
  if (const int rc = check_camera(cam, "render_hits_camera")) return rc;
  if (T <= 0 || n_max <= 0) return set_error(NSVF_EINVAL, "render_hits_camera: bad size");
  if (const int rc = check_tree_depth(depth, "render_hits_camera")) return rc;
  if (!records || !ray_start || !ray_dir || !hits || !count || !tot || !hit_flag || !bound)
    return set_error(NSVF_EINVAL, "render_hits_camera: null pointer");
  if (T >= (1 << 27)) return set_error(NSVF_EINVAL, "render_hits_camera: more than 2^27 nodes");
  const float su = float(double(cam->W) / cam->w), sv = float(double(cam->H) / cam->h);
  const int64_t R = int64_t(cam->h) * cam->w;
  const int64_t threads = ceil_div(cam->h, 8) * ceil_div(cam->w, 8) * 64;
  hipLaunchKernelGGL(hits_kernel<true>, dim3(unsigned(ceil_div(threads, kHitsBlock))),
                     dim3(kHitsBlock), 0, as_stream(stream), nullptr, nullptr, R, cam->w, *cam,
                     su, sv, ray_start, ray_dir, static_cast<const float4*>(records), T, n_max,
                     step_size, static_cast<HitRec*>(hits), count, tot, hit_flag, bound,
                     nullptr);
  return check_launch("render_hits_camera");
}

int nsvf_render_compact(const int32_t* hit_flag, const int32_t* rank, int64_t R, int32_t* list,
                        void* stream) {
  clear_error();
  if (R <= 0) return NSVF_OK;
  hipLaunchKernelGGL(compact_kernel, dim3(unsigned(ceil_div(R, kBlock))), dim3(kBlock), 0,
                     as_stream(stream), hit_flag, rank, R, list);
  return check_launch("render_compact");
}

int nsvf_render_sample_coarse(const int32_t* hit_list, const int32_t* n_hit_dev, int64_t n_hit_max,
                              const void* hits, const int32_t* count, const float* tot,
                              const int32_t* bound, const int32_t* c_off, int n_max,
                              float step_size, int compat, float fine_steps, void* samples,
                              int32_t* c_cnt, int32_t* f_bound, int32_t* max_count,
                              const int32_t* seg_bounds, int n_seg, int32_t* seg_max,
                              int64_t noise_seed, void* stream) {
  clear_error();
  if (n_hit_max <= 0) return NSVF_OK;
  if (n_seg < 0 || n_seg > kMaxSegments || (n_seg > 0 && (!seg_bounds || !seg_max)))
    return set_error(NSVF_EINVAL, "render_sample_coarse: %d segments (at most %d, with bounds "
                     "and maxima)", n_seg, kMaxSegments);
  if (noise_seed > int64_t(UINT32_MAX))
    return set_error(NSVF_EINVAL, "render_sample_coarse: noise seed %lld beyond 32 bits",
                     static_cast<long long>(noise_seed));
  auto kernel = noise_seed >= 0 ? coarse_sample_kernel<true> : coarse_sample_kernel<false>;
  hipLaunchKernelGGL(kernel,
                     dim3(unsigned(std::min<int64_t>(ceil_div(n_hit_max, kSampleWaves), 8192))),
                     dim3(kBlock), 0, as_stream(stream), hit_list, n_hit_dev,
                     static_cast<const HitRec*>(hits), count, tot, bound, c_off, n_max, step_size,
                     compat, fine_steps, static_cast<SampleRec*>(samples), c_cnt, f_bound,
                     max_count, seg_bounds, n_seg, seg_max,
                     uint32_t(noise_seed >= 0 ? noise_seed : 0));
  return check_launch("render_sample_coarse");
}

int nsvf_render_sample_fine(const int32_t* hit_list, const int32_t* n_hit_dev, int64_t n_hit_max,
                            const int32_t* c_off, const int32_t* c_cnt, const int32_t* c_max,
                            const void* coarse, const void* c_field, float* c_prob,
                            const int32_t* f_bound, const int32_t* f_off, float fine_steps,
                            int compat, void* samples, int32_t* f_cnt, int16_t* src,
                            const int32_t* seg_bounds, int n_seg, const int32_t* seg_max,
                            int64_t noise_seed, void* stream) {
  clear_error();
  if (n_hit_max <= 0) return NSVF_OK;
  if (n_seg < 0 || n_seg > kMaxSegments || (n_seg > 0 && (!seg_bounds || !seg_max)))
    return set_error(NSVF_EINVAL, "render_sample_fine: %d segments (at most %d, with bounds "
                     "and maxima)", n_seg, kMaxSegments);
  if (noise_seed > int64_t(UINT32_MAX))
    return set_error(NSVF_EINVAL, "render_sample_fine: noise seed %lld beyond 32 bits",
                     static_cast<long long>(noise_seed));
  const dim3 grid(unsigned(std::min<int64_t>(ceil_div(n_hit_max, kSampleWaves), 8192)));
  const bool noisy = noise_seed >= 0;
  const uint32_t seed = uint32_t(noisy ? noise_seed : 0);
  auto kernel = c_field ? (noisy ? fine_sample_kernel<true, true> : fine_sample_kernel<true, false>)
                        : (noisy ? fine_sample_kernel<false, true> : fine_sample_kernel<false, false>);
  hipLaunchKernelGGL(kernel, grid, dim3(kBlock), 0, as_stream(stream), hit_list, n_hit_dev, c_off,
                     c_cnt, c_max, static_cast<const SampleRec*>(coarse),
                     static_cast<const float4*>(c_field), c_prob, f_bound, f_off, fine_steps,
                     compat, static_cast<SampleRec*>(samples), f_cnt, src, seg_bounds, n_seg,
                     seg_max, seed);
  return check_launch("render_sample_fine");
}

int64_t nsvf_render_eval_list_blocks(int64_t n_max) {
  return n_max > 0 ? ceil_div(n_max, kMarkPerBlock) : 0;
}

int nsvf_render_eval_list_count(const int16_t* src, const int32_t* n_dev, int64_t n_max,
                                int32_t* block_cnt, void* stream) {
  clear_error();
  if (n_max <= 0) return NSVF_OK;
  if (!src || !n_dev || !block_cnt) return set_error(NSVF_EINVAL, "eval_list_count: null pointer");
  hipLaunchKernelGGL(mark_count_kernel, dim3(unsigned(nsvf_render_eval_list_blocks(n_max))),
                     dim3(kBlock), 0, as_stream(stream), src, n_dev, n_max, block_cnt);
  return check_launch("render_eval_list_count");
}

int nsvf_render_eval_list(const int16_t* src, const int32_t* n_dev, int64_t n_max,
                          const int32_t* block_off, int32_t* list, void* stream) {
  clear_error();
  if (n_max <= 0) return NSVF_OK;
  if (!src || !n_dev || !block_off || !list)
    return set_error(NSVF_EINVAL, "render_eval_list: null pointer");
  hipLaunchKernelGGL(mark_compact_kernel, dim3(unsigned(nsvf_render_eval_list_blocks(n_max))),
                     dim3(kBlock), 0, as_stream(stream), src, n_dev, n_max, block_off, list);
  return check_launch("render_eval_list");
}

int nsvf_render_reuse_fill(const int32_t* hit_list, const int32_t* n_hit_dev, int64_t n_hit_max,
                           const int32_t* c_off, const int32_t* f_off, const int32_t* f_cnt,
                           const int16_t* src, const float* c_field, float* f_field,
                           void* stream) {
  clear_error();
  if (n_hit_max <= 0) return NSVF_OK;
  if (!hit_list || !n_hit_dev || !c_off || !f_off || !f_cnt || !src || !c_field || !f_field)
    return set_error(NSVF_EINVAL, "render_reuse_fill: null pointer");
  const int per_block = kBlock / 64;
  hipLaunchKernelGGL(reuse_fill_kernel, dim3(unsigned(ceil_div(n_hit_max, per_block))),
                     dim3(kBlock), 0, as_stream(stream), hit_list, n_hit_dev, c_off, f_off, f_cnt,
                     src, reinterpret_cast<const float4*>(c_field),
                     reinterpret_cast<float4*>(f_field));
  return check_launch("render_reuse_fill");
}

int nsvf_render_composite(const int32_t* hit_list, const int32_t* n_hit_dev, int64_t n_hit_max,
                          const int32_t* off, const int32_t* cnt, const void* samples,
                          const float* field, float* prob_out, float* ray_rgb, float* ray_depth,
                          float* ray_missed, void* stream) {
  clear_error();
  if (n_hit_max <= 0) return NSVF_OK;
  const int per_block = kBlock / 64;
  hipLaunchKernelGGL(composite_kernel<false>, dim3(unsigned(ceil_div(n_hit_max, per_block))),
                     dim3(kBlock), 0, as_stream(stream), hit_list, n_hit_dev, off, cnt,
                     static_cast<const SampleRec*>(samples), reinterpret_cast<const float4*>(field),
                     prob_out, ray_rgb, ray_depth, ray_missed, nullptr, nullptr, nullptr);
  return check_launch("render_composite");
}

int nsvf_render_composite_reuse(const int32_t* hit_list, const int32_t* n_hit_dev,
                                int64_t n_hit_max, const int32_t* off, const int32_t* cnt,
                                const void* samples, const float* field, const int16_t* src,
                                const int32_t* c_off, const float* c_field, float* prob_out,
                                float* ray_rgb, float* ray_depth, float* ray_missed,
                                void* stream) {
  clear_error();
  if (n_hit_max <= 0) return NSVF_OK;
  if (!src || !c_off || !c_field)
    return set_error(NSVF_EINVAL, "render_composite_reuse: null pointer");
  const int per_block = kBlock / 64;
  hipLaunchKernelGGL(composite_kernel<true>, dim3(unsigned(ceil_div(n_hit_max, per_block))),
                     dim3(kBlock), 0, as_stream(stream), hit_list, n_hit_dev, off, cnt,
                     static_cast<const SampleRec*>(samples), reinterpret_cast<const float4*>(field),
                     prob_out, ray_rgb, ray_depth, ray_missed, src, c_off,
                     reinterpret_cast<const float4*>(c_field));
  return check_launch("render_composite_reuse");
}

int nsvf_track_voxel_probs(const int32_t* hit_list, const int32_t* n_hit_dev, int64_t n_hit_max,
                           const int32_t* off, const int32_t* cnt, const void* samples,
                           const float* probs, float* max_voxel_probs, int64_t N, void* stream) {
  clear_error();
  if (n_hit_max <= 0) return NSVF_OK;
  if (N <= 0) return set_error(NSVF_EINVAL, "track_voxel_probs: no voxels");
  if (!hit_list || !n_hit_dev || !off || !cnt || !samples || !probs || !max_voxel_probs)
    return set_error(NSVF_EINVAL, "track_voxel_probs: null pointer");
  hipLaunchKernelGGL(track_probs_kernel,
                     dim3(unsigned(std::min<int64_t>(ceil_div(n_hit_max, kSampleWaves), 8192))),
                     dim3(kBlock), 0, as_stream(stream), hit_list, n_hit_dev, off, cnt,
                     static_cast<const SampleRec*>(samples), probs, max_voxel_probs, N);
  return check_launch("track_voxel_probs");
}

int nsvf_render_composite_backward(const int32_t* hit_list, const int32_t* n_hit_dev,
                                   int64_t n_hit_max, const int32_t* off, const int32_t* cnt,
                                   const void* samples, const float* field, const float* grad_rgb,
                                   const float* grad_depth, const float* grad_missed,
                                   float* grad_field, void* stream) {
  clear_error();
  if (n_hit_max <= 0) return NSVF_OK;
  if (!hit_list || !n_hit_dev || !off || !cnt || !samples || !field || !grad_rgb ||
      !grad_depth || !grad_missed || !grad_field)
    return set_error(NSVF_EINVAL, "render_composite_backward: null pointer");
  const int per_block = kBlock / 64;
  hipLaunchKernelGGL(composite_backward_kernel, dim3(unsigned(ceil_div(n_hit_max, per_block))),
                     dim3(kBlock), 0, as_stream(stream), hit_list, n_hit_dev, off, cnt,
                     static_cast<const SampleRec*>(samples), reinterpret_cast<const float4*>(field),
                     grad_rgb, grad_depth, grad_missed, reinterpret_cast<float4*>(grad_field));
  return check_launch("render_composite_backward");
}

int nsvf_render_finish(int64_t R, const int32_t* hit_flag, const float* ray_rgb,
                       const float* ray_depth, const float* ray_missed, float bg_r, float bg_g,
                       float bg_b, float bg_depth, float* rgb, float* depth, float* missed,
                       void* stream) {
  clear_error();
  if (R <= 0) return NSVF_OK;
  hipLaunchKernelGGL(finish_kernel, dim3(unsigned(ceil_div(R, kBlock))), dim3(kBlock), 0,
                     as_stream(stream), R, hit_flag, ray_rgb, ray_depth, ray_missed, bg_r, bg_g,
                     bg_b, bg_depth, rgb, depth, missed);
  return check_launch("render_finish");
}

}  // extern "C"
