// Backward of the trilinear embedding gather: the gradient of the corner
// embedding table from the gradient of the gathered embeddings.
//
// Forward (encoder.py:582-590, geometry.py:195-200): out[m] = sum_c w_c(m) *
// values[feats[v_m][c]].  Its adjoint, which torch forms for the reference as
// the backward of F.embedding (an index-add into values.weight), is
//   grad_values[k] = sum over (v, c) with feats[v][c] == k of
//                    sum over samples m in voxel v of w_c(m) * grad_out[m].
// The reference's index-add uses atomics, so its summation order (and last
// bits) vary run to run.  Here the sums are deterministic and atomic-free:
//
//   1. sort the samples by voxel (stable radix sort; slot order within a
//      voxel) and find each voxel's segment by binary search;
//   2. one wavefront per voxel: lane (c, j) accumulates w_c(m) * grad_out[m]
//      over the voxel's samples for corner c and features 4j.. (+32 ...),
//      in sample order, into G[v][c] -- the per-(voxel, corner) partials;
//   3. sort the (voxel, corner) incidences by corner key the same way, and
//      sum each key's <= 8 partials G[v][c] in incidence order.  The
//      incidence order depends only on the scene's corner keys, so a caller
//      that differentiates the same scene repeatedly computes it once (a
//      "plan": the sorted incidences and the key segments) and passes it.
//
// The corner weights are the forward gather's, computed with the same
// operations, so the backward is the exact adjoint of the forward's weights.
// Traffic: the sample gradients are read once (plus 12 bytes of position
// and 8 of sort keys per sample); G is N * 8 * D floats written once and
// read once; the table gradient is written once.
#include <hipcub/hipcub.hpp>

#include "common.h"

namespace nsvf {
namespace {

constexpr int kBlock = 256;

inline int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

// bits needed for keys in [0, n]
inline int key_bits(int64_t n) {
  int b = 1;
  while (b < 31 && (int64_t(1) << b) <= n) ++b;
  return b;
}

struct GradWorkspace {
  uint32_t* keys_in;   // [M] voxel of each sample (N for unused)
  uint32_t* keys_out;
  int32_t* vals_in;    // [M] sample ids
  int32_t* vals_out;
  int32_t* seg;        // [N + 1] voxel segment starts in the sorted samples
  uint32_t* ikeys_in;  // [8N] corner key of each (voxel, corner)
  uint32_t* ikeys_out;
  int32_t* ivals_in;   // [8N] incidence ids v * 8 + c
  int32_t* ivals_out;
  int32_t* kseg;       // [K + 1] key segment starts in the sorted incidences
  float* partial;      // [8N, D] per-(voxel, corner) partial gradients
  void* sort_tmp;
  size_t sort_tmp_bytes;
  int64_t total;
};

int sort_temp_bytes(int64_t n, int bits, size_t* bytes, hipStream_t stream) {
  *bytes = 0;
  const hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      nullptr, *bytes, static_cast<const uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
      static_cast<const int32_t*>(nullptr), static_cast<int32_t*>(nullptr), int(n), 0, bits,
      stream);
  if (e != hipSuccess)
    return set_error(NSVF_ELAUNCH, "radix sort size query: %s", hipGetErrorString(e));
  return NSVF_OK;
}

int layout(int64_t M, int64_t N, int64_t K, int D, hipStream_t stream, char* base,
           GradWorkspace* w) {
  size_t t1 = 0, t2 = 0;
  int rc = sort_temp_bytes(M, key_bits(N), &t1, stream);
  if (rc) return rc;
  rc = sort_temp_bytes(8 * N, key_bits(K), &t2, stream);
  if (rc) return rc;
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    char* p = base ? base + o : nullptr;
    o += align256(bytes);
    return p;
  };
  w->keys_in = reinterpret_cast<uint32_t*>(take(M * 4));
  w->keys_out = reinterpret_cast<uint32_t*>(take(M * 4));
  w->vals_in = reinterpret_cast<int32_t*>(take(M * 4));
  w->vals_out = reinterpret_cast<int32_t*>(take(M * 4));
  w->seg = reinterpret_cast<int32_t*>(take((N + 1) * 4));
  w->ikeys_in = reinterpret_cast<uint32_t*>(take(8 * N * 4));
  w->ikeys_out = reinterpret_cast<uint32_t*>(take(8 * N * 4));
  w->ivals_in = reinterpret_cast<int32_t*>(take(8 * N * 4));
  w->ivals_out = reinterpret_cast<int32_t*>(take(8 * N * 4));
  w->kseg = reinterpret_cast<int32_t*>(take((K + 1) * 4));
  w->partial = reinterpret_cast<float*>(take(8 * N * int64_t(D) * 4));
  w->sort_tmp_bytes = std::max(t1, t2);
  w->sort_tmp = take(int64_t(w->sort_tmp_bytes));
  w->total = o;
  return NSVF_OK;
}

__global__ __launch_bounds__(kBlock) void sample_keys_kernel(const int32_t* __restrict__ s_idx,
                                                             int64_t M, int32_t N,
                                                             uint32_t* __restrict__ keys,
                                                             int32_t* __restrict__ ids) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= M) return;
  const int32_t v = s_idx[i];
  keys[i] = uint32_t((v < 0 || v >= N) ? N : v);
  ids[i] = int32_t(i);
}

__global__ __launch_bounds__(kBlock) void incidence_keys_kernel(
    const int32_t* __restrict__ feats, int64_t n8, int32_t K, uint32_t* __restrict__ keys,
    int32_t* __restrict__ ids) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n8) return;
  const int32_t k = feats[i];
  keys[i] = uint32_t((k < 0 || k >= K) ? K : k);
  ids[i] = int32_t(i);
}

// seg[v] = first position of a key >= v in sorted[0, n), v = 0..nseg-1.
__global__ __launch_bounds__(kBlock) void segment_starts_kernel(
    const uint32_t* __restrict__ sorted, int64_t n, int64_t nseg, int32_t* __restrict__ seg) {
  const int64_t v = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (v >= nseg) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (int64_t(sorted[mid]) < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  seg[v] = int32_t(lo);
}

// One wavefront per voxel; lane = 8 * corner + j.  Feature float4 groups
// q = j, j + 8, ... < D / 4 (D <= 128).
__global__ __launch_bounds__(kBlock) void voxel_corner_kernel(
    const int32_t* __restrict__ seg, const int32_t* __restrict__ order,
    const float* __restrict__ xyz, const float* __restrict__ centers, float voxel_size,
    const float* __restrict__ grad_out, int D, int64_t N, float* __restrict__ partial) {
  const int64_t v = int64_t(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  if (v >= N) return;
  const int lane = threadIdx.x & 63;
  const int c = lane >> 3, j = lane & 7;
  const int DQ = D >> 2;
  const float cx = centers[v * 3 + 0], cy = centers[v * 3 + 1], cz = centers[v * 3 + 2];
  float4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  const int32_t s0 = seg[v], s1 = seg[v + 1];
  for (int32_t i = s0; i < s1; ++i) {
    const int64_t m = order[i];
    // corner weight of the forward gather: p = (x - c) / vs + 0.5,
    // w = (wx * wy) * wz
    const float px = (xyz[m * 3 + 0] - cx) / voxel_size + 0.5f;
    const float py = (xyz[m * 3 + 1] - cy) / voxel_size + 0.5f;
    const float pz = (xyz[m * 3 + 2] - cz) / voxel_size + 0.5f;
    const float wx = (c & 4) ? px : 1.0f - px;
    const float wy = (c & 2) ? py : 1.0f - py;
    const float wz = (c & 1) ? pz : 1.0f - pz;
    const float w = (wx * wy) * wz;
    const float4* g = reinterpret_cast<const float4*>(grad_out + m * D);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int q = j + 8 * t;
      if (q < DQ) {
        const float4 x = g[q];
        acc[t].x = acc[t].x + w * x.x;
        acc[t].y = acc[t].y + w * x.y;
        acc[t].z = acc[t].z + w * x.z;
        acc[t].w = acc[t].w + w * x.w;
      }
    }
  }
  float4* out = reinterpret_cast<float4*>(partial + (v * 8 + c) * D);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int q = j + 8 * t;
    if (q < DQ) out[q] = acc[t];
  }
}

// Thread per (key, float4 group): sum of the key's (voxel, corner) partials
// in incidence order.
__global__ __launch_bounds__(kBlock) void key_sum_kernel(const int32_t* __restrict__ kseg,
                                                         const int32_t* __restrict__ inc,
                                                         const float* __restrict__ partial,
                                                         int64_t K, int D,
                                                         float* __restrict__ grad_values) {
  const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int DQ = D >> 2;
  if (t >= K * DQ) return;
  const int64_t k = t / DQ;
  const int q = int(t - k * DQ);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int32_t i = kseg[k]; i < kseg[k + 1]; ++i) {
    const float4 x = reinterpret_cast<const float4*>(partial + int64_t(inc[i]) * D)[q];
    acc.x = acc.x + x.x;
    acc.y = acc.y + x.y;
    acc.z = acc.z + x.z;
    acc.w = acc.w + x.w;
  }
  reinterpret_cast<float4*>(grad_values + k * D)[q] = acc;
}

// Plan layout: [8N] sorted incidence ids, then [K + 1] key segment starts.
inline int64_t plan_bytes(int64_t N, int64_t K) { return align256(8 * N * 4) + (K + 1) * 4; }

// Step 3's incidence order into plan (ivals: 8N ids, kseg: K + 1 starts),
// using the workspace's incidence buffers and sort space.
int build_plan(const int32_t* feats, int64_t N, int64_t K, const GradWorkspace& w,
               int32_t* ivals, int32_t* kseg, hipStream_t st) {
  hipLaunchKernelGGL(incidence_keys_kernel, dim3(unsigned(ceil_div(8 * N, kBlock))), dim3(kBlock),
                     0, st, feats, 8 * N, int32_t(K), w.ikeys_in, w.ivals_in);
  size_t tb = w.sort_tmp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, tb, w.ikeys_in, w.ikeys_out,
                                                    w.ivals_in, ivals, int(8 * N), 0,
                                                    key_bits(K), st);
  if (e != hipSuccess)
    return set_error(NSVF_ELAUNCH, "gather backward: incidence sort: %s", hipGetErrorString(e));
  hipLaunchKernelGGL(segment_starts_kernel, dim3(unsigned(ceil_div(K + 1, kBlock))), dim3(kBlock),
                     0, st, w.ikeys_out, 8 * N, K + 1, kseg);
  return check_launch("gather backward plan");
}

int check_sizes(const char* what, int64_t M, int64_t N, int64_t K, int D) {
  if (M < 0 || N <= 0 || K <= 0 || D <= 0) return set_error(NSVF_EINVAL, "%s: bad size", what);
  if (D % 4 != 0 || D > 128)
    return set_error(NSVF_EINVAL, "%s: D=%d (a multiple of 4, <= 128)", what, D);
  if (M >= (int64_t(1) << 31) || 8 * N >= (int64_t(1) << 31) || K >= (int64_t(1) << 30))
    return set_error(NSVF_EINVAL, "%s: size beyond 32-bit indexing", what);
  return NSVF_OK;
}

}  // namespace
}  // namespace nsvf

using namespace nsvf;

extern "C" {

int64_t nsvf_trilinear_gather_backward_plan_bytes(int64_t N, int64_t K) {
  clear_error();
  if (N <= 0 || K <= 0) {
    set_error(NSVF_EINVAL, "trilinear_gather_backward_plan_bytes: bad size");
    return -1;
  }
  return plan_bytes(N, K);
}

int nsvf_trilinear_gather_backward_plan(const int32_t* feats, int64_t N, int64_t K, void* plan,
                                        int64_t plan_bytes_given, void* workspace,
                                        int64_t workspace_bytes, void* stream) {
  clear_error();
  if (int rc = check_sizes("trilinear_gather_backward_plan", 0, N, K, 4)) return rc;
  if (!feats || !plan || !workspace)
    return set_error(NSVF_EINVAL, "trilinear_gather_backward_plan: null pointer");
  if (plan_bytes_given < plan_bytes(N, K))
    return set_error(NSVF_EINVAL, "trilinear_gather_backward_plan: plan buffer too small");
  hipStream_t st = as_stream(stream);
  GradWorkspace w;
  // the incidence buffers and sort space do not depend on M or D
  if (int rc = layout(0, N, K, 4, st, static_cast<char*>(workspace), &w)) return rc;
  if (workspace_bytes < w.total)
    return set_error(NSVF_EINVAL, "trilinear_gather_backward_plan: workspace too small");
  char* pb = static_cast<char*>(plan);
  return build_plan(feats, N, K, w, reinterpret_cast<int32_t*>(pb),
                    reinterpret_cast<int32_t*>(pb + align256(8 * N * 4)), st);
}

int64_t nsvf_trilinear_gather_backward_workspace(int64_t M, int64_t N, int64_t K, int D,
                                                 void* stream) {
  clear_error();
  if (M < 0 || N <= 0 || K <= 0 || D <= 0) {
    set_error(NSVF_EINVAL, "trilinear_gather_backward_workspace: bad size");
    return -1;
  }
  GradWorkspace w;
  if (layout(M, N, K, D, as_stream(stream), nullptr, &w)) return -1;
  return w.total;
}

int nsvf_trilinear_gather_backward_planned(const int32_t* s_idx, const float* xyz, int64_t M,
                                           const float* centers, const int32_t* feats, int64_t N,
                                           int64_t K, int D, float voxel_size,
                                           const float* grad_out, float* grad_values,
                                           const void* plan, void* workspace,
                                           int64_t workspace_bytes, void* stream) {
  clear_error();
  if (int rc = check_sizes("trilinear_gather_backward", M, N, K, D)) return rc;
  if ((M > 0 && (!s_idx || !xyz || !grad_out)) || !centers || !feats || !grad_values ||
      !workspace)
    return set_error(NSVF_EINVAL, "trilinear_gather_backward: null pointer");
  if ((reinterpret_cast<uintptr_t>(grad_out) | reinterpret_cast<uintptr_t>(grad_values)) & 15)
    return set_error(NSVF_EINVAL, "trilinear_gather_backward: gradients must be 16-byte aligned");
  hipStream_t st = as_stream(stream);
  GradWorkspace w;
  int rc = layout(M, N, K, D, st, static_cast<char*>(workspace), &w);
  if (rc) return rc;
  if (workspace_bytes < w.total)
    return set_error(NSVF_EINVAL, "trilinear_gather_backward: workspace too small (%lld < %lld)",
                     (long long)workspace_bytes, (long long)w.total);
  // 1. samples sorted by voxel
  if (M > 0) {
    hipLaunchKernelGGL(sample_keys_kernel, dim3(unsigned(ceil_div(M, kBlock))), dim3(kBlock), 0,
                       st, s_idx, M, int32_t(N), w.keys_in, w.vals_in);
    size_t tb = w.sort_tmp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, tb, w.keys_in, w.keys_out,
                                                      w.vals_in, w.vals_out, int(M), 0,
                                                      key_bits(N), st);
    if (e != hipSuccess)
      return set_error(NSVF_ELAUNCH, "gather backward: sample sort: %s", hipGetErrorString(e));
  }
  hipLaunchKernelGGL(segment_starts_kernel, dim3(unsigned(ceil_div(N + 1, kBlock))), dim3(kBlock),
                     0, st, w.keys_out, M, N + 1, w.seg);
  // 2. per-(voxel, corner) partials
  hipLaunchKernelGGL(voxel_corner_kernel, dim3(unsigned(ceil_div(N, kBlock / 64))), dim3(kBlock),
                     0, st, w.seg, w.vals_out, xyz, centers, voxel_size, grad_out, D, N,
                     w.partial);
  // 3. incidences in corner-key order (the plan, or computed here), then
  // one sum per key
  const int32_t* ivals = w.ivals_out;
  const int32_t* kseg = w.kseg;
  if (plan) {
    ivals = static_cast<const int32_t*>(plan);
    kseg = reinterpret_cast<const int32_t*>(static_cast<const char*>(plan) + align256(8 * N * 4));
  } else if (int rc2 = build_plan(feats, N, K, w, w.ivals_out, w.kseg, st)) {
    return rc2;
  }
  hipLaunchKernelGGL(key_sum_kernel, dim3(unsigned(ceil_div(K * (D / 4), kBlock))), dim3(kBlock),
                     0, st, kseg, ivals, w.partial, K, D, grad_values);
  return check_launch("trilinear_gather_backward");
}

int nsvf_trilinear_gather_backward(const int32_t* s_idx, const float* xyz, int64_t M,
                                   const float* centers, const int32_t* feats, int64_t N,
                                   int64_t K, int D, float voxel_size, const float* grad_out,
                                   float* grad_values, void* workspace, int64_t workspace_bytes,
                                   void* stream) {
  return nsvf_trilinear_gather_backward_planned(s_idx, xyz, M, centers, feats, N, K, D, voxel_size,
                                                grad_out, grad_values, nullptr, workspace,
                                                workspace_bytes, stream);
}

}  // extern "C"
