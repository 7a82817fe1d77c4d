// Early ray termination as the reference's chunked ray marching does it
// (fairnr/modules/renderer.py:136-174, with forward():234-247 splitting the
// rows): the hit rays of a pass go in groups of `chunk` rows; within a
// group the sample columns go in chunks of at most `chunk` hits; after each
// chunk every row adds its chunk's free energy to a running total, and rows
// past -log(tolerance) stop -- their later samples are not evaluated and no
// longer count toward chunk sizes.
//
// One step advances every group by one chunk, in six launches and with no
// host synchronisation: a histogram of the active rows' sample counts per
// group, a plan (one thread per group: the chunk's end column from the
// histogram), the chunk's per-row sample counts, their scan (scan.hip), the
// evaluation list, and -- after the field kernels ran over the list -- the
// free-energy totals and stop flags.  Row n of a pass is hit ray
// hit_list[n]; its samples are stream slots off[r], off[r] + 1, ... (column
// j = slot off[r] + j).
#include <algorithm>

#include "common.h"

namespace nsvf {
namespace {

struct MarchRec {  // a stream record {voxel id, depth, length, ray}
  int32_t idx;
  float depth;
  float dist;
  int32_t ray;
};

__global__ __launch_bounds__(256) void march_hist_kernel(const int32_t* __restrict__ hit_list,
                                                         int64_t n_hit,
                                                         const int32_t* __restrict__ cnt,
                                                         const int32_t* __restrict__ active,
                                                         int chunk, int S,
                                                         int32_t* __restrict__ hist) {
  const int64_t n = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (n >= n_hit || !active[n]) return;
  const int c = min(cnt[hit_list[n]], S);
  atomicAdd(&hist[(n / chunk) * (S + 1) + c], 1);
}

// span[g] = [start, end) of group g's next chunk (start == end == S when the
// group is done); next[g] advances to end.  hist[g] is consumed and cleared.
__global__ __launch_bounds__(256) void march_plan_kernel(int32_t* __restrict__ hist, int groups,
                                                         int S, int64_t chunk,
                                                         int32_t* __restrict__ span,
                                                         int32_t* __restrict__ next) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= groups) return;
  int32_t* h = hist + int64_t(g) * (S + 1);
  const int start = next[g];
  int end = S;
  if (start < S) {
    // hits in column i = active rows with more than i samples
    int64_t col = 0;
    for (int c = start + 1; c <= S; ++c) col += h[c];
    int64_t size = 0;
    int i = start;
    for (; i < S; ++i) {
      if (size + col > chunk && i > start) break;
      size += col;
      col -= h[i + 1];
    }
    end = i;
  } else {
    end = S;
  }
  span[2 * g] = start < S ? start : S;
  span[2 * g + 1] = end;
  next[g] = end;
  for (int c = 0; c <= S; ++c) h[c] = 0;
}

__global__ __launch_bounds__(256) void march_counts_kernel(const int32_t* __restrict__ hit_list,
                                                           int64_t n_hit,
                                                           const int32_t* __restrict__ cnt,
                                                           const int32_t* __restrict__ active,
                                                           int chunk,
                                                           const int32_t* __restrict__ span,
                                                           int32_t* __restrict__ ecnt) {
  const int64_t n = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (n >= n_hit) return;
  const int64_t g = n / chunk;
  const int s = span[2 * g], e = min(span[2 * g + 1], cnt[hit_list[n]]);
  ecnt[n] = (active[n] && e > s) ? e - s : 0;
}

__global__ __launch_bounds__(256) void march_list_kernel(
    const int32_t* __restrict__ hit_list, int64_t n_hit, const int32_t* __restrict__ off,
    int chunk, const int32_t* __restrict__ span, const int32_t* __restrict__ ecnt,
    const int32_t* __restrict__ eoff, int32_t* __restrict__ list, int8_t* __restrict__ mark) {
  const int64_t n = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (n >= n_hit) return;
  const int c = ecnt[n];
  if (c == 0) return;
  const int32_t base = off[hit_list[n]] + span[2 * (n / chunk)];
  const int32_t o = eoff[n];
  for (int t = 0; t < c; ++t) {
    list[o + t] = base + t;
    if (mark) mark[base + t] = 1;
  }
}

// Row totals += the chunk's free energies relu(sigma) * dist * 7, added left
// to right in float32 (renderer.py:119-121,162-168); rows past `threshold`
// stop.
__global__ __launch_bounds__(256) void march_accumulate_kernel(
    const int32_t* __restrict__ hit_list, int64_t n_hit, const int32_t* __restrict__ off,
    const int32_t* __restrict__ ecnt, int chunk, const int32_t* __restrict__ span,
    const MarchRec* __restrict__ samples, const float4* __restrict__ field,
    float* __restrict__ total, int32_t* __restrict__ active, float threshold) {
  const int64_t n = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (n >= n_hit) return;
  const int c = ecnt[n];
  if (c == 0) return;
  const int64_t base = int64_t(off[hit_list[n]]) + span[2 * (n / chunk)];
  float part = 0.0f;
  for (int t = 0; t < c; ++t) {
    const float sg = field[base + t].x;
    part = part + ((sg > 0.0f ? sg : 0.0f) * samples[base + t].dist) * 7.0f;
  }
  const float acc = total[n] + part;
  total[n] = acc;
  if (acc > threshold) active[n] = 0;
}

inline unsigned blocks_of(int64_t n) { return unsigned(std::max<int64_t>(1, ceil_div(n, 256))); }

}  // namespace
}  // namespace nsvf

using namespace nsvf;

extern "C" {

int nsvf_march_step_begin(const int32_t* hit_list, int64_t n_hit, const int32_t* cnt,
                          const int32_t* active, int chunk, int S, int32_t* hist, int32_t* span,
                          int32_t* next, int32_t* ecnt, void* stream) {
  clear_error();
  if (n_hit <= 0) return NSVF_OK;
  if (chunk <= 0 || S < 0)
    return set_error(NSVF_EINVAL, "march_step_begin: chunk %d, S %d", chunk, S);
  if (!hit_list || !cnt || !active || !hist || !span || !next || !ecnt)
    return set_error(NSVF_EINVAL, "march_step_begin: null pointer");
  const int groups = int(ceil_div(n_hit, chunk));
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(march_hist_kernel, dim3(blocks_of(n_hit)), dim3(256), 0, st, hit_list,
                     n_hit, cnt, active, chunk, S, hist);
  hipLaunchKernelGGL(march_plan_kernel, dim3(blocks_of(groups)), dim3(256), 0, st, hist, groups,
                     S, int64_t(chunk), span, next);
  hipLaunchKernelGGL(march_counts_kernel, dim3(blocks_of(n_hit)), dim3(256), 0, st, hit_list,
                     n_hit, cnt, active, chunk, span, ecnt);
  return check_launch("march_step_begin");
}

int nsvf_march_step_list(const int32_t* hit_list, int64_t n_hit, const int32_t* off, int chunk,
                         const int32_t* span, const int32_t* ecnt, const int32_t* eoff,
                         int32_t* list, int8_t* mark, void* stream) {
  clear_error();
  if (n_hit <= 0) return NSVF_OK;
  if (chunk <= 0) return set_error(NSVF_EINVAL, "march_step_list: chunk %d", chunk);
  if (!hit_list || !off || !span || !ecnt || !eoff || !list)
    return set_error(NSVF_EINVAL, "march_step_list: null pointer");
  hipLaunchKernelGGL(march_list_kernel, dim3(blocks_of(n_hit)), dim3(256), 0, as_stream(stream),
                     hit_list, n_hit, off, chunk, span, ecnt, eoff, list, mark);
  return check_launch("march_step_list");
}

int nsvf_march_step_end(const int32_t* hit_list, int64_t n_hit, const int32_t* off,
                        const int32_t* ecnt, int chunk, const int32_t* span, const void* samples,
                        const float* field, float* total, int32_t* active, float threshold,
                        void* stream) {
  clear_error();
  if (n_hit <= 0) return NSVF_OK;
  if (chunk <= 0) return set_error(NSVF_EINVAL, "march_step_end: chunk %d", chunk);
  if (!hit_list || !off || !ecnt || !span || !samples || !field || !total || !active)
    return set_error(NSVF_EINVAL, "march_step_end: null pointer");
  hipLaunchKernelGGL(march_accumulate_kernel, dim3(blocks_of(n_hit)), dim3(256), 0,
                     as_stream(stream), hit_list, n_hit, off, ecnt, chunk, span,
                     static_cast<const MarchRec*>(samples), reinterpret_cast<const float4*>(field),
                     total, active, threshold);
  return check_launch("march_step_end");
}

}  // extern "C"
