// Early ray termination as the reference's chunked ray marching decides it
// (fairnr/modules/renderer.py:136-174, with forward():234-247 splitting the
// rows): the hit rays of a pass go in groups of `chunk` rows; within a
// group the sample columns go in chunks of at most `chunk` hits of rows
// still marching; after each chunk every row adds its chunk's free energy
// to a running total, and rows past -log(tolerance) stop -- their later
// samples are not evaluated (their field is 0) and no longer count toward
// chunk sizes.
//
// The outputs depend only on which samples are evaluated and on their
// values, and a sample's field value does not depend on the others.  So the
// field runs once over every sample of the pass (one launch at full
// occupancy, where the reference runs a forward pass per chunk), then the
// marching is replayed in steps of two small launches -- each group's next
// chunk end from its histogram of marching rows (one block per group), then
// every marching row's chunk sum and stop test (one thread per row, stopped
// rows leaving the histogram) -- marking the samples the reference
// evaluates; finally the field of the others is zeroed.  Row n of a pass is
// hit ray hit_list[n]; its samples are stream slots off[r], off[r] + 1, ...
// (column j = slot off[r] + j).
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

// Every row marching -- active[n] = its sample count (at most S) plus one,
// so that a step reads a row's count without the gather through hit_list --
// total 0; hist[g][c] = rows of group g with c samples (hist zeroed by the
// caller).
__global__ __launch_bounds__(256) void march_init_kernel(const int32_t* __restrict__ hit_list,
                                                         int64_t n_hit,
                                                         const int32_t* __restrict__ cnt,
                                                         int chunk, int S,
                                                         int32_t* __restrict__ active,
                                                         float* __restrict__ total,
                                                         int32_t* __restrict__ hist) {
  const int64_t n = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (n >= n_hit) return;
  const int c = min(cnt[hit_list[n]], S);
  active[n] = c + 1;
  total[n] = 0.0f;
  atomicAdd(&hist[(n / chunk) * (S + 1) + c], 1);
}

// march_init_kernel with the histogram privatised: a block takes kInitRows
// rows, counts them in LDS and adds its nonzero bins to hist once (hist is
// a few cache lines, and 400k rows adding to it one by one serialise on
// them: ~0.5 ms).  Needs chunk >= kInitRows, so a block spans at most two
// groups, and 2 (S + 1) ints of LDS.
constexpr int kInitRows = 4096;
constexpr int kInitLdsInts = 12288;

__global__ __launch_bounds__(256) void march_init_lds_kernel(const int32_t* __restrict__ hit_list,
                                                             int64_t n_hit,
                                                             const int32_t* __restrict__ cnt,
                                                             int chunk, int S, int64_t groups,
                                                             int32_t* __restrict__ active,
                                                             float* __restrict__ total,
                                                             int32_t* __restrict__ hist) {
  extern __shared__ int32_t lh[];  // [2][S + 1]
  const int bins = 2 * (S + 1);
  const int64_t n0 = int64_t(blockIdx.x) * kInitRows;
  const int64_t g0 = n0 / chunk;
  for (int i = threadIdx.x; i < bins; i += 256) lh[i] = 0;
  __syncthreads();
  const int64_t n1 = min(n0 + kInitRows, n_hit);
  for (int64_t n = n0 + threadIdx.x; n < n1; n += 256) {
    const int c = min(cnt[hit_list[n]], S);
    active[n] = c + 1;
    total[n] = 0.0f;
    atomicAdd(&lh[int(n / chunk - g0) * (S + 1) + c], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < bins; i += 256) {
    const int v = lh[i];
    const int64_t g = g0 + i / (S + 1);
    if (v != 0 && g < groups) atomicAdd(&hist[g * (S + 1) + i % (S + 1)], v);
  }
}

constexpr int kPlanThreads = 256;

// Inclusive prefix sum over the block's threads (buf: kPlanThreads slots).
__device__ __forceinline__ int64_t block_scan(int64_t v, int64_t* buf) {
  const int t = threadIdx.x;
  buf[t] = v;
  __syncthreads();
  for (int d = 1; d < kPlanThreads; d <<= 1) {
    const int64_t x = t >= d ? buf[t - d] : 0;
    __syncthreads();
    buf[t] += x;
    __syncthreads();
  }
  const int64_t r = buf[t];
  __syncthreads();
  return r;
}

// span[g] = [start, end) of group g's next chunk ([S, S) once the group is
// done); next[g] advances to end.  A chunk ends before the column that would
// take it past `chunk` hits (hits in column i = marching rows with more than
// i samples) and holds at least one column: end = the first i > start with
// size(i) = sum_{j = start..i} col(j) > chunk, else S.  With T = the hits
// past start, col(j) = T - sum_{c = start+1..j} h[c], so size(i) =
// (i - start + 1) T - sum_{j = start..i} sum_{c = start+1..j} h[c]: a block
// per group computes T by a reduction and size by two scans over column
// tiles of kPlanThreads, exact in 64-bit integers (the one-thread walk over
// the columns that this replaces cost ~20 us per step).
__global__ __launch_bounds__(kPlanThreads) void march_plan_kernel(const int32_t* __restrict__ hist,
                                                                  int S, int64_t chunk,
                                                                  int32_t* __restrict__ span,
                                                                  int32_t* __restrict__ next) {
  __shared__ int64_t buf[kPlanThreads];
  __shared__ int first;
  const int g = blockIdx.x, t = threadIdx.x;
  const int32_t* h = hist + int64_t(g) * (S + 1);
  const int start = next[g];
  if (start >= S) {
    if (t == 0) {
      span[2 * g] = S;
      span[2 * g + 1] = S;
    }
    return;
  }
  int64_t part = 0;
  for (int c = start + 1 + t; c <= S; c += kPlanThreads) part += h[c];
  const int64_t T = block_scan(part, buf);  // thread kPlanThreads - 1 holds the total
  if (t == kPlanThreads - 1) buf[0] = T;
  __syncthreads();
  const int64_t total = buf[0];
  __syncthreads();
  int end = S;
  int64_t hp_carry = 0, gs_carry = 0;
  for (int base = start; base < S; base += kPlanThreads) {
    const int i = base + t;
    const int64_t hp = block_scan(i < S && i > start ? int64_t(h[i]) : 0, buf) + hp_carry;
    const int64_t gs = block_scan(i < S ? hp : 0, buf) + gs_carry;
    const bool over = i < S && i > start && int64_t(i - start + 1) * total - gs > chunk;
    if (t == 0) first = S;
    __syncthreads();
    if (over) atomicMin(&first, i);
    if (t == kPlanThreads - 1) {
      buf[0] = hp;
      buf[1] = gs;
    }
    __syncthreads();
    end = first;
    hp_carry = buf[0];
    gs_carry = buf[1];
    __syncthreads();
    if (end < S) break;
  }
  if (t == 0) {
    span[2 * g] = start;
    span[2 * g + 1] = end;
    next[g] = end;
  }
}

// Each marching row: its chunk's free energies relu(sigma) * dist * 7 added
// left to right in float32 (renderer.py:119-121,162-168), the samples
// marked, then the stop test; a row that stops leaves its group's
// histogram.
__global__ __launch_bounds__(256) void march_accumulate_kernel(
    const int32_t* __restrict__ hit_list, int64_t n_hit, const int32_t* __restrict__ off,
    const MarchRec* __restrict__ samples, const float4* __restrict__ field, int chunk, int S,
    const int32_t* __restrict__ span,
    float threshold, int32_t* __restrict__ active, float* __restrict__ total,
    int32_t* __restrict__ hist, int8_t* __restrict__ mark) {
  const int64_t n = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (n >= n_hit) return;
  // the row's own loads issued together, ahead of the test that needs them
  // (the step is a chain of dependent loads: these, the row's offset, its
  // samples)
  const int c = active[n] - 1;  // -1: stopped
  const int32_t r = hit_list[n];
  const float tot = total[n];
  const int64_t g = n / chunk;
  const int start = span[2 * g];
  const int e = min(span[2 * g + 1], c);
  if (e <= start) return;
  const int64_t base = off[r];
  float part = 0.0f;
  for (int j = start; j < e; ++j) {
    const float sg = field[base + j].x;
    part = part + ((sg > 0.0f ? sg : 0.0f) * samples[base + j].dist) * 7.0f;
    mark[base + j] = 1;
  }
  const float acc = tot + part;
  total[n] = acc;
  if (acc > threshold) {
    active[n] = 0;
    atomicSub(&hist[g * (S + 1) + c], 1);
  }
}

// field[slot] = 0 for the slots of the pass's rows that were not marked.
__global__ __launch_bounds__(256) void march_zero_kernel(const int32_t* __restrict__ hit_list,
                                                         int64_t n_hit,
                                                         const int32_t* __restrict__ off,
                                                         const int32_t* __restrict__ cnt,
                                                         const int8_t* __restrict__ mark,
                                                         float4* __restrict__ field) {
  // one wave per row, lanes over its samples
  const int64_t n = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (n >= n_hit) return;
  const int32_t r = hit_list[n];
  const int64_t base = off[r];
  const int c = cnt[r];
  for (int j = threadIdx.x & 63; j < c; j += 64)
    if (!mark[base + j]) field[base + j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace
}  // namespace nsvf

using namespace nsvf;

extern "C" {

int64_t nsvf_march_hist_ints(int64_t n_hit, int chunk, int S) {
  if (n_hit < 0 || chunk <= 0 || S < 0) return -1;
  return ceil_div(std::max<int64_t>(n_hit, 1), chunk) * (S + 1);
}

int nsvf_march_begin(const int32_t* hit_list, int64_t n_hit, const int32_t* cnt, int chunk,
                     int S, int32_t* active, float* total, int32_t* hist, void* stream) {
  clear_error();
  if (n_hit <= 0) return NSVF_OK;
  if (chunk <= 0 || S < 0) return set_error(NSVF_EINVAL, "march_begin: chunk %d, S %d", chunk, S);
  if (!hit_list || !cnt || !active || !total || !hist)
    return set_error(NSVF_EINVAL, "march_begin: null pointer");
  if (chunk >= kInitRows && 2 * (S + 1) <= kInitLdsInts) {
    hipLaunchKernelGGL(march_init_lds_kernel, dim3(unsigned(ceil_div(n_hit, kInitRows))),
                       dim3(256), size_t(2 * (S + 1)) * sizeof(int32_t), as_stream(stream),
                       hit_list, n_hit, cnt, chunk, S, ceil_div(n_hit, chunk), active, total,
                       hist);
  } else {
    hipLaunchKernelGGL(march_init_kernel, dim3(unsigned(ceil_div(n_hit, 256))), dim3(256), 0,
                       as_stream(stream), hit_list, n_hit, cnt, chunk, S, active, total, hist);
  }
  return check_launch("march_begin");
}

int nsvf_march_step(const int32_t* hit_list, int64_t n_hit, const int32_t* off,
                    const int32_t* cnt, const void* samples, const float* field, int chunk, int S,
                    float threshold, int32_t* active, float* total, int32_t* hist, int32_t* span,
                    int32_t* next, int8_t* mark, void* stream) {
  clear_error();
  if (n_hit <= 0) return NSVF_OK;
  if (chunk <= 0 || S < 0) return set_error(NSVF_EINVAL, "march_step: chunk %d, S %d", chunk, S);
  if (!hit_list || !off || !cnt || !samples || !field || !active || !total || !hist || !span ||
      !next || !mark)
    return set_error(NSVF_EINVAL, "march_step: null pointer");
  const int64_t groups = ceil_div(n_hit, chunk);
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(march_plan_kernel, dim3(unsigned(groups)), dim3(kPlanThreads), 0, st, hist,
                     S, int64_t(chunk), span, next);
  hipLaunchKernelGGL(march_accumulate_kernel, dim3(unsigned(ceil_div(n_hit, 256))), dim3(256), 0,
                     st, hit_list, n_hit, off, static_cast<const MarchRec*>(samples),
                     reinterpret_cast<const float4*>(field), chunk, S, span, threshold, active,
                     total, hist, mark);
  return check_launch("march_step");
}

int nsvf_march_end(const int32_t* hit_list, int64_t n_hit, const int32_t* off, const int32_t* cnt,
                   const int8_t* mark, float* field, void* stream) {
  clear_error();
  if (n_hit <= 0) return NSVF_OK;
  if (!hit_list || !off || !cnt || !mark || !field)
    return set_error(NSVF_EINVAL, "march_end: null pointer");
  hipLaunchKernelGGL(march_zero_kernel, dim3(unsigned(ceil_div(n_hit, 4))), dim3(256), 0,
                     as_stream(stream), hit_list, n_hit, off, cnt, mark,
                     reinterpret_cast<float4*>(field));
  return check_launch("march_end");
}

}  // extern "C"
