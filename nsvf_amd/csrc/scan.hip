// Device-wide exclusive prefix sum of int32 counts (three passes: block sums,
// scan of the block sums in one workgroup, block-local scans with offsets).
// Used to turn per-ray hit flags into compacted ray ranks and per-ray sample
// counts into sample-stream offsets, without a host round trip.
#include "common.h"

namespace nsvf {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;                      // per thread
constexpr int kScanTile = kScanBlock * kScanItems;  // elements per block
constexpr int kPartialItems = 64;                 // per thread, block-sum pass
constexpr int kMaxPartials = kScanBlock * kPartialItems;

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// Inclusive scan of one value per thread across the workgroup.
__device__ __forceinline__ int block_incl_scan(int v, int* smem) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_incl_scan(v);
  if (lane == 63) smem[wave] = v;
  __syncthreads();
  int add = 0;
  for (int w = 0; w < wave; ++w) add += smem[w];
  __syncthreads();
  return v + add;
}

__global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(const int32_t* __restrict__ in,
                                                                 int64_t n,
                                                                 int32_t* __restrict__ partial) {
  __shared__ int smem[kScanBlock / 64];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) s += in[base + i];
  s = block_incl_scan(s, smem);
  if (threadIdx.x == kScanBlock - 1) partial[blockIdx.x] = s;
}

// One workgroup scans up to kMaxPartials block sums (exclusive, in place) and
// writes the grand total.
__global__ __launch_bounds__(kScanBlock) void scan_partials_kernel(int32_t* __restrict__ partial,
                                                                   int nb,
                                                                   int32_t* __restrict__ total) {
  __shared__ int smem[kScanBlock / 64];
  const int base = threadIdx.x * kPartialItems;
  int s = 0;
  for (int i = 0; i < kPartialItems; ++i)
    if (base + i < nb) s += partial[base + i];
  const int incl = block_incl_scan(s, smem);
  int run = incl - s;
  for (int i = 0; i < kPartialItems; ++i) {
    if (base + i < nb) {
      const int v = partial[base + i];
      partial[base + i] = run;
      run += v;
    }
  }
  if (threadIdx.x == kScanBlock - 1 && total) *total = incl;
}

__global__ __launch_bounds__(kScanBlock) void scan_apply_kernel(const int32_t* __restrict__ in,
                                                                int64_t n,
                                                                const int32_t* __restrict__ partial,
                                                                int32_t* __restrict__ out) {
  __shared__ int smem[kScanBlock / 64];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    s += v[i];
  }
  const int incl = block_incl_scan(s, smem);
  int run = partial[blockIdx.x] + incl - s;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
}

}  // namespace

// Exclusive scan of n int32 values into `out` (may alias `in`), grand total
// into *total (device).  `partial` needs scan_partials_size(n) ints.
int scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* partial,
                       int32_t* total, hipStream_t stream) {
  if (n < 0) return set_error(NSVF_EINVAL, "scan: negative n");
  const int64_t nb = ceil_div(n, kScanTile);
  if (nb > kMaxPartials) return set_error(NSVF_EINVAL, "scan: n too large (%lld)", (long long)n);
  if (n == 0) {
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanBlock), 0, stream, partial, 0,
                       total);
    return check_launch("scan");
  }
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(unsigned(nb)), dim3(kScanBlock), 0, stream, in, n,
                     partial);
  hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanBlock), 0, stream, partial, int(nb),
                     total);
  hipLaunchKernelGGL(scan_apply_kernel, dim3(unsigned(nb)), dim3(kScanBlock), 0, stream, in, n,
                     partial, out);
  return check_launch("scan");
}

int64_t scan_partials_size(int64_t n) { return ceil_div(n, kScanTile) + 1; }

}  // namespace nsvf

extern "C" {

int64_t nsvf_scan_workspace_ints(int64_t n) { return nsvf::scan_partials_size(n); }

int nsvf_scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n,
                            int32_t* workspace, int32_t* total, void* stream) {
  nsvf::clear_error();
  if ((n > 0 && (!in || !out)) || !workspace)
    return nsvf::set_error(NSVF_EINVAL, "scan_exclusive_i32: null pointer");
  return nsvf::scan_exclusive_i32(in, out, n, workspace, total, nsvf::as_stream(stream));
}

}  // extern "C"
