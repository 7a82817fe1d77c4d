// The field MLP for training: forward and backward on v_mfma_f32_16x16x4_f32
// (FP32 operands, FP32 accumulation), the differentiable counterpart of
// fairnr/modules/field.py:217-279 (Linear + ReLU layers, then sigma and
// sigmoid RGB) that the reference trains through torch autograd.
//
// Layout.  As in the field kernels, a layer runs transposed, Y^T = W X^T:
// weight fragments are operand A (rows = output features), activations
// operand B (columns = the wave's 16 samples).  Lane (g = lane / 16,
// c = lane % 16) holds, for k-tile 4t + i, feature 16t + 4g + i of sample c,
// and a layer's accumulator for m-tile mt holds features 16mt + 4g + r in
// its four registers -- the operand order of the next layer's k-tiles, so
// activations stay in registers through all layers.  The backward products
// dX^T = W^T dY^T use fragments of W^T packed in the same order.
//
// Backward kernel.  A block of 8 waves takes 128 samples at a time (16 per
// wave).  Each wave recomputes its tile's forward pass (every layer's output
// kept in registers), then walks the layers back: the head's dlogits (the
// sigmoid's derivative fused), then per layer dz = dh * (h > 0), the input
// gradient dh_prev = W^T dz and the weight gradient dW = dz^T h_prev.  The
// weight gradient sums over samples, which sit in the wrong operand position
// (across the accumulator's columns); so each wave writes its tile's dz and
// h_prev to LDS, and after a barrier every wave accumulates its own rows of
// dW (one m-tile of each hidden layer, one column tile of the head) over the
// block's 128 samples, 4 samples per MFMA step, with a column of ones for the
// bias gradient.  These accumulators stay in registers for the whole kernel;
// at the end each block writes its partial sums, and a second kernel adds
// the blocks' partials in block order.  The sums therefore run in a fixed
// order for a given sample count: the gradients are deterministic, with no
// atomics.
#include <algorithm>

#include "common.h"

namespace nsvf {
namespace {

typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kTrainWaves = 8;
constexpr int kTrainBlock = 64 * kTrainWaves;
constexpr int kTrainChunk = 16 * kTrainWaves;  // samples per block step
constexpr int kLdsRow = 144;  // floats per LDS row: 144 = 16 mod 64 banks
constexpr int kMaxTrainBlocks = 256;

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Closes one m-tile's loads and MFMAs: left to itself the scheduler hoists
// the weight loads of every later m-tile (they depend on nothing) and runs
// out of registers.
__device__ __forceinline__ void tile_fence() { __builtin_amdgcn_sched_barrier(0); }

// The weights' address as a value the compiler cannot see through: the
// weight loads of a tile loop are loop-invariant, and hoisted out of it they
// would need hundreds of registers.
__device__ __forceinline__ const float* opaque(const float* p) {
  asm volatile("" : "+s"(p));
  return p;
}

// acc += sum over k-tiles kt of A[kt * 64 + lane] x[kt]: the fragment loads go
// at most 8 ahead of their MFMAs (a fence every 8 k-tiles).
template <int KT>
__device__ __forceinline__ f32x4 mfma_chain(const float* __restrict__ A, const float (&x)[KT],
                                            int lane) {
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kt = 0; kt < KT; ++kt) {
    acc = mfma4(A[kt * 64 + lane], x[kt], acc);
    if (kt % 8 == 7) tile_fence();
  }
  return acc;
}

// The logistic function, 1 / (1 + e^-x), in IEEE single precision.
__device__ __forceinline__ float logistic(float x) { return 1.0f / (1.0f + expf(-x)); }

// Offsets (in 64-float fragments) of the packed weights of an MLP with input
// width D, L hidden layers of width H and a 4-wide head.
template <int D, int H, int L>
struct TrainDims {
  static constexpr int KT0 = D / 4, KTH = H / 4, MT = H / 16, MT0 = D / 16;
  // forward fragments: layer 0 [MT][KT0], layers 1..L-1 [MT][KTH], head [KTH]
  static constexpr int F_HIDDEN = MT * KT0;  // first fragment of layer 1
  static constexpr int F_HEAD = MT * KT0 + (L - 1) * MT * KTH;
  static constexpr int F_FRAGS = F_HEAD + KTH;
  // transposed fragments: layer 0 [MT0][KTH], layers 1..L-1 [MT][KTH], head [MT]
  static constexpr int T_HIDDEN = MT0 * KTH;
  static constexpr int T_HEAD = MT0 * KTH + (L - 1) * MT * KTH;
  static constexpr int T_FRAGS = T_HEAD + MT;
  static constexpr int64_t T_BASE = int64_t(F_FRAGS) * 64;
  static constexpr int64_t BIAS_BASE = T_BASE + int64_t(T_FRAGS) * 64;
  static constexpr int64_t PACKED = BIAS_BASE + int64_t(L) * H + 4;
  // parameter gradients in torch order: W_0 [H][D], b_0 [H], ..., W_L [4][H], b_L [4]
  static constexpr int64_t P_W0 = 0, P_B0 = int64_t(H) * D;
  static constexpr int64_t P_LAYER1 = P_B0 + H;
  static constexpr int64_t P_HEAD = P_LAYER1 + int64_t(L - 1) * (int64_t(H) * H + H);
  static constexpr int64_t PARAMS = P_HEAD + 4 * H + 4;
};

// out[mt] = relu(W x + b) over the MT m-tiles: fragments A [MT][KT], input x
// in k-tile order, bias b (this layer's H values).
template <int KT, int MT>
__device__ __forceinline__ void dense_relu(const float* __restrict__ A, const float (&x)[KT],
                                           const float* __restrict__ b, float (&y)[4 * MT],
                                           int lane, int g) {
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const f32x4 acc = mfma_chain<KT>(A + mt * KT * 64, x, lane);
    const float4 bb = *reinterpret_cast<const float4*>(b + 16 * mt + 4 * g);
    y[4 * mt + 0] = fmaxf(acc[0] + bb.x, 0.0f);
    y[4 * mt + 1] = fmaxf(acc[1] + bb.y, 0.0f);
    y[4 * mt + 2] = fmaxf(acc[2] + bb.z, 0.0f);
    y[4 * mt + 3] = fmaxf(acc[3] + bb.w, 0.0f);
    tile_fence();
  }
}

// One wave's 16 samples [base, base + 16) through the MLP: x0 (the input in
// k-tile order, zero past M), every hidden layer's output h[l], and the
// head's accumulator z (lanes of group 0 hold the 4 logits of sample c,
// biases not yet added).
template <int D, int H, int L>
__device__ __forceinline__ void forward_tile(const float* __restrict__ emb, int64_t M,
                                             const float* __restrict__ packed, int64_t s,
                                             float (&x0)[D / 4], float (&h)[L][H / 4], f32x4& z,
                                             int lane, int g) {
  using T = TrainDims<D, H, L>;
  const float* bias = packed + T::BIAS_BASE;
#pragma unroll
  for (int t = 0; t < D / 16; ++t) {
    float4 v = {0.0f, 0.0f, 0.0f, 0.0f};
    if (s < M) v = *reinterpret_cast<const float4*>(emb + s * D + 16 * t + 4 * g);
    x0[4 * t + 0] = v.x;
    x0[4 * t + 1] = v.y;
    x0[4 * t + 2] = v.z;
    x0[4 * t + 3] = v.w;
  }
  dense_relu<T::KT0, T::MT>(packed, x0, bias, h[0], lane, g);
#pragma unroll
  for (int l = 1; l < L; ++l)
    dense_relu<T::KTH, T::MT>(packed + (T::F_HIDDEN + (l - 1) * T::MT * T::KTH) * 64, h[l - 1],
                              bias + l * H, h[l], lane, g);
  z = mfma_chain<T::KTH>(packed + T::F_HEAD * 64, h[L - 1], lane);
}

// out[s] = (sigma, sigmoid(r), sigmoid(g), sigmoid(b)) of emb[s] for s < M.
template <int D, int H, int L>
__global__ __launch_bounds__(kTrainBlock) void mlp_train_forward_kernel(
    const float* __restrict__ emb, int64_t M, const float* __restrict__ packed,
    float* __restrict__ out) {
  using T = TrainDims<D, H, L>;
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const int64_t tiles = (M + 15) / 16;
  for (int64_t t = int64_t(blockIdx.x) * kTrainWaves + (threadIdx.x >> 6); t < tiles;
       t += int64_t(gridDim.x) * kTrainWaves) {
    const int64_t s = t * 16 + col;
    float x0[T::KT0];
    float h[L][T::KTH];
    f32x4 z;
    forward_tile<D, H, L>(emb, M, opaque(packed), s, x0, h, z, lane, g);
    if (g == 0 && s < M) {
      const float4 bb = *reinterpret_cast<const float4*>(packed + T::BIAS_BASE + L * H);
      float4 o;
      o.x = z[0] + bb.x;
      o.y = logistic(z[1] + bb.y);
      o.z = logistic(z[2] + bb.z);
      o.w = logistic(z[3] + bb.w);
      *reinterpret_cast<float4*>(out + s * 4) = o;
    }
  }
}

// Backward: demb [M][D] and this block's partial parameter gradients
// (partial + blockIdx.x * PARAMS, torch order), from dout [M][4] = the
// gradient of (sigma, r, g, b) after the sigmoid.  Needs H == 16 * waves.
template <int D, int H, int L>
__global__ __launch_bounds__(kTrainBlock) void mlp_train_backward_kernel(
    const float* __restrict__ emb, int64_t M, const float* __restrict__ packed,
    const float* __restrict__ dout, float* __restrict__ demb, float* __restrict__ partial) {
  using T = TrainDims<D, H, L>;
  static_assert(T::MT == kTrainWaves, "one hidden m-tile per wave");
  static_assert(D % 16 == 0 && D <= H, "input width");
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  float* lds_dz = reinterpret_cast<float*>(smem_raw);    // [128 samples][kLdsRow]
  float* lds_x = lds_dz + kTrainChunk * kLdsRow;          // [128 samples][kLdsRow]
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const int w = threadIdx.x >> 6;
  const float* bias = packed + T::BIAS_BASE;
  // this wave's rows of the weight gradients: layer 0 rows 16w.. over the
  // D / 16 input tiles, hidden layers 1..L-1 rows 16w.. over H / 16 tiles,
  // the head's columns 16w..; a bias tile per layer (wave 0: the head's)
  f32x4 gw0[T::MT0], gw[L > 1 ? L - 1 : 1][T::MT], gb[L], gh, ghb;
#pragma unroll
  for (int i = 0; i < T::MT0; ++i) gw0[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int l = 0; l < (L > 1 ? L - 1 : 1); ++l)
#pragma unroll
    for (int i = 0; i < T::MT; ++i) gw[l][i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int l = 0; l < L; ++l) gb[l] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  gh = ghb = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  float* my_dz = lds_dz + (w * 16 + col) * kLdsRow;
  float* my_x = lds_x + (w * 16 + col) * kLdsRow;
  const int64_t chunks = (M + kTrainChunk - 1) / kTrainChunk;
  for (int64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const int64_t s = ch * kTrainChunk + w * 16 + col;
    const float* pk = opaque(packed);
    const float* At = pk + T::T_BASE;
    float x0[T::KT0];
    float h[L][T::KTH];
    f32x4 z;
    forward_tile<D, H, L>(emb, M, pk, s, x0, h, z, lane, g);
    // dlogits on the lanes of group 0: d sigma as given, d rgb through the
    // sigmoid's derivative y (1 - y)
    f32x4 dl = {0.0f, 0.0f, 0.0f, 0.0f};
    if (g == 0 && s < M) {
      const float4 d = *reinterpret_cast<const float4*>(dout + s * 4);
      const float4 bb = *reinterpret_cast<const float4*>(bias + L * H);
      const float y1 = logistic(z[1] + bb.y), y2 = logistic(z[2] + bb.z),
                  y3 = logistic(z[3] + bb.w);
      dl[0] = d.x;
      dl[1] = d.y * (y1 * (1.0f - y1));
      dl[2] = d.z * (y2 * (1.0f - y2));
      dl[3] = d.w * (y3 * (1.0f - y3));
    }
    // ---- head: dW_head = dl^T h[L-1], dh = W_head^T dl
    if (g == 0) *reinterpret_cast<f32x4*>(my_dz) = dl;
#pragma unroll
    for (int mt = 0; mt < T::MT; ++mt)
      *reinterpret_cast<f32x4*>(my_x + 16 * mt + 4 * g) =
          f32x4{h[L - 1][4 * mt], h[L - 1][4 * mt + 1], h[L - 1][4 * mt + 2], h[L - 1][4 * mt + 3]};
    __syncthreads();
#pragma unroll 1
    for (int kk = 0; kk < kTrainChunk / 4; ++kk) {
      const int row = (4 * kk + g) * kLdsRow;
      const float a = col < 4 ? lds_dz[row + col] : 0.0f;
      gh = mfma4(a, lds_x[row + 16 * w + col], gh);
      if (w == 0) ghb = mfma4(a, 1.0f, ghb);
    }
    __syncthreads();
    float dh[T::KTH];
    {
      // operand B of W_head^T: lane group g takes logit g of sample c
      const float t0 = __shfl(dl[0], col, 64), t1 = __shfl(dl[1], col, 64),
                  t2 = __shfl(dl[2], col, 64), t3 = __shfl(dl[3], col, 64);
      const float b = g == 0 ? t0 : (g == 1 ? t1 : (g == 2 ? t2 : t3));
#pragma unroll
      for (int mt = 0; mt < T::MT; ++mt) {
        const f32x4 acc =
            mfma4(At[(T::T_HEAD + mt) * 64 + lane], b, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
        for (int r = 0; r < 4; ++r) dh[4 * mt + r] = acc[r];
        tile_fence();
      }
    }
    // ---- hidden layers, last to first: layer l maps h[l-1] (x0 for l = 0)
    // to h[l]
#pragma unroll
    for (int l = L - 1; l >= 0; --l) {
      float dz[T::KTH];
#pragma unroll
      for (int k = 0; k < T::KTH; ++k) dz[k] = h[l][k] > 0.0f ? dh[k] : 0.0f;
#pragma unroll
      for (int mt = 0; mt < T::MT; ++mt)
        *reinterpret_cast<f32x4*>(my_dz + 16 * mt + 4 * g) =
            f32x4{dz[4 * mt], dz[4 * mt + 1], dz[4 * mt + 2], dz[4 * mt + 3]};
      if (l > 0) {
        // the first layer's output again, from x0: keeping it from the
        // forward pass would hold 32 more registers through every later phase
        if (l == 1) dense_relu<T::KT0, T::MT>(opaque(packed), x0, bias, h[0], lane, g);
#pragma unroll
        for (int mt = 0; mt < T::MT; ++mt)
          *reinterpret_cast<f32x4*>(my_x + 16 * mt + 4 * g) =
              f32x4{h[l - 1][4 * mt], h[l - 1][4 * mt + 1], h[l - 1][4 * mt + 2],
                    h[l - 1][4 * mt + 3]};
      } else {
#pragma unroll
        for (int t = 0; t < T::MT0; ++t)
          *reinterpret_cast<f32x4*>(my_x + 16 * t + 4 * g) =
              f32x4{x0[4 * t], x0[4 * t + 1], x0[4 * t + 2], x0[4 * t + 3]};
      }
      __syncthreads();
      // weight gradient rows 16w..16w+15 of layer l, 4 samples per step
#pragma unroll 1
      for (int kk = 0; kk < kTrainChunk / 4; ++kk) {
        const int row = (4 * kk + g) * kLdsRow;
        const float a = lds_dz[row + 16 * w + col];
        if (l > 0) {
#pragma unroll
          for (int mi = 0; mi < T::MT; ++mi)
            gw[l > 0 ? l - 1 : 0][mi] = mfma4(a, lds_x[row + 16 * mi + col],
                                              gw[l > 0 ? l - 1 : 0][mi]);
        } else {
#pragma unroll
          for (int mi = 0; mi < T::MT0; ++mi)
            gw0[mi] = mfma4(a, lds_x[row + 16 * mi + col], gw0[mi]);
        }
        gb[l] = mfma4(a, 1.0f, gb[l]);
      }
      __syncthreads();
      if (l > 0) {
        const float* A = At + (T::T_HIDDEN + (l - 1) * T::MT * T::KTH) * 64;
#pragma unroll
        for (int mt = 0; mt < T::MT; ++mt) {
          const f32x4 acc = mfma_chain<T::KTH>(A + mt * T::KTH * 64, dz, lane);
#pragma unroll
          for (int r = 0; r < 4; ++r) dh[4 * mt + r] = acc[r];
          tile_fence();
        }
      } else {
#pragma unroll
        for (int mt = 0; mt < T::MT0; ++mt) {
          const f32x4 acc = mfma_chain<T::KTH>(At + mt * T::KTH * 64, dz, lane);
          if (s < M) *reinterpret_cast<f32x4*>(demb + s * D + 16 * mt + 4 * g) = acc;
          tile_fence();
        }
      }
    }
  }
  // this block's partial sums, in torch order
  float* P = partial + int64_t(blockIdx.x) * T::PARAMS;
  const int o = 16 * w + 4 * g;  // hidden output rows o..o+3 of this lane
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int mi = 0; mi < T::MT0; ++mi) P[T::P_W0 + int64_t(o + r) * D + 16 * mi + col] = gw0[mi][r];
    if (col == 0) P[T::P_B0 + o + r] = gb[0][r];
#pragma unroll
    for (int l = 1; l < L; ++l) {
      const int64_t base = T::P_LAYER1 + int64_t(l - 1) * (int64_t(H) * H + H);
#pragma unroll
      for (int mi = 0; mi < T::MT; ++mi) P[base + int64_t(o + r) * H + 16 * mi + col] = gw[l - 1][mi][r];
      if (col == 0) P[base + int64_t(H) * H + o + r] = gb[l][r];
    }
    if (g == 0) {
      P[T::P_HEAD + int64_t(r) * H + 16 * w + col] = gh[r];
      if (w == 0 && col == 0) P[T::P_HEAD + 4 * H + r] = ghb[r];
    }
  }
}

// dparams[p] = sum over blocks b = 0, 1, ... of partial[b][p].
__global__ __launch_bounds__(256) void sum_partials_kernel(const float* __restrict__ partial,
                                                           int nb, int64_t P,
                                                           float* __restrict__ dparams) {
  const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (p >= P) return;
  float acc = 0.0f;
  for (int b = 0; b < nb; ++b) acc = acc + partial[int64_t(b) * P + p];
  dparams[p] = acc;
}

// Packs one layer W [Hout][Hin] (row-major, torch's Linear weight):
// mode 0: forward fragments [ceil(Hout/16)][Hin/4][64], lane (g, c) of
//   fragment (mt, 4t + i) = W[16 mt + c][16 t + 4 g + i] (0 past Hout);
// mode 1: transposed fragments [Hin/16][Hout/4][64], lane (g, c) of
//   fragment (mt, 4t + i) = W[16 t + 4 g + i][16 mt + c];
// mode 2 (the 4-wide head): [Hin/16][64], lane (g, c) of fragment mt =
//   W[g][16 mt + c].
__global__ __launch_bounds__(256) void pack_layer_kernel(const float* __restrict__ W, int Hout,
                                                         int Hin, int mode,
                                                         float* __restrict__ dst, int64_t n) {
  const int64_t f = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (f >= n) return;
  const int lane = int(f & 63), g = lane >> 4, c = lane & 15;
  const int64_t frag = f >> 6;
  float v = 0.0f;
  if (mode == 0) {
    const int KT = Hin / 4;
    const int mt = int(frag / KT), kt = int(frag % KT), t = kt >> 2, i = kt & 3;
    const int row = 16 * mt + c;
    if (row < Hout) v = W[int64_t(row) * Hin + 16 * t + 4 * g + i];
  } else if (mode == 1) {
    const int KT = Hout / 4;
    const int mt = int(frag / KT), kt = int(frag % KT), t = kt >> 2, i = kt & 3;
    v = W[int64_t(16 * t + 4 * g + i) * Hin + 16 * mt + c];
  } else {
    v = W[int64_t(g) * Hin + 16 * frag + c];
  }
  dst[f] = v;
}

template <int D, int H, int L>
int pack_impl(const float* const* weights, const float* const* biases, float* packed,
              hipStream_t st) {
  using T = TrainDims<D, H, L>;
  auto launch = [&](const float* W, int Hout, int Hin, int mode, float* dst, int64_t n) {
    hipLaunchKernelGGL(pack_layer_kernel, dim3(unsigned(ceil_div(n, 256))), dim3(256), 0, st, W,
                       Hout, Hin, mode, dst, n);
  };
  launch(weights[0], H, D, 0, packed, int64_t(T::MT) * T::KT0 * 64);
  for (int l = 1; l < L; ++l)
    launch(weights[l], H, H, 0, packed + (int64_t(T::F_HIDDEN) + (l - 1) * T::MT * T::KTH) * 64,
           int64_t(T::MT) * T::KTH * 64);
  launch(weights[L], 4, H, 0, packed + int64_t(T::F_HEAD) * 64, int64_t(T::KTH) * 64);
  float* Tp = packed + T::T_BASE;
  launch(weights[0], H, D, 1, Tp, int64_t(T::MT0) * T::KTH * 64);
  for (int l = 1; l < L; ++l)
    launch(weights[l], H, H, 1, Tp + (int64_t(T::T_HIDDEN) + (l - 1) * T::MT * T::KTH) * 64,
           int64_t(T::MT) * T::KTH * 64);
  launch(weights[L], 4, H, 2, Tp + int64_t(T::T_HEAD) * 64, int64_t(T::MT) * 64);
  if (check_launch("mlp_train_pack") != NSVF_OK) return NSVF_ELAUNCH;
  float* bias = packed + T::BIAS_BASE;
  for (int l = 0; l <= L; ++l) {
    const hipError_t e = hipMemcpyAsync(bias + int64_t(l) * H, biases[l],
                                        sizeof(float) * (l < L ? H : 4),
                                        hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess)
      return set_error(NSVF_ELAUNCH, "mlp_train_pack: bias copy: %s", hipGetErrorString(e));
  }
  return NSVF_OK;
}

int backward_blocks(int64_t M) {
  return int(std::max<int64_t>(1, std::min<int64_t>(ceil_div(M, kTrainChunk), kMaxTrainBlocks)));
}

constexpr size_t kBackwardLds = size_t(2) * kTrainChunk * kLdsRow * sizeof(float);

bool train_shape_ok(int D, int H, int L) { return D == 32 && H == 128 && L == 3; }

}  // namespace
}  // namespace nsvf

using namespace nsvf;

extern "C" {

int nsvf_mlp_train_supported(int D, int H, int L) { return train_shape_ok(D, H, L) ? 1 : 0; }

int64_t nsvf_mlp_train_packed_floats(int D, int H, int L) {
  if (!train_shape_ok(D, H, L)) return -1;
  return TrainDims<32, 128, 3>::PACKED;
}

int64_t nsvf_mlp_train_param_floats(int D, int H, int L) {
  if (!train_shape_ok(D, H, L)) return -1;
  return TrainDims<32, 128, 3>::PARAMS;
}

int64_t nsvf_mlp_train_partial_floats(int D, int H, int L, int64_t M) {
  if (!train_shape_ok(D, H, L) || M < 0) return -1;
  return int64_t(backward_blocks(M)) * TrainDims<32, 128, 3>::PARAMS;
}

int nsvf_mlp_train_pack(int D, int H, int L, const float* const* weights,
                        const float* const* biases, float* packed, void* stream) {
  clear_error();
  if (!train_shape_ok(D, H, L))
    return set_error(NSVF_EINVAL, "mlp_train_pack: shape D=%d H=%d L=%d (built: 32, 128, 3)", D,
                     H, L);
  if (!weights || !biases || !packed) return set_error(NSVF_EINVAL, "mlp_train_pack: null");
  for (int l = 0; l <= L; ++l)
    if (!weights[l] || !biases[l]) return set_error(NSVF_EINVAL, "mlp_train_pack: layer %d null", l);
  return pack_impl<32, 128, 3>(weights, biases, packed, as_stream(stream));
}

int nsvf_mlp_train_forward(int D, int H, int L, const float* packed, const float* emb, int64_t M,
                           float* out, void* stream) {
  clear_error();
  if (!train_shape_ok(D, H, L))
    return set_error(NSVF_EINVAL, "mlp_train_forward: shape D=%d H=%d L=%d (built: 32, 128, 3)",
                     D, H, L);
  if (M < 0) return set_error(NSVF_EINVAL, "mlp_train_forward: M=%lld", static_cast<long long>(M));
  if (M == 0) return NSVF_OK;
  if (!packed || !emb || !out) return set_error(NSVF_EINVAL, "mlp_train_forward: null pointer");
  const int64_t tiles = ceil_div(M, 16);
  const unsigned grid = unsigned(std::min<int64_t>(ceil_div(tiles, kTrainWaves), 4096));
  hipLaunchKernelGGL((mlp_train_forward_kernel<32, 128, 3>), dim3(grid), dim3(kTrainBlock), 0,
                     as_stream(stream), emb, M, packed, out);
  return check_launch("mlp_train_forward");
}

int nsvf_mlp_train_backward(int D, int H, int L, const float* packed, const float* emb,
                            int64_t M, const float* dout, float* demb, float* partial,
                            int64_t partial_floats, float* dparams, void* stream) {
  clear_error();
  if (!train_shape_ok(D, H, L))
    return set_error(NSVF_EINVAL, "mlp_train_backward: shape D=%d H=%d L=%d (built: 32, 128, 3)",
                     D, H, L);
  if (M < 0) return set_error(NSVF_EINVAL, "mlp_train_backward: M=%lld", static_cast<long long>(M));
  if (!packed || !partial || !dparams || (M > 0 && (!emb || !dout || !demb)))
    return set_error(NSVF_EINVAL, "mlp_train_backward: null pointer");
  using T = TrainDims<32, 128, 3>;
  const int nb = backward_blocks(M);
  if (partial_floats < int64_t(nb) * T::PARAMS)
    return set_error(NSVF_EINVAL, "mlp_train_backward: partial buffer of %lld floats, %lld needed",
                     static_cast<long long>(partial_floats),
                     static_cast<long long>(int64_t(nb) * T::PARAMS));
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL((mlp_train_backward_kernel<32, 128, 3>), dim3(nb), dim3(kTrainBlock),
                     kBackwardLds, st, emb, M, packed, dout, demb, partial);
  if (check_launch("mlp_train_backward") != NSVF_OK) return NSVF_ELAUNCH;
  hipLaunchKernelGGL(sum_partials_kernel, dim3(unsigned(ceil_div(T::PARAMS, 256))), dim3(256), 0,
                     st, partial, nb, T::PARAMS, dparams);
  return check_launch("mlp_train_sum_partials");
}

}  // extern "C"
