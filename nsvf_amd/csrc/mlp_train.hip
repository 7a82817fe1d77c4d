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
//
// Split-BF16 backward (mlp_train_backward_kernel<..., true>).  The same
// kernel with the forward recomputation of the hidden layers, the input
// gradients dX^T = W^T dz^T and the weight and bias gradients on
// v_mfma_f32_16x16x32_bf16 in split BF16: an
// FP32 operand f is hi + lo with hi = bf16(f), lo = bf16(f - hi), and a
// product is Wl Xh + Wh Xl + Wh Xh (FP32 accumulation; the dropped Wl Xl is
// 2^-16 of it), so these products keep about FP32's accuracy at three BF16
// MFMAs per 16x16x32 tile against eight FP32 16x16x4 MFMAs.  The weight
// fragments (forward and transposed, hi and lo) come from global memory,
// 288 KB that stay in L2; a k-tile's operand lane (g, c) holds features
// 32 kt + 4 g + i and 32 kt + 16 + 4 g + i (i < 4) of sample c -- the
// accumulators of m-tiles 2 kt and 2 kt + 1, so a layer's FP32 output
// splits into the next layer's operands in registers.  The weight
// gradients sum over samples, the MFMA's k: dz and the layer input are
// staged in LDS as hi and lo planes of [feature][sample] bf16, so an
// operand (8 samples of one feature) is one 16-byte read, and each wave
// keeps its rows of every dW in registers as in the FP32 kernel.  The
// head's recomputed logits and its input gradient (4 wide) stay in FP32.
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

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// (a, b) as packed bf16 pairs hi = bf16(a, b) and lo = bf16(a - hi, b - hi)
// (round to nearest even; the differences are exact in FP32).
__device__ __forceinline__ void split2(float a, float b, unsigned int& hi, unsigned int& lo) {
  const bf16x2 h = __builtin_convertvector(f32x2{a, b}, bf16x2);
  const f32x2 hf = __builtin_convertvector(h, f32x2);
  const bf16x2 l = __builtin_convertvector(f32x2{a - hf[0], b - hf[1]}, bf16x2);
  hi = __builtin_bit_cast(unsigned int, h);
  lo = __builtin_bit_cast(unsigned int, l);
}

// The split operands of k-tile kt from a lane's FP32 values v[4 mt + r]
// (feature 16 mt + 4 g + r): positions 0-3 from m-tile 2 kt, 4-7 from
// m-tile 2 kt + 1.
template <int N>
__device__ __forceinline__ void split_ktile(const float (&v)[N], int kt, bf16x8& hi,
                                            bf16x8& lo) {
  unsigned int h[4], l[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) split2(v[8 * kt + 2 * p], v[8 * kt + 2 * p + 1], h[p], l[p]);
  hi = __builtin_bit_cast(bf16x8, u32x4{h[0], h[1], h[2], h[3]});
  lo = __builtin_bit_cast(bf16x8, u32x4{l[0], l[1], l[2], l[3]});
}

// acc + the split product over KT k-tiles with fragments F (per fragment
// 64 hi then 64 lo bf16x8 lane vectors), fragment kt at F + 128 kt; per
// k-tile in the order Wl Xh, Wh Xl, Wh Xh -- the renderer's split-BF16 MLP
// kernel's order, with the bias as the initial accumulator there too, so a
// recomputed layer equals that kernel's forward bit for bit.
template <int KT>
__device__ __forceinline__ f32x4 mfma3_chain(const bf16x8* __restrict__ F,
                                             const bf16x8 (&xh)[KT], const bf16x8 (&xl)[KT],
                                             int lane, f32x4 acc) {
#pragma unroll
  for (int kt = 0; kt < KT; ++kt) {
    const bf16x8 wh = F[128 * kt + lane], wl = F[128 * kt + 64 + lane];
    acc = mfma32(wl, xh[kt], acc);
    acc = mfma32(wh, xl[kt], acc);
    acc = mfma32(wh, xh[kt], acc);
  }
  return acc;
}

// Closes one m-tile's loads and MFMAs: left to itself the scheduler hoists
// the weight loads of every later m-tile (they depend on nothing) and runs
// out of registers.
__device__ __forceinline__ void tile_fence() { __builtin_amdgcn_sched_barrier(0); }

// The weights' address as a value the compiler cannot see through: the
// weight loads of a tile loop are loop-invariant, and hoisted out of it they
// would need hundreds of registers.
template <class P>
__device__ __forceinline__ P opaque(P p) {
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
  // split-BF16 fragments (1024 bf16 each: 512 hi, 512 lo), 32-wide k-tiles:
  // layer 0 [MT][D/32], layers 1..L-1 [MT][H/32], then transposed: layer 0
  // [MT0][H/32], layers 1..L-1 [MT][H/32]
  static constexpr int KX0 = D / 32, KXH = H / 32;
  static constexpr int X_HIDDEN = MT * KX0;
  static constexpr int X_T0 = X_HIDDEN + (L - 1) * MT * KXH;
  static constexpr int X_THIDDEN = X_T0 + MT0 * KXH;
  static constexpr int X_FRAGS = X_THIDDEN + (L - 1) * MT * KXH;
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

// Split-BF16 staging of the weight-gradient operands in LDS: per operand a
// hi and a lo plane of [feature rows][kXRow] bf16 with the block's 128
// samples along a row, so an MFMA operand (8 samples of one feature) is
// one 16-byte read; 136 = 128 + 8 staggers the rows over the banks.
constexpr int kXRow = 136;
constexpr int kXPlane = kTrainChunk * kXRow;  // bf16 per plane

// The split of a lane's values v[4 mt + r] (feature 16 mt + 4 g + r of
// block sample s) for m-tiles mt < NMT, into rows of the planes ph / pl.
template <int NMT, int N>
__device__ __forceinline__ void stage_split(unsigned short* ph, unsigned short* pl,
                                            const float (&v)[N], int g, int s) {
  static_assert(4 * NMT <= N, "values");
  // one lane address per plane; the rows are constant offsets from it
  unsigned short* qh = ph + 4 * g * kXRow + s;
  unsigned short* ql = pl + 4 * g * kXRow + s;
#pragma unroll
  for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      unsigned int hw, lw;
      split2(v[4 * mt + 2 * p], v[4 * mt + 2 * p + 1], hw, lw);
      const int row = 16 * mt + 2 * p;
      qh[row * kXRow] = static_cast<unsigned short>(hw);
      qh[(row + 1) * kXRow] = static_cast<unsigned short>(hw >> 16);
      ql[row * kXRow] = static_cast<unsigned short>(lw);
      ql[(row + 1) * kXRow] = static_cast<unsigned short>(lw >> 16);
    }
}

__device__ __forceinline__ bf16x8 lds8(const unsigned short* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}

// Weight and bias gradients over the block's samples in split BF16:
// acc[mi] += sum over samples of A row ra x B row rb0 + 16 mi, and (kBias)
// accb += the sum of A row ra (its product with a row of ones, whose lo
// part is 0), with A rows from planes (pah, pal) -- zero where !a_on -- and
// B rows from (pbh, pbl); the lane reads samples 32 ks + 8 g.. of each
// 32-sample step.  One step at a time with a fence per m-tile, so that the
// operand reads are not hoisted into registers the kernel does not have.
template <int NMI, bool kBias>
__device__ __forceinline__ void dw_x3(const unsigned short* pah, const unsigned short* pal,
                                      int ra, const unsigned short* pbh,
                                      const unsigned short* pbl, int rb0, int g, bool a_on,
                                      f32x4 (&acc)[NMI], f32x4& accb) {
  const bf16x8 zero = {};
  const bf16x8 ones = __builtin_bit_cast(
      bf16x8, u32x4{0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u});
#pragma unroll 1
  for (int ks = 0; ks < kTrainChunk / 32; ++ks) {
    const int off = 32 * ks + 8 * g;
    const bf16x8 ah = a_on ? lds8(pah + ra * kXRow + off) : zero;
    const bf16x8 al = a_on ? lds8(pal + ra * kXRow + off) : zero;
#pragma unroll
    for (int mi = 0; mi < NMI; ++mi) {
      const int rb = rb0 + 16 * mi;
      const bf16x8 bh = lds8(pbh + rb * kXRow + off), bl = lds8(pbl + rb * kXRow + off);
      acc[mi] = mfma32(al, bh, acc[mi]);
      acc[mi] = mfma32(ah, bl, acc[mi]);
      acc[mi] = mfma32(ah, bh, acc[mi]);
      tile_fence();
    }
    if (kBias) {
      accb = mfma32(al, ones, accb);
      accb = mfma32(ah, ones, accb);
    }
  }
}

// ReLU as an integer max on the float's bits (the split-BF16 MLP kernel's).
__device__ __forceinline__ float relu_bits(float x) {
  return __int_as_float(max(__float_as_int(x), 0));
}

// A split-BF16 layer over MT m-tiles: y[4 mt + r] = (relu of, kRelu) the
// chain of mfma3_chain over fragments A [MT][KT] (fragment (mt, kt) at
// A + 128 (mt KT + kt)) on operands xh / xl, from the bias b (kBias) or 0.
// kAhead: the next m-tile's fragments are loaded before the current one's
// MFMAs -- they come from L2, and waiting for them at the start of every
// m-tile cost 15% of the backward (4.8 -> 4.05 ms on 3.09 M samples with
// the recomputation and the hidden layers' input gradients loading ahead).
template <int KT, int MT, bool kBias, bool kRelu, bool kAhead>
__device__ __forceinline__ void layer_x3(const bf16x8* __restrict__ A, const bf16x8 (&xh)[KT],
                                         const bf16x8 (&xl)[KT], const float* __restrict__ b,
                                         float (&y)[4 * MT], int lane, int g) {
  bf16x8 ch[KT], cl[KT];
  if (kAhead) {
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
      ch[kt] = A[128 * kt + lane];
      cl[kt] = A[128 * kt + 64 + lane];
    }
  }
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if (kBias) acc = *reinterpret_cast<const f32x4*>(b + 16 * mt + 4 * g);
    if (kAhead) {
      bf16x8 nh[KT], nl[KT];
      if (mt + 1 < MT) {
#pragma unroll
        for (int kt = 0; kt < KT; ++kt) {
          nh[kt] = A[128 * ((mt + 1) * KT + kt) + lane];
          nl[kt] = A[128 * ((mt + 1) * KT + kt) + 64 + lane];
        }
      }
#pragma unroll
      for (int kt = 0; kt < KT; ++kt) {
        acc = mfma32(cl[kt], xh[kt], acc);
        acc = mfma32(ch[kt], xl[kt], acc);
        acc = mfma32(ch[kt], xh[kt], acc);
      }
      if (mt + 1 < MT) {
#pragma unroll
        for (int kt = 0; kt < KT; ++kt) {
          ch[kt] = nh[kt];
          cl[kt] = nl[kt];
        }
      }
    } else {
      acc = mfma3_chain<KT>(A + 128 * mt * KT, xh, xl, lane, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) y[4 * mt + r] = kRelu ? relu_bits(acc[r]) : acc[r];
    tile_fence();
  }
}

// dense_relu in split BF16 (relu(W x + b), the bias as the initial
// accumulator), fragments loaded an m-tile ahead.
template <int KT, int MT>
__device__ __forceinline__ void dense_relu_x3(const bf16x8* __restrict__ A,
                                              const bf16x8 (&xh)[KT], const bf16x8 (&xl)[KT],
                                              const float* __restrict__ b, float (&y)[4 * MT],
                                              int lane, int g) {
  layer_x3<KT, MT, true, true, true>(A, xh, xl, b, y, lane, g);
}

// Layer 0 in split BF16 from the input x0 (k-tile order of forward_tile).
template <int D, int H, int L>
__device__ __forceinline__ void layer0_x3(const bf16x8* __restrict__ X, const float (&x0)[D / 4],
                                          const float* __restrict__ bias, float (&h0)[H / 4],
                                          int lane, int g) {
  using T = TrainDims<D, H, L>;
  bf16x8 xh[T::KX0], xl[T::KX0];
#pragma unroll
  for (int kt = 0; kt < T::KX0; ++kt) split_ktile(x0, kt, xh[kt], xl[kt]);
  dense_relu_x3<T::KX0, T::MT>(X, xh, xl, bias, h0, lane, g);
}

// forward_tile with the hidden layers in split BF16 (fragments X); the head
// stays FP32.
template <int D, int H, int L>
__device__ __forceinline__ void forward_tile_x3(const float* __restrict__ emb, int64_t M,
                                                const float* __restrict__ packed,
                                                const bf16x8* __restrict__ X, int64_t s,
                                                float (&x0)[D / 4], float (&h)[L][H / 4],
                                                f32x4& z, int lane, int g) {
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
  layer0_x3<D, H, L>(X, x0, bias, h[0], lane, g);
#pragma unroll
  for (int l = 1; l < L; ++l) {
    bf16x8 xh[T::KXH], xl[T::KXH];
#pragma unroll
    for (int kt = 0; kt < T::KXH; ++kt) split_ktile(h[l - 1], kt, xh[kt], xl[kt]);
    dense_relu_x3<T::KXH, T::MT>(X + 128 * (T::X_HIDDEN + (l - 1) * T::MT * T::KXH), xh, xl,
                                 bias + l * H, h[l], lane, g);
  }
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
// X3: the forward recomputation and the input gradients of layers 0..L-1
// in split BF16 with the fragments x3w (TrainDims::X_FRAGS of them).
template <int D, int H, int L, bool X3>
__global__ __launch_bounds__(kTrainBlock) void mlp_train_backward_kernel(
    const float* __restrict__ emb, int64_t M, const float* __restrict__ packed,
    const bf16x8* __restrict__ x3w, const float* __restrict__ dout, float* __restrict__ demb,
    float* __restrict__ partial) {
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
  // split-BF16 planes (X3): dz hi, dz lo, x hi, x lo
  unsigned short* pzh = reinterpret_cast<unsigned short*>(smem_raw);
  unsigned short* pzl = pzh + kXPlane;
  unsigned short* pxh = pzl + kXPlane;
  unsigned short* pxl = pxh + kXPlane;
  const int sb = w * 16 + col;  // this lane's sample in the block
  const int64_t chunks = (M + kTrainChunk - 1) / kTrainChunk;
  for (int64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const int64_t s = ch * kTrainChunk + w * 16 + col;
    const float* pk = opaque(packed);
    const float* At = pk + T::T_BASE;
    float x0[T::KT0];
    float h[L][T::KTH];
    f32x4 z;
    if constexpr (X3)
      forward_tile_x3<D, H, L>(emb, M, pk, opaque(x3w), s, x0, h, z, lane, g);
    else
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
    if constexpr (X3) {
      if (g == 0) {
        const float dlv[4] = {dl[0], dl[1], dl[2], dl[3]};
        stage_split<1>(pzh, pzl, dlv, 0, sb);
      }
      stage_split<T::MT>(pxh, pxl, h[L - 1], g, sb);
      __syncthreads();
      f32x4 ghv[1] = {gh};
      if (w == 0)
        dw_x3<1, true>(pzh, pzl, col, pxh, pxl, 16 * w + col, g, col < 4, ghv, ghb);
      else
        dw_x3<1, false>(pzh, pzl, col, pxh, pxl, 16 * w + col, g, col < 4, ghv, ghb);
      gh = ghv[0];
      __syncthreads();
    } else {
      if (g == 0) *reinterpret_cast<f32x4*>(my_dz) = dl;
#pragma unroll
      for (int mt = 0; mt < T::MT; ++mt)
        *reinterpret_cast<f32x4*>(my_x + 16 * mt + 4 * g) = f32x4{
            h[L - 1][4 * mt], h[L - 1][4 * mt + 1], h[L - 1][4 * mt + 2], h[L - 1][4 * mt + 3]};
      __syncthreads();
#pragma unroll 1
      for (int kk = 0; kk < kTrainChunk / 4; ++kk) {
        const int row = (4 * kk + g) * kLdsRow;
        const float a = col < 4 ? lds_dz[row + col] : 0.0f;
        gh = mfma4(a, lds_x[row + 16 * w + col], gh);
        if (w == 0) ghb = mfma4(a, 1.0f, ghb);
      }
      __syncthreads();
    }
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
      if constexpr (X3) {
        stage_split<T::MT>(pzh, pzl, dz, g, sb);
      } else {
#pragma unroll
        for (int mt = 0; mt < T::MT; ++mt)
          *reinterpret_cast<f32x4*>(my_dz + 16 * mt + 4 * g) =
              f32x4{dz[4 * mt], dz[4 * mt + 1], dz[4 * mt + 2], dz[4 * mt + 3]};
      }
      if (l > 0) {
        // the first layer's output again, from x0: keeping it from the
        // forward pass would hold 32 more registers through every later phase
        if (l == 1) {
          if constexpr (X3)
            layer0_x3<D, H, L>(opaque(x3w), x0, bias, h[0], lane, g);
          else
            dense_relu<T::KT0, T::MT>(opaque(packed), x0, bias, h[0], lane, g);
        }
        if constexpr (X3) {
          stage_split<T::MT>(pxh, pxl, h[l - 1], g, sb);
        } else {
#pragma unroll
          for (int mt = 0; mt < T::MT; ++mt)
            *reinterpret_cast<f32x4*>(my_x + 16 * mt + 4 * g) =
                f32x4{h[l - 1][4 * mt], h[l - 1][4 * mt + 1], h[l - 1][4 * mt + 2],
                      h[l - 1][4 * mt + 3]};
        }
      } else {
        if constexpr (X3) {
          stage_split<T::MT0>(pxh, pxl, x0, g, sb);
        } else {
#pragma unroll
          for (int t = 0; t < T::MT0; ++t)
            *reinterpret_cast<f32x4*>(my_x + 16 * t + 4 * g) =
                f32x4{x0[4 * t], x0[4 * t + 1], x0[4 * t + 2], x0[4 * t + 3]};
        }
      }
      __syncthreads();
      // weight gradient rows 16w..16w+15 of layer l: split BF16 over the
      // block's 128 samples, or FP32 4 samples per step
      if constexpr (X3) {
        if (l > 0)
          dw_x3<T::MT, true>(pzh, pzl, 16 * w + col, pxh, pxl, col, g, true,
                             gw[l > 0 ? l - 1 : 0], gb[l]);
        else
          dw_x3<T::MT0, true>(pzh, pzl, 16 * w + col, pxh, pxl, col, g, true, gw0, gb[l]);
      }
#pragma unroll 1
      for (int kk = 0; kk < (X3 ? 0 : kTrainChunk / 4); ++kk) {
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
      if constexpr (X3) {
        bf16x8 dzh[T::KXH], dzl[T::KXH];
#pragma unroll
        for (int kt = 0; kt < T::KXH; ++kt) split_ktile(dz, kt, dzh[kt], dzl[kt]);
        const bf16x8* Xt = opaque(x3w);
        if (l > 0) {
          layer_x3<T::KXH, T::MT, false, false, true>(
              Xt + 128 * (T::X_THIDDEN + (l - 1) * T::MT * T::KXH), dzh, dzl, nullptr, dh, lane,
              g);
        } else {
#pragma unroll
          for (int mt = 0; mt < T::MT0; ++mt) {
            const f32x4 acc = mfma3_chain<T::KXH>(Xt + 128 * (T::X_T0 + mt * T::KXH), dzh,
                                                  dzl, lane, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            if (s < M) *reinterpret_cast<f32x4*>(demb + s * D + 16 * mt + 4 * g) = acc;
            tile_fence();
          }
        }
      } else if (l > 0) {
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
constexpr size_t kBackwardLdsX3 = size_t(4) * kXPlane * sizeof(unsigned short);

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

int64_t nsvf_mlp_train_x3_frags(int D, int H, int L) {
  if (!train_shape_ok(D, H, L)) return -1;
  return TrainDims<32, 128, 3>::X_FRAGS;
}

static int train_backward(const char* name, int D, int H, int L, const float* packed,
                          const int16_t* x3w, const float* emb, int64_t M, const float* dout,
                          float* demb, float* partial, int64_t partial_floats, float* dparams,
                          void* stream) {
  if (!train_shape_ok(D, H, L))
    return set_error(NSVF_EINVAL, "%s: shape D=%d H=%d L=%d (built: 32, 128, 3)", name, D, H, L);
  if (M < 0) return set_error(NSVF_EINVAL, "%s: M=%lld", name, static_cast<long long>(M));
  if (!packed || !partial || !dparams || (M > 0 && (!emb || !dout || !demb)))
    return set_error(NSVF_EINVAL, "%s: null pointer", name);
  using T = TrainDims<32, 128, 3>;
  const int nb = backward_blocks(M);
  if (partial_floats < int64_t(nb) * T::PARAMS)
    return set_error(NSVF_EINVAL, "%s: partial buffer of %lld floats, %lld needed", name,
                     static_cast<long long>(partial_floats),
                     static_cast<long long>(int64_t(nb) * T::PARAMS));
  hipStream_t st = as_stream(stream);
  const bf16x8* x3 = reinterpret_cast<const bf16x8*>(x3w);
  if (x3w)
    hipLaunchKernelGGL((mlp_train_backward_kernel<32, 128, 3, true>), dim3(nb),
                       dim3(kTrainBlock), kBackwardLdsX3, st, emb, M, packed, x3, dout, demb,
                       partial);
  else
    hipLaunchKernelGGL((mlp_train_backward_kernel<32, 128, 3, false>), dim3(nb),
                       dim3(kTrainBlock), kBackwardLds, st, emb, M, packed, x3, dout, demb,
                       partial);
  if (check_launch(name) != NSVF_OK) return NSVF_ELAUNCH;
  hipLaunchKernelGGL(sum_partials_kernel, dim3(unsigned(ceil_div(T::PARAMS, 256))), dim3(256), 0,
                     st, partial, nb, T::PARAMS, dparams);
  return check_launch("mlp_train_sum_partials");
}

int nsvf_mlp_train_backward(int D, int H, int L, const float* packed, const float* emb,
                            int64_t M, const float* dout, float* demb, float* partial,
                            int64_t partial_floats, float* dparams, void* stream) {
  clear_error();
  return train_backward("mlp_train_backward", D, H, L, packed, nullptr, emb, M, dout, demb,
                        partial, partial_floats, dparams, stream);
}

int nsvf_mlp_train_backward_x3(int D, int H, int L, const float* packed, const int16_t* x3w,
                               const float* emb, int64_t M, const float* dout, float* demb,
                               float* partial, int64_t partial_floats, float* dparams,
                               void* stream) {
  clear_error();
  if (!x3w || reinterpret_cast<uintptr_t>(x3w) % 16 != 0)
    return set_error(NSVF_EINVAL, "mlp_train_backward_x3: fragments null or not 16-byte aligned");
  return train_backward("mlp_train_backward_x3", D, H, L, packed, x3w, emb, M, dout, demb,
                        partial, partial_floats, dparams, stream);
}

}  // extern "C"
