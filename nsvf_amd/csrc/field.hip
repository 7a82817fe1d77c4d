// Field evaluation over a compacted sample stream: trilinear gather of the
// voxel-corner embeddings fused with the field MLP on the matrix cores.
//
// Gather (encoder.py:582-590, geometry.py:195-200): x = o + d*z; p = (x - c)
// / vs + 0.5; w_c = prod over axes of (bit ? p : 1 - p), corners c = 4x+2y+z;
// emb = sum_c w_c * values[feats[v][c]], in corner order, each product and
// sum rounded separately.
//
// MLP: L hidden layers (Linear + ReLU) of width H, then a 4-wide head giving
// sigma and RGB logits (RGB through a sigmoid).  The layers run in the
// transposed form Y^T = W X^T on 16x16 MFMA tiles: the weights are operand A
// (rows = output features), the activations operand B (columns = samples).
// Each lane's accumulator rows of one layer are then exactly the operand-B
// elements it needs for the next layer, provided the next layer's weights are
// stored with the input features in the matching permuted order -- which the
// host packer does once per scene -- so activations never leave registers.
//
//   BF16: v_mfma_f32_16x16x32_bf16, FP32 accumulation.  K-tile of 32 input
//         features; lane group g = lane/16 holds features
//         {4g..4g+3, 16+4g..16+4g+3} of its K-tile.
//   FP32: v_mfma_f32_16x16x4_f32.  K-tile of 4 features; lane group g holds
//         feature 16t + 4g + i of K-tile (t, i).
// Both layouts make a lane gather the same 8 embedding features per 32.
//
// Weight fragments: [layer][m-tile][k-tile][64 lanes][8 bf16 | 1 f32],
// biases fp32 [layer][rows padded to 16].  BF16 fragments are staged in LDS.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "render.h"

namespace nsvf {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kFieldBlock = 256;  // 4 waves

struct FieldScene {
  const float* centers;   // [N,3] voxel centres (x-shifted)
  const int32_t* feats;   // [N,8] corner key ids
  const float* values;    // [K,D] corner embeddings
  float voxel_size;
  float inv_voxel;        // 1 / voxel_size, used only when exact
  int voxel_pow2;         // voxel_size is a power of two: x / vs == x * (1/vs)
};

struct FieldRays {
  const float* start;  // [R,3]
  const float* dir;    // [R,3]
};

// Trilinear weights of a point x in voxel v (corner order 4x + 2y + z).
__device__ __forceinline__ void corner_weights(const FieldScene& sc, int32_t v, const float x[3],
                                               float w[8], int32_t key[8]) {
  float p[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) p[a] = (x[a] - sc.centers[int64_t(v) * 3 + a]) / sc.voxel_size + 0.5f;
  const float q[3] = {1.0f - p[0], 1.0f - p[1], 1.0f - p[2]};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float wx = (c & 4) ? p[0] : q[0];
    const float wy = (c & 2) ? p[1] : q[1];
    const float wz = (c & 1) ? p[2] : q[2];
    w[c] = (wx * wy) * wz;
    key[c] = sc.feats[int64_t(v) * 8 + c];
  }
}

// The 8 features {32kt + 4g + 0..3, 32kt + 16 + 4g + 0..3} of one sample's
// embedding (zero past D).
template <int D>
__device__ __forceinline__ void gather8(const FieldScene& sc, const float w[8],
                                        const int32_t key[8], int kt, int g, float out[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = 0.0f;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int f0 = 32 * kt + 16 * half + 4 * g;
    if (f0 < D) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float4 e = *reinterpret_cast<const float4*>(sc.values + int64_t(key[c]) * D + f0);
        out[4 * half + 0] = out[4 * half + 0] + w[c] * e.x;
        out[4 * half + 1] = out[4 * half + 1] + w[c] * e.y;
        out[4 * half + 2] = out[4 * half + 2] + w[c] * e.z;
        out[4 * half + 3] = out[4 * half + 3] + w[c] * e.w;
      }
    }
  }
}

// Corner weights for the BF16 field kernel: the same formula, with the
// division by the voxel size done as an exact multiply when the size is a
// power of two.
__device__ __forceinline__ void corner_weights_fast(const FieldScene& sc, int32_t v,
                                                    const float x[3], float w[8],
                                                    int32_t key[8]) {
  float p[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float d = x[a] - sc.centers[int64_t(v) * 3 + a];
    p[a] = (sc.voxel_pow2 ? d * sc.inv_voxel : d / sc.voxel_size) + 0.5f;
  }
  const float q[3] = {1.0f - p[0], 1.0f - p[1], 1.0f - p[2]};
  const float xy[4] = {q[0] * q[1], q[0] * p[1], p[0] * q[1], p[0] * p[1]};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    w[c] = xy[c >> 1] * ((c & 1) ? p[2] : q[2]);
    key[c] = sc.feats[int64_t(v) * 8 + c];
  }
}

// gather8 with fused multiply-adds (the BF16 path rounds to bf16 afterwards).
template <int D>
__device__ __forceinline__ void gather8_fma(const FieldScene& sc, const float w[8],
                                            const int32_t key[8], int kt, int g, float out[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = 0.0f;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int f0 = 32 * kt + 16 * half + 4 * g;
    if (f0 < D) {
      float4 e[8];
#pragma unroll
      for (int c = 0; c < 8; ++c)
        e[c] = *reinterpret_cast<const float4*>(sc.values + int64_t(key[c]) * D + f0);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        out[4 * half + 0] = __fmaf_rn(w[c], e[c].x, out[4 * half + 0]);
        out[4 * half + 1] = __fmaf_rn(w[c], e[c].y, out[4 * half + 1]);
        out[4 * half + 2] = __fmaf_rn(w[c], e[c].z, out[4 * half + 2]);
        out[4 * half + 3] = __fmaf_rn(w[c], e[c].w, out[4 * half + 3]);
      }
    }
  }
}

typedef __attribute__((ext_vector_type(2))) short s16x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;

// ReLU then round to bf16 == round then ReLU (rounding keeps the sign); the
// ReLU runs on the packed bf16 pair as a signed 16-bit max with 0.
typedef __attribute__((ext_vector_type(2))) float f32x2;

__device__ __forceinline__ unsigned int relu_pack(float a, float b) {
  const bf16x2 v = __builtin_convertvector(f32x2{a, b}, bf16x2);
  s16x2 t = __builtin_bit_cast(s16x2, v);
  t = __builtin_elementwise_max(t, s16x2{0, 0});
  return __builtin_bit_cast(unsigned int, t);
}

__device__ __forceinline__ bf16x8 relu_pack8(const f32x4& lo, const f32x4& hi) {
  u32x4 u = {relu_pack(lo[0], lo[1]), relu_pack(lo[2], lo[3]), relu_pack(hi[0], hi[1]),
             relu_pack(hi[2], hi[3])};
  return __builtin_bit_cast(bf16x8, u);
}

// Split-BF16 operands: f = hi + lo with hi = bf16(f) and lo = bf16(f - hi)
// (the difference is exact in FP32), so a product with hi and lo carries
// about 16 significant bits of f.  Packs features {lo4[0..3], hi4[0..3]} in
// the operand order of relu_pack8.
__device__ __forceinline__ void split_pack2(float a, float b, unsigned int& hi,
                                            unsigned int& lo) {
  const bf16x2 h = __builtin_convertvector(f32x2{a, b}, bf16x2);
  const f32x2 hf = __builtin_convertvector(h, f32x2);
  const bf16x2 l = __builtin_convertvector(f32x2{a - hf[0], b - hf[1]}, bf16x2);
  hi = __builtin_bit_cast(unsigned int, h);
  lo = __builtin_bit_cast(unsigned int, l);
}

__device__ __forceinline__ void split_pack8(const f32x4& lo4, const f32x4& hi4, bf16x8& hi,
                                            bf16x8& lo) {
  unsigned int h0, h1, h2, h3, l0, l1, l2, l3;
  split_pack2(lo4[0], lo4[1], h0, l0);
  split_pack2(lo4[2], lo4[3], h1, l1);
  split_pack2(hi4[0], hi4[1], h2, l2);
  split_pack2(hi4[2], hi4[3], h3, l3);
  hi = __builtin_bit_cast(bf16x8, u32x4{h0, h1, h2, h3});
  lo = __builtin_bit_cast(bf16x8, u32x4{l0, l1, l2, l3});
}

// 1 / (1 + e^-x) from the hardware base-2 exponential and reciprocal (a few
// ulp; the head's three sigmoids otherwise cost as much VALU as a hidden
// layer's epilogue).  Saturates to 0 / 1 for large |x|.
__device__ __forceinline__ float sigmoidf(float x) {
  const float e = __builtin_amdgcn_exp2f(x * -1.44269504088896341f);
  return __builtin_amdgcn_rcpf(1.0f + e);
}

// ---------------------------------------------------------------------------
// Sample sources: sample s -> (voxel, corner weights and keys), false for a
// slot without a sample.

// Samples of a sample stream: x = o + d * z of the record's ray
// (geometry.py:160-161), each operation rounded.
// With a list, stream sample s is the record at slot list[s].
struct StreamSource {
  FieldScene sc;
  FieldRays rays;
  const SampleRec* samples;
  const int32_t* list;
  template <bool kFast>
  __device__ __forceinline__ bool weights(int64_t s, float w[8], int32_t key[8]) const {
    const SampleRec rec = samples[list ? list[s] : s];
    if (rec.idx < 0) return false;  // an unused slot of a ray's range
    float x[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
      x[a] = rays.start[int64_t(rec.ray) * 3 + a] + rays.dir[int64_t(rec.ray) * 3 + a] * rec.depth;
    if (kFast)
      corner_weights_fast(sc, rec.idx, x, w, key);
    else
      corner_weights(sc, rec.idx, x, w, key);
    return true;
  }
};

// The pruning lattice (encoder.py:626-631, geometry.py:229-238): bits^3
// points per voxel v0 + s / bits^3, point (i, j, k) (i slowest) at
// c + fp32((2i + 1 - bits) / (bits - 1)) * (vs / 2) per axis, corners
// included.
struct LatticeSource {
  FieldScene sc;
  int64_t voxel0;
  int bits;
  float half_voxel;
  template <bool kFast>
  __device__ __forceinline__ bool weights(int64_t s, float w[8], int32_t key[8]) const {
    const int64_t per = int64_t(bits) * bits * bits;
    const int64_t v = voxel0 + s / per;
    const int l = int(s - (s / per) * per);
    const int ijk[3] = {l / (bits * bits), (l / bits) % bits, l % bits};
    float x[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float off = float(2 * ijk[a] + 1 - bits) / float(bits - 1);
      x[a] = sc.centers[v * 3 + a] + off * half_voxel;
    }
    if (kFast)
      corner_weights_fast(sc, int32_t(v), x, w, key);
    else
      corner_weights(sc, int32_t(v), x, w, key);
    return true;
  }
};

// Output sinks: begin_tile() runs on every lane of the wave when a tile of n
// samples at base starts; put() receives each head result (sigma and the
// three RGB logits, biases added) of sample s = base + 16 j + (lane % 16) on
// every lane, to be kept where ok; end_tile() runs on every lane after each
// tile.

// Writes (sigma, sigmoid(rgb)) per sample: to slot s, or with a list to
// slot list[s].  A tile's slots are loaded when it starts (one per lane:
// sample base + lane) and handed to the storing lanes by a lane shuffle, so
// the stores never wait on the list.
struct StoreSink {
  float4* out;
  const int32_t* list;
  int32_t mine = 0;
  __device__ __forceinline__ void begin_tile(int64_t base, int64_t M, int n) {
    const int lane = threadIdx.x & 63;
    if (list) mine = (lane < n && base + lane < M) ? list[base + lane] : 0;
  }
  __device__ __forceinline__ void put(int j, int64_t s, bool ok, float sg, float r, float g,
                                      float b) {
    int64_t o = s;
    if (list) o = __shfl(mine, 16 * j + (threadIdx.x & 15), 64);
    if (ok) out[o] = make_float4(sg, sigmoidf(r), sigmoidf(g), sigmoidf(b));
  }
  __device__ __forceinline__ void end_tile(int64_t) {}
};

// Per-voxel max of relu(sigma) over the lattice (encoder.py:633-638 keeps a
// voxel when 1 - min exp(-relu sigma) > th, i.e. from this max).  A tile's
// samples lie in one voxel (bits^3 is a multiple of the tile size), so the
// wave reduces the tile's max and one lane folds it into the voxel's value
// with an integer atomicMax (non-negative floats order as their bits).
struct MaxSigmaSink {
  float* vmax;   // [voxels of the launch]
  int64_t per;   // samples per voxel
  float m;
  __device__ __forceinline__ void begin_tile(int64_t, int64_t, int) {}
  __device__ __forceinline__ void put(int, int64_t, bool ok, float sg, float, float, float) {
    if (ok && sg > m) m = sg;
  }
  __device__ __forceinline__ void end_tile(int64_t base) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0)
      atomicMax(reinterpret_cast<int*>(vmax + base / per), __float_as_int(m));
    m = 0.0f;
  }
};

template <int D, int H, int L>
struct Dims {
  static constexpr int KT0 = (D + 31) / 32;   // k-tiles of the input (bf16)
  static constexpr int KTH = H / 32;          // k-tiles of a hidden layer (bf16)
  static constexpr int MTH = H / 16;          // m-tiles of a hidden layer
  static constexpr int NT = H <= 128 ? 4 : 2; // 16-sample n-tiles per wave
  static constexpr int SAMPLES = 16 * NT;     // samples per wave step
  // bf16 fragment counts (1 KB each)
  static constexpr int F0 = MTH * KT0;
  static constexpr int FH = MTH * KTH;
  static constexpr int FHEAD = KTH;
  static constexpr int FRAGS = F0 + (L - 1) * FH + FHEAD;
  static constexpr int BIASES = H * L + 16;
  // blocks per CU the MLP kernel is built for: two waves per SIMD when the
  // layer width lets the activations fit half the register file
  static constexpr int MIN_BLOCKS = H <= 128 ? 2 : 1;
};

template <int D, int H, int L>
constexpr bool weights_in_lds() {
  return Dims<D, H, L>::FRAGS * 1024 <= 96 * 1024;
}

// The MLP on one wave's tile of SAMPLES samples: operand B of layer 0 in
// b0, weights lw / biases lb (LDS or global), head results of samples < M
// to the sink.
// One wave's layer-0 operand tile: [k-tile][n-tile] fragments.
template <int D, int H, int L>
using OperandTile = bf16x8[Dims<D, H, L>::KT0][Dims<D, H, L>::NT];

// Work a caller interleaves with the MLP's layers (hook->layer_done(l)
// after layer l, hook->before_head() before the head); none by default.
struct NoHook {
  __device__ __forceinline__ void layer_done(int) {}
  __device__ __forceinline__ void before_head() {}
};

template <int D, int H, int L, class Sink, class Hook>
__device__ __forceinline__ void mlp_tile_hooked(OperandTile<D, H, L>& b0, const bf16x8* lw,
                                                const float* lb, int lane, int g, int col,
                                                int64_t base, int64_t M, Sink& sink,
                                                const bf16x8* __restrict__ next_b0, Hook& hook);

template <int D, int H, int L, class Sink>
__device__ __forceinline__ void mlp_tile(OperandTile<D, H, L>& b0,
                                         const bf16x8* lw, const float* lb, int lane, int g,
                                         int col, int64_t base, int64_t M, Sink& sink,
                                         const bf16x8* __restrict__ next_b0 = nullptr) {
  NoHook none;
  mlp_tile_hooked<D, H, L>(b0, lw, lb, lane, g, col, base, M, sink, next_b0, none);
}

// One layer's matrix work, out = relu(A in + bias), over m-tile pairs: each
// weight fragment feeds NT MFMAs.  With kPin (weights in LDS) the fragments
// are read two ahead and a scheduling barrier closes each fragment's MFMAs,
// which pins this order: left to itself the compiler runs the k-tiles
// outermost over all m-tiles, which keeps every accumulator live and pushes
// each fragment read down to just before its use, so the wave waits out an
// LDS read every NT MFMAs.  (With the weights in global memory the pinned
// form makes the compiler hoist the loads of later layers and spill, so
// there the fragments go one ahead and the order is left to the compiler.)
// The accumulation order per output is k-tile 0, 1, ... either way.
template <int KT, int MT, int NT, bool kPin>
__device__ __forceinline__ void mfma_relu_layer(const bf16x8* A, const float* B, int lane, int g,
                                                const bf16x8 (&in)[KT][NT],
                                                bf16x8 (&out)[MT / 2][NT]) {
  constexpr int kFragLead = kPin ? 2 : 1;
  constexpr int NF = MT * KT, NB = kFragLead + 1;
  bf16x8 a[NB];
#pragma unroll
  for (int p = 0; p < kFragLead; ++p)
    if (p < NF) a[p] = A[p * 64 + lane];
  f32x4 bias_next = *reinterpret_cast<const f32x4*>(B + 4 * g);
#pragma unroll
  for (int kt2 = 0; kt2 < MT / 2; ++kt2) {
    f32x4 acc[2][NT];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int mt = 2 * kt2 + half;
      const f32x4 bias4 = bias_next;
      if (mt + 1 < MT) bias_next = *reinterpret_cast<const f32x4*>(B + 16 * (mt + 1) + 4 * g);
#pragma unroll
      for (int kt = 0; kt < KT; ++kt) {
        const int f = mt * KT + kt;
        if (f + kFragLead < NF) a[(f + kFragLead) % NB] = A[(f + kFragLead) * 64 + lane];
#pragma unroll
        for (int j = 0; j < NT; ++j)
          acc[half][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              a[f % NB], in[kt][j], kt == 0 ? bias4 : acc[half][j], 0, 0, 0);
        if constexpr (kPin) __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) out[kt2][j] = relu_pack8(acc[0][j], acc[1][j]);
  }
}

template <int D, int H, int L, class Sink, class Hook>
__device__ __forceinline__ void mlp_tile_hooked(OperandTile<D, H, L>& b0, const bf16x8* lw,
                                                const float* lb, int lane, int g, int col,
                                                int64_t base, int64_t M, Sink& sink,
                                                const bf16x8* __restrict__ next_b0, Hook& hook) {
  using Dm = Dims<D, H, L>;
  constexpr int NT = Dm::NT, KTH = Dm::KTH, MTH = Dm::MTH;
  constexpr bool kPin = weights_in_lds<D, H, L>();
  static_assert(MTH == 2 * KTH, "a layer's output m-tile pairs are the next layer's k-tiles");
  bf16x8 bh[KTH][NT];
  // ---- layer 0: bias enters as the first MFMA's C operand
  mfma_relu_layer<Dm::KT0, MTH, NT, kPin>(lw, lb, lane, g, b0, bh);
  hook.layer_done(0);
  // ---- hidden layers 1..L-1
  if constexpr (kPin) {
    // two per trip, so that the activations alternate between bh and nb
    // without copies
    bf16x8 nb[KTH][NT];
#pragma unroll 1
    for (int l = 1; l + 1 < L; l += 2) {
      mfma_relu_layer<KTH, MTH, NT, true>(lw + (Dm::F0 + (l - 1) * Dm::FH) * 64, lb + H * l,
                                          lane, g, bh, nb);
      hook.layer_done(l);
      mfma_relu_layer<KTH, MTH, NT, true>(lw + (Dm::F0 + l * Dm::FH) * 64, lb + H * (l + 1),
                                          lane, g, nb, bh);
      hook.layer_done(l + 1);
    }
    if ((L - 1) % 2) {
      mfma_relu_layer<KTH, MTH, NT, true>(lw + (Dm::F0 + (L - 2) * Dm::FH) * 64,
                                          lb + H * (L - 1), lane, g, bh, nb);
#pragma unroll
      for (int kt = 0; kt < KTH; ++kt)
#pragma unroll
        for (int j = 0; j < NT; ++j) bh[kt][j] = nb[kt][j];
      hook.layer_done(L - 1);
    }
  } else {
#pragma unroll 1
    for (int l = 1; l < L; ++l) {
      bf16x8 nb[KTH][NT];
      mfma_relu_layer<KTH, MTH, NT, false>(lw + (Dm::F0 + (l - 1) * Dm::FH) * 64, lb + H * l,
                                           lane, g, bh, nb);
#pragma unroll
      for (int kt = 0; kt < KTH; ++kt)
#pragma unroll
        for (int j = 0; j < NT; ++j) bh[kt][j] = nb[kt][j];
      hook.layer_done(l);
    }
  }
  // b0 is dead since layer 0 and the hidden layers' double buffer now too:
  // the next tile's operands load into b0 while the head and the stores run
  if (next_b0) {
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int kt = 0; kt < Dm::KT0; ++kt) b0[kt][j] = next_b0[(j * Dm::KT0 + kt) * 64 + lane];
  }
  hook.before_head();
  // ---- head (4 outputs in rows 0..3 of one m-tile)
  {
    const bf16x8* A = lw + (Dm::F0 + (L - 1) * Dm::FH) * 64;
    const f32x4 bias4 = *reinterpret_cast<const f32x4*>(lb + H * L + 4 * g);
    constexpr int kFragLead = kPin ? 2 : 1;
    constexpr int NB = kFragLead + 1;
    f32x4 acc[NT];
    bf16x8 a[NB];
#pragma unroll
    for (int p = 0; p < kFragLead; ++p)
      if (p < KTH) a[p] = A[p * 64 + lane];
#pragma unroll
    for (int kt = 0; kt < KTH; ++kt) {
      if (kt + kFragLead < KTH) a[(kt + kFragLead) % NB] = A[(kt + kFragLead) * 64 + lane];
#pragma unroll
      for (int j = 0; j < NT; ++j)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kt % NB], bh[kt][j],
                                                         kt == 0 ? bias4 : acc[j], 0, 0, 0);
      if constexpr (kPin) __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int64_t s = base + 16 * j + col;
      sink.put(j, s, g == 0 && s < M, acc[j][0], acc[j][1], acc[j][2], acc[j][3]);
    }
    sink.end_tile(base);
  }
}

template <int D, int H, int L, class Src, class Sink>
__global__ __launch_bounds__(kFieldBlock) void field_bf16_kernel(
    Src src, const int32_t* m_dev, int64_t m_max, const bf16x8* __restrict__ wfrag,
    const float* __restrict__ bias, Sink sink) {
  using Dm = Dims<D, H, L>;
  constexpr int NT = Dm::NT;
  constexpr bool kLds = weights_in_lds<D, H, L>();
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const bf16x8* lw = wfrag;
  float* lb = reinterpret_cast<float*>(smem_raw);
  if constexpr (kLds) {
    bf16x8* w = reinterpret_cast<bf16x8*>(smem_raw);
    for (int i = threadIdx.x; i < Dm::FRAGS * 64; i += kFieldBlock) w[i] = wfrag[i];
    lw = w;
    lb = reinterpret_cast<float*>(smem_raw + size_t(Dm::FRAGS) * 64 * sizeof(bf16x8));
  }
  for (int i = threadIdx.x; i < Dm::BIASES; i += kFieldBlock) lb[i] = bias[i];
  __syncthreads();

  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  // Per-wave staging of the gathered layer-0 operands: each lane writes and
  // reads back only its own slots, so no barrier is needed; staging keeps
  // the gather loop rolled (one sample's loads in flight per lane).
  bf16x8* stage = reinterpret_cast<bf16x8*>(
                      reinterpret_cast<unsigned char*>(lb) +
                      ((size_t(Dm::BIASES) * sizeof(float) + 15) & ~size_t(15))) +
                  (threadIdx.x >> 6) * (NT * Dm::KT0 * 64);
  const int64_t M = m_dev ? min(int64_t(*m_dev), m_max) : m_max;
  const int64_t tiles = (M + Dm::SAMPLES - 1) / Dm::SAMPLES;
  for (int64_t t = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); t < tiles;
       t += int64_t(gridDim.x) * 4) {
    const int64_t base = t * Dm::SAMPLES;
    sink.begin_tile(base, M, Dm::SAMPLES);
    // ---- gather -> operand B of layer 0
#pragma unroll 1
    for (int j = 0; j < NT; ++j) {
      const int64_t s = base + 16 * j + col;
      float w[8];
      int32_t key[8];
      const bool valid = s < M && src.template weights<true>(s, w, key);
#pragma unroll
      for (int kt = 0; kt < Dm::KT0; ++kt) {
        float f[8];
        if (valid) {
          gather8_fma<D>(src.sc, w, key, kt, g, f);
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) f[i] = 0.0f;
        }
        const bf16x2 p0 = __builtin_convertvector(f32x2{f[0], f[1]}, bf16x2);
        const bf16x2 p1 = __builtin_convertvector(f32x2{f[2], f[3]}, bf16x2);
        const bf16x2 p2 = __builtin_convertvector(f32x2{f[4], f[5]}, bf16x2);
        const bf16x2 p3 = __builtin_convertvector(f32x2{f[6], f[7]}, bf16x2);
        const u32x4 u = {__builtin_bit_cast(unsigned int, p0), __builtin_bit_cast(unsigned int, p1),
                         __builtin_bit_cast(unsigned int, p2), __builtin_bit_cast(unsigned int, p3)};
        stage[(j * Dm::KT0 + kt) * 64 + lane] = __builtin_bit_cast(bf16x8, u);
      }
    }
    bf16x8 b0[Dm::KT0][NT];
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int kt = 0; kt < Dm::KT0; ++kt) b0[kt][j] = stage[(j * Dm::KT0 + kt) * 64 + lane];
    mlp_tile<D, H, L>(b0, lw, lb, lane, g, col, base, M, sink);
  }
}

// Gather kernel of the split BF16 path.  A wavefront takes 64 consecutive
// slots; consecutive samples of a ray mostly share their voxel, so the wave
// finds its runs of equal voxel ids (about 8 per 64 samples), stages each
// run's 8 corner rows in LDS with one coalesced load per row block, and each
// lane then accumulates its sample from LDS -- instead of every lane pulling
// 1 KB of rows through the texture path.  Per sample and feature the sum is
// the same chain of fused multiply-adds in corner order as the fused kernel.
// Output: the MLP kernel's tile layout -- for k-tile kt and lane group g,
// features {32kt + 4g + 0..3, 32kt + 16 + 4g + 0..3} as one 16-byte bf16x8
// at ((t * NT + j) * KT0 + kt) * 64 + 16g + col, for sample
// s = t * 16 NT + 16j + col (NT n-tiles of 16 samples per MLP tile; 64 is a
// multiple of 16 NT), written with nontemporal stores: the tiles are read
// once, by the MLP kernel, after far more has passed through the caches
// (the kernel is bound by these writes).  With kLo the split-BF16 MLP's
// second operand, the rounding residual bf16(f - bf16(f)) of each feature,
// goes to the same position lo_off elements further on.
template <int D, int H, int L, int NT = Dims<D, H, L>::NT, bool kLo = false>
__global__ __launch_bounds__(256) void embed_bf16_kernel(FieldScene sc, FieldRays rays,
                                                         const SampleRec* __restrict__ samples,
                                                         const int32_t* __restrict__ list,
                                                         const int32_t* m_dev, int64_t m_max,
                                                         bf16x8* __restrict__ emb,
                                                         int64_t lo_off = 0) {
  using Dm = Dims<D, H, L>;
  static_assert(D % 4 == 0, "embedding width must be a multiple of 4");
  static_assert(64 % (16 * NT) == 0, "MLP tiles must divide the 64-slot gather tiles");
  constexpr int SAMPLES = 16 * NT;
  constexpr int RUNS = 8;        // voxels staged per round
  constexpr int ROW4 = 2 * D;    // float4 per voxel: 8 corners x D / 4
  // per-run stride padded by one float4, so the same element of different
  // runs falls in different LDS banks
  constexpr int RUN_STRIDE = ROW4 + 1;
  __shared__ float4 rows_lds[4][RUNS * RUN_STRIDE];
  __shared__ int32_t run_vox[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t M = min(int64_t(*m_dev), m_max);
  const int64_t total = ((M + 63) / 64) * 64;
  // persistent waves over 64-slot tiles; each tile's sample records are
  // loaded one tile ahead, so a tile's dependent chain starts at its keys
  const int64_t ntile = total / 64, stride = int64_t(gridDim.x) * 4;
  int64_t tile = int64_t(blockIdx.x) * 4 + wave;
  if (tile >= ntile) return;  // the whole wave
  SampleRec nrec{-1, 0.0f, 0.0f, 0};
  if (tile * 64 + lane < M) {
    const int64_t s0 = tile * 64 + lane;
    nrec = samples[list ? list[s0] : s0];
  }
  for (; tile < ntile; tile += stride) {
  const int64_t s = tile * 64 + lane;
  const SampleRec rec = nrec;
  if (tile + stride < ntile && s + stride * 64 < M) {
    const int64_t sn = s + stride * 64;
    nrec = samples[list ? list[sn] : sn];
  }
  // the sample's voxel; the loads its corner weights need (ray and centre)
  // are issued here and used after the row staging below, so their latency
  // overlaps the staging's
  const int32_t v = s < M ? rec.idx : -1;
  float ro[3] = {0.f, 0.f, 0.f}, rdir[3] = {0.f, 0.f, 0.f}, ctr[3] = {0.f, 0.f, 0.f};
  const float z = rec.depth;
  if (v >= 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      ro[a] = rays.start[int64_t(rec.ray) * 3 + a];
      rdir[a] = rays.dir[int64_t(rec.ray) * 3 + a];
      ctr[a] = sc.centers[int64_t(v) * 3 + a];
    }
  }
  // runs of equal voxel ids among the wave's lanes
  const int32_t vprev = __shfl_up(v, 1, 64);
  const bool head = lane == 0 || v != vprev;
  const uint64_t hmask = __ballot(head);
  const int r = __popcll(hmask & ((2ull << lane) - 1)) - 1;
  const int U = __popcll(hmask);
  if (head) run_vox[wave][r] = v;
  wave_sync();
  float acc[D];
#pragma unroll
  for (int f = 0; f < D; ++f) acc[f] = 0.0f;
  float w[8];
  bool have_w = false;
  constexpr int SLOTS = RUNS * ROW4 / 64;  // staging slots per lane and round
  for (int u0 = 0; u0 < U; u0 += RUNS) {
    const int nr = min(RUNS, U - u0);
    // one unrolled, branch-free pass over the round's staging slots (slots of
    // absent runs load a valid row and are never read), so the corner-key
    // and row loads of all slots are in flight together
#pragma unroll
    for (int t = 0; t < SLOTS; ++t) {
      const int i = lane + 64 * t;
      const int j = i / ROW4, e = i - j * ROW4;
      const int32_t vr = run_vox[wave][(u0 + j) & 63];  // in bounds for any j
      const int32_t vu = (j < nr && vr >= 0) ? vr : 0;
      const int32_t k = sc.feats[int64_t(vu) * 8 + e / (D / 4)];
      rows_lds[wave][j * RUN_STRIDE + e] =
          reinterpret_cast<const float4*>(sc.values + int64_t(k) * D)[e % (D / 4)];
    }
    if (!have_w) {
      // corner weights: x = o + d * z, p = (x - c) / vs + 0.5 (the fused
      // kernel's corner_weights_fast arithmetic)
      float p[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float x = ro[a] + rdir[a] * z;
        const float d = x - ctr[a];
        p[a] = (sc.voxel_pow2 ? d * sc.inv_voxel : d / sc.voxel_size) + 0.5f;
      }
      const float q[3] = {1.0f - p[0], 1.0f - p[1], 1.0f - p[2]};
      const float xy[4] = {q[0] * q[1], q[0] * p[1], p[0] * q[1], p[0] * p[1]};
#pragma unroll
      for (int c = 0; c < 8; ++c) w[c] = xy[c >> 1] * ((c & 1) ? p[2] : q[2]);
      have_w = true;
    }
    wave_sync();
    if (v >= 0 && r >= u0 && r < u0 + nr) {
      const float4* rows = &rows_lds[wave][(r - u0) * RUN_STRIDE];
#pragma unroll
      for (int qd = 0; qd < D / 4; ++qd) {
        float4 e[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) e[c] = rows[c * (D / 4) + qd];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          acc[4 * qd + 0] = __fmaf_rn(w[c], e[c].x, acc[4 * qd + 0]);
          acc[4 * qd + 1] = __fmaf_rn(w[c], e[c].y, acc[4 * qd + 1]);
          acc[4 * qd + 2] = __fmaf_rn(w[c], e[c].z, acc[4 * qd + 2]);
          acc[4 * qd + 3] = __fmaf_rn(w[c], e[c].w, acc[4 * qd + 3]);
        }
      }
    }
    wave_sync();
  }
  const int64_t t = s / SAMPLES;
  const int within = int(s - t * SAMPLES);
  const int j = within >> 4, col = within & 15;
  bf16x8* base = emb + ((t * NT + j) * Dm::KT0) * 64 + col;
#pragma unroll
  for (int kt = 0; kt < Dm::KT0; ++kt) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      float f[8];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int feat = 32 * kt + 16 * h + 4 * g + i;
          f[4 * h + i] = feat < D ? acc[feat < D ? feat : 0] : 0.0f;
        }
      if constexpr (kLo) {
        bf16x8 hi, lo;
        split_pack8(f32x4{f[0], f[1], f[2], f[3]}, f32x4{f[4], f[5], f[6], f[7]}, hi, lo);
        __builtin_nontemporal_store(hi, &base[kt * 64 + 16 * g]);
        __builtin_nontemporal_store(lo, &base[lo_off + kt * 64 + 16 * g]);
      } else {
        const bf16x2 p0 = __builtin_convertvector(f32x2{f[0], f[1]}, bf16x2);
        const bf16x2 p1 = __builtin_convertvector(f32x2{f[2], f[3]}, bf16x2);
        const bf16x2 p2 = __builtin_convertvector(f32x2{f[4], f[5]}, bf16x2);
        const bf16x2 p3 = __builtin_convertvector(f32x2{f[6], f[7]}, bf16x2);
        const u32x4 u = {__builtin_bit_cast(unsigned int, p0),
                         __builtin_bit_cast(unsigned int, p1),
                         __builtin_bit_cast(unsigned int, p2),
                         __builtin_bit_cast(unsigned int, p3)};
        __builtin_nontemporal_store(__builtin_bit_cast(bf16x8, u), &base[kt * 64 + 16 * g]);
      }
    }
  }
  }
}

// MLP kernel of the split BF16 path: operand B of layer 0 read from the
// gather kernel's output.
template <int D, int H, int L>
__global__ __launch_bounds__(kFieldBlock, (Dims<D, H, L>::MIN_BLOCKS)) void mlp_bf16_kernel(
    const bf16x8* __restrict__ emb, const int32_t* m_dev, int64_t m_max,
    const bf16x8* __restrict__ wfrag, const float* __restrict__ bias, float4* __restrict__ out,
    const int32_t* __restrict__ list) {
  using Dm = Dims<D, H, L>;
  constexpr int NT = Dm::NT;
  constexpr bool kLds = weights_in_lds<D, H, L>();
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int64_t M = min(int64_t(*m_dev), m_max);
  const int64_t tiles = (M + Dm::SAMPLES - 1) / Dm::SAMPLES;
  if (int64_t(blockIdx.x) * 4 >= tiles) return;  // no tile: no weight staging
  const bf16x8* lw = wfrag;
  float* lb = reinterpret_cast<float*>(smem_raw);
  if constexpr (kLds) {
    bf16x8* w = reinterpret_cast<bf16x8*>(smem_raw);
    for (int i = threadIdx.x; i < Dm::FRAGS * 64; i += kFieldBlock) w[i] = wfrag[i];
    lw = w;
    lb = reinterpret_cast<float*>(smem_raw + size_t(Dm::FRAGS) * 64 * sizeof(bf16x8));
  }
  for (int i = threadIdx.x; i < Dm::BIASES; i += kFieldBlock) lb[i] = bias[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const int64_t stride = int64_t(gridDim.x) * 4;
  int64_t t = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (t >= tiles) return;
  bf16x8 b0[Dm::KT0][NT];
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int kt = 0; kt < Dm::KT0; ++kt) b0[kt][j] = emb[((t * NT + j) * Dm::KT0 + kt) * 64 + lane];
  for (; t < tiles; t += stride) {
    const int64_t tn = t + stride;
    StoreSink sink{out, list};
    sink.begin_tile(t * Dm::SAMPLES, M, Dm::SAMPLES);
    mlp_tile<D, H, L>(b0, lw, lb, lane, g, col, t * Dm::SAMPLES, M, sink,
                      tn < tiles ? emb + tn * NT * Dm::KT0 * 64 : nullptr);
  }
}

// ---------------------------------------------------------------------------
// Split-BF16 MLP ("bf16x3"): each layer's product W X runs as
// Wl Xh + Wh Xl + Wh Xh on v_mfma_f32_16x16x32_bf16 with FP32 accumulation,
// where W = Wh + Wl and X = Xh + Xl are split as in split_pack2 (Henry, Tang
// and Heinecke, "Leveraging the bfloat16 artificial intelligence datatype
// for higher-precision computations", ARITH 2019).  The dropped Wl Xl term is
// about 2^-16 of the product, so a layer keeps about 16 significant bits of
// its operands where plain BF16 keeps 8; three MFMAs per product.  Weight
// fragment f is stored as a pair: hi part at [2f][64 lanes], lo part at
// [2f + 1][64 lanes].  For the shapes it supports both parts fill most of the
// LDS (D = 32, H = 128, L = 3: 152 KB), so one block of eight waves -- two
// per SIMD -- runs per CU, each wave on tiles of 32 samples (two n-tiles:
// the hi and lo activations of four layers' worth of registers fit two waves
// per SIMD); the layer-0 operands are the gather kernel's hi / lo tiles.
template <int D, int H, int L>
struct Dims3 : Dims<D, H, L> {
  static constexpr int NT = 2;
  static constexpr int SAMPLES = 32;
  static constexpr size_t LDS_BYTES =
      size_t(Dims<D, H, L>::FRAGS) * 2048 + size_t(Dims<D, H, L>::BIASES) * sizeof(float);
};
constexpr int kX3Waves = 8;

template <int D, int H, int L>
constexpr bool x3_supported() {
  return H <= 128 && Dims3<D, H, L>::LDS_BYTES <= 160 * 1024;
}

// One split-BF16 layer, out = relu(W in + bias) split into hi / lo, over
// m-tile pairs with the weight pairs read two ahead (as mfma_relu_layer).
// Per output the accumulation order is k-tile 0, 1, ..., and within a
// k-tile Wl Xh, Wh Xl, Wh Xh.
// ReLU on the float's bits: a non-negative float is a non-negative int32 and
// a negative one a negative int32, so max with 0 clears exactly the negative
// values -- one integer max, where fmaxf costs a NaN-quieting max more (a NaN
// with the sign bit clear passes through, as torch.relu passes NaN).
__device__ __forceinline__ float relu_bits(float x) {
  return __int_as_float(max(__float_as_int(x), 0));
}

// The epilogue of one m-tile pair and sample n-tile j: ReLU, split into hi /
// lo BF16 operands of the next layer.
template <int NT>
__device__ __forceinline__ void x3_epilogue(const f32x4 (&acc)[2][NT], int j, bf16x8& oh,
                                            bf16x8& ol) {
  f32x4 r0, r1;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    r0[i] = relu_bits(acc[0][j][i]);
    r1[i] = relu_bits(acc[1][j][i]);
  }
  split_pack8(r0, r1, oh, ol);
}

// A quarter of x3_epilogue: word q of the hi and lo operands of n-tile j,
// from outputs 2q' and 2q' + 1 of half h = q / 2 (q' = q % 2) -- the same
// values and arithmetic as split_pack8's pair q.
template <int NT>
__device__ __forceinline__ void x3_epilogue_unit(const f32x4 (&acc)[2][NT], int j, int q,
                                                 bf16x8& oh, bf16x8& ol) {
  const int h = q >> 1, p = (q & 1) * 2;
  unsigned int hw, lw;
  split_pack2(relu_bits(acc[h][j][p]), relu_bits(acc[h][j][p + 1]), hw, lw);
  u32x4 hv = __builtin_bit_cast(u32x4, oh), lv = __builtin_bit_cast(u32x4, ol);
  hv[q] = hw;
  lv[q] = lw;
  oh = __builtin_bit_cast(bf16x8, hv);
  ol = __builtin_bit_cast(bf16x8, lv);
}

// One split-BF16 layer, out = relu(W in + bias) split into hi / lo, over
// m-tile pairs with the weight pairs read two ahead (as mfma_relu_layer).
// Per output the accumulation order is k-tile 0, 1, ..., and within a
// k-tile Wl Xh, Wh Xl, Wh Xh.  A pair's epilogue runs between the next
// pair's MFMAs, spread evenly over its k-tile groups in quarters (4 NT
// quarters over 2 KT groups: one per group of a hidden layer, where each
// group issues 6 MFMAs), so its VALU work issues while they execute
// (spread, against whole n-tiles in the first NT groups: -2% per MLP
// launch).  The last pair has no later MFMAs in the layer: its
// accumulators go out in pout, and the next layer (kPendIn: pin) or the
// head finishes them into its input's last k-tile during its first KT - 1
// k-tile groups, before that k-tile is read (-1% per launch).  The wave
// raises its priority for the layer so that the SIMD's other wave, if it
// is between layers, does not take its issue slots.
template <int KT, int MT, int NT, bool kPendIn>
__device__ __forceinline__ void mfma3_relu_layer(const bf16x8* A, const float* B, int lane, int g,
                                                 bf16x8 (&xh)[KT][NT], bf16x8 (&xl)[KT][NT],
                                                 const f32x4 (&pin)[2][NT],
                                                 bf16x8 (&oh)[MT / 2][NT],
                                                 bf16x8 (&ol)[MT / 2][NT],
                                                 f32x4 (&pout)[2][NT]) {
  constexpr int kLead = 2, NF = MT * KT, NB = kLead + 1;
  static_assert(!kPendIn || KT > 1, "a pending input k-tile needs earlier k-tile groups");
  bf16x8 ah[NB], al[NB];
#pragma unroll
  for (int p = 0; p < kLead; ++p)
    if (p < NF) {
      ah[p] = A[(2 * p) * 64 + lane];
      al[p] = A[(2 * p + 1) * 64 + lane];
    }
  f32x4 bias_next = *reinterpret_cast<const f32x4*>(B + 4 * g);
  __builtin_amdgcn_s_setprio(1);
  f32x4 prev[2][NT];  // the previous pair's accumulators, epilogue pending
#pragma unroll
  for (int kt2 = 0; kt2 < MT / 2; ++kt2) {
    f32x4 acc[2][NT];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int mt = 2 * kt2 + half;
      const f32x4 bias4 = bias_next;
      if (mt + 1 < MT) bias_next = *reinterpret_cast<const f32x4*>(B + 16 * (mt + 1) + 4 * g);
#pragma unroll
      for (int kt = 0; kt < KT; ++kt) {
        const int f = mt * KT + kt;
        if (f + kLead < NF) {
          ah[(f + kLead) % NB] = A[(2 * (f + kLead)) * 64 + lane];
          al[(f + kLead) % NB] = A[(2 * (f + kLead) + 1) * 64 + lane];
        }
#pragma unroll
        for (int j = 0; j < NT; ++j)
          acc[half][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              al[f % NB], xh[kt][j], kt == 0 ? bias4 : acc[half][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < NT; ++j)
          acc[half][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[f % NB], xl[kt][j],
                                                                 acc[half][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < NT; ++j)
          acc[half][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[f % NB], xh[kt][j],
                                                                 acc[half][j], 0, 0, 0);
        const int slot = half * KT + kt;
        if (kt2 > 0) {
          constexpr int U = 4 * NT, S = 2 * KT;
#pragma unroll
          for (int u = slot * U / S; u < (slot + 1) * U / S; ++u)
            x3_epilogue_unit<NT>(prev, u >> 2, u & 3, oh[kt2 - 1][u >> 2], ol[kt2 - 1][u >> 2]);
        } else if (kPendIn && half == 0 && kt + 1 < KT) {
          constexpr int U = 4 * NT, S = KT > 1 ? KT - 1 : 1;
#pragma unroll
          for (int u = kt * U / S; u < (kt + 1) * U / S; ++u)
            x3_epilogue_unit<NT>(pin, u >> 2, u & 3, xh[KT - 1][u >> 2], xl[KT - 1][u >> 2]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int j = 0; j < NT; ++j) prev[h][j] = acc[h][j];
  }
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int j = 0; j < NT; ++j) pout[h][j] = prev[h][j];
  __builtin_amdgcn_s_setprio(0);
}

// MLP kernel of the split-BF16 path: operand tiles (hi at emb, lo at
// emb + lo_off) from embed_bf16_kernel<D, H, L, 2, true>.
template <int D, int H, int L>
__global__ __launch_bounds__(64 * kX3Waves, 1) void mlp_bf16x3_kernel(
    const bf16x8* __restrict__ emb, int64_t lo_off, const int32_t* m_dev, int64_t m_max,
    const bf16x8* __restrict__ wfrag, const float* __restrict__ bias, float4* __restrict__ out,
    const int32_t* __restrict__ list) {
  using Dm = Dims3<D, H, L>;
  constexpr int NT = Dm::NT, KT0 = Dm::KT0, KTH = Dm::KTH, MTH = Dm::MTH;
  static_assert(MTH == 2 * KTH, "a layer's output m-tile pairs are the next layer's k-tiles");
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int64_t M = min(int64_t(*m_dev), m_max);
  const int64_t tiles = (M + Dm::SAMPLES - 1) / Dm::SAMPLES;
  // a block without tiles (the device count below this launch's bound)
  // leaves before staging the weights
  if (int64_t(blockIdx.x) * kX3Waves >= tiles) return;
  bf16x8* lw = reinterpret_cast<bf16x8*>(smem_raw);
  for (int i = threadIdx.x; i < Dm::FRAGS * 2 * 64; i += 64 * kX3Waves) lw[i] = wfrag[i];
  float* lb = reinterpret_cast<float*>(smem_raw + size_t(Dm::FRAGS) * 2048);
  for (int i = threadIdx.x; i < Dm::BIASES; i += 64 * kX3Waves) lb[i] = bias[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const int64_t stride = int64_t(gridDim.x) * kX3Waves;
  int64_t t = int64_t(blockIdx.x) * kX3Waves + (threadIdx.x >> 6);
  if (t >= tiles) return;
  bf16x8 xh[KT0][NT], xl[KT0][NT];
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int kt = 0; kt < KT0; ++kt) {
      xh[kt][j] = emb[((t * NT + j) * KT0 + kt) * 64 + lane];
      xl[kt][j] = emb[lo_off + ((t * NT + j) * KT0 + kt) * 64 + lane];
    }
  for (; t < tiles; t += stride) {
    const int64_t tn = t + stride;
    const int64_t base = t * Dm::SAMPLES;
    StoreSink sink{out, list};
    sink.begin_tile(base, M, Dm::SAMPLES);
    bf16x8 bh[KTH][NT], bl[KTH][NT];
    f32x4 pa[2][NT], pb[2][NT];
    mfma3_relu_layer<KT0, MTH, NT, false>(lw, lb, lane, g, xh, xl, pb, bh, bl, pa);
    // the layer-0 operands are dead: the next tile's load into them
    if (tn < tiles) {
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int kt = 0; kt < KT0; ++kt) {
          xh[kt][j] = emb[((tn * NT + j) * KT0 + kt) * 64 + lane];
          xl[kt][j] = emb[lo_off + ((tn * NT + j) * KT0 + kt) * 64 + lane];
        }
    }
    // hidden layers 1..L-1, two per trip so the activations alternate
    // between two register sets without copies
    bf16x8 nh[KTH][NT], nl[KTH][NT];
#pragma unroll 1
    for (int l = 1; l + 1 < L; l += 2) {
      mfma3_relu_layer<KTH, MTH, NT, true>(lw + 2 * (Dm::F0 + (l - 1) * Dm::FH) * 64, lb + H * l,
                                           lane, g, bh, bl, pa, nh, nl, pb);
      mfma3_relu_layer<KTH, MTH, NT, true>(lw + 2 * (Dm::F0 + l * Dm::FH) * 64, lb + H * (l + 1),
                                           lane, g, nh, nl, pb, bh, bl, pa);
    }
    if ((L - 1) % 2) {
      mfma3_relu_layer<KTH, MTH, NT, true>(lw + 2 * (Dm::F0 + (L - 2) * Dm::FH) * 64,
                                           lb + H * (L - 1), lane, g, bh, bl, pa, nh, nl, pb);
#pragma unroll
      for (int kt = 0; kt < KTH; ++kt)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          bh[kt][j] = nh[kt][j];
          bl[kt][j] = nl[kt][j];
        }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < NT; ++j) pa[h][j] = pb[h][j];
    }
    // head: 4 outputs in rows 0..3 of one m-tile
    const bf16x8* A = lw + 2 * (Dm::F0 + (L - 1) * Dm::FH) * 64;
    const f32x4 bias4 = *reinterpret_cast<const f32x4*>(lb + H * L + 4 * g);
    f32x4 acc[NT];
#pragma unroll
    for (int kt = 0; kt < KTH; ++kt) {
      const bf16x8 ah = A[(2 * kt) * 64 + lane], al = A[(2 * kt + 1) * 64 + lane];
#pragma unroll
      for (int j = 0; j < NT; ++j)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[kt][j], kt == 0 ? bias4 : acc[j],
                                                         0, 0, 0);
#pragma unroll
      for (int j = 0; j < NT; ++j)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[kt][j], acc[j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < NT; ++j)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[kt][j], acc[j], 0, 0, 0);
      // the last hidden layer's last pair, into k-tile KTH - 1 before it is read
      if (kt + 1 < KTH) {
        constexpr int U = 4 * NT, S = KTH > 1 ? KTH - 1 : 1;
#pragma unroll
        for (int u = kt * U / S; u < (kt + 1) * U / S; ++u)
          x3_epilogue_unit<NT>(pa, u >> 2, u & 3, bh[KTH - 1][u >> 2], bl[KTH - 1][u >> 2]);
      }
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int64_t s = base + 16 * j + col;
      sink.put(j, s, g == 0 && s < M, acc[j][0], acc[j][1], acc[j][2], acc[j][3]);
    }
  }
}

// Fused BF16 field of the render path: the split path's run-staged gather
// and the MLP in one kernel, so the layer-0 operands never go through HBM.
// A block of 8 waves shares one LDS copy of the weights.  Each wave takes 64
// consecutive slots (one MLP tile), stages the corner rows of their runs of
// equal voxel ids in its own LDS area, and each lane computes its own
// operand-B elements -- features {32kt + 4g + 0..3, 32kt + 16 + 4g + 0..3}
// of samples 16j + col -- with the gather kernel's corner-order chain of
// fused multiply-adds, so operands and field equal the split path's bit for
// bit.
constexpr int kFusedWaves = 8;
constexpr int kFusedRuns = 7;   // voxels staged per round

template <int D>
constexpr size_t fused_wave_bytes() {
  // rows (padded stride), per-sample weights, per-sample run, run voxels,
  // the corner keys of a staging round
  return size_t(kFusedRuns * (2 * D + 1)) * 16 + 64 * 8 * 4 + 64 * 4 + 64 * 4 + 64 * 4;
}

template <int D, int H, int L>
constexpr size_t fused_lds_bytes() {
  using Dm = Dims<D, H, L>;
  return ((size_t(Dm::FRAGS) * 1024 + size_t(Dm::BIASES) * 4 + 15) & ~size_t(15)) +
         kFusedWaves * fused_wave_bytes<D>();
}

// The corner weights of a sample at depth z on ray (o, d) in voxel c (the
// gather kernel's arithmetic: x = o + d z, p = (x - c) / vs + 0.5).
__device__ __forceinline__ void fused_weights(const FieldScene& sc, const float o[3],
                                              const float d[3], const float c[3], float z,
                                              float w[8]) {
  float p[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float x = o[a] + d[a] * z;
    const float e = x - c[a];
    p[a] = (sc.voxel_pow2 ? e * sc.inv_voxel : e / sc.voxel_size) + 0.5f;
  }
  const float q[3] = {1.0f - p[0], 1.0f - p[1], 1.0f - p[2]};
  const float xy[4] = {q[0] * q[1], q[0] * p[1], p[0] * q[1], p[0] * p[1]};
#pragma unroll
  for (int c = 0; c < 8; ++c) w[c] = xy[c >> 1] * ((c & 1) ? p[2] : q[2]);
}

// The next tile's gather, in stages around the current tile's MLP so that
// each stage's loads land while matrix work runs and nothing it holds stays
// live across the hidden layers: the tile's runs, ray / centre loads and the
// corner keys of its first staging round before the MLP (from sample
// records loaded one tile earlier), its corner rows straight into LDS (a
// run's rows are one 64-lane load when D = 32, so they take no registers)
// and its weights into LDS after layer 0, and the records of the tile after
// it before the head.
template <int D, int H, int L>
struct FusedPrefetch {
  static constexpr int ROW4 = 2 * D, RUN_STRIDE = ROW4 + 1, RUNS = kFusedRuns, Q = D / 4;
  static constexpr int SLOTS = (RUNS * ROW4 + 63) / 64;
  static_assert(ROW4 == 64, "one run per 64-lane load");
  static_assert(RUNS <= 8, "a round's corner keys in one 64-lane load");
  FieldScene sc;
  FieldRays rays;
  const SampleRec* samples;
  const int32_t* list;  // stream sample s is the record at slot list[s]; null: s
  float4* rows;
  float* wts;
  int32_t* rix;
  int32_t* run_vox;
  int32_t* keys;  // corner c of the round's run j at keys[8 j + c]
  int* in_mlp;  // MLPs of this wave's SIMD pair in flight (LDS)
  int64_t M, tiles;
  int lane;
  int64_t rtile = -1;  // the tile whose records are in rec
  int64_t tile = -1;   // the tile being staged, -1 for none
  SampleRec rec{-1, 0.0f, 0.0f, 0};
  int32_t v = -1;
  int U = 1;
  float z = 0.0f;
  float ro[3] = {0.f, 0.f, 0.f}, rd[3] = {0.f, 0.f, 0.f}, ctr[3] = {0.f, 0.f, 0.f};

  __device__ __forceinline__ void issue_rec(int64_t t) {
    rtile = t;
    rec = SampleRec{-1, 0.0f, 0.0f, 0};
    const int64_t s = t * 64 + lane;
    if (t < tiles && s < M) rec = samples[list ? list[s] : s];
  }
  // runs into LDS; ray / centre loads issued, the first round's corner keys
  // straight into LDS
  __device__ __forceinline__ void stage_a() {
    tile = rtile < tiles ? rtile : -1;
    if (tile < 0) return;
    v = rec.idx;
    z = rec.depth;
    if (v >= 0) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        ro[a] = rays.start[int64_t(rec.ray) * 3 + a];
        rd[a] = rays.dir[int64_t(rec.ray) * 3 + a];
        ctr[a] = sc.centers[int64_t(v) * 3 + a];
      }
    }
    const int32_t vprev = __shfl_up(v, 1, 64);
    const bool head = lane == 0 || v != vprev;
    const uint64_t hmask = __ballot(head);
    const int r = __popcll(hmask & ((2ull << lane) - 1)) - 1;
    U = __popcll(hmask);
    if (head) run_vox[r] = v;
    rix[lane] = v >= 0 ? r : -1;
    wave_sync();
    load_keys(0);
  }
  // lane l: corner l % 8 of run u0 + l / 8 (padding runs read voxel 0's)
  __device__ __forceinline__ void load_keys(int u0) {
    const int jr = lane >> 3;
    const int32_t vr = run_vox[(u0 + jr) & 63];
    const int32_t vu = (jr < RUNS && u0 + jr < U && vr >= 0) ? vr : 0;
    __builtin_amdgcn_global_load_lds(
        (const __attribute__((address_space(1))) void*)(sc.feats + int64_t(vu) * 8 + (lane & 7)),
        (__attribute__((address_space(3))) void*)(keys), 4, 0, 0);
  }
  // the first round's rows straight into LDS (lane e of load tt: element e
  // of run tt at rows + tt * RUN_STRIDE + e); the weights into LDS
  __device__ __forceinline__ void stage_b() {
    if (tile < 0) return;
    __builtin_amdgcn_s_waitcnt(0);  // the keys have landed
    wave_sync();
#pragma unroll
    for (int tt = 0; tt < SLOTS; ++tt) {
      const int32_t k = keys[tt * 8 + lane / Q];
      const float4* src =
          reinterpret_cast<const float4*>(sc.values + int64_t(k) * D) + lane % Q;
      __builtin_amdgcn_global_load_lds(
          (const __attribute__((address_space(1))) void*)(src),
          (__attribute__((address_space(3))) void*)(rows + tt * RUN_STRIDE), 16, 0, 0);
    }
    // a tile of more runs than a round holds: the second round's corner keys
    // into the same buffer (the first round's are read: the row loads above
    // take their addresses from them), so that round waits on its row loads
    // only
    if (U > RUNS) load_keys(RUNS);
    float w[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = 0.0f;
    if (v >= 0) fused_weights(sc, ro, rd, ctr, z, w);
#pragma unroll
    for (int c = 0; c < 8; ++c) wts[lane * 8 + c] = w[c];
  }
  // the direct row loads have landed
  __device__ __forceinline__ void rows_ready() { __builtin_amdgcn_s_waitcnt(0); }

  int64_t next_rec = -1;  // the records to load before the head
  __device__ __forceinline__ void layer_done(int l) {
    if (l == 0) stage_b();
    // the partner's MLP has ended: this one is alone on the matrix pipe
    if (__builtin_amdgcn_readfirstlane(*(volatile int*)in_mlp) == 1) __builtin_amdgcn_s_setprio(2);
  }
  __device__ __forceinline__ void before_head() {
    if (next_rec >= 0) issue_rec(next_rec);
  }
};

template <int D, int H, int L>
__global__ __launch_bounds__(64 * kFusedWaves, 1) void field_fused_kernel(
    FieldScene sc, FieldRays rays, const SampleRec* __restrict__ samples,
    const int32_t* __restrict__ list, const int32_t* m_dev, int64_t m_max,
    const bf16x8* __restrict__ wfrag, const float* __restrict__ bias, float4* __restrict__ out) {
  using Dm = Dims<D, H, L>;
  using PF = FusedPrefetch<D, H, L>;
  static_assert(Dm::SAMPLES == 64, "one MLP tile per wave");
  static_assert(weights_in_lds<D, H, L>(), "weights staged in LDS");
  constexpr int NT = Dm::NT, KT0 = Dm::KT0, Q = D / 4;
  constexpr int RUN_STRIDE = PF::RUN_STRIDE, RUNS = PF::RUNS;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  bf16x8* lw = reinterpret_cast<bf16x8*>(smem_raw);
  float* lb = reinterpret_cast<float*>(smem_raw + size_t(Dm::FRAGS) * 1024);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned char* wb = smem_raw + ((size_t(Dm::FRAGS) * 1024 + size_t(Dm::BIASES) * 4 + 15) &
                                  ~size_t(15)) + size_t(wave) * fused_wave_bytes<D>();
  float4* rows = reinterpret_cast<float4*>(wb);
  float* wts = reinterpret_cast<float*>(wb + size_t(RUNS * RUN_STRIDE) * 16);
  int32_t* rix = reinterpret_cast<int32_t*>(wts + 64 * 8);
  int32_t* run_vox = rix + 64;
  int32_t* keys = run_vox + 64;
  // MLPs in flight per SIMD pair (waves w and w + 4): a wave that starts
  // its MLP while its partner's runs takes the lower priority, so the
  // earlier MLP keeps the matrix pipe and that wave returns to its gather
  // sooner (a hint only: nothing waits on it)
  __shared__ int mlp_count[kFusedWaves / 2];
  for (int i = threadIdx.x; i < Dm::FRAGS * 64; i += 64 * kFusedWaves) lw[i] = wfrag[i];
  for (int i = threadIdx.x; i < Dm::BIASES; i += 64 * kFusedWaves) lb[i] = bias[i];
  if (threadIdx.x < kFusedWaves / 2) mlp_count[threadIdx.x] = 0;
  __syncthreads();
  const int64_t M = min(int64_t(*m_dev), m_max);
  const int64_t tiles = (M + 63) / 64;
  const int64_t stride = int64_t(gridDim.x) * kFusedWaves;
  int64_t t = int64_t(blockIdx.x) * kFusedWaves + wave;
  if (t >= tiles) return;  // the whole wave
  int* in_mlp = mlp_count + wave % (kFusedWaves / 2);
  PF pf{sc, rays, samples, list, rows, wts, rix, run_vox, keys, in_mlp, M, tiles, lane};
  // the first tile's gather stages back to back, the second tile's records
  pf.issue_rec(t);
  pf.stage_a();
  pf.stage_b();
  pf.issue_rec(t + stride);
  // a wave in its MLP issues first: the other wave's gather (VALU and LDS)
  // fills the matrix pipe's gaps (the MLP priorities below also keep the two
  // waves of a SIMD out of step, so they start together)
  for (; t < tiles; t += stride) {
    // the lane index made opaque once per tile: addresses derived from it are
    // recomputed in the loop (a few ALU ops) instead of being hoisted out of
    // it and spilled
    int ln = lane;
    asm volatile("" : "+v"(ln));
    pf.lane = ln;
    const int g = ln >> 4, col = ln & 15;
    const int U = pf.U;
    StoreSink sink{out, list};
    sink.begin_tile(t * 64, M, 64);
    pf.rows_ready();
    wave_sync();
    float f[NT][KT0][8];
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int kt = 0; kt < KT0; ++kt)
#pragma unroll
        for (int i = 0; i < 8; ++i) f[j][kt][i] = 0.0f;
    for (int u0 = 0; u0 < U; u0 += RUNS) {
      const int nr = min(RUNS, U - u0);
      if (u0 > 0) {
        // further staging rounds of this tile (more runs than one round
        // holds), rows straight into LDS as in the first; the second round's
        // keys came with the first round's rows
        __builtin_amdgcn_s_waitcnt(0);  // the last round's row reads are done
        if (u0 > RUNS) {
          pf.load_keys(u0);
          __builtin_amdgcn_s_waitcnt(0);
        }
        wave_sync();
#pragma unroll
        for (int tt = 0; tt < RUNS; ++tt) {
          const int32_t k = keys[tt * 8 + ln / Q];
          __builtin_amdgcn_global_load_lds(
              (const __attribute__((address_space(1))) void*)(
                  reinterpret_cast<const float4*>(sc.values + int64_t(k) * D) + ln % Q),
              (__attribute__((address_space(3))) void*)(rows + tt * RUN_STRIDE), 16, 0, 0);
        }
        __builtin_amdgcn_s_waitcnt(0);
        wave_sync();
      }
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const int qs = 16 * j + col;
        const int rq = rix[qs];
        if (rq >= u0 && rq < u0 + nr) {
          float wq[8];
#pragma unroll
          for (int c = 0; c < 8; ++c) wq[c] = wts[qs * 8 + c];
          const float4* rw = rows + (rq - u0) * RUN_STRIDE;
#pragma unroll
          for (int kt = 0; kt < KT0; ++kt)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int quad = 8 * kt + 4 * h + g;
              if (quad < Q) {
                float4 e[8];
#pragma unroll
                for (int c = 0; c < 8; ++c) e[c] = rw[c * Q + quad];
                float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                  a0 = __fmaf_rn(wq[c], e[c].x, a0);
                  a1 = __fmaf_rn(wq[c], e[c].y, a1);
                  a2 = __fmaf_rn(wq[c], e[c].z, a2);
                  a3 = __fmaf_rn(wq[c], e[c].w, a3);
                }
                f[j][kt][4 * h + 0] = a0;
                f[j][kt][4 * h + 1] = a1;
                f[j][kt][4 * h + 2] = a2;
                f[j][kt][4 * h + 3] = a3;
              }
            }
        }
      }
      wave_sync();
    }
    OperandTile<D, H, L> b0;
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int kt = 0; kt < KT0; ++kt) {
        const float* ff = f[j][kt];
        const bf16x2 p0 = __builtin_convertvector(f32x2{ff[0], ff[1]}, bf16x2);
        const bf16x2 p1 = __builtin_convertvector(f32x2{ff[2], ff[3]}, bf16x2);
        const bf16x2 p2 = __builtin_convertvector(f32x2{ff[4], ff[5]}, bf16x2);
        const bf16x2 p3 = __builtin_convertvector(f32x2{ff[6], ff[7]}, bf16x2);
        const u32x4 u = {__builtin_bit_cast(unsigned int, p0), __builtin_bit_cast(unsigned int, p1),
                         __builtin_bit_cast(unsigned int, p2), __builtin_bit_cast(unsigned int, p3)};
        b0[kt][j] = __builtin_bit_cast(bf16x8, u);
      }
    // the next tile's gather runs in stages around this tile's MLP
    pf.stage_a();
    pf.next_rec = t + 2 * stride;
    int before = 0;
    if (lane == 0) before = atomicAdd(in_mlp, 1);
    if (__builtin_amdgcn_readfirstlane(before) == 0)
      __builtin_amdgcn_s_setprio(2);
    else
      __builtin_amdgcn_s_setprio(1);
    mlp_tile_hooked<D, H, L>(b0, lw, lb, ln, g, col, t * 64, M, sink, nullptr, pf);
    if (lane == 0) atomicSub(in_mlp, 1);
    __builtin_amdgcn_s_setprio(0);
  }
}

// FP32 variant: same structure on v_mfma_f32_16x16x4_f32, weights read from
// global memory (fragments of 64 floats).
template <int D, int H, int L, class Src, class Sink>
__global__ __launch_bounds__(kFieldBlock) void field_f32_kernel(
    Src src, const int32_t* m_dev, int64_t m_max, const float* __restrict__ wfrag,
    const float* __restrict__ bias, Sink sink) {
  constexpr int NT = 2;
  constexpr int KT0 = 4 * ((D + 15) / 16);  // k-tiles (t, i) of the input
  constexpr int KTH = H / 4;
  constexpr int MTH = H / 16;
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const int64_t M = m_dev ? min(int64_t(*m_dev), m_max) : m_max;
  const int64_t tiles = (M + 16 * NT - 1) / (16 * NT);
  for (int64_t t = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); t < tiles;
       t += int64_t(gridDim.x) * 4) {
    const int64_t base = t * 16 * NT;
    sink.begin_tile(base, M, 16 * NT);
    float x0[KT0][NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int64_t s = base + 16 * j + col;
      float w[8];
      int32_t key[8];
      const bool valid = s < M && src.template weights<false>(s, w, key);
#pragma unroll
      for (int kt32 = 0; kt32 < (KT0 + 7) / 8; ++kt32) {
        float f[8];
        if (valid) {
          gather8<D>(src.sc, w, key, kt32, g, f);
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) f[i] = 0.0f;
        }
        // features 32*kt32 + 4g + i -> k-tile (t = 2*kt32, i); 32*kt32 + 16 +
        // 4g + i -> k-tile (t = 2*kt32 + 1, i)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if ((2 * kt32) * 4 + i < KT0) x0[(2 * kt32) * 4 + i][j] = f[i];
          if ((2 * kt32 + 1) * 4 + i < KT0) x0[(2 * kt32 + 1) * 4 + i][j] = f[4 + i];
        }
      }
    }
    float xh[KTH][NT];
    {
      const float* A = wfrag;
#pragma unroll
      for (int mt = 0; mt < MTH; ++mt) {
        f32x4 acc[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < KT0; ++kt) {
          const float a = A[(mt * KT0 + kt) * 64 + lane];
#pragma unroll
          for (int j = 0; j < NT; ++j)
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, x0[kt][j], acc[j], 0, 0, 0);
        }
        const float4 bb = *reinterpret_cast<const float4*>(bias + 16 * mt + 4 * g);
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          xh[4 * mt + 0][j] = fmaxf(acc[j][0] + bb.x, 0.0f);
          xh[4 * mt + 1][j] = fmaxf(acc[j][1] + bb.y, 0.0f);
          xh[4 * mt + 2][j] = fmaxf(acc[j][2] + bb.z, 0.0f);
          xh[4 * mt + 3][j] = fmaxf(acc[j][3] + bb.w, 0.0f);
        }
      }
    }
    const float* A = wfrag + MTH * KT0 * 64;
#pragma unroll 1
    for (int l = 1; l < L; ++l) {
      float nx[KTH][NT];
#pragma unroll
      for (int mt = 0; mt < MTH; ++mt) {
        f32x4 acc[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < KTH; ++kt) {
          const float a = A[(mt * KTH + kt) * 64 + lane];
#pragma unroll
          for (int j = 0; j < NT; ++j)
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xh[kt][j], acc[j], 0, 0, 0);
        }
        const float4 bb = *reinterpret_cast<const float4*>(bias + H * l + 16 * mt + 4 * g);
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          nx[4 * mt + 0][j] = fmaxf(acc[j][0] + bb.x, 0.0f);
          nx[4 * mt + 1][j] = fmaxf(acc[j][1] + bb.y, 0.0f);
          nx[4 * mt + 2][j] = fmaxf(acc[j][2] + bb.z, 0.0f);
          nx[4 * mt + 3][j] = fmaxf(acc[j][3] + bb.w, 0.0f);
        }
      }
#pragma unroll
      for (int kt = 0; kt < KTH; ++kt)
#pragma unroll
        for (int j = 0; j < NT; ++j) xh[kt][j] = nx[kt][j];
      A += MTH * KTH * 64;
    }
    {
      f32x4 acc[NT];
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < KTH; ++kt) {
        const float a = A[kt * 64 + lane];
#pragma unroll
        for (int j = 0; j < NT; ++j)
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xh[kt][j], acc[j], 0, 0, 0);
      }
      const float* B = bias + H * L;
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const int64_t s = base + 16 * j + col;
        sink.put(j, s, g == 0 && s < M, acc[j][0] + B[0], acc[j][1] + B[1], acc[j][2] + B[2],
                 acc[j][3] + B[3]);
      }
      sink.end_tile(base);
    }
  }
}

// Op-level trilinear gather: one thread per (sample, feature); consecutive
// threads cover consecutive features, so embedding reads and output writes
// are coalesced.
__global__ __launch_bounds__(256) void gather_kernel(const int32_t* __restrict__ s_idx,
                                                     const float* __restrict__ xyz, int64_t M,
                                                     int D, FieldScene sc,
                                                     float* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= M * D) return;
  const int64_t s = i / D;
  const int f = int(i - s * D);
  const int32_t v = s_idx[s];
  if (v < 0) {  // unused sample slot
    out[i] = 0.0f;
    return;
  }
  float p[3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
    p[a] = (xyz[s * 3 + a] - sc.centers[int64_t(v) * 3 + a]) / sc.voxel_size + 0.5f;
  const float q[3] = {1.0f - p[0], 1.0f - p[1], 1.0f - p[2]};
  float acc = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float w = (((c & 4) ? p[0] : q[0]) * ((c & 2) ? p[1] : q[1])) * ((c & 1) ? p[2] : q[2]);
    acc = acc + w * sc.values[int64_t(sc.feats[int64_t(v) * 8 + c]) * D + f];
  }
  out[i] = acc;
}

// True when v is a (normal) power of two, so that x / v == x * (1 / v)
// exactly for every x.
int is_pow2(float v) {
  int e;
  const float m = std::frexp(v, &e);
  return (m == 0.5f && v > 0.0f && std::isnormal(v) && std::isnormal(1.0f / v)) ? 1 : 0;
}

// Operand tiles of the split BF16 path: the gather kernel writes every slot
// of its 64-slot tiles, KT0 bf16x8 per lane and 16-sample column (64 bytes
// per sample and k-tile).
template <int D, int H, int L>
int64_t bf16_workspace_bytes(int64_t m_max) {
  using Dm = Dims<D, H, L>;
  return ceil_div(m_max, 64) * 64 * Dm::KT0 * 64;
}

// Blocks of `kernel` (`threads` per block, `lds` bytes of dynamic LDS) that
// the current device holds at once: its occupancy per CU times the CU count.
template <class K>
int64_t resident_blocks(K kernel, int threads = 256, size_t lds = 0) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return 256;
  return int64_t(cus) * per_cu;
}

template <int D, int H, int L>
int launch_embed_bf16(FieldScene sc, FieldRays rays, const SampleRec* samples,
                      const int32_t* list, const int32_t* m_dev, int64_t m_max, void* workspace,
                      hipStream_t stream) {
  // persistent waves: one resident round of blocks, each wave looping over
  // 64-slot tiles
  static const int64_t resident = resident_blocks(embed_bf16_kernel<D, H, L>);
  const int64_t tiles64 = ceil_div(m_max, 64);
  const int64_t blocks = std::max<int64_t>(1, std::min(ceil_div(tiles64, 4), resident));
  hipLaunchKernelGGL((embed_bf16_kernel<D, H, L>), dim3(unsigned(blocks)), dim3(256), 0, stream,
                     sc, rays, samples, list, m_dev, m_max, static_cast<bf16x8*>(workspace),
                     int64_t(0));
  return check_launch("field_embed_bf16");
}

template <int D, int H, int L>
int launch_mlp_bf16(const int32_t* m_dev, int64_t m_max, const void* wfrag, const float* bias,
                    float* out, const int32_t* list, const void* workspace, hipStream_t stream) {
  using Dm = Dims<D, H, L>;
  const int64_t tiles = ceil_div(m_max, Dm::SAMPLES);
  const size_t lds = (weights_in_lds<D, H, L>() ? size_t(Dm::FRAGS) * 1024 : 0) +
                     size_t(Dm::BIASES) * sizeof(float);
  const int64_t blocks = std::min<int64_t>(ceil_div(tiles, 4), 2048);
  hipLaunchKernelGGL((mlp_bf16_kernel<D, H, L>), dim3(unsigned(blocks)), dim3(kFieldBlock), lds,
                     stream, static_cast<const bf16x8*>(workspace), m_dev, m_max,
                     static_cast<const bf16x8*>(wfrag), bias, reinterpret_cast<float4*>(out), list);
  return check_launch("field_mlp_bf16");
}

template <int D, int H, int L>
int launch_bf16_split(FieldScene sc, FieldRays rays, const SampleRec* samples,
                      const int32_t* list, const int32_t* m_dev, int64_t m_max, const void* wfrag,
                      const float* bias, float* out, void* workspace, hipStream_t stream) {
  const int rc =
      launch_embed_bf16<D, H, L>(sc, rays, samples, list, m_dev, m_max, workspace, stream);
  if (rc) return rc;
  return launch_mlp_bf16<D, H, L>(m_dev, m_max, wfrag, bias, out, list, workspace, stream);
}

// Split-BF16 path: hi and lo operand tiles of 32-sample MLP tiles, each
// region covering every slot of the gather kernel's 64-slot tiles.
template <int D, int H, int L>
int64_t x3_lo_offset(int64_t m_max) {  // in bf16x8 elements
  return ceil_div(m_max, 64) * 64 * Dims<D, H, L>::KT0 * 4;
}

template <int D, int H, int L>
int64_t bf16x3_workspace_bytes(int64_t m_max) {
  return 2 * x3_lo_offset<D, H, L>(m_max) * 16;
}

// Tiles each wave of the split-BF16 kernels takes before its block retires.
// The grid holds as many blocks as that needs (not one resident round), so a
// kernel of another stream -- a concurrent request's -- gets compute units
// as blocks retire instead of waiting for the whole launch; a block's
// weight staging is amortised over its waves' 32 tiles.
constexpr int64_t kX3TilesPerWave = 32;

template <int D, int H, int L>
int launch_embed_bf16x3(FieldScene sc, FieldRays rays, const SampleRec* samples,
                        const int32_t* list, const int32_t* m_dev, int64_t m_max,
                        void* workspace, hipStream_t stream) {
  if constexpr (x3_supported<D, H, L>()) {
    static const int64_t resident = resident_blocks(embed_bf16_kernel<D, H, L, 2, true>);
    const int64_t blocks =
        std::max<int64_t>(std::min<int64_t>(resident, ceil_div(ceil_div(m_max, 64), 4)),
                          ceil_div(ceil_div(m_max, 64), 4 * kX3TilesPerWave));
    hipLaunchKernelGGL((embed_bf16_kernel<D, H, L, 2, true>), dim3(unsigned(blocks)), dim3(256),
                       0, stream, sc, rays, samples, list, m_dev, m_max,
                       static_cast<bf16x8*>(workspace), x3_lo_offset<D, H, L>(m_max));
    return check_launch("field_embed_bf16x3");
  } else {
    return set_error(NSVF_EINVAL, "field bf16x3: shape D=%d H=%d L=%d not supported", D, H, L);
  }
}

template <int D, int H, int L>
int launch_mlp_bf16x3(const int32_t* m_dev, int64_t m_max, const void* wfrag, const float* bias,
                      float* out, const int32_t* list, const void* workspace,
                      hipStream_t stream) {
  if constexpr (x3_supported<D, H, L>()) {
    using Dm = Dims3<D, H, L>;
    static const int64_t resident =
        resident_blocks(mlp_bf16x3_kernel<D, H, L>, 64 * kX3Waves, Dm::LDS_BYTES);
    const int64_t tiles = ceil_div(m_max, Dm::SAMPLES);
    const int64_t blocks =
        std::max<int64_t>(std::min<int64_t>(resident, ceil_div(tiles, kX3Waves)),
                          ceil_div(tiles, kX3Waves * kX3TilesPerWave));
    hipLaunchKernelGGL((mlp_bf16x3_kernel<D, H, L>), dim3(unsigned(blocks)), dim3(64 * kX3Waves),
                       Dm::LDS_BYTES, stream, static_cast<const bf16x8*>(workspace),
                       x3_lo_offset<D, H, L>(m_max), m_dev, m_max,
                       static_cast<const bf16x8*>(wfrag), bias, reinterpret_cast<float4*>(out),
                       list);
    return check_launch("field_mlp_bf16x3");
  } else {
    return set_error(NSVF_EINVAL, "field bf16x3: shape D=%d H=%d L=%d not supported", D, H, L);
  }
}

// The split-BF16 MLP's operand tiles from features computed elsewhere
// (training's gather: feats [M, D] row-major float32), laid out and split as
// embed_bf16_kernel<D, H, L, 2, true> writes gathered ones: a thread per
// sample, k-tile and lane group g writes features {32kt + 4g + 0..3,
// 32kt + 16 + 4g + 0..3} as one hi and one lo bf16x8.  Samples up to the
// next multiple of 64 (the tiles the MLP kernel may read) are written, the
// ones past M as zeros.
template <int D, int H, int L>
__global__ __launch_bounds__(256) void tiles_bf16x3_kernel(const float* __restrict__ feats,
                                                           int64_t M, int64_t total,
                                                           bf16x8* __restrict__ emb,
                                                           int64_t lo_off) {
  using Dm = Dims3<D, H, L>;
  constexpr int KT0 = Dm::KT0, NT = Dm::NT, SAMPLES = Dm::SAMPLES;
  const int64_t tid = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int g = int(tid & 3);
  const int64_t q = tid >> 2;
  const int kt = int(q % KT0);
  const int64_t s = q / KT0;
  if (s >= total) return;
  float f[8];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int feat = 32 * kt + 16 * h + 4 * g + i;
      f[4 * h + i] = (s < M && feat < D) ? feats[s * D + feat] : 0.0f;
    }
  bf16x8 hi, lo;
  split_pack8(f32x4{f[0], f[1], f[2], f[3]}, f32x4{f[4], f[5], f[6], f[7]}, hi, lo);
  const int64_t t = s / SAMPLES;
  const int within = int(s - t * SAMPLES);
  const int64_t idx = ((t * NT + (within >> 4)) * KT0 + kt) * 64 + 16 * g + (within & 15);
  __builtin_nontemporal_store(hi, &emb[idx]);
  __builtin_nontemporal_store(lo, &emb[lo_off + idx]);
}

template <int D, int H, int L>
int launch_tiles_bf16x3(const float* feats, int64_t M, void* workspace, hipStream_t stream) {
  if constexpr (x3_supported<D, H, L>()) {
    const int64_t total = ceil_div(M, 64) * 64;
    const int64_t threads = total * Dims3<D, H, L>::KT0 * 4;
    hipLaunchKernelGGL((tiles_bf16x3_kernel<D, H, L>), dim3(unsigned(ceil_div(threads, 256))),
                       dim3(256), 0, stream, feats, M, total, static_cast<bf16x8*>(workspace),
                       x3_lo_offset<D, H, L>(M));
    return check_launch("field_tiles_bf16x3");
  } else {
    return set_error(NSVF_EINVAL, "field bf16x3: shape D=%d H=%d L=%d not supported", D, H, L);
  }
}

// The fused BF16 field over a sample stream: the run-staged gather + MLP
// kernel where the weights fit in LDS, else the per-lane gather kernel.
template <int D, int H, int L>
int launch_bf16_fused(FieldScene sc, FieldRays rays, const SampleRec* samples,
                      const int32_t* list, const int32_t* m_dev, int64_t m_max, const void* wfrag,
                      const float* bias, float* out, hipStream_t stream);

template <int D, int H, int L, class Src, class Sink>
int launch_bf16(const Src& src, const int32_t* m_dev, int64_t m_max, const void* wfrag,
                const float* bias, const Sink& sink, hipStream_t stream) {
  using Dm = Dims<D, H, L>;
  const size_t lds = (weights_in_lds<D, H, L>() ? size_t(Dm::FRAGS) * 1024 : 0) +
                     ((size_t(Dm::BIASES) * sizeof(float) + 15) & ~size_t(15)) +
                     size_t(4) * Dm::NT * Dm::KT0 * 64 * 16;
  const int64_t tiles = ceil_div(m_max, Dm::SAMPLES);
  const int64_t blocks = std::min<int64_t>(ceil_div(tiles, 4), 2048);
  hipLaunchKernelGGL((field_bf16_kernel<D, H, L, Src, Sink>), dim3(unsigned(blocks)),
                     dim3(kFieldBlock), lds, stream, src, m_dev, m_max,
                     static_cast<const bf16x8*>(wfrag), bias, sink);
  return check_launch("field_bf16");
}

template <int D, int H, int L, class Src, class Sink>
int launch_f32(const Src& src, const int32_t* m_dev, int64_t m_max, const void* wfrag,
               const float* bias, const Sink& sink, hipStream_t stream) {
  const int64_t tiles = ceil_div(m_max, 32);
  const int64_t blocks = std::min<int64_t>(ceil_div(tiles, 4), 2048);
  hipLaunchKernelGGL((field_f32_kernel<D, H, L, Src, Sink>), dim3(unsigned(blocks)),
                     dim3(kFieldBlock), 0, stream, src, m_dev, m_max,
                     static_cast<const float*>(wfrag), bias, sink);
  return check_launch("field_f32");
}

template <int D, int H, int L>
int launch_bf16_fused(FieldScene sc, FieldRays rays, const SampleRec* samples,
                      const int32_t* list, const int32_t* m_dev, int64_t m_max, const void* wfrag,
                      const float* bias, float* out, hipStream_t stream) {
  if constexpr (weights_in_lds<D, H, L>() && Dims<D, H, L>::SAMPLES == 64) {
    constexpr size_t lds = fused_lds_bytes<D, H, L>();
    static const int64_t resident =
        resident_blocks(field_fused_kernel<D, H, L>, 64 * kFusedWaves, lds);
    const int64_t tiles = ceil_div(m_max, 64);
    const int64_t blocks = std::max<int64_t>(1, std::min(ceil_div(tiles, kFusedWaves), resident));
    hipLaunchKernelGGL((field_fused_kernel<D, H, L>), dim3(unsigned(blocks)),
                       dim3(64 * kFusedWaves), lds, stream, sc, rays, samples, list, m_dev, m_max,
                       static_cast<const bf16x8*>(wfrag), bias, reinterpret_cast<float4*>(out));
    return check_launch("field_fused_bf16");
  } else {
    const StreamSource src{sc, rays, samples, list};
    const StoreSink sink{reinterpret_cast<float4*>(out), list};
    return launch_bf16<D, H, L>(src, m_dev, m_max, wfrag, bias, sink, stream);
  }
}

__global__ __launch_bounds__(256) void fill_zero_kernel(float* __restrict__ p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// Pruning scores of voxels [v0, v0 + n): max over each voxel's lattice of
// relu(sigma), into vmax[0, n).
template <int D, int H, int L>
int launch_prune(int precision, const FieldScene& sc, int64_t v0, int64_t n, int bits,
                 const void* wfrag, const float* bias, float* vmax, hipStream_t stream) {
  hipLaunchKernelGGL(fill_zero_kernel, dim3(unsigned(ceil_div(n, 256))), dim3(256), 0, stream,
                     vmax, n);
  const int64_t per = int64_t(bits) * bits * bits;
  const LatticeSource src{sc, v0, bits, sc.voxel_size * 0.5f};
  const MaxSigmaSink sink{vmax, per, 0.0f};
  if (precision == 0) return launch_bf16<D, H, L>(src, nullptr, n * per, wfrag, bias, sink, stream);
  return launch_f32<D, H, L>(src, nullptr, n * per, wfrag, bias, sink, stream);
}

}  // namespace
}  // namespace nsvf

namespace nsvf {

// Scratch of the allocation-free field entry point (nsvf_field_eval_static):
// the split paths' operand tiles, 64 MiB on every GPU the library runs on,
// held until the process ends.  Host code never uses the array's name as an
// address: each device's own copy is found once with hipGetSymbolAddress.
constexpr int64_t kStaticScratchBytes = int64_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxScratchDevices] = {};

// The current device's address of `scratch`, looked up on the first call
// on that device (with it current) and kept.
static int scratch_address(void** out) {
  int dev = 0;
  const hipError_t de = hipGetDevice(&dev);
  if (de != hipSuccess)
    return set_error(NSVF_ELAUNCH, "field_eval_static: hipGetDevice: %s", hipGetErrorString(de));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return set_error(NSVF_EINVAL, "field_eval_static: device %d beyond the first %d", dev,
                     kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(NSVF_ELAUNCH, "field_eval_static: hipGetSymbolAddress: %s",
                       hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return NSVF_OK;
}

}  // namespace nsvf

using namespace nsvf;

extern "C" {

// Supported (D, H, L) shapes of the fused field kernels.
int nsvf_field_supported(int D, int H, int L) {
  return (D == 32 && H == 128 && L == 3) || (D == 48 && H == 256 && L == 4) ||
         (D == 32 && H == 256 && L == 4);
}

int nsvf_field_fused_supported(int D, int H, int L) {
#define NSVF_FUSED_CASE(d, h, l) \
  if (D == d && H == h && L == l) \
    return (weights_in_lds<d, h, l>() && Dims<d, h, l>::SAMPLES == 64) ? 1 : 0;
  NSVF_FUSED_CASE(32, 128, 3)
  NSVF_FUSED_CASE(48, 256, 4)
  NSVF_FUSED_CASE(32, 256, 4)
#undef NSVF_FUSED_CASE
  return 0;
}

int nsvf_field_bf16x3_supported(int D, int H, int L) {
#define NSVF_X3_CASE(d, h, l) \
  if (D == d && H == h && L == l) return x3_supported<d, h, l>() ? 1 : 0;
  NSVF_X3_CASE(32, 128, 3)
  NSVF_X3_CASE(48, 256, 4)
  NSVF_X3_CASE(32, 256, 4)
#undef NSVF_X3_CASE
  return 0;
}

int64_t nsvf_field_workspace_bytes(int precision, int D, int H, int L, int64_t m_max) {
  if ((precision != 0 && precision != 2) || m_max <= 0) return 0;
#define NSVF_WS_CASE(d, h, l)                                                  \
  if (D == d && H == h && L == l)                                              \
    return precision == 2 ? (x3_supported<d, h, l>() ? bf16x3_workspace_bytes<d, h, l>(m_max) \
                                                     : 0)                      \
                          : bf16_workspace_bytes<d, h, l>(m_max);
  NSVF_WS_CASE(32, 128, 3)
  NSVF_WS_CASE(48, 256, 4)
  NSVF_WS_CASE(32, 256, 4)
#undef NSVF_WS_CASE
  return 0;
}

int nsvf_field_eval(int precision, int D, int H, int L, const float* centers,
                    const int32_t* feats, const float* values, float voxel_size,
                    const float* ray_start, const float* ray_dir, const void* samples,
                    const int32_t* list, const int32_t* m_dev, int64_t m_max, const void* wfrag,
                    const float* bias, float* out, void* workspace, int64_t workspace_bytes,
                    void* stream) {
  clear_error();
  if (m_max <= 0) return NSVF_OK;
  if (!centers || !feats || !values || !ray_start || !ray_dir || !samples || !m_dev || !wfrag ||
      !bias || !out)
    return set_error(NSVF_EINVAL, "field_eval: null pointer");
  FieldScene sc{centers, feats, values, voxel_size, 1.0f / voxel_size, is_pow2(voxel_size)};
  FieldRays rays{ray_start, ray_dir};
  hipStream_t st = as_stream(stream);
  const StreamSource src{sc, rays, static_cast<const SampleRec*>(samples), list};
  const StoreSink sink{reinterpret_cast<float4*>(out), list};
  if (precision < 0 || precision > 2)
    return set_error(NSVF_EINVAL, "field_eval: precision %d (0 bf16, 1 fp32, 2 bf16x3)",
                     precision);
  if (precision == 2) {
    if (!nsvf_field_bf16x3_supported(D, H, L))
      return set_error(NSVF_EINVAL, "field_eval: bf16x3 does not support D=%d H=%d L=%d", D, H,
                       L);
    if (!workspace || workspace_bytes < nsvf_field_workspace_bytes(2, D, H, L, m_max))
      return set_error(NSVF_EINVAL, "field_eval: bf16x3 needs a workspace of %lld bytes",
                       (long long)nsvf_field_workspace_bytes(2, D, H, L, m_max));
  }
#define NSVF_FIELD_CASE(d, h, l)                                                               \
  if (D == d && H == h && L == l) {                                                            \
    if (precision == 2) {                                                                      \
      const int rc = launch_embed_bf16x3<d, h, l>(                                             \
          sc, rays, static_cast<const SampleRec*>(samples), list, m_dev, m_max, workspace, st); \
      if (rc) return rc;                                                                       \
      return launch_mlp_bf16x3<d, h, l>(m_dev, m_max, wfrag, bias, out, list, workspace, st);  \
    }                                                                                          \
    if (precision == 0 && workspace &&                                                        \
        workspace_bytes >= bf16_workspace_bytes<d, h, l>(m_max))                               \
      return launch_bf16_split<d, h, l>(sc, rays, static_cast<const SampleRec*>(samples),      \
                                        list, m_dev, m_max, wfrag, bias, out, workspace, st);  \
    if (precision == 0)                                                                        \
      return launch_bf16_fused<d, h, l>(sc, rays, static_cast<const SampleRec*>(samples),      \
                                        list, m_dev, m_max, wfrag, bias, out, st);             \
    return launch_f32<d, h, l>(src, m_dev, m_max, wfrag, bias, sink, st);                      \
  }
  NSVF_FIELD_CASE(32, 128, 3)
  NSVF_FIELD_CASE(48, 256, 4)
  NSVF_FIELD_CASE(32, 256, 4)
#undef NSVF_FIELD_CASE
  return set_error(NSVF_EINVAL, "field_eval: unsupported shape D=%d H=%d L=%d", D, H, L);
}

int64_t nsvf_field_static_scratch_bytes(void) { return kStaticScratchBytes; }

int nsvf_field_eval_static(int precision, int D, int H, int L, const float* centers,
                           const int32_t* feats, const float* values, float voxel_size,
                           const float* ray_start, const float* ray_dir, const void* samples,
                           const int32_t* list, const int32_t* m_dev, int64_t m_max,
                           const void* wfrag, const float* bias, float* out, void* stream) {
  clear_error();
  if (m_max <= 0) return NSVF_OK;
  if (precision < 0 || precision > 2)
    return set_error(NSVF_EINVAL, "field_eval_static: precision %d (0 bf16, 1 fp32, 2 bf16x3)",
                     precision);
  if (precision == 2 && !nsvf_field_bf16x3_supported(D, H, L))
    return set_error(NSVF_EINVAL, "field_eval_static: bf16x3 does not support D=%d H=%d L=%d", D,
                     H, L);
  // FP32 runs without scratch; BF16 and split BF16 stage their operand tiles
  const int64_t need = precision == 1 ? 0 : nsvf_field_workspace_bytes(precision, D, H, L, m_max);
  if (precision != 1 && need <= 0)
    return set_error(NSVF_EINVAL, "field_eval_static: unsupported shape D=%d H=%d L=%d", D, H, L);
  if (need > kStaticScratchBytes) {
    const int64_t per = need / m_max > 0 ? need / m_max : 1;
    return set_error(NSVF_EINVAL,
                     "field_eval_static: %lld samples need %lld bytes of scratch; the static "
                     "scratch holds %lld (about %lld samples at this precision and shape)",
                     (long long)m_max, (long long)need, (long long)kStaticScratchBytes,
                     (long long)(kStaticScratchBytes / per));
  }
  void* ws = nullptr;
  if (need > 0)
    if (const int rc = scratch_address(&ws)) return rc;
  return nsvf_field_eval(precision, D, H, L, centers, feats, values, voxel_size, ray_start,
                         ray_dir, samples, list, m_dev, m_max, wfrag, bias, out, ws,
                         need > 0 ? kStaticScratchBytes : 0, stream);
}

int nsvf_field_embed_bf16(int D, int H, int L, const float* centers, const int32_t* feats,
                          const float* values, float voxel_size, const float* ray_start,
                          const float* ray_dir, const void* samples, const int32_t* list,
                          const int32_t* m_dev, int64_t m_max, void* workspace,
                          int64_t workspace_bytes, void* stream) {
  clear_error();
  if (m_max <= 0) return NSVF_OK;
  if (!centers || !feats || !values || !ray_start || !ray_dir || !samples || !m_dev || !workspace)
    return set_error(NSVF_EINVAL, "field_embed_bf16: null pointer");
  if (workspace_bytes < nsvf_field_workspace_bytes(0, D, H, L, m_max))
    return set_error(NSVF_EINVAL, "field_embed_bf16: workspace too small");
  FieldScene sc{centers, feats, values, voxel_size, 1.0f / voxel_size, is_pow2(voxel_size)};
  FieldRays rays{ray_start, ray_dir};
#define NSVF_EMB_CASE(d, h, l)                                                                \
  if (D == d && H == h && L == l)                                                             \
    return launch_embed_bf16<d, h, l>(sc, rays, static_cast<const SampleRec*>(samples), list, \
                                      m_dev, m_max, workspace, as_stream(stream));
  NSVF_EMB_CASE(32, 128, 3)
  NSVF_EMB_CASE(48, 256, 4)
  NSVF_EMB_CASE(32, 256, 4)
#undef NSVF_EMB_CASE
  return set_error(NSVF_EINVAL, "field_embed_bf16: unsupported shape D=%d H=%d L=%d", D, H, L);
}

int nsvf_field_mlp_bf16(int D, int H, int L, const void* workspace, int64_t workspace_bytes,
                        const int32_t* m_dev, int64_t m_max, const void* wfrag, const float* bias,
                        float* out, const int32_t* list, void* stream) {
  clear_error();
  if (m_max <= 0) return NSVF_OK;
  if (!workspace || !m_dev || !wfrag || !bias || !out)
    return set_error(NSVF_EINVAL, "field_mlp_bf16: null pointer");
  if (workspace_bytes < nsvf_field_workspace_bytes(0, D, H, L, m_max))
    return set_error(NSVF_EINVAL, "field_mlp_bf16: workspace too small");
#define NSVF_MLP_CASE(d, h, l) \
  if (D == d && H == h && L == l) \
    return launch_mlp_bf16<d, h, l>(m_dev, m_max, wfrag, bias, out, list, workspace,  \
                                    as_stream(stream));
  NSVF_MLP_CASE(32, 128, 3)
  NSVF_MLP_CASE(48, 256, 4)
  NSVF_MLP_CASE(32, 256, 4)
#undef NSVF_MLP_CASE
  return set_error(NSVF_EINVAL, "field_mlp_bf16: unsupported shape D=%d H=%d L=%d", D, H, L);
}

int nsvf_field_embed_bf16x3(int D, int H, int L, const float* centers, const int32_t* feats,
                            const float* values, float voxel_size, const float* ray_start,
                            const float* ray_dir, const void* samples, const int32_t* list,
                            const int32_t* m_dev, int64_t m_max, void* workspace,
                            int64_t workspace_bytes, void* stream) {
  clear_error();
  if (m_max <= 0) return NSVF_OK;
  if (!centers || !feats || !values || !ray_start || !ray_dir || !samples || !m_dev || !workspace)
    return set_error(NSVF_EINVAL, "field_embed_bf16x3: null pointer");
  if (!nsvf_field_bf16x3_supported(D, H, L))
    return set_error(NSVF_EINVAL, "field_embed_bf16x3: unsupported shape D=%d H=%d L=%d", D, H,
                     L);
  if (workspace_bytes < nsvf_field_workspace_bytes(2, D, H, L, m_max))
    return set_error(NSVF_EINVAL, "field_embed_bf16x3: workspace too small");
  FieldScene sc{centers, feats, values, voxel_size, 1.0f / voxel_size, is_pow2(voxel_size)};
  FieldRays rays{ray_start, ray_dir};
#define NSVF_EMB3_CASE(d, h, l)                                                                 \
  if (D == d && H == h && L == l)                                                               \
    return launch_embed_bf16x3<d, h, l>(sc, rays, static_cast<const SampleRec*>(samples), list, \
                                        m_dev, m_max, workspace, as_stream(stream));
  NSVF_EMB3_CASE(32, 128, 3)
  NSVF_EMB3_CASE(48, 256, 4)
  NSVF_EMB3_CASE(32, 256, 4)
#undef NSVF_EMB3_CASE
  return set_error(NSVF_EINVAL, "field_embed_bf16x3: unsupported shape D=%d H=%d L=%d", D, H, L);
}

int nsvf_field_mlp_bf16x3(int D, int H, int L, const void* workspace, int64_t workspace_bytes,
                          const int32_t* m_dev, int64_t m_max, const void* wfrag,
                          const float* bias, float* out, const int32_t* list, void* stream) {
  clear_error();
  if (m_max <= 0) return NSVF_OK;
  if (!workspace || !m_dev || !wfrag || !bias || !out)
    return set_error(NSVF_EINVAL, "field_mlp_bf16x3: null pointer");
  if (!nsvf_field_bf16x3_supported(D, H, L))
    return set_error(NSVF_EINVAL, "field_mlp_bf16x3: unsupported shape D=%d H=%d L=%d", D, H, L);
  if (workspace_bytes < nsvf_field_workspace_bytes(2, D, H, L, m_max))
    return set_error(NSVF_EINVAL, "field_mlp_bf16x3: workspace too small");
#define NSVF_MLP3_CASE(d, h, l) \
  if (D == d && H == h && L == l) \
    return launch_mlp_bf16x3<d, h, l>(m_dev, m_max, wfrag, bias, out, list, workspace, \
                                      as_stream(stream));
  NSVF_MLP3_CASE(32, 128, 3)
  NSVF_MLP3_CASE(48, 256, 4)
  NSVF_MLP3_CASE(32, 256, 4)
#undef NSVF_MLP3_CASE
  return set_error(NSVF_EINVAL, "field_mlp_bf16x3: unsupported shape D=%d H=%d L=%d", D, H, L);
}

int nsvf_field_tiles_bf16x3(int D, int H, int L, const float* feats, int64_t M,
                            void* workspace, int64_t workspace_bytes, void* stream) {
  clear_error();
  if (M < 0) return set_error(NSVF_EINVAL, "field_tiles_bf16x3: M=%lld", static_cast<long long>(M));
  if (M == 0) return NSVF_OK;
  if (!feats || !workspace) return set_error(NSVF_EINVAL, "field_tiles_bf16x3: null pointer");
  if (!nsvf_field_bf16x3_supported(D, H, L))
    return set_error(NSVF_EINVAL, "field_tiles_bf16x3: unsupported shape D=%d H=%d L=%d", D, H, L);
  if (workspace_bytes < nsvf_field_workspace_bytes(2, D, H, L, M))
    return set_error(NSVF_EINVAL, "field_tiles_bf16x3: workspace too small");
#define NSVF_TILES3_CASE(d, h, l) \
  if (D == d && H == h && L == l) \
    return launch_tiles_bf16x3<d, h, l>(feats, M, workspace, as_stream(stream));
  NSVF_TILES3_CASE(32, 128, 3)
  NSVF_TILES3_CASE(48, 256, 4)
  NSVF_TILES3_CASE(32, 256, 4)
#undef NSVF_TILES3_CASE
  return set_error(NSVF_EINVAL, "field_tiles_bf16x3: unsupported shape D=%d H=%d L=%d", D, H, L);
}

int nsvf_field_prune(int precision, int D, int H, int L, const float* centers,
                     const int32_t* feats, const float* values, int64_t N, float voxel_size,
                     int64_t v0, int64_t n, int bits, const void* wfrag, const float* bias,
                     float* vmax, void* stream) {
  clear_error();
  if (n < 0 || v0 < 0 || v0 + n > N) return set_error(NSVF_EINVAL, "field_prune: bad voxel range");
  if (n == 0) return NSVF_OK;
  if (bits < 4 || bits % 4 != 0 || bits > 64)
    return set_error(NSVF_EINVAL, "field_prune: bits=%d (a multiple of 4 in [4, 64])", bits);
  if (precision != 0 && precision != 1)
    return set_error(NSVF_EINVAL, "field_prune: precision %d", precision);
  if (!centers || !feats || !values || !wfrag || !bias || !vmax)
    return set_error(NSVF_EINVAL, "field_prune: null pointer");
  FieldScene sc{centers, feats, values, voxel_size, 1.0f / voxel_size, is_pow2(voxel_size)};
  hipStream_t st = as_stream(stream);
#define NSVF_PRUNE_CASE(d, h, l) \
  if (D == d && H == h && L == l) \
    return launch_prune<d, h, l>(precision, sc, v0, n, bits, wfrag, bias, vmax, st);
  NSVF_PRUNE_CASE(32, 128, 3)
  NSVF_PRUNE_CASE(48, 256, 4)
  NSVF_PRUNE_CASE(32, 256, 4)
#undef NSVF_PRUNE_CASE
  return set_error(NSVF_EINVAL, "field_prune: unsupported shape D=%d H=%d L=%d", D, H, L);
}

int nsvf_trilinear_gather(const int32_t* s_idx, const float* xyz, int64_t M, const float* centers,
                          const int32_t* feats, const float* values, int D, float voxel_size,
                          float* out, void* stream) {
  clear_error();
  if (M < 0 || D <= 0) return set_error(NSVF_EINVAL, "trilinear_gather: bad size");
  if (M == 0) return NSVF_OK;
  if (!s_idx || !xyz || !centers || !feats || !values || !out)
    return set_error(NSVF_EINVAL, "trilinear_gather: null pointer");
  FieldScene sc{centers, feats, values, voxel_size, 1.0f / voxel_size, is_pow2(voxel_size)};
  hipLaunchKernelGGL(gather_kernel, dim3(unsigned(ceil_div(M * D, 256))), dim3(256), 0,
                     as_stream(stream), s_idx, xyz, M, D, sc, out);
  return check_launch("trilinear_gather");
}

}  // extern "C"
