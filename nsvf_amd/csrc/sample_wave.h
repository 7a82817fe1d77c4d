// Inverse-CDF sampling of one ray by one wavefront, with the same outputs,
// bit for bit, as the sequential walk of sample_core.h (icdf_ray into a
// StreamSink, with the same noise).  The sequential walk interleaves two
// sorted sequences -- the steps' cdf values (i + u_i) / steps, u_i in (0, 1)
// (0.5 when deterministic), and the bins' cumulative probabilities C_k -- so
// every output has a closed-form place:
//
//   b_i   = the first bin k with cdf_i <= C_k (binary search; C in LDS);
//   step i's sample goes to position i + b_i (i steps and b_i bin tails
//   precede it), with z = fma(u, hi - lo, lo), u = (cdf - C_{b-1}) /
//   (C_b - C_{b-1}), and z_low = z_{i-1} if b_{i-1} == b_i, else lo_b;
//   the tails of the bins crossed between steps i-1 and i, k in
//   [b_{i-1}, b_i), go to positions i + k, z_low = z_{i-1} for
//   k == b_{i-1}, else lo_k -- computed per bin: i = i_k, the number of
//   steps with cdf <= C_k;
//   the first step with cdf beyond C of the last bin ends the walk after
//   the remaining tails ("done"); otherwise the trailing tails follow the
//   last step, bin after bin, while each bin passes the same tests as the
//   sequential loop.
//
// So each lane takes one step, then one bin (64 per pass), positions are
// contiguous across lanes (coalesced 16-byte stores), and only the prefix C
// -- whose FP32 rounding order is the reference's left-to-right sum -- is a
// serial chain.  Rows with more bins than the LDS holds, a NaN in C, or a step count
// beyond the slot bound take the sequential walk on lane 0 instead.
#pragma once

#include <type_traits>

#include "render.h"
#include "sample_core.h"

namespace nsvf {

// Writes the samples that carry a voxel id into the ray's slot range (dists
// clamped at 0: encoder.py:547).
struct StreamSink {
  SampleRec* out;
  int32_t ray_id;
  int cap;  // the ray's slot count (never exceeded for valid bins; a guard)
  int n = 0;
  __device__ __forceinline__ void operator()(int, int32_t id, float d, float mid) {
    if (id == -1 || n >= cap) return;
    out[n] = SampleRec{id, mid, d > 0.0f ? d : 0.0f, ray_id};
    ++n;
  }
  // Marks the slots past the written samples as unused.
  __device__ __forceinline__ void close(int bound) {
    for (int k = n; k < bound; ++k) out[k] = SampleRec{-1, 0.0f, 0.0f, ray_id};
  }
};

// A wave's LDS staging of one ray's bins: cumulative probability, bounds and
// voxel ids of bins [0, cap); rows with more bins take the sequential walk.
struct WaveBinsLds {
  float* C;
  float* lo;
  float* hi;
  int32_t* id;
  int cap;
};

__device__ __forceinline__ int first_lane(uint64_t mask) {
  return mask ? __ffsll(static_cast<unsigned long long>(mask)) - 1 : 64;
}

// Writes one sample record (distance clamped at 0: encoder.py:547); slots
// past the ray's bound are never written (a guard: valid bins stay within).
__device__ __forceinline__ void put_rec(SampleRec* out, int bound, int pos, int32_t id, float d,
                                        float mid, int32_t ray) {
  if (pos < bound) out[pos] = SampleRec{id, mid, d > 0.0f ? d : 0.0f, ray};
}

// Work done beside each record a sampler writes: put(pos, bound, k) for the
// sample at slot pos, k >= 0 when that sample is bin k's own sample again
// (kBins hooks: the whole bin [lo_k, hi_k] with bins.depth(k) == (lo_k +
// hi_k) / 2 bit for bit, so voxel id and depth equal bin k's), else -1;
// seq(n) after the sequential walk wrote slots [0, n); unused(pos) for each
// slot past the samples.  NoSlotHook does nothing.
struct NoSlotHook {
  static constexpr bool kBins = false;
  __device__ __forceinline__ void put(int, int, int) {}
  __device__ __forceinline__ void seq(int) {}
  __device__ __forceinline__ void unused(int) {}
};

// Samples ray `ray` with bins `bins` (P columns, the first Pv valid) into
// out[0, bound); slots past the samples are marked unused.  Returns the
// sample count (same value in every lane).  lds: this wave's staging arrays;
// hook: called beside every slot written (the caller's lane decides nothing);
// noise: u_i of step i (its cdf is (i + u_i) / steps; 0.5 when
// deterministic).  Any u_i in (0, 1) keeps the steps' cdf increasing, so the
// closed-form places hold for training-mode noise too.
template <class Bins, class Hook = NoSlotHook, class Noise = ConstNoise>
__device__ int icdf_wave(const Bins& bins, int P, int Pv, float steps, int compat, int bound,
                         int32_t ray, SampleRec* out, const WaveBinsLds& lds,
                         Hook&& hook = NoSlotHook{}, const Noise& noise = ConstNoise{0.5f}) {
  const int lane = threadIdx.x & 63;
  // the row's sizes are the same in every lane; scalar copies keep the
  // loops below uniform (no per-lane exit masks)
  P = __builtin_amdgcn_readfirstlane(P);
  Pv = __builtin_amdgcn_readfirstlane(Pv);
  bound = __builtin_amdgcn_readfirstlane(bound);
  steps = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(steps)));
  using HookT = std::remove_reference_t<Hook>;
  int n = 0;
  // bit p of lane l: bin 64 p + l is its own sample when whole (kBins)
  uint32_t same = 0;
  bool seq = Pv <= 0 || Pv > lds.cap || !(steps <= float(bound));
  if (!seq) {
    // stage the bins; the prefix C_k = C_{k-1} + p_k runs left to right (the
    // reference kernel's order) across the lanes: a whole-wave shift right by
    // one lane (DPP wave_shr:1) hands lane k the running sum of lane k-1, so
    // after step i lanes 0..i hold C_k, each sum rounded once in the order of
    // a sequential loop; lane 0 starts from the previous 64 bins' carry
    float c = 0.0f;
    for (int t0 = 0; t0 < Pv; t0 += 64) {
      const int k = t0 + lane;
      float pk = 0.0f;
      if (k < Pv) {
        pk = bins.staged_prob(k);
        const float lk = bins.lo(k), hk = bins.hi(k);
        lds.lo[k] = lk;
        lds.hi[k] = hk;
        lds.id[k] = bins.id(k);
        if constexpr (HookT::kBins)
          if (__float_as_int((hk + lk) * 0.5f) == __float_as_int(bins.depth(k)))
            same |= 1u << (t0 >> 6);
      }
      const int m = Pv - t0 < 64 ? Pv - t0 : 64;
      const float q = (lane == 0 && t0 > 0) ? c + pk : pk;
      float run = q;
      for (int i = 1; i < m; ++i) {
        // lane 0 has no left neighbour and reads 0: 0 + q == q
        const float left = __int_as_float(
            __builtin_amdgcn_update_dpp(0, __float_as_int(run), 0x138 /* wave_shr:1 */, 0xf, 0xf,
                                        false));
        run = left + q;
      }
      c = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(run), m - 1));
      if (k < Pv) lds.C[k] = run;
    }
    wave_sync();
    seq = isnan(c);
  }
  if (seq) {
    bins.before_seq();
    if (lane == 0) {
      StreamSink sink{out, ray, bound};
      icdf_ray(bins, P, steps, -1.0f, noise, compat, bound, sink);
      n = sink.n;
    }
    n = __shfl(n, 0, 64);
    hook.seq(n);
  } else {
    const float* C = lds.C;
    const float ss = float(1.0 / double(steps));
    const int T = int(ceilf(steps));
    int bprev = 0;
    float zprev = lds.lo[0];
    bool done = false;
    int cb = 0;
    float zl = 0.0f, cmax = 0.0f;
    for (int base = 0; base < T; base += 64) {
      const int i = base + lane;
      const bool act = i < T;
      int b = Pv;
      float z = 0.0f;
      if (act) {
        const float cdf = (float(i) + noise.at(i)) * ss;
        int lo = 0, hi = Pv;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (C[mid] < cdf)
            lo = mid + 1;
          else
            hi = mid;
        }
        b = lo;
        if (b < Pv) {
          const float cmin = b > 0 ? C[b - 1] : 0.0f;
          const float u = (cdf - cmin) / (C[b] - cmin);
          const float lb = lds.lo[b];
          z = __fmaf_rn(u, lds.hi[b] - lb, lb);
        }
      }
      int bp = __shfl_up(b, 1, 64);
      float zp = __shfl_up(z, 1, 64);
      if (lane == 0) {
        bp = bprev;
        zp = zprev;
      }
      const uint64_t dmask = __ballot(act && b >= Pv);
      const int dl = first_lane(dmask);
      if (act && lane <= dl && b < Pv) {
        const float zk = (b == bp) ? zp : lds.lo[b];
        put_rec(out, bound, i + b, lds.id[b], z - zk, (z + zk) * 0.5f, ray);
        hook.put(i + b, bound, -1);
      }
      if (dmask) {
        const int bpd = __shfl(bp, dl, 64);
        const float zpd = __shfl(zp, dl, 64);
        done = true;
        n = base + dl + Pv;
        cb = Pv;
        zl = (bpd == Pv - 1) ? zpd : lds.lo[Pv - 1];
        cmax = lds.hi[Pv - 1];
        break;
      }
      const int last = (T - 1 - base) < 63 ? (T - 1 - base) : 63;
      bprev = __shfl(b, last, 64);
      zprev = __shfl(z, last, 64);
    }
    // the tails of the bins crossed by a step, one lane per bin: bin k's
    // tail follows the i_k steps with cdf <= C_k and the k earlier tails
    // (position i_k + k); its lower end is the depth of step i_k - 1 when
    // that step fell in bin k (C_{k-1} < cdf <= C_k), else the bin's entry.
    // i_k comes from an estimate corrected against the exact, monotone cdf.
    for (int k0 = 0; k0 < Pv; k0 += 64) {
      const int k = k0 + lane;
      if (k < Pv) {
        const float ck = C[k];
        float e = floorf(ck * steps - 0.5f) + 1.0f;
        e = fminf(fmaxf(e, 0.0f), float(T));
        int ik = int(e);
        while (ik > 0 && (float(ik - 1) + noise.at(ik - 1)) * ss > ck) --ik;
        while (ik < T && (float(ik) + noise.at(ik)) * ss <= ck) ++ik;
        if (ik < T) {
          const float lk = lds.lo[k], hk = lds.hi[k];
          float zk = lk;
          if (ik > 0) {
            const float cdfp = (float(ik - 1) + noise.at(ik - 1)) * ss;
            const float cmin = k > 0 ? C[k - 1] : 0.0f;
            if (k == 0 || cdfp > cmin) zk = __fmaf_rn((cdfp - cmin) / (ck - cmin), hk - lk, lk);
          }
          put_rec(out, bound, ik + k, lds.id[k], hk - zk, (hk + zk) * 0.5f, ray);
          const bool whole = __float_as_int(zk) == __float_as_int(lk);
          const bool own = whole && ((same >> ((k0 >> 6) & 31)) & 1u);
          hook.put(ik + k, bound, own ? k : -1);
        }
      }
    }
    if (!done) {
      n = T > 0 ? T + bprev : 0;
      cb = T > 0 ? bprev : 0;
      zl = zprev;
      cmax = lds.hi[cb];
    }
    // trailing tails (sample_core.h: the loop after the steps); bins past
    // Pv are read through `bins` (pads)
    if (!(done && !compat) && zl < cmax) {
      const int32_t id0 = cb < Pv ? lds.id[cb] : (cb < P ? bins.id(cb) : bins.next_first());
      if (id0 != -1) {
        // every lane takes part in the shuffles (their source lanes must be active)
        const uint32_t s_cb = __shfl(same, cb & 63, 64);
        const bool same_cb = cb < Pv && ((s_cb >> ((cb >> 6) & 31)) & 1u);
        if (lane == 0) {
          put_rec(out, bound, n, id0, cmax - zl, (cmax + zl) * 0.5f, ray);
          const bool whole = same_cb && __float_as_int(zl) == __float_as_int(lds.lo[cb]) &&
                             __float_as_int(cmax) == __float_as_int(lds.hi[cb]);
          hook.put(n, bound, whole ? cb : -1);
        }
        ++n;
      }
      const int k0 = cb + 1;
      for (int base = k0; base < P; base += 64) {
        const int k = base + lane;
        bool ok = false;
        float lk = 0.0f, hk = 0.0f;
        int32_t idk = -1;
        if (k < P) {
          const int32_t probe = compat ? bins.ray0_id(k) : (k < Pv ? lds.id[k] : bins.id(k));
          lk = k < Pv ? lds.lo[k] : bins.lo(k);
          hk = k < Pv ? lds.hi[k] : bins.hi(k);
          idk = k < Pv ? lds.id[k] : bins.id(k);
          ok = probe != -1 && lk < hk;
        }
        const int stop = first_lane(__ballot(!ok));
        const uint32_t s_k = __shfl(same, k & 63, 64);
        const bool same_k = k < Pv && ((s_k >> ((k >> 6) & 31)) & 1u);
        if (lane < stop) {
          put_rec(out, bound, n + lane, idk, hk - lk, (hk + lk) * 0.5f, ray);
          hook.put(n + lane, bound, same_k ? k : -1);
        }
        n += stop < 64 ? stop : 64;
        if (stop < 64) break;
      }
    }
  }
  for (int k = n + lane; k < bound; k += 64) {
    out[k] = SampleRec{-1, 0.0f, 0.0f, ray};
    hook.unused(k);
  }
  wave_sync();  // the next ray's staging must not overtake this ray's reads
  return n;
}

}  // namespace nsvf
