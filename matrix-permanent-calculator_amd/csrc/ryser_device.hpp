// ryser_device.hpp -- device-side building blocks shared by the Ryser kernels
// (gfx950 / CDNA4, wave64).
//
//  * double-double ("dd") accumulation: error-free TwoSum, used for every
//    sum that leaves a lane (block folds, wave / workgroup / grid reductions);
//  * wave reductions built on ds_swizzle (xor butterflies inside each
//    32-lane half) plus one v_readlane across the halves -- fixed order, so
//    every reduction is deterministic run to run;
//  * the partial-sum record every Ryser kernel writes per workgroup.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace perm {

constexpr int kWave = 64;

// One workgroup's contribution: signed sum as hi+lo, sum of |terms|, and two
// sums over runs of 2^(R+5) consecutive terms of one lane that size the
// statistical error estimate: sq_sum = sum of (sum of |terms| in the run)^2
// (an upper bound on sum term^2: the product roundings) and sq_run = sum of
// (signed sum of the run)^2 (the X drift: it moves a run's terms together,
// so its effect follows the run's cancelled sum).  pad: the dense walk kernel
// stores its workgroup's shader cycles per 10 ns real-time tick (summed by
// the grid reduction: mean clock = 0.1 GHz * sum / workgroups).  48 bytes.
struct Partial {
  double hi, lo, abs_sum, sq_sum, sq_run, pad;
};

// The plain-summed statistics carried through every reduction.
struct Moments {
  double a, q, r;  // abs_sum, sq_sum, sq_run
};

struct DD {
  double hi, lo;
};

__device__ __forceinline__ DD two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

__device__ __forceinline__ DD quick_two_sum(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}

// IEEE-style double-double addition (TwoSum on both words, renormalise).
__device__ __forceinline__ DD dd_add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi);
  DD t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}

// Fold one double into a dd accumulator.
__device__ __forceinline__ DD dd_add_d(DD a, double b) {
  DD s = two_sum(a.hi, b);
  s.lo += a.lo;
  return quick_two_sum(s.hi, s.lo);
}

// ds_swizzle in bit-mask mode: lane l reads lane (l ^ XOR) within its
// 32-lane half.  Offset layout: and_mask [4:0] = 0x1f, or_mask [9:5] = 0,
// xor_mask [14:10].
template <int XOR>
__device__ __forceinline__ double swizzle_xor(double v) {
  constexpr int pattern = (XOR << 10) | 0x1f;
  int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), pattern);
  int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), pattern);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Wave-wide dd sum (and plain sums of the moments); the full result is
// valid in lane 0.
__device__ __forceinline__ void wave_reduce(DD& s, Moments& m) {
#define PERM_XOR_STEP(M)                                 \
  {                                                      \
    DD o = {swizzle_xor<M>(s.hi), swizzle_xor<M>(s.lo)}; \
    s = dd_add(s, o);                                    \
    m.a += swizzle_xor<M>(m.a);                          \
    m.q += swizzle_xor<M>(m.q);                          \
    m.r += swizzle_xor<M>(m.r);                          \
  }
  PERM_XOR_STEP(1)
  PERM_XOR_STEP(2)
  PERM_XOR_STEP(4)
  PERM_XOR_STEP(8)
  PERM_XOR_STEP(16)
#undef PERM_XOR_STEP
  DD up = {readlane_d(s.hi, 32), readlane_d(s.lo, 32)};
  const Moments mu = {readlane_d(m.a, 32), readlane_d(m.q, 32), readlane_d(m.r, 32)};
  s = dd_add(s, up);
  m.a += mu.a;
  m.q += mu.q;
  m.r += mu.r;
}

__device__ __forceinline__ Partial make_partial(DD s, const Moments& m) {
  return {s.hi, s.lo, m.a, m.q, m.r, 0.0};
}

// Workgroup reduction of (s, m) into one Partial (thread 0 sums the waves).
// `scratch` must hold blockDim.x/64 entries.
// pad: the record's last word as thread 0 holds it, or, with pad_terms,
// the sum of pad_terms[0, blockDim.x) (written by the block before the call).
__device__ __forceinline__ void block_reduce_store(DD s, Moments m, Partial* scratch,
                                                   Partial* out, double pad = 0.0,
                                                   const double* pad_terms = nullptr) {
  wave_reduce(s, m);
  const int wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) scratch[wid] = make_partial(s, m);
  __syncthreads();
  if (threadIdx.x == 0) {
    DD t = {scratch[0].hi, scratch[0].lo};
    Moments tm = {scratch[0].abs_sum, scratch[0].sq_sum, scratch[0].sq_run};
    for (int w = 1; w < nw; ++w) {
      t = dd_add(t, DD{scratch[w].hi, scratch[w].lo});
      tm.a += scratch[w].abs_sum;
      tm.q += scratch[w].sq_sum;
      tm.r += scratch[w].sq_run;
    }
    scratch[0] = make_partial(t, tm);
    if (pad_terms) {
      pad = 0.0;
      for (int i = 0; i < (int)blockDim.x; ++i) pad += pad_terms[i];
    }
    scratch[0].pad = pad;
  }
  __syncthreads();
  // the record goes out as one store instruction of six lanes (one 48-byte
  // request) instead of three 16-byte stores of one lane, each a separate
  // partial-line write to HBM
  constexpr int kWords = (int)(sizeof(Partial) / sizeof(double));
  if (threadIdx.x < kWords)
    reinterpret_cast<double*>(out)[threadIdx.x] =
        reinterpret_cast<const double*>(&scratch[0])[threadIdx.x];
}

}  // namespace perm
