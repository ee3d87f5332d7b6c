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

// One workgroup's contribution: signed sum as hi+lo, sum of |terms|, and
// sum over runs of consecutive terms of (sum of |terms| in the run)^2 -- an
// upper bound on sum term^2 that sizes the statistical error estimate.
struct Partial {
  double hi, lo, abs_sum, sq_sum;
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

// Wave-wide dd sum (and plain sums of `a` and `q`); the full result is
// valid in lane 0.
__device__ __forceinline__ void wave_reduce(DD& s, double& a, double& q) {
#define PERM_XOR_STEP(M)                                 \
  {                                                      \
    DD o = {swizzle_xor<M>(s.hi), swizzle_xor<M>(s.lo)}; \
    s = dd_add(s, o);                                    \
    a += swizzle_xor<M>(a);                              \
    q += swizzle_xor<M>(q);                              \
  }
  PERM_XOR_STEP(1)
  PERM_XOR_STEP(2)
  PERM_XOR_STEP(4)
  PERM_XOR_STEP(8)
  PERM_XOR_STEP(16)
#undef PERM_XOR_STEP
  DD up = {readlane_d(s.hi, 32), readlane_d(s.lo, 32)};
  double aup = readlane_d(a, 32);
  double qup = readlane_d(q, 32);
  s = dd_add(s, up);
  a += aup;
  q += qup;
}

// Workgroup reduction of (s, a, q) into one Partial written by thread 0.
// `scratch` must hold blockDim.x/64 entries.
__device__ __forceinline__ void block_reduce_store(DD s, double a, double q, Partial* scratch,
                                                   Partial* out) {
  wave_reduce(s, a, q);
  const int wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) scratch[wid] = {s.hi, s.lo, a, q};
  __syncthreads();
  if (threadIdx.x == 0) {
    DD t = {scratch[0].hi, scratch[0].lo};
    double ta = scratch[0].abs_sum, tq = scratch[0].sq_sum;
    for (int w = 1; w < nw; ++w) {
      t = dd_add(t, DD{scratch[w].hi, scratch[w].lo});
      ta += scratch[w].abs_sum;
      tq += scratch[w].sq_sum;
    }
    *out = {t.hi, t.lo, ta, tq};
  }
}

}  // namespace perm
