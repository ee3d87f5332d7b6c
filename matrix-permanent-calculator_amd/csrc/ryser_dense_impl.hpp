// ryser_dense_impl.hpp -- the dense Ryser kernel family for gfx950
// (Nijenhuis-Wilf form of Ryser's formula, Gray-code subset order).
//
// Reference semantics: gpu_exact_dense.cu:129-399 (kernels) and :466-699
// (wrappers) of berkyaglioglu/matrix-permanent-calculator.  The sum is the
// same -- over all subsets S of the first n-1 columns of
// (-1)^|S| prod_j x_j(S), x(S) = x0 + sum_{k in S} a_k, scaled by
// (4(n&1)-2) -- walked in a way designed for CDNA4:
//
//  * Every lane owns an ALIGNED block of 2^w subsets: its "high" columns
//    [w, n-1) are fixed by the lane id (x built in closed form) and it walks
//    the low columns [0, w) in standard Gray order.  All 64 lanes of a wave
//    flip the SAME column with the SAME sign at every step: the column index
//    is wave-uniform.
//  * Because it is uniform, the column is fetched with scalar loads
//    (s_load_dwordx16 through the scalar cache) into SGPRs and feeds
//    v_add_f64 as the SGPR operand; each lane keeps x in VGPRs.  The store
//    holds +a_k and -a_k, so a signed update is one add.
//  * One column is prefetched per step while the previous step's product
//    runs (sched_barrier separates the phases).
//  * The innermost 2^R steps are unrolled with compile-time column indices.
//  * Per step: n adds, n-2 multiplies, one FMA that multiplies in the last x
//    and accumulates the signed term, one FMA for sum |term|: 2n FP64 VALU
//    instructions per subset.
//  * Nested variants (shape Nest<K0, K1, K2>, K0 <= K1 <= K2 <= N): Gray
//    column 0 flips on every odd step, column 1 on steps 2 mod 4, column 2
//    on steps 4 mod 8.  The host places sparse columns there and orders the
//    rows so that column l touches only rows [0, K_l).  A step of level l
//    (the flipped column's position; 3 for any higher column) then adds
//    K_l rows and the products of the untouched tails are reused:
//    T3 = prod x[K2..N) changes only at level 3, T2 = prod x[K1..K2) T3 at
//    levels >= 2, T1 = prod x[K0..K1) T2 at levels >= 1, and every term is
//    prod x[0..K0) T1.  Per 8 steps: 4 K0 + 2 K1 + K2 + N adds and as many
//    multiplies instead of 8 N each.  Nest<K0, N, N> is the pair variant
//    (only the passive product of column 0's untouched rows is shared);
//    Nest<N, N, N> is the plain walk.  Adds of zero are exact, so the
//    skipped adds change no bits of any x.
//  * Lane sums are folded into double-double every 2^(R+5) steps; lanes,
//    waves and workgroups are reduced in double-double in a fixed order.
#pragma once
#include <utility>

#include "ryser_device.hpp"
#include "ryser_kernels.hpp"

namespace perm {

// Pointer into the column store in the constant address space: loads through
// it are known not to alias any store of the kernel, so uniform loads are
// selected as scalar (SMEM) loads.
typedef const double __attribute__((address_space(4)))* ColPtr;

// Kernel modes.
//  kFast: FP64 X, FP64 products (the production kernel).
//  kDDProd: FP64 X, double-double products and lane sums -- exact products
//    for integer matrices (whose X is exact), used to round 0/1 results to
//    the exact integer.
//  kDDAll: X itself kept as a double-double (every column addition through
//    TwoSum), double-double products and sums: no X drift; a verification
//    mode for large n where no CPU oracle finishes.
enum Mode : int { kFast = 0, kDDProd = 1, kDDAll = 2 };

#ifndef PERM_FAST_WAVES_PER_EU
#define PERM_FAST_WAVES_PER_EU 1
#endif
// Minimum waves per SIMD the fast kernel is compiled for (caps its VGPRs at
// 512/waves).  1 = no cap: unconstrained, n <= 40 already fits in 128 VGPRs
// (4 waves/SIMD) without spills; measured, 3 and 4 waves run at the same
// VALU issue rate.
constexpr int kFastWavesPerEU = PERM_FAST_WAVES_PER_EU;

// Load COUNT doubles of column block `blk` (wave-uniform address) into
// SGPRs.  The pointer is laundered through an empty asm so that the compiler
// neither merges repeated loads of one column across steps (which would pin
// N SGPR pairs and spill) nor hoists them out of the loop.
template <int N, int COUNT>
__device__ __forceinline__ void load_col(double (&cv)[N], ColPtr base, int blk) {
  ColPtr c = base + (size_t)blk * N;
  asm volatile("" : "+s"(c));
#pragma unroll
  for (int j = 0; j < COUNT; ++j) cv[j] = c[j];
}

// Per-lane accumulator.  kFast: plain double `p` folded into double-double
// every 32 groups.  Other modes: the signed sum is kept as an unevaluated
// pair (p, plo) updated with TwoSum on every term.
struct LaneAcc {
  double p = 0.0, plo = 0.0, a = 0.0;
};

// Product of the factors {seed if SEED} + x[LO..HI) as four interleaved
// chains (independent work for the scheduler): F-1 multiplies for F factors.
template <int N, int LO, int HI, bool SEED>
__device__ __forceinline__ double chain_prod(const double (&x)[N], double seed) {
  constexpr int F = HI - LO + (SEED ? 1 : 0);
  static_assert(F >= 1, "empty product");
  double c[4] = {1.0, 1.0, 1.0, 1.0};
#pragma unroll
  for (int f = 0; f < F; ++f) {
    const double v = (SEED && f == 0) ? seed : x[LO + f - (SEED ? 1 : 0)];
    if (f < 4) c[f] = v;
    else c[f & 3] *= v;
  }
  if constexpr (F == 1) return c[0];
  else if constexpr (F == 2) return c[0] * c[1];
  else if constexpr (F == 3) return (c[0] * c[1]) * c[2];
  else return (c[0] * c[1]) * (c[2] * c[3]);
}

// Signed accumulation of the term q * last: one FMA for the sum, one for
// sum |term|.  The asm pins the |term| FMA to this step (left alone, the
// compiler defers it to the end of the group and keeps 2 registers per step
// live).
template <bool NEG>
__device__ __forceinline__ void add_term(double q, double last, LaneAcc& acc) {
  acc.p = __builtin_fma(NEG ? -q : q, last, acc.p);
  acc.a = __builtin_fma(__builtin_fabs(q), __builtin_fabs(last), acc.a);
  asm volatile("" : "+v"(acc.a));
}

// p += sign * prod(x);  a += |prod(x)|  (all N factors).
// kDDProd / kDDAll: the product is a double-double (TwoProd through FMA:
// hi*x = hi' + e exactly), two interleaved chains joined by a double-double
// multiply; the term enters (p, plo) through TwoSum.  In kDDAll each factor
// is the pair (x[j], xl[j]).
template <int N, bool NEG, int MODE>
__device__ __forceinline__ void accumulate_term(const double (&x)[N], const double (&xl)[N],
                                                LaneAcc& acc) {
  if constexpr (N == 1) {
    const double v = MODE == kDDAll ? x[0] + xl[0] : x[0];
    acc.p += NEG ? -v : v;
    acc.a += __builtin_fabs(v);
  } else if constexpr (MODE == kFast) {
    add_term<NEG>(chain_prod<N, 0, N - 1, false>(x, 0.0), x[N - 1], acc);
  } else {
    // chain 0 over even j, chain 1 over odd j
    double h0 = x[0], l0 = MODE == kDDAll ? xl[0] : 0.0;
    double h1 = x[1], l1 = MODE == kDDAll ? xl[1] : 0.0;
#pragma unroll
    for (int j = 2; j < N; ++j) {
      double& h = (j & 1) ? h1 : h0;
      double& l = (j & 1) ? l1 : l0;
      const double ph = h * x[j];
      double e = __builtin_fma(h, x[j], -ph);
      if constexpr (MODE == kDDAll) e = __builtin_fma(h, xl[j], e);
      l = __builtin_fma(l, x[j], e);
      h = ph;
    }
    // (h0, l0) * (h1, l1)
    double th = h0 * h1;
    double tl = __builtin_fma(h0, h1, -th);
    tl = __builtin_fma(h0, l1, __builtin_fma(l0, h1, tl));
    if (NEG) {
      th = -th;
      tl = -tl;
    }
    // (p, plo) += (th, tl)
    const double s = acc.p + th;
    const double bb = s - acc.p;
    const double err = (acc.p - (s - bb)) + (th - bb);
    acc.p = s;
    acc.plo += err + tl;
    acc.a += __builtin_fabs(th);
  }
}

// x[0..COUNT) += column.  kDDAll: (x, xl) += c through TwoSum; the rounding
// error of every addition is carried in xl (exact, never renormalised).
template <int N, int MODE, int COUNT>
__device__ __forceinline__ void add_column(double (&x)[N], double (&xl)[N], const double (&cv)[N]) {
#pragma unroll
  for (int j = 0; j < COUNT; ++j) {
    if constexpr (MODE == kDDAll) {
      const double s = x[j] + cv[j];
      const double bb = s - x[j];
      xl[j] += (x[j] - (s - bb)) + (cv[j] - bb);
      x[j] = s;
    } else {
      x[j] += cv[j];
    }
  }
}

// Walk shape: column l < 3 of the Gray walk touches rows [0, K_l) only.
template <int K0_, int K1_, int K2_>
struct Nest {
  static constexpr int K0 = K0_, K1 = K1_, K2 = K2_;
};

// Cached tail products of the nested variants (see the file comment).
// (kDDProd: the tails are double-double pairs, low parts in l1..l3.)
struct Tails {
  double t1 = 1.0, t2 = 1.0, t3 = 1.0;
  double l1 = 0.0, l2 = 0.0, l3 = 0.0;
};

// Rows touched by a column at Gray level `lv` (3: any column >= 3).
template <int N, class S>
constexpr int level_rows(int lv) {
  return lv == 0 ? S::K0 : lv == 1 ? S::K1 : lv == 2 ? S::K2 : N;
}

// Product x[LO..HI) times `seed` (SEED) -- `seed` itself for an empty range.
template <int N, int LO, int HI, bool SEED>
__device__ __forceinline__ double tail_prod(const double (&x)[N], double seed) {
  if constexpr (LO >= HI) {
    static_assert(SEED, "empty tail without a seed");
    return seed;
  } else {
    return chain_prod<N, LO, HI, SEED>(x, seed);
  }
}

// Double-double product (oh, ol) of x[LO..HI) times the seed (sh, sl) when
// SEED -- the seed itself for an empty range.  One TwoProd chain: (h, l) * x
// = h x (rounded) + (the FMA-exact error of h x + l x).
template <int N, int LO, int HI, bool SEED>
__device__ __forceinline__ void dd_tail_prod(const double (&x)[N], double sh, double sl,
                                             double& oh, double& ol) {
  if constexpr (LO >= HI) {
    static_assert(SEED, "empty tail without a seed");
    oh = sh;
    ol = sl;
  } else {
    double h = SEED ? sh : x[LO], l = SEED ? sl : 0.0;
#pragma unroll
    for (int j = SEED ? LO : LO + 1; j < HI; ++j) {
      const double ph = h * x[j];
      const double e = __builtin_fma(h, x[j], -ph);
      l = __builtin_fma(l, x[j], e);
      h = ph;
    }
    oh = h;
    ol = l;
  }
}

// (p, plo) += sign * (th, tl) through TwoSum; a += |th|.
template <bool NEG>
__device__ __forceinline__ void dd_add_term(double th, double tl, LaneAcc& acc) {
  if (NEG) {
    th = -th;
    tl = -tl;
  }
  const double s = acc.p + th;
  const double bb = s - acc.p;
  const double err = (acc.p - (s - bb)) + (th - bb);
  acc.p = s;
  acc.plo += err + tl;
  acc.a += __builtin_fabs(th);
}

// Step v of group g (step index t = g*2^R + v).  Phase A adds the column
// that the previous step prefetched into SGPRs (the rows its level touches);
// phase B issues the scalar loads of the next step's column and, while they
// are in flight, refreshes the tail products the step's level invalidated,
// forms the product and accumulates the term.  sched_barrier keeps the
// phases apart so the loads always overlap a product.  NOADD0: step 0 of the
// group adds nothing (x already holds its subset; cv is not read).
template <int N, int R, int MODE, class S, int V, bool NOADD0 = false>
__device__ __forceinline__ void group_steps(double (&x)[N], double (&xl)[N], double (&cv)[N],
                                            ColPtr cols, uint32_t g, uint32_t groups,
                                            int zero_blk, LaneAcc& acc, Tails& tl) {
  constexpr int K0 = S::K0, K1 = S::K1, K2 = S::K2;
  constexpr bool kNested = K0 < N;
  if constexpr (V < (1 << R)) {
    constexpr int lv = V == 0 ? 3 : __builtin_ctz(V);
    if constexpr (NOADD0 && V == 0) {
    } else {
      add_column<N, MODE, level_rows<N, S>(lv)>(x, xl, cv);
    }
    __builtin_amdgcn_sched_barrier(0);
    int blk;
    if constexpr (V + 1 < (1 << R)) {
      // Next step inside the group: column k = ctz(v+1); the new gray bit k
      // is set iff bit k+1 of the step index is clear.  For k+1 < R that
      // bit is inside the group, for k+1 == R it is bit 0 of g.
      constexpr int k = __builtin_ctz(V + 1);
      int neg;
      if constexpr (k + 1 < R) neg = ((V + 1) >> (k + 1)) & 1;
      else neg = (int)(g & 1u);
      blk = 2 * k + neg;
      load_col<N, level_rows<N, S>(k)>(cv, cols, blk);
    } else {
      // Next step opens group g+1: column R + ctz(g+1).  Past the last group
      // the (unused) prefetch reads the zero column to stay in bounds.
      const uint32_t gn = g + 1;
      const int c = __builtin_ctz(gn);
      const int neg = (int)((gn >> (c + 1)) & 1u);
      blk = gn < groups ? 2 * (R + c) + neg : zero_blk;
      load_col<N, N>(cv, cols, blk);
    }
    if constexpr (kNested && MODE == kDDProd) {
      if constexpr (lv >= 3 && K2 < N) dd_tail_prod<N, K2, N, false>(x, 0.0, 0.0, tl.t3, tl.l3);
      if constexpr (lv >= 2 && K1 < N)
        dd_tail_prod<N, K1, K2, (K2 < N)>(x, tl.t3, tl.l3, tl.t2, tl.l2);
      if constexpr (lv >= 1) dd_tail_prod<N, K0, K1, (K1 < N)>(x, tl.t2, tl.l2, tl.t1, tl.l1);
      double th, tlo;
      dd_tail_prod<N, 0, K0, true>(x, tl.t1, tl.l1, th, tlo);
      dd_add_term<(V & 1) != 0>(th, tlo, acc);
    } else if constexpr (kNested) {
      if constexpr (lv >= 3 && K2 < N) tl.t3 = tail_prod<N, K2, N, false>(x, 0.0);
      if constexpr (lv >= 2 && K1 < N) tl.t2 = tail_prod<N, K1, K2, (K2 < N)>(x, tl.t3);
      if constexpr (lv >= 1) tl.t1 = tail_prod<N, K0, K1, (K1 < N)>(x, tl.t2);
      add_term<(V & 1) != 0>(chain_prod<N, 0, K0 - 1, true>(x, tl.t1), x[K0 - 1], acc);
    } else {
      accumulate_term<N, (V & 1) != 0, MODE>(x, xl, acc);
    }
    __builtin_amdgcn_sched_barrier(0);
    group_steps<N, R, MODE, S, V + 1>(x, xl, cv, cols, g, groups, zero_blk, acc, tl);
  }
}

// The Gray walk of one lane over its groups.  Every 32 groups (a run of
// 2^(R+5) terms) the FP64 lane sum is folded into the double-double s, the
// square of the run's sum of |terms| is added to q and (FP64 mode) the
// square of its signed sum to r.
template <int N, int R, int MODE, class S>
__device__ __forceinline__ void walk_groups(double (&x)[N], double (&xl)[N], ColPtr cols,
                                            uint32_t groups, int zero_blk, DD& s, LaneAcc& acc,
                                            double& q, double& r) {
  Tails tl;
  double a_run = 0.0;  // acc.a at the start of the current run
  double cv[N];
  load_col<N, N>(cv, cols, zero_blk);  // step 0 of the walk changes nothing
  for (uint32_t g = 0; g < groups; ++g) {
    group_steps<N, R, MODE, S, 0>(x, xl, cv, cols, g, groups, zero_blk, acc, tl);
    if ((g & 31u) == 31u) {
      if constexpr (MODE == kFast) {  // fold into double-double
        s = dd_add_d(s, acc.p);
        r = __builtin_fma(acc.p, acc.p, r);
        acc.p = 0.0;
      }
      const double d = acc.a - a_run;
      q = __builtin_fma(d, d, q);
      a_run = acc.a;
    }
  }
  const double d = acc.a - a_run;
  q = __builtin_fma(d, d, q);
}

// The walk of one lane: closed-form start from its high columns, then 2^w
// subsets of its low columns.  Returns the lane's signed double-double sum
// (sign of the high pattern applied) and its moments: sum of |terms|, sums
// over runs of the squared absolute and signed run sums (see walk_groups).
// colpm: 2(n-1)+1 blocks of N doubles: block 2k = +a_k, 2k+1 = -a_k for the
// Gray columns k < n-1, block 2(n-1) = zeros.  Shape S: column l < 3 is zero
// in rows [K_l, N).
template <int N, int R, int MODE, class S>
__device__ __forceinline__ void lane_walk(const double* __restrict__ colpm,
                                          const double* __restrict__ x0, uint64_t pattern,
                                          int n_walk_bits, int n_high_bits, DD& s,
                                          Moments& m) {
  static_assert(S::K0 <= S::K1 && S::K1 <= S::K2 && S::K2 <= N, "nested shape");
  static_assert(S::K0 == N || (MODE != kDDAll && R >= 1 && S::K0 >= 1),
                "nested: FP64 or double-double products");
  static_assert(S::K1 == N || R >= 2, "level-1 tail needs a 4-step group");
  static_assert(S::K2 == N || R >= 3, "level-2 tail needs an 8-step group");
  const ColPtr cols = (ColPtr)colpm;
  const int zero_blk = 2 * (N - 1);

  double x[N], xl[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    x[j] = x0[j];
    // kDDAll starts from the double-double start vector (low parts after the
    // n high ones); the other modes keep X in FP64
    xl[j] = MODE == kDDAll ? x0[N + j] : 0.0;
  }

  // Closed-form start: add the high columns selected by the pattern.  Bits
  // 0..5 differ between the lanes of a wave: adding sel*c with sel in {0, 1}
  // is exact selection, no branch.  Bits 6 and up are the same in every lane
  // of a wave (waves cover aligned runs of 64 lane ids, and pattern 0 past
  // the last lane only occurs when that run is all 0 above bit 5), so those
  // columns are added under a wave-uniform branch; skipping a column leaves
  // x as adding its zero multiple does (x + 0 = x, up to the sign of a zero).
  int parity = 0;
  const int lane_bits = n_high_bits < 6 ? n_high_bits : 6;
  for (int b = 0; b < lane_bits; ++b) {
    ColPtr c = cols + (size_t)(2 * (n_walk_bits + b)) * N;
    const int bit = (int)((pattern >> b) & 1);
    parity ^= bit;
    const double sel = bit ? 1.0 : 0.0;
    double cv[N];
#pragma unroll
    for (int j = 0; j < N; ++j) cv[j] = sel * c[j];
    add_column<N, MODE, N>(x, xl, cv);
  }
  // (readfirstlane returns int: each half goes back through uint32_t, or a
  // set bit 31 -- pattern bit 37 -- would sign-extend over the upper half)
  const uint64_t wave_bits =
      ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(pattern >> 38)) << 32) |
      (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(pattern >> 6));
  for (int b = lane_bits; b < n_high_bits; ++b) {
    if (!((wave_bits >> (b - 6)) & 1)) continue;
    parity ^= 1;
    ColPtr c = cols + (size_t)(2 * (n_walk_bits + b)) * N;
    double cv[N];
#pragma unroll
    for (int j = 0; j < N; ++j) cv[j] = c[j];
    add_column<N, MODE, N>(x, xl, cv);
  }

  s = {0.0, 0.0};
  m = {0.0, 0.0, 0.0};
  LaneAcc acc;
  const uint32_t groups = 1u << (n_walk_bits - R);
  walk_groups<N, R, MODE, S>(x, xl, cols, groups, zero_blk, s, acc, m.q, m.r);
  if constexpr (MODE == kFast) {  // the last (partial) run
    s = dd_add_d(s, acc.p);
    m.r = __builtin_fma(acc.p, acc.p, m.r);
  } else {
    s = two_sum(acc.p, acc.plo);
  }
  m.a = acc.a;
  if (parity) {
    s.hi = -s.hi;
    s.lo = -s.lo;
  }
}

// One pass (or a block range of it) of one permanent: lane = lane_base +
// global thread id, one partial per workgroup.
template <int N, int R, int MODE, class S>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MODE == kFast ? kFastWavesPerEU : 1)))
void ryser_dense_kernel(const double* __restrict__ colpm, const double* __restrict__ x0,
                        Partial* __restrict__ partials, uint64_t lane_base, uint64_t lanes,
                        int n_walk_bits, int n_high_bits, const uint32_t* __restrict__ block_ids) {
  __shared__ Partial scratch[256 / kWave];
  // shader-clock and 100 MHz real-time counters around the workgroup's work:
  // their ratio is the clock the workgroup ran at (scalar reads, per wave)
  const uint64_t c0 = __builtin_amdgcn_s_memtime();
  const uint64_t r0 = __builtin_amdgcn_s_memrealtime();
  // workgroup b walks lanes [lane_base + 256 b, +256); b = the listed index
  // when a list is given (the record still goes to slot blockIdx.x)
  const uint64_t b = block_ids ? (uint64_t)block_ids[blockIdx.x] : (uint64_t)blockIdx.x;
  const uint64_t lane = lane_base + b * blockDim.x + threadIdx.x;
  const uint64_t pattern = lane < lanes ? lane : 0ull;  // high columns fixed by the lane id
  DD s;
  Moments m;
  lane_walk<N, R, MODE, S>(colpm, x0, pattern, n_walk_bits, n_high_bits, s, m);
  if (lane >= lanes) {
    s = {0.0, 0.0};
    m = {0.0, 0.0, 0.0};
  }
  // cycles per 10 ns tick over the walk (thread 0's); the grid reduction
  // sums it (pad) over workgroups
  const uint64_t c1 = __builtin_amdgcn_s_memtime();
  const uint64_t r1 = __builtin_amdgcn_s_memrealtime();
  const double ratio = r1 > r0 ? (double)(c1 - c0) / (double)(r1 - r0) : 0.0;
  block_reduce_store(s, m, scratch, partials + blockIdx.x, ratio);
}

template <int N>
constexpr int unroll_bits_for() {
  return N - 1 < 3 ? N - 1 : 3;
}

// Workgroups [block_begin, block_end) of pass P (all of them when
// block_end == 0); partial of workgroup b goes to
// P.partials[b - P.record_begin].
template <int N, int MODE, class S>
hipError_t launch_pass(const DenseLaunch& L, const LanePass& P, hipStream_t stream) {
  constexpr int R = unroll_bits_for<N>();
  if (P.lanes == 0) return hipSuccess;
  if (P.walk_bits < R) return hipErrorInvalidValue;
  if (P.block_ids) {  // listed workgroups (indices checked by the caller)
    if (P.block_count == 0) return hipSuccess;
    ryser_dense_kernel<N, R, MODE, S><<<dim3(P.block_count), dim3(256), 0, stream>>>(
        L.colpm, L.x0, P.partials, 0, P.lanes, P.walk_bits, P.high_bits, P.block_ids);
    return hipGetLastError();
  }
  const uint64_t b0 = P.block_begin, b1 = P.block_end ? P.block_end : P.blocks;
  if (b1 <= b0) return hipSuccess;
  if (b1 > P.blocks || P.record_begin > b0 || b1 - b0 > 0xffffffffull)
    return hipErrorInvalidValue;
  ryser_dense_kernel<N, R, MODE, S><<<dim3((unsigned)(b1 - b0)), dim3(256), 0, stream>>>(
      L.colpm, L.x0, P.partials + (b0 - P.record_begin), b0 * 256, P.lanes, P.walk_bits,
      P.high_bits, nullptr);
  return hipGetLastError();
}

// One run: the kernel of mode MODE and shape S.
template <int N, int MODE, class S>
hipError_t launch_mode(const DenseLaunch& L, hipStream_t stream, hipEvent_t main_done) {
  hipError_t e = launch_pass<N, MODE, S>(L, L.main, stream);
  if (e == hipSuccess && main_done) e = hipEventRecord(main_done, stream);
  return e;
}

// Shape j of K0 = 4(q+1) in the compiled menu (nest_menu_shape).
template <int N, int MODE, int Q, int J>
hipError_t launch_shape(const DenseLaunch& L, hipStream_t stream, hipEvent_t main_done,
                        bool& hit) {
  constexpr int K0 = 4 * (Q + 1);
  constexpr NestShape sh = nest_menu_shape(N, K0, J);
  if constexpr (sh.k0 == 0) {
    return hipErrorInvalidValue;
  } else {
    if (L.nu != sh.k0 || L.nu1 != sh.k1 || L.nu2 != sh.k2) return hipErrorInvalidValue;
    hit = true;
    return launch_mode<N, MODE, Nest<sh.k0, sh.k1, sh.k2>>(L, stream, main_done);
  }
}

template <int N, int MODE, int Q, int... J>
hipError_t launch_k0(const DenseLaunch& L, hipStream_t stream, hipEvent_t main_done, bool& hit,
                     std::integer_sequence<int, J...>) {
  hipError_t e = hipErrorInvalidValue;
  ((hit ? 0 : (e = launch_shape<N, MODE, Q, J>(L, stream, main_done, hit), 0)), ...);
  return e;
}

template <int N, int MODE, int... Q>
hipError_t launch_nested(const DenseLaunch& L, hipStream_t stream, hipEvent_t main_done,
                         std::integer_sequence<int, Q...>) {
  hipError_t e = hipErrorInvalidValue;
  bool hit = false;
  ((L.nu == 4 * (Q + 1) && !hit
        ? (e = launch_k0<N, MODE, Q>(L, stream, main_done, hit,
                                     std::make_integer_sequence<int, kNestShapesPerK0>{}),
           0)
        : 0),
   ...);
  return hit ? e : hipErrorInvalidValue;
}

template <int N, int MODE>
hipError_t launch_shaped(const DenseLaunch& L, hipStream_t stream, hipEvent_t main_done) {
  if (L.nu == N || L.nu <= 0) return launch_mode<N, MODE, Nest<N, N, N>>(L, stream, main_done);
  if constexpr (N >= kMinPairN && N <= kMaxNestN) {
    if (L.nu % 4 == 0 && L.nu < N)
      return launch_nested<N, MODE>(L, stream, main_done,
                                    std::make_integer_sequence<int, (N - 1) / 4>{});
  }
  return hipErrorInvalidValue;
}

// kFast and kDDProd run the walk shape in L (nu, nu1, nu2); kDDAll the plain
// walk.
template <int N>
hipError_t launch_n(const DenseLaunch& L, hipStream_t stream, hipEvent_t main_done) {
  if (L.mode == kDDAll) return launch_mode<N, kDDAll, Nest<N, N, N>>(L, stream, main_done);
  if (L.mode == kDDProd) return launch_shaped<N, kDDProd>(L, stream, main_done);
  if (L.mode == kFast) return launch_shaped<N, kFast>(L, stream, main_done);
  return hipErrorInvalidValue;
}

}  // namespace perm

// Instantiate launch_n for a list of N in one translation unit; the
// dispatcher in ryser_dense.hip calls launch_dense_range_<TAG>.
#define PERM_DENSE_RANGE(TAG, LO, HI)                                                         \
  namespace perm {                                                                            \
  template <int... I>                                                                         \
  static hipError_t dispatch_##TAG(const DenseLaunch& L, hipStream_t s, hipEvent_t ev,        \
                                   std::integer_sequence<int, I...>) {                        \
    hipError_t e = hipErrorInvalidValue;                                                      \
    ((L.n == LO + I ? (e = launch_n<LO + I>(L, s, ev), 0) : 0), ...);                         \
    return e;                                                                                 \
  }                                                                                           \
  hipError_t launch_dense_range_##TAG(const DenseLaunch& L, hipStream_t s, hipEvent_t ev) {   \
    return dispatch_##TAG(L, s, ev, std::make_integer_sequence<int, HI - LO + 1>{});          \
  }                                                                                           \
  }
