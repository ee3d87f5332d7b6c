// ryser_skip_impl.hpp -- the Ryser walk with zero-run skipping (the
// reference's sparse skipper, gpu_exact_sparse.cu:555-670 and :1123-1190,
// CLI sparse ID 7), designed for wave-uniform lanes.
//
// For an integer matrix every x_j is exact, so a term is zero exactly when
// some x_j(S) is zero.  The reference's skipper lets each thread jump past
// the Gray steps that cannot change a zero row.  Lanes here must keep
// flipping the same column, so the skip is taken per wave and per aligned
// block of the walk:
//
//  * Within an aligned block of 2^m steps [t, t + 2^m) only the walk columns
//    [0, m) change.  A row r whose first entry among the walk columns is at
//    column lev[r] >= m keeps its x_r throughout the block; if x_r is zero at
//    the block's first subset, all 2^m terms of the block are zero.
//  * The host orders the rows by lev, ascending.  Every 2^S steps each lane
//    finds the level of its highest zero row (the largest level it could
//    skip; rows are in level order, so it is the largest level among its
//    zero rows), the wave takes the minimum over its lanes (six ballots of a
//    binary search: no LDS traffic, no memory load on the path), and skips
//    the block of level m = min(that level, alignment of t) if m >= S: two
//    column additions replace 2^m steps (in Gray order the subset before the
//    next block differs from the block's first subset only in column m-1;
//    the next block's first step flips column m + ctz(t/2^m + 1)).  The row
//    levels are packed four to a 32-bit word and held in SGPRs.
//  * Otherwise the wave walks 2^S steps with the plain kernel's unrolled
//    groups (ryser_dense_impl.hpp) and checks again.
//
// The host orders the walk columns so the low ones touch few rows (greedy:
// next the column adding the fewest new rows), which makes lev large for
// most rows.  Skipped terms are exactly zero, so the sum, its rounding and
// the stated error are those of the plain walk over the same column order.
#pragma once
#include "ryser_dense_impl.hpp"

namespace perm {

typedef const uint32_t __attribute__((address_space(4)))* LevPtr;

// Minimum over the wave of v in [0, 64) (every lane active): the largest t
// with v >= t in every lane, by binary search on ballots.
__device__ __forceinline__ uint32_t wave_min_u6(uint32_t v) {
  const uint64_t all = __builtin_amdgcn_ballot_w64(true);
  uint32_t lo = 0;
#pragma unroll
  for (int b = 5; b >= 0; --b) {
    const uint32_t t = lo | (1u << b);
    if (__builtin_amdgcn_ballot_w64(v >= t) == all) lo = t;
  }
  return lo;
}

template <int N>
__device__ __forceinline__ void load_col_n(double (&cv)[N], ColPtr base, int blk) {
  load_col<N, N>(cv, base, blk);
}

// Column block of step t (t >= 1) of the Gray walk: column ctz(t), added
// when bit ctz(t)+1 of t is clear, else subtracted.
__device__ __forceinline__ int step_block(uint32_t t) {
  const int c = __builtin_ctz(t);
  return 2 * c + (int)((t >> (c + 1)) & 1u);
}

// The walk of one wave over steps [t_begin, t_begin + 2^t_bits) (t_begin a
// multiple of 2^t_bits), with skips.  budget > 0: the wave stops once it has
// walked `budget` groups of 2^R steps, at the start of its next check; stop_t
// is then that step (0 when the range was finished) and the rest of the
// range is left to later waves (a continuation starts from the closed form
// of x at its first step).  PIECES = false: the whole walk without a budget
// (t_begin = 0, t_bits = n_walk_bits, budget = 0 fixed at compile time, so
// the loop carries none of the range bookkeeping).  n_walk_bits < 32: every
// step index fits 32 bits.
template <int N, int R, int MODE, bool PIECES>
__device__ __forceinline__ void skip_walk(const double* __restrict__ colpm,
                                          const double* __restrict__ x0,
                                          const uint32_t* __restrict__ row_level,
                                          uint64_t pattern, bool valid, int n_walk_bits,
                                          int n_high_bits, int check_bits, uint32_t t_begin_arg,
                                          int t_bits_arg, uint32_t budget_arg, DD& s, double& a,
                                          double& q, uint32_t& walked_groups, uint32_t& stop_t) {
  const uint32_t t_begin = PIECES ? t_begin_arg : 0u;
  const int t_bits = PIECES ? t_bits_arg : n_walk_bits;
  const uint32_t budget = PIECES ? budget_arg : 0u;
  const ColPtr cols = (ColPtr)colpm;
  // row levels, byte j of word j/4 = level of row j (scalar loads, SGPRs)
  uint32_t lev[(N + 3) / 4];
  {
    const LevPtr lp = (LevPtr)row_level;
#pragma unroll
    for (int i = 0; i < (N + 3) / 4; ++i) lev[i] = lp[i];
  }
  const int zero_blk = 2 * (N - 1);
  double x[N], xl[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    x[j] = x0[j];
    xl[j] = 0.0;
  }
  int parity = 0;
  for (int b = 0; b < n_high_bits; ++b) {
    ColPtr c = cols + (size_t)(2 * (n_walk_bits + b)) * N;
    const int bit = (int)((pattern >> b) & 1);
    parity ^= bit;
    const double sel = bit ? 1.0 : 0.0;
    double cv[N];
#pragma unroll
    for (int j = 0; j < N; ++j) cv[j] = sel * c[j];
    add_column<N, MODE, N>(x, xl, cv);
  }
  double cv[N];
  if (t_begin > 0) {
    // x at step t_begin - 1: the walk columns in gray(t_begin - 1), added
    // (wave-uniform: the same for every lane)
    const uint32_t gp = (t_begin - 1) ^ ((t_begin - 1) >> 1);
    for (int b = 0; b < n_walk_bits; ++b)
      if ((gp >> b) & 1u) {
        load_col_n<N>(cv, cols, 2 * b);
        add_column<N, MODE, N>(x, xl, cv);
      }
  }

  s = {0.0, 0.0};
  q = 0.0;
  walked_groups = 0;
  stop_t = 0;
  LaneAcc acc;
  const uint32_t t_end = t_begin + (1u << t_bits);
  const uint32_t groups = t_end >> R;
  const uint32_t chunk_groups = 1u << (check_bits - R);
  Tails tl;
  double a_run = 0.0;
  // the change of step t_begin (step 0 changes nothing)
  load_col_n<N>(cv, cols, t_begin == 0 ? zero_blk : step_block(t_begin));
  uint32_t t = t_begin;
  while (t < t_end) {
    // x = x(gray(t-1)), cv = the change of step t: apply it, x = x(gray(t)).
    add_column<N, MODE, N>(x, xl, cv);
    // 1 + level of this lane's highest zero row (rows are in ascending
    // level order), 0 without a zero row; lanes past the walk do not limit
    uint32_t ml = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const uint32_t lj = ((lev[j >> 2] >> (8 * (j & 3))) & 0xffu) + 1u;
      ml = x[j] == 0.0 ? lj : ml;
    }
    if (!valid) ml = 63;
    const int align = t == t_begin ? t_bits : __builtin_ctz(t);
    int m = (int)wave_min_u6(ml) - 1;
    m = m < align ? m : align;
    if (m >= check_bits) {
      // skip [t, t + 2^m): x <- x(gray(t + 2^m - 1)), cv <- step t + 2^m
      const uint32_t b = t >> m;
      load_col_n<N>(cv, cols, 2 * (m - 1) + (int)(b & 1u));  // bit m-1 becomes !b_0
      add_column<N, MODE, N>(x, xl, cv);
      t += 1u << m;
      load_col_n<N>(cv, cols, t < t_end ? step_block(t) : zero_blk);
      continue;
    }
    // walk 2^check_bits steps; the first adds nothing (x is already at t)
    const uint32_t g0 = t >> R;
    group_steps<N, R, MODE, Nest<N, N, N>, 0, true>(x, xl, cv, cols, g0, groups, zero_blk, acc,
                                                   tl);
    for (uint32_t g = g0 + 1; g < g0 + chunk_groups; ++g) {
      group_steps<N, R, MODE, Nest<N, N, N>, 0>(x, xl, cv, cols, g, groups, zero_blk, acc, tl);
      if constexpr (MODE == kFast) {
        if ((g & 31u) == 31u) {  // at least every 32 groups
          s = dd_add_d(s, acc.p);
          acc.p = 0.0;
        }
      }
    }
    if constexpr (MODE == kFast) {  // and at the end of every chunk
      s = dd_add_d(s, acc.p);
      acc.p = 0.0;
    }
    walked_groups += chunk_groups;
    const double d = acc.a - a_run;
    q = __builtin_fma(d, d, q);
    a_run = acc.a;
    t += 1u << check_bits;
    if (PIECES && budget != 0 && walked_groups >= budget && t < t_end) {
      stop_t = t;  // the rest of the range is a continuation's
      break;
    }
  }
  if constexpr (MODE == kFast) s = dd_add_d(s, acc.p);
  else s = dd_add(s, two_sum(acc.p, acc.plo));
  a = acc.a;
  if (parity) {
    s.hi = -s.hi;
    s.lo = -s.lo;
  }
}

// One skipping walk per lane.  PIECES = false: no work list and no budget
// (the arguments are ignored).  Without a work list (work == nullptr): the
// plain split, lane = global thread id, the whole walk [0, 2^walk_bits),
// stopping after `budget` walked groups (cont[wave] = the step it stopped
// at, 0 if it finished).  With a work list: wave j of the launch walks
// entry j = (wave w of the plain split, first step, log2 steps) for the 64
// lanes of wave w -- the continuations, to the end of their range.
// partials: one record per workgroup; walked: groups of 2^R steps walked per
// wave (the rest was skipped), one slot per wave, for the statistics.
template <int N, int R, int MODE, bool PIECES>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MODE == kFast ? kFastWavesPerEU : 1)))
void ryser_skip_kernel(const double* __restrict__ colpm, const double* __restrict__ x0,
                       const uint32_t* __restrict__ row_level, Partial* __restrict__ partials,
                       uint32_t* __restrict__ walked, uint64_t lanes, int n_walk_bits,
                       int n_high_bits, int check_bits, int lane_reverse,
                       const uint32_t* __restrict__ work, uint32_t n_work, uint32_t budget,
                       uint32_t* __restrict__ cont) {
  __shared__ Partial scratch[256 / kWave];
  const uint32_t wave_g = blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
  const uint32_t lane_w = threadIdx.x & (kWave - 1);
  uint64_t lane;
  uint32_t t_begin = 0;
  int t_bits = n_walk_bits;
  bool has = true;
  if (PIECES && work) {
    has = wave_g < n_work;
    const uint32_t e = has ? wave_g : 0u;
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)work[4 * e]);
    t_begin = (uint32_t)__builtin_amdgcn_readfirstlane((int)work[4 * e + 1]);
    t_bits = __builtin_amdgcn_readfirstlane((int)work[4 * e + 2]);
    lane = (uint64_t)w * kWave + lane_w;
  } else {
    lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  }
  const bool valid = has && lane < lanes;
  // lanes of one wave differ in the low bits of the lane id: reversed, they
  // differ in the last high columns instead
  uint64_t pattern = valid ? lane : 0ull;
  if (lane_reverse && n_high_bits > 0)
    pattern = __builtin_bitreverse64(pattern) >> (64 - n_high_bits);
  DD s = {0.0, 0.0};
  double a = 0.0, q = 0.0;
  uint32_t wg = 0, stop_t = 0;
  if (has)
    skip_walk<N, R, MODE, PIECES>(colpm, x0, row_level, pattern, valid, n_walk_bits, n_high_bits,
                                  check_bits, t_begin, t_bits, work ? 0u : budget, s, a, q, wg,
                                  stop_t);
  if (!valid) {
    s = {0.0, 0.0};
    a = 0.0;
    q = 0.0;
  }
  // integer matrices only: X is exact, no drift statistic (sq_run = 0)
  block_reduce_store(s, Moments{a, q, 0.0}, scratch, partials + blockIdx.x);
  if (lane_w == 0) {
    walked[wave_g] = wg;
    if (PIECES && cont) cont[wave_g] = stop_t;
  }
}

template <int N>
hipError_t launch_skip_n(const SkipLaunch& L, hipStream_t stream) {
  constexpr int R = unroll_bits_for<N>();
  if (L.walk_bits < R || L.check_bits < R || L.check_bits > L.walk_bits)
    return hipErrorInvalidValue;
  if (L.walk_bits >= 32) return hipErrorInvalidValue;
  const bool pieces = L.work != nullptr || L.budget != 0;
#define PERM_SKIP_LAUNCH(MODE_, PIECES_)                                                   \
  ryser_skip_kernel<N, R, MODE_, PIECES_><<<dim3((unsigned)L.blocks), dim3(256), 0, stream>>>( \
      L.colpm, L.x0, L.row_level, L.partials, L.walked, L.lanes, L.walk_bits, L.high_bits,  \
      L.check_bits, L.lane_reverse, L.work, L.n_work, L.budget, L.cont)
  if (L.mode == kFast) {
    if (pieces) PERM_SKIP_LAUNCH(kFast, true);
    else PERM_SKIP_LAUNCH(kFast, false);
  } else if (L.mode == kDDProd) {
    if (pieces) PERM_SKIP_LAUNCH(kDDProd, true);
    else PERM_SKIP_LAUNCH(kDDProd, false);
  } else {
    return hipErrorInvalidValue;
  }
#undef PERM_SKIP_LAUNCH
  return hipGetLastError();
}

}  // namespace perm

#define PERM_SKIP_RANGE(TAG, LO, HI)                                                          \
  namespace perm {                                                                            \
  template <int... I>                                                                         \
  static hipError_t skip_dispatch_##TAG(const SkipLaunch& L, hipStream_t s,                   \
                                        std::integer_sequence<int, I...>) {                   \
    hipError_t e = hipErrorInvalidValue;                                                      \
    ((L.n == LO + I ? (e = launch_skip_n<LO + I>(L, s), 0) : 0), ...);                        \
    return e;                                                                                 \
  }                                                                                           \
  hipError_t launch_skip_range_##TAG(const SkipLaunch& L, hipStream_t s) {                    \
    return skip_dispatch_##TAG(L, s, std::make_integer_sequence<int, HI - LO + 1>{});         \
  }                                                                                           \
  }
