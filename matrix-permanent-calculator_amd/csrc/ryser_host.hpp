// ryser_host.hpp -- host-side preparation shared by every entry point:
// Nijenhuis-Wilf start vector, signed column store, work split, and the
// error statement attached to each result.
#pragma once
#include <stdint.h>

#include <cmath>
#include <vector>

namespace perm {

// Start vector of the Nijenhuis-Wilf form, x0_j = a[j][n-1] - rs_j/2 with
// rs_j = sum_k a[j][k] (reference: gpu_exact_dense.cu:647-654; sparse
// gpu_exact_sparse.cu:861-871, whose loop skips zero entries and so has the
// same row sum).  The reference accumulates rs_j in plain double; here it is
// summed with TwoSum compensation and x0_j rounded once, so x0_j carries at
// most half an ulp of error.  That matters: the sum's derivative with
// respect to x0_j is a minor of the matrix, so an error d_j in x0_j moves
// the result by about d_j * perm / a[j][n-1] -- with the plain row sum about
// n u |perm| in total, here about sqrt(n) u |perm|.  Integer matrices get
// the exact x0 either way.
inline void nw_start_vector(const double* a, int n, double* x0) {
  for (int j = 0; j < n; ++j) {
    double hi = 0.0, lo = 0.0;
    for (int k = 0; k < n; ++k) {
      const double v = a[(size_t)j * n + k];
      const double s = hi + v;
      const double bb = s - hi;
      lo += (hi - (s - bb)) + (v - bb);
      hi = s;
    }
    // x0 = a[j][n-1] - hi/2 - lo/2 (halving is exact), one final rounding
    const double c = a[(size_t)j * n + (n - 1)];
    const double h = -0.5 * hi;
    const double s = c + h;
    const double bb = s - c;
    const double e = (c - (s - bb)) + (h - bb);
    x0[j] = s + (e - 0.5 * lo);
  }
}

// Signed column store for the Gray walk over columns 0..n-2:
// 2(n-1)+1 blocks of n doubles, block 2k = +a[.][k], block 2k+1 = -a[.][k],
// last block zeros.  `col(k, j)` supplies a[j][k].
template <class ColFn>
inline void build_column_store(int n, ColFn col, std::vector<double>& out) {
  const int B = n - 1;
  out.assign((size_t)(2 * B + 1) * n, 0.0);
  for (int k = 0; k < B; ++k)
    for (int j = 0; j < n; ++j) {
      const double v = col(k, j);
      out[(size_t)(2 * k) * n + j] = v;
      out[(size_t)(2 * k + 1) * n + j] = -v;
    }
}

// Work split of the 2^(n-1) Gray subsets: every lane walks 2^walk_bits
// subsets; 2^high_bits lanes.  Default: lanes = 2^min(21, n-11) (2^21
// lanes fill the 256 CUs of an MI355X eight times over at 4 waves/SIMD),
// with at least 2^10 steps per lane so the closed-form lane start stays a
// small fraction of the walk.
struct Split {
  int walk_bits;
  int high_bits;
};

inline Split choose_split(int n, int unroll_bits, int walk_override) {
  const int B = n - 1;
  int h = B - 10;
  if (h > 21) h = 21;
  if (h < 0) h = 0;
  int w = B - h;
  if (walk_override > 0) {
    w = walk_override > B ? B : walk_override;
    h = B - w;
  }
  if (w < unroll_bits) {
    w = unroll_bits;
    h = B - w;
  }
  return {w, h};
}

// Extreme-subset cut of the FP64 walk (summed in double-double instead, see
// ryser_extreme.hip): the largest k with 2 sum_{i<=k} C(n-1, i) subsets
// within min(2^21, 2^(n-7)) and 2k < n-1, for n >= 24; -1 (no cut) below.
// `override_k` >= -1 forces a value (-1: off); -2 means automatic.
constexpr int kMinExtN = 24;

inline uint64_t binom_u64(int c, int p) {
  if (p < 0 || p > c) return 0;
  uint64_t r = 1;
  for (int i = 1; i <= p; ++i) r = r * (uint64_t)(c - p + i) / (uint64_t)i;
  return r;
}

inline uint64_t extreme_per_side(int n, int k) {
  uint64_t m = 0;
  for (int i = 0; i <= k; ++i) m += binom_u64(n - 1, i);
  return m;
}

inline int choose_ext_k(int n, int max_k, int override_k) {
  const int B = n - 1;
  if (override_k >= -1) {
    int k = override_k > max_k ? max_k : override_k;
    while (k >= 0 && 2 * k >= B) --k;
    return k;
  }
  if (n < kMinExtN) return -1;
  const uint64_t budget = 1ull << (B - 6 < 21 ? B - 6 : 21);
  int k = -1;
  while (k + 1 <= max_k && 2 * (k + 1) < B && 2 * extreme_per_side(n, k + 1) <= budget) ++k;
  return k;
}

// Lanes of the two FP64 passes under an extreme-subset cut k, for the split
// (w, h) of n-1 = w + h columns.  A lane whose pattern has popcount c walks
// subsets of sizes c .. c+w, so it can meet an extreme subset iff c <= k or
// c + w >= n-1-k ("side" classes); the other classes form the main pass.
// The side pass splits each side lane's walk into 2^free_bits lanes (free
// bits taken from the top of the walk) so that its few lanes still fill the
// GPU.
struct PassGeometry {
  uint64_t lanes, class_mask;
  int walk_bits, high_bits, free_bits;
};

inline void cut_passes(int n, int w, int h, int k, int unroll_bits, PassGeometry& main,
                       PassGeometry& side) {
  const int B = n - 1;
  uint64_t mid_mask = 0, side_mask = 0, mid = 0, sd = 0;
  for (int c = 0; c <= h; ++c) {
    if (c <= k || c + w >= B - k) {
      side_mask |= 1ull << c;
      sd += binom_u64(h, c);
    } else {
      mid_mask |= 1ull << c;
      mid += binom_u64(h, c);
    }
  }
  int f = w - unroll_bits;
  if (f > 6) f = 6;
  if (f < 0) f = 0;
  main = {mid, mid_mask, w, h, 0};
  side = {sd << f, side_mask, w - f, h + f, f};
}

// x_all = x0 + sum of all walk columns (block 2c of the column store), per
// row as an unevaluated pair hi + lo formed with TwoSum: out[j] = hi,
// out[n + j] = lo.
inline void extreme_total_vector(int n, const std::vector<double>& cs,
                                 const std::vector<double>& x0, std::vector<double>& out) {
  out.assign(2 * (size_t)n, 0.0);
  for (int j = 0; j < n; ++j) {
    double hi = x0[j], lo = 0.0;
    for (int c = 0; c < n - 1; ++c) {
      const double v = cs[(size_t)(2 * c) * n + j];
      const double s = hi + v;
      const double bb = s - hi;
      lo += (hi - (s - bb)) + (v - bb);
      hi = s;
    }
    const double s = hi + lo;
    out[j] = s;
    out[n + j] = lo - (s - hi);
  }
}

// Stated error of a result, with u = 2^-53 and A = |scale| * sum of |terms|
// (A/|value| is the condition number of the alternating sum):
//  * product rounding: N-2 multiplies and one FMA per term, <= (n+1) u A;
//  * in-lane accumulation: 2^(R+5) terms in plain double between
//    double-double folds, <= 2^(R+5) u A (worst case; the folds, wave,
//    workgroup and grid sums are double-double);
//  * X drift (non-integer matrices only): each x_j is its closed-form start
//    plus up to 2^w signed column additions.  Estimated as a random walk of
//    rounding errors, n sqrt(2^w) u A.  This term is an estimate, not a
//    bound; for integer matrices X is exact and it vanishes;
//  * rounding of the start vector x0 (non-integer matrices): n u |value|,
//    every row's half-ulp error moving the result the same way (see
//    nw_start_vector).
// With double-double products (mode 1) the first two terms shrink to
// (n + 8 + 2^(w+1)) 2^-104 A; with X in double-double as well (mode 2) the
// drift term vanishes too.
inline double error_statement(double abs_sum_scaled, double value, int n, int walk_bits,
                              int unroll_bits, bool integral, int mode) {
  const double u = std::ldexp(1.0, -53);
  // rounding of x0 (non-integer matrices): every row coherent, n u |value|
  const double start = integral ? 0.0 : n * u * std::fabs(value);
  const double dd_terms = std::ldexp((double)n + 8.0 + std::ldexp(2.0, walk_bits), -104);
  if (mode == 2) return abs_sum_scaled * 4.0 * dd_terms + start;
  const double drift = integral ? 0.0 : n * std::sqrt(std::ldexp(1.0, walk_bits));
  if (mode == 1) return abs_sum_scaled * (dd_terms + u * drift) + start;
  const double fold = std::ldexp(1.0, unroll_bits + 5);
  return abs_sum_scaled * u * ((double)n + 1.0 + drift + fold) + start;
}

// Expected error when the roundings of different terms behave as
// independent random errors: per term about (n + sqrt(n)) u |T| (n product
// roundings and the X drift), in total (n + sqrt(n)) u sqrt(sum T^2).  The
// kernels report Q = sum over runs of 2^(R+5) consecutive terms of
// (sum of |T| in the run)^2 >= sum T^2 (equal when one term dominates each
// run), so the estimate uses sqrt(Q): it can exceed the typical error when
// all terms are alike, never fall short of it by the run structure.
// sq_scaled = scale^2 Q.  Double-double products (mode 1) leave only the X
// drift of non-integer matrices; mode 2 has neither.  The rounding of x0
// (non-integer matrices) adds about 2 sqrt(n) u |value| (rows independent).
inline double error_estimate(double sq_scaled, double value, int n, int mode, bool integral) {
  const double u = std::ldexp(1.0, -53);
  const double dd = ((double)n + 8.0) * std::ldexp(1.0, -104);
  const double rms = std::sqrt(sq_scaled);
  const double start = integral ? 0.0 : 2.0 * std::sqrt((double)n) * u * std::fabs(value);
  if (mode == 2 || (mode == 1 && integral)) return dd * rms + start;
  if (mode == 1) return (std::sqrt((double)n) * u + dd) * rms + start;
  return ((double)n + std::sqrt((double)n)) * u * rms + start;
}

// Integer matrices: every x_j is an exact half-integer and the permanent an
// integer.  When the stated (worst-case) error is below 1/2 and the value is
// below 2^53, rounding to the nearest integer gives the exact permanent;
// the error fields then read 0.  Returns true if rounded.
inline bool round_integral(bool integral, double& value, double& err_bound, double& err_est) {
  if (!integral || !(err_bound < 0.5) || !(std::fabs(value) < 9007199254740992.0)) return false;
  value = std::nearbyint(value);
  err_bound = 0.0;
  err_est = 0.0;
  return true;
}

}  // namespace perm
