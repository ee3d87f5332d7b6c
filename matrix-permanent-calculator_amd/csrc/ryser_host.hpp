// ryser_host.hpp -- host-side preparation shared by every entry point:
// Nijenhuis-Wilf start vector, signed column store, work split, and the
// error statement attached to each result.
#pragma once
#include <stdint.h>

#include <cmath>
#include <vector>

namespace perm {

// Start vector x0 and the constant factor of the Nijenhuis-Wilf form,
// computed as the reference's double-X GPU wrappers do
// (gpu_exact_dense.cu:647-654; sparse: gpu_exact_sparse.cu:861-871 -- that
// loop skips zero entries, which leaves the double row sum unchanged):
//   rs_j = sum_k a[j][k] (k ascending, double), x0_j = a[j][n-1] - rs_j/2.
inline void nw_start_vector(const double* a, int n, double* x0) {
  for (int j = 0; j < n; ++j) {
    double rs = 0.0;
    for (int k = 0; k < n; ++k) rs += a[(size_t)j * n + k];
    x0[j] = a[(size_t)j * n + (n - 1)] - rs / 2;
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

// Stated error of a result, with u = 2^-53 and A = |scale| * sum of |terms|
// (A/|value| is the condition number of the alternating sum):
//  * product rounding: N-2 multiplies and one FMA per term, <= (n+1) u A;
//  * in-lane accumulation: 2^(R+5) terms in plain double between
//    double-double folds, <= 2^(R+5) u A (worst case; the folds, wave,
//    workgroup and grid sums are double-double);
//  * X drift (non-integer matrices only): each x_j is its closed-form start
//    plus up to 2^w signed column additions.  Estimated as a random walk of
//    rounding errors, n sqrt(2^w) u A.  This term is an estimate, not a
//    bound; for integer matrices X is exact and it vanishes.
// With double-double products (mode 1) the first two terms shrink to
// (n + 8 + 2^(w+1)) 2^-104 A; with X in double-double as well (mode 2) the
// drift term vanishes too.
inline double error_statement(double abs_sum_scaled, int n, int walk_bits, int unroll_bits,
                              bool integral, int mode) {
  const double u = std::ldexp(1.0, -53);
  const double dd_terms = std::ldexp((double)n + 8.0 + std::ldexp(2.0, walk_bits), -104);
  if (mode == 2) return abs_sum_scaled * 4.0 * dd_terms;
  const double drift = integral ? 0.0 : n * std::sqrt(std::ldexp(1.0, walk_bits));
  if (mode == 1) return abs_sum_scaled * (dd_terms + u * drift);
  const double fold = std::ldexp(1.0, unroll_bits + 5);
  return abs_sum_scaled * u * ((double)n + 1.0 + drift + fold);
}

// Expected error when the roundings behave as independent random errors
// (the alternating Gray sum turns the correlated X drift of a lane into an
// effectively independent error per step): about (n + sqrt(n)) u A / sqrt(N)
// with N = 2^(n-1) terms; reported with a factor 3.  Calibrated against the
// double-double mode at n = 36 and 40 and the exact oracle for n <= 26.
inline double error_estimate(double abs_sum_scaled, int n, int mode) {
  const double u = mode == 0 ? std::ldexp(1.0, -53) : std::ldexp(1.0, -104);
  const double terms = std::ldexp(1.0, n - 1);
  return 3.0 * ((double)n + std::sqrt((double)n)) * u * abs_sum_scaled / std::sqrt(terms);
}

}  // namespace perm
