// ryser_host.hpp -- host-side preparation shared by every entry point:
// Nijenhuis-Wilf start vector, signed column store, work split, and the
// error statement attached to each result.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace perm {

// Start vector of the Nijenhuis-Wilf form (and, in x0_lo when given, the
// low part of its double-double value), x0_j = a[j][n-1] - rs_j/2 with
// rs_j = sum_k a[j][k] (reference: gpu_exact_dense.cu:647-654; sparse
// gpu_exact_sparse.cu:861-871, whose loop skips zero entries and so has the
// same row sum).  The reference accumulates rs_j in plain double; here it is
// summed with TwoSum compensation and x0_j rounded once, so x0_j carries at
// most half an ulp of error.  That matters: the sum's derivative with
// respect to x0_j is a minor of the matrix, so an error d_j in x0_j moves
// the result by about d_j * perm / a[j][n-1] -- with the plain row sum about
// n u |perm| in total, here about sqrt(n) u |perm|.  Integer matrices get
// the exact x0 either way.
inline void nw_start_vector(const double* a, int n, double* x0, double* x0_lo = nullptr) {
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
    // the remainder, for kernels that keep X in double-double (mode 2): x0 +
    // x0_lo is the start value to about u^2 (the row sum's compensation error)
    if (x0_lo) x0_lo[j] = (s - x0[j]) + (e - 0.5 * lo);
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
// subsets; 2^high_bits lanes.  The FP64 error of the walk is its X drift
// (each x_j carries the roundings of the additions since the lane's start),
// and the drift of all lanes moves the result the same way, scaled by the
// lane sums -- so shorter walks give a more accurate sum.  Measured at
// n = 40 (FP64 against the double-double mode): 2^18-step walks 2.3e-10,
// 2^12-step walks 6e-12 relative error, in the same time; the drift term of
// the error estimate grows with the walk (2^(w-1) steps on average), and
// 2^10-step walks halve it against 2^12 for ~2 % more lane starts.  Default:
//   n <= 32: lanes = 2^min(n-1-R, max(n-11, 18)) (at most 2^21): walks of
//            2^R steps up to n = 22, 2^10 steps from n = 29 on;
//   n >= 33: 2^10 steps per lane, at most 2^30 lanes (2^22 workgroup
//            partials), longer walks beyond n = 41.
struct Split {
  int walk_bits;
  int high_bits;
};

inline Split choose_split(int n, int unroll_bits, int walk_override) {
  const int B = n - 1;
  int h;
  if (B <= 31) {
    // many short walks: up to 2^18 lanes (the chip holds ~2^18-2^19 in
    // flight), so a small permanent is one short wave-parallel walk rather
    // than a few long serial ones; 2^(B-10) lanes when that is more
    h = std::min(B - unroll_bits, std::max(B - 10, 18));
    if (h > 21) h = 21;
    if (h < 0) h = 0;
  } else {
    h = B - 10;
    if (h > 30) h = 30;
  }
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

// Split of the skipping kernel: long walks (the skips are aligned runs
// inside a lane's walk) -- lanes = 2^min(21, n-11), but at least 2^(n-32)
// (a walk counts its steps in 32 bits).  Integer matrices only, where X is
// exact and there is no drift.
inline Split choose_skip_split(int n, int unroll_bits, int walk_override) {
  const int B = n - 1;
  int h = B - 10;
  if (h > 21) h = 21;
  if (h < B - 31) h = B - 31;  // walks of at most 2^31 steps (32-bit step counter)
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

// Stated bound of the rounding errors of a result (err_bound), with
// u = 2^-53 and A = |scale| * sum of |terms| (A/|value| is the condition
// number of the alternating sum), to first order in u:
//  * product rounding: N-2 multiplies and one FMA per term, <= (n+1) u A;
//  * in-lane accumulation: 2^(R+5) terms in plain double between
//    double-double folds, <= 2^(R+5) u A (the folds, wave, workgroup and
//    grid sums are double-double);
//  * rounding of the start vector x0 (non-integer matrices): n u |value|,
//    every row's half-ulp error moving the result the same way (see
//    nw_start_vector); mode 2 starts from the double-double x0 (n u^2).
// With double-double products (mode 1) the first two terms shrink to
// (n + 8 + 2^(w+1)) 2^-104 A; mode 2 (X in double-double too) is bounded by
// four times that.
// NOT included: the drift of X in the FP64 walk of a non-integer matrix
// (modes 0 and 1).  Each x_j is its closed-form start plus up to 2^w signed
// column additions, and how their roundings move a term depends on how close
// x_j comes to zero, for which no useful worst-case bound exists; the
// statistical estimate (error_estimate) covers it.  For integer matrices X
// is exact, so this is a bound on the whole error.
inline double error_statement(double abs_sum_scaled, double value, int n, int walk_bits,
                              int unroll_bits, bool integral, int mode) {
  const double u = std::ldexp(1.0, -53);
  // mode 2 starts from x0 in double-double (error ~u^2)
  const double start = integral ? 0.0 : (mode == 2 ? n * u * u : n * u) * std::fabs(value);
  const double dd_terms = std::ldexp((double)n + 8.0 + std::ldexp(2.0, walk_bits), -104);
  if (mode == 2) return abs_sum_scaled * 4.0 * dd_terms + start;
  if (mode == 1) return abs_sum_scaled * dd_terms + start;
  const double fold = std::ldexp(1.0, unroll_bits + 5);
  return abs_sum_scaled * u * ((double)n + 1.0 + fold) + start;
}

// Expected error (err_est, three standard deviations) when roundings behave
// as independent random errors, from the moments the kernels measure:
// Q = sum over runs of 2^(R+5) consecutive terms of (sum of |T| in the run)^2
// (>= sum T^2) and Qr = sum over runs of (signed run sum)^2, both scaled by
// scale^2 (sq_scaled, run_scaled).
//  * Product and accumulation roundings: about n roundings of relative size
//    u/sqrt(12) per term, independent across terms: variance (n/12) u^2 Q.
//  * X drift (non-integer matrices, FP64 walk): the roundings of the column
//    additions accumulate in x_j as a random walk of up to h + 2^w steps and
//    move all terms of a run together; a run's error is then about
//    sum_j e_j d(run sum)/dx_j, which follows the run's signed (cancelled)
//    sum rather than its sum of |T|: variance (n/12) u^2 (h + 2^(w-1)) Qr,
//    with the start error (h additions) for the whole lane and the walk's
//    drift at its mean over the lane (2^(w-1) steps).
//  * The rounding of x0 (non-integer matrices): about 2 sqrt(n) u |value|
//    (rows independent), added as is.
// Calibrated against the double-double Ryser oracle on an emulation of the
// lane walk (n = 18..30, walks of 2^6..2^12 steps, densities 0.2..0.9,
// U(0,1], U(-1,1] and log-uniform entries) and checked on the GPU against
// the exact oracle (n <= 26) and the double-double X mode (n = 28..40).
// Mode 1 (double-double products) leaves the drift of non-integer matrices,
// estimated there without the signed run sums (not measured in that mode)
// as sqrt(n) sqrt(2^w + h) u sqrt(Q); mode 2 has no drift.
inline double error_estimate(double sq_scaled, double run_scaled, double value, int n,
                             int walk_bits, int high_bits, int mode, bool integral) {
  const double u = std::ldexp(1.0, -53);
  const double dd = ((double)n + 8.0) * std::ldexp(1.0, -104);
  const double rms = std::sqrt(sq_scaled);
  const double rn = std::sqrt((double)n);
  const double start = integral ? 0.0 : 2.0 * rn * u * std::fabs(value);
  if (mode == 2) return 3.0 * dd * rms + (integral ? 0.0 : 2.0 * rn * u * u * std::fabs(value));
  if (mode == 1) {
    const double drift =
        integral ? 0.0 : rn * std::sqrt(std::ldexp(1.0, walk_bits) + (double)high_bits) * u;
    return 3.0 * (drift + dd) * rms + start;
  }
  const double steps = integral ? 0.0 : std::ldexp(1.0, walk_bits - 1) + (double)high_bits;
  const double var = (double)n / 12.0 * (sq_scaled + steps * run_scaled);
  return 3.0 * u * std::sqrt(var) + start;
}

// Relative accuracy every result is held to: a result whose estimated error
// exceeds kTargetRel |value| is computed again in a more accurate mode.
constexpr double kTargetRel = 1e-10;

// Double-double addition on the host (same steps as the device dd_add).
struct HostDD {
  double hi, lo;
};
inline HostDD host_two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
inline HostDD host_dd_add(HostDD a, HostDD b) {
  HostDD s = host_two_sum(a.hi, b.hi);
  const HostDD t = host_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  double h = s.hi + s.lo;
  s.lo = s.lo - (h - s.hi);
  s.hi = h;
  s.lo += t.lo;
  h = s.hi + s.lo;
  return {h, s.lo - (h - s.hi)};
}

// Integer matrices: m_j = 2 x_j(S) = 2 x0_j + 2 sum_{k in S} c_kj is an
// integer, and over all subsets S it stays within [2 x0_j + 2 sum_k
// min(c_kj, 0), 2 x0_j + 2 sum_k max(c_kj, 0)]: |m_j| <= r_j, the larger end
// in magnitude (the row's sum of |entries| for a dense row).  Every partial
// product the walk forms is (a product of some m_j) / 2^k, its numerator at
// most prod_j max(r_j, 1); below 2^53 every product is exact.  Returns log2
// of that bound (cols[k*n + j], k < n-1).
inline double integer_product_bits(const double* x0, const double* cols, int n) {
  double bits = 0.0;
  for (int j = 0; j < n; ++j) {
    double lo = 2.0 * x0[j], hi = 2.0 * x0[j];
    for (int k = 0; k < n - 1; ++k) {
      const double c = 2.0 * cols[(size_t)k * n + j];
      (c < 0.0 ? lo : hi) += c;
    }
    const double r = std::max(std::fabs(lo), std::fabs(hi));
    if (r > 1.0) bits += std::log2(r);
  }
  return bits;
}

// With exact products every term is a multiple of 2^-n; the lane sums
// (magnitudes below A = sum |term|) are then exact while A 2^n < 2^53, and
// the double-double folds and reductions of exact values are exact: the
// computed sum is the permanent itself.  abs_sum_scaled = |scale| A.
inline bool integer_sum_exact(bool integral, double product_bits, double abs_sum_scaled,
                              double scale, int n) {
  return integral && product_bits < 52.5 &&
         abs_sum_scaled / std::fabs(scale) * std::ldexp(1.0, n) <= std::ldexp(1.0, 52);
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
