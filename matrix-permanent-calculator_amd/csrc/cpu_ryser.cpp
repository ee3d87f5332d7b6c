// cpu_ryser.cpp -- CPU path of `perman -c` (the reference's parallel_perman64,
// algo.h:662-746, parallel_perman64_sparse, algo.h:568-660, and the skip
// variants parallel_skip_perman64_w / _w_balanced, algo.h:748-1023).
//
// Same sum as the GPU kernels, computed with OpenMP on the host:
//  * FP64 X and products (the reference keeps X in float, which loses the
//    heavily cancelling alternating sum);
//  * the Gray range is cut into 2^h aligned blocks of 2^w subsets; a block
//    starts from its closed-form x and walks its low columns in Gray order;
//  * per-block sums are folded into double-double, blocks are combined in
//    block order (deterministic for any thread count).
#include <omp.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/perman_hip.h"
#include "perman_internal.hpp"
#include "ryser_host.hpp"

namespace {

struct DDh {
  double hi = 0.0, lo = 0.0;
};

inline DDh dd_add_d(DDh a, double b) {
  double s = a.hi + b, bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b - bb) + a.lo;
  double h = s + e;
  return {h, e - (h - s)};
}
inline DDh dd_add(DDh a, DDh b) {
  DDh s = dd_add_d(a, b.hi);
  return dd_add_d(s, b.lo);
}

// Sum over all subsets of the first n-1 columns, columns given as
// cols[k*n + j] = a[j][k] (k < n-1), start vector x0.
double cpu_ryser(const std::vector<double>& cols, const std::vector<double>& x0, int n,
                 int threads) {
  const int B = n - 1;
  int w = B < 12 ? B : 12;  // 4096 subsets per block
  const int h = B - w;
  const uint64_t blocks = 1ull << h;
  std::vector<DDh> part(blocks);
  if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int64_t blk = 0; blk < (int64_t)blocks; ++blk) {
    double x[64];
    std::memcpy(x, x0.data(), sizeof(double) * n);
    int parity = 0;
    for (int b = 0; b < h; ++b)
      if ((blk >> b) & 1) {
        parity ^= 1;
        const double* c = &cols[(size_t)(w + b) * n];
        for (int j = 0; j < n; ++j) x[j] += c[j];
      }
    DDh acc;
    double p = 0.0;
    const uint64_t steps = 1ull << w;
    for (uint64_t t = 0; t < steps; ++t) {
      if (t) {
        const int k = __builtin_ctzll(t);
        const double* c = &cols[(size_t)k * n];
        if ((t >> (k + 1)) & 1) {
          for (int j = 0; j < n; ++j) x[j] -= c[j];
        } else {
          for (int j = 0; j < n; ++j) x[j] += c[j];
        }
      }
      double prod = x[0];
      for (int j = 1; j < n; ++j) prod *= x[j];
      p += (t & 1) ? -prod : prod;
      if ((t & 255) == 255) {
        acc = dd_add_d(acc, p);
        p = 0.0;
      }
    }
    acc = dd_add_d(acc, p);
    if (parity) {
      acc.hi = -acc.hi;
      acc.lo = -acc.lo;
    }
    part[blk] = acc;
  }
  DDh tot;
  for (uint64_t b = 0; b < blocks; ++b) tot = dd_add(tot, part[b]);
  return (double)(4 * (n & 1) - 2) * (tot.hi + tot.lo);
}

template <class T>
double dense_cpu(const T* mat, int n, int threads) {
  if (!mat || n < 1 || n > 63) return std::numeric_limits<double>::quiet_NaN();
  std::vector<double> a((size_t)n * n), x0(n), cols((size_t)n * n);
  for (size_t i = 0; i < a.size(); ++i) a[i] = (double)mat[i];
  perm::nw_start_vector(a.data(), n, x0.data());
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j) cols[(size_t)k * n + j] = a[(size_t)j * n + k];
  return cpu_ryser(cols, x0, n, threads);
}

template <class T>
double sparse_cpu(const T* mat, const int* cptrs, const int* rows, const T* cvals, int n,
                  int threads) {
  if (!mat || !cptrs || n < 1 || n > 63) return std::numeric_limits<double>::quiet_NaN();
  std::vector<double> a((size_t)n * n), x0(n), cols((size_t)n * n, 0.0);
  for (size_t i = 0; i < a.size(); ++i) a[i] = (double)mat[i];
  perm::nw_start_vector(a.data(), n, x0.data());
  for (int k = 0; k < n; ++k)
    for (int e = cptrs[k]; e < cptrs[k + 1]; ++e) {
      if (rows[e] < 0 || rows[e] >= n) return std::numeric_limits<double>::quiet_NaN();
      cols[(size_t)k * n + rows[e]] += (double)cvals[e];
    }
  return cpu_ryser(cols, x0, n, threads);
}

// ---- CPU skip walk (parallel_skip_perman64_w / _w_balanced, algo.h:748-1023)
//
// The Gray range is cut into 2^h aligned blocks of 2^w subsets as above.
// Within a block, a row j with x_j == 0 keeps x_j (bitwise: other columns
// add 0.0 to it) until a walk column touching row j flips; column c flips at
// the steps t with ctz(t) == c.  When the product is zero the walk jumps to
// the latest of the zero rows' next flips (every term before it has a zero
// factor) and moves x to the jump's subset by the columns in which the two
// Gray codes differ.  A zero row touched by no low column zeroes the rest of
// the block.  The reference jumps by the highest-index zero row only; any
// zero row is a valid witness and the latest flip skips the most.
struct SkipSum {
  DDh s;
  double abs = 0.0;
  uint64_t evaluated = 0;
};

// Next step t' > t at which column c flips (ctz(t') == c).
inline uint64_t next_flip(uint64_t t, int c) {
  const uint64_t period = 2ull << c;
  uint64_t f = (t & ~(period - 1)) + (1ull << c);
  return f > t ? f : f + period;
}

SkipSum skip_block(const std::vector<double>& cols, const std::vector<double>& x0,
                   const std::vector<uint64_t>& rowmask, int n, int w, uint64_t blk) {
  double x[64];
  std::memcpy(x, x0.data(), sizeof(double) * n);
  int parity = 0;
  for (int b = 0; b < n - 1 - w; ++b)
    if ((blk >> b) & 1) {
      parity ^= 1;
      const double* c = &cols[(size_t)(w + b) * n];
      for (int j = 0; j < n; ++j) x[j] += c[j];
    }
  const uint64_t steps = 1ull << w, low = steps - 1;
  SkipSum r;
  double p = 0.0;
  uint64_t t = 0, g_prev = 0;
  while (t < steps) {
    const uint64_t g = t ^ (t >> 1);
    for (uint64_t d = g ^ g_prev; d; d &= d - 1) {  // columns where the codes differ
      const int k = __builtin_ctzll(d);
      const double* c = &cols[(size_t)k * n];
      if ((g >> k) & 1) {
        for (int j = 0; j < n; ++j) x[j] += c[j];
      } else {
        for (int j = 0; j < n; ++j) x[j] -= c[j];
      }
    }
    g_prev = g;
    uint64_t jump = t + 1;
    bool zero = false;
    for (int j = 0; j < n; ++j)
      if (x[j] == 0.0) {
        zero = true;
        const uint64_t m = rowmask[j] & low;
        uint64_t f = steps;
        for (uint64_t mm = m; mm; mm &= mm - 1) {
          const uint64_t c = next_flip(t, __builtin_ctzll(mm));
          f = c < f ? c : f;
        }
        jump = f > jump ? f : jump;
      }
    if (!zero) {
      double prod = x[0];
      for (int j = 1; j < n; ++j) prod *= x[j];
      p += (t & 1) ? -prod : prod;
      r.abs += std::fabs(prod);
      if ((++r.evaluated & 255) == 0) {
        r.s = dd_add_d(r.s, p);
        p = 0.0;
      }
    }
    t = jump;
  }
  r.s = dd_add_d(r.s, p);
  if (parity) {
    r.s.hi = -r.s.hi;
    r.s.lo = -r.s.lo;
  }
  return r;
}

template <class T>
int skip_cpu(int dtype_unused, const int* rptrs, const int* cols_idx, const T* rvals,
             const int* cptrs, const int* rows, const T* cvals, int n, int threads,
             perm_result* out) {
  (void)dtype_unused, (void)rvals;
  if (!out) return perm::fail(PERM_ERR_ARG, "null result");
  if (!rptrs || !cols_idx || !cptrs || !rows || !cvals)
    return perm::fail(PERM_ERR_ARG, "null CRS/CCS array");
  if (n < 1 || n > 63) return perm::fail(PERM_ERR_SIZE, "n must be in [1, 63] on the CPU");
  // the CCS drives the walk; the CRS must describe the same entry positions
  std::vector<double> a((size_t)n * n, 0.0), x0(n), cs((size_t)n * n, 0.0);
  std::vector<uint64_t> rowmask(n, 0);
  for (int k = 0; k < n; ++k) {
    if (cptrs[k + 1] < cptrs[k]) return perm::fail(PERM_ERR_ARG, "cptrs not ascending");
    for (int e = cptrs[k]; e < cptrs[k + 1]; ++e) {
      if (rows[e] < 0 || rows[e] >= n) return perm::fail(PERM_ERR_ARG, "row index out of range");
      a[(size_t)rows[e] * n + k] += (double)cvals[e];
      cs[(size_t)k * n + rows[e]] += (double)cvals[e];
      if (k < n - 1) rowmask[rows[e]] |= 1ull << k;
    }
  }
  for (int j = 0; j < n; ++j) {
    if (rptrs[j + 1] < rptrs[j]) return perm::fail(PERM_ERR_ARG, "rptrs not ascending");
    for (int e = rptrs[j]; e < rptrs[j + 1]; ++e)
      if (cols_idx[e] < 0 || cols_idx[e] >= n || a[(size_t)j * n + cols_idx[e]] == 0.0)
        return perm::fail(PERM_ERR_ARG, "CRS entry not in the CCS");
  }
  bool integral = true;
  for (double v : a) integral = integral && v == std::floor(v) && std::fabs(v) < 0x1p31;
  perm::nw_start_vector(a.data(), n, x0.data());
  const int B = n - 1;
  const int h = B - 12 < 0 ? 0 : (B - 12 > 10 ? 10 : B - 12);
  const int w = B - h;
  const uint64_t blocks = 1ull << h;
  std::vector<SkipSum> part(blocks);
  if (threads < 1) threads = 1;
  const double t0 = omp_get_wtime();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int64_t b = 0; b < (int64_t)blocks; ++b) part[b] = skip_block(cs, x0, rowmask, n, w, b);
  DDh tot;
  double abs = 0.0, evaluated = 0.0;
  for (const SkipSum& s : part) {
    tot = dd_add(tot, s.s);
    abs += s.abs;
    evaluated += (double)s.evaluated;
  }
  const double scale = (double)(4 * (n & 1) - 2);
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  out->sum_hi = tot.hi;
  out->sum_lo = tot.lo;
  out->value = scale * (tot.hi + tot.lo);
  out->abs_sum = std::fabs(scale) * abs;
  out->err_bound = perm::error_statement(out->abs_sum, out->value, n, w, 3, integral, 0);
  // no run moments on this path: the X drift of a non-integer matrix is
  // estimated without cancellation, sqrt(n) sqrt(2^w + h) u A
  out->err_est = out->err_bound +
                 (integral ? 0.0
                           : std::sqrt((double)n * (std::ldexp(1.0, w) + h)) *
                                 std::ldexp(out->abs_sum, -53));
  out->subsets = std::ldexp(1.0, n - 1);
  out->main_subsets = evaluated;
  out->kernel_ms = (omp_get_wtime() - t0) * 1e3;
  out->main_ms = out->kernel_ms;
  out->integral = integral ? 1 : 0;
  out->nu = out->nu1 = out->nu2 = n;
  out->status = PERM_OK;
  if (integral &&
      perm::integer_sum_exact(true, perm::integer_product_bits(x0.data(), cs.data(), n),
                              out->abs_sum, scale, n))
    out->err_bound = out->err_est = 0.0;  // exact arithmetic throughout
  perm::round_integral(integral, out->value, out->err_bound, out->err_est);
  out->certified = perm::is_certified(*out) ? 1 : 0;
  return perm::ok();
}

template <class T>
double skip_dropin(const int* rptrs, const int* cols, const T* rvals, const int* cptrs,
                   const int* rows, const T* cvals, int n, int threads) {
  perm_result r;
  if (skip_cpu<T>(0, rptrs, cols, rvals, cptrs, rows, cvals, n, threads, &r) != PERM_OK)
    return std::numeric_limits<double>::quiet_NaN();
  return r.value;
}

}  // namespace

namespace perm {

// Lanes [l0, l1) of the GPU walk, computed on the host with `threads`
// OpenMP threads: the same subsets as the kernel's lanes (closed-form start
// from the column store blocks 2(w+b) selected by the lane id, Gray walk over
// the low w columns), FP64 X and products, lane sums folded into
// double-double every 256 steps and combined in lane order.
// out = {hi, lo, sum |term|, sum over runs of (run sum of |term|)^2,
// sum over runs of (signed run sum)^2}.
void cpu_lanes(const std::vector<double>& cs, const std::vector<double>& x0, int n,
               int walk_bits, int high_bits, uint64_t l0, uint64_t l1, int threads,
               double out[5]) {
  const uint64_t count = l1 > l0 ? l1 - l0 : 0;
  std::vector<DDh> part(count);
  std::vector<double> abs_part(count), sq_part(count), run_part(count);
  if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int64_t i = 0; i < (int64_t)count; ++i) {
    const uint64_t pattern = l0 + i;
    double x[64];
    std::memcpy(x, x0.data(), sizeof(double) * n);
    int parity = 0;
    for (int b = 0; b < high_bits; ++b)
      if ((pattern >> b) & 1) {
        parity ^= 1;
        const double* c = &cs[(size_t)(2 * (walk_bits + b)) * n];
        for (int j = 0; j < n; ++j) x[j] += c[j];
      }
    DDh acc;
    double p = 0.0, a = 0.0, a_run = 0.0, q = 0.0, r = 0.0;
    const uint64_t steps = 1ull << walk_bits;
    for (uint64_t t = 0; t < steps; ++t) {
      if (t) {
        const int k = __builtin_ctzll(t);
        const double* c = &cs[(size_t)(2 * k + ((t >> (k + 1)) & 1)) * n];
        for (int j = 0; j < n; ++j) x[j] += c[j];
      }
      double prod = x[0];
      for (int j = 1; j < n; ++j) prod *= x[j];
      p += (t & 1) ? -prod : prod;
      a += std::fabs(prod);
      if ((t & 255) == 255) {
        acc = dd_add_d(acc, p);
        r += p * p;
        p = 0.0;
        q += (a - a_run) * (a - a_run);
        a_run = a;
      }
    }
    acc = dd_add_d(acc, p);
    r += p * p;
    q += (a - a_run) * (a - a_run);
    if (parity) {
      acc.hi = -acc.hi;
      acc.lo = -acc.lo;
    }
    part[i] = acc;
    abs_part[i] = a;
    sq_part[i] = q;
    run_part[i] = r;
  }
  DDh tot;
  double a = 0.0, q = 0.0, r = 0.0;
  for (uint64_t i = 0; i < count; ++i) {
    tot = dd_add(tot, part[i]);
    a += abs_part[i];
    q += sq_part[i];
    r += run_part[i];
  }
  out[0] = tot.hi;
  out[1] = tot.lo;
  out[2] = a;
  out[3] = q;
  out[4] = r;
}

}  // namespace perm

extern "C" {
double perm_parallel_perman64_i32(const int* mat, int nov, int threads) {
  return dense_cpu(mat, nov, threads);
}
double perm_parallel_perman64_f32(const float* mat, int nov, int threads) {
  return dense_cpu(mat, nov, threads);
}
double perm_parallel_perman64_f64(const double* mat, int nov, int threads) {
  return dense_cpu(mat, nov, threads);
}
double perm_parallel_perman64_sparse_i32(const int* mat, const int* cptrs, const int* rows,
                                         const int* cvals, int nov, int threads) {
  return sparse_cpu(mat, cptrs, rows, cvals, nov, threads);
}
double perm_parallel_perman64_sparse_f32(const float* mat, const int* cptrs, const int* rows,
                                         const float* cvals, int nov, int threads) {
  return sparse_cpu(mat, cptrs, rows, cvals, nov, threads);
}
double perm_parallel_perman64_sparse_f64(const double* mat, const int* cptrs, const int* rows,
                                         const double* cvals, int nov, int threads) {
  return sparse_cpu(mat, cptrs, rows, cvals, nov, threads);
}

int perm_cpu_skip(int dtype, const int* rptrs, const int* cols, const void* rvals,
                  const int* cptrs, const int* rows, const void* cvals, int nov, int threads,
                  perm_result* out) {
  switch (dtype) {
    case PERM_DTYPE_I32:
      return skip_cpu<int>(dtype, rptrs, cols, (const int*)rvals, cptrs, rows,
                           (const int*)cvals, nov, threads, out);
    case PERM_DTYPE_F32:
      return skip_cpu<float>(dtype, rptrs, cols, (const float*)rvals, cptrs, rows,
                             (const float*)cvals, nov, threads, out);
    case PERM_DTYPE_F64:
      return skip_cpu<double>(dtype, rptrs, cols, (const double*)rvals, cptrs, rows,
                              (const double*)cvals, nov, threads, out);
    default:
      return perm::fail(PERM_ERR_ARG, "unknown dtype");
  }
}

#define PERM_CPU_SKIP_DROPIN(SFX, T)                                                        \
  double perm_parallel_skip_perman64_w_##SFX(const int* rptrs, const int* cols,             \
                                             const T* rvals, const int* cptrs,              \
                                             const int* rows, const T* cvals, int nov,      \
                                             int threads) {                                 \
    return skip_dropin<T>(rptrs, cols, rvals, cptrs, rows, cvals, nov, threads);            \
  }                                                                                         \
  double perm_parallel_skip_perman64_w_balanced_##SFX(                                      \
      const int* rptrs, const int* cols, const T* rvals, const int* cptrs, const int* rows, \
      const T* cvals, int nov, int threads) {                                               \
    return skip_dropin<T>(rptrs, cols, rvals, cptrs, rows, cvals, nov, threads);            \
  }
PERM_CPU_SKIP_DROPIN(i32, int)
PERM_CPU_SKIP_DROPIN(f32, float)
PERM_CPU_SKIP_DROPIN(f64, double)
}  // extern "C"
