// cpu_ryser.cpp -- CPU path of `perman -c` (the reference's parallel_perman64,
// algo.h:662-746, and parallel_perman64_sparse, algo.h:568-660).
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

}  // namespace

namespace perm {

// Lanes [l0, l1) of the GPU walk, computed on the host with `threads`
// OpenMP threads: the same subsets as the kernel's lanes (closed-form start
// from the column store blocks 2(w+b) selected by the lane id, Gray walk over
// the low w columns), FP64 X and products, lane sums folded into
// double-double every 256 steps and combined in lane order.
// out = {hi, lo, sum |term|, sum over runs of (run sum of |term|)^2}.
void cpu_lanes(const std::vector<double>& cs, const std::vector<double>& x0, int n,
               int walk_bits, int high_bits, uint64_t l0, uint64_t l1, int threads,
               double out[4]) {
  const uint64_t count = l1 > l0 ? l1 - l0 : 0;
  std::vector<DDh> part(count);
  std::vector<double> abs_part(count), sq_part(count);
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
    double p = 0.0, a = 0.0, a_run = 0.0, q = 0.0;
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
        p = 0.0;
        q += (a - a_run) * (a - a_run);
        a_run = a;
      }
    }
    acc = dd_add_d(acc, p);
    q += (a - a_run) * (a - a_run);
    if (parity) {
      acc.hi = -acc.hi;
      acc.lo = -acc.lo;
    }
    part[i] = acc;
    abs_part[i] = a;
    sq_part[i] = q;
  }
  DDh tot;
  double a = 0.0, q = 0.0;
  for (uint64_t i = 0; i < count; ++i) {
    tot = dd_add(tot, part[i]);
    a += abs_part[i];
    q += sq_part[i];
  }
  out[0] = tot.hi;
  out[1] = tot.lo;
  out[2] = a;
  out[3] = q;
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
}  // extern "C"
