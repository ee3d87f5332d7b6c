/*
 * ryser_oracle.c -- CPU restatement of the reference's exact-permanent
 * arithmetic, used ONLY as a checker by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg.  Nothing in the product links it.
 *
 * Reference: berkyaglioglu/matrix-permanent-calculator.  Each function cites
 * the reference file:line whose arithmetic it restates.  The code is written
 * from the algorithm (Nijenhuis-Wilf form of Ryser's formula with Gray-code
 * subset order), not copied.
 *
 * Contents
 *   orc_nw_setup_f64       host x-vector / p0 / transpose of the GPU wrappers
 *   orc_perman64           serial double-X Ryser           (algo.h:1031-1088)
 *   orc_cpu_perman64       double-X range evaluator        (gpu_exact_dense.cu:6-69)
 *   orc_cpu_perman64_sparse double-X CCS range evaluator   (gpu_exact_sparse.cu:6-87)
 *   orc_parallel_perman64_{i32,f32,f64}   float-X OpenMP baseline (algo.h:662-746)
 *   orc_parallel_perman64_sparse_{i32,f32,f64}  float-X CCS (algo.h:568-660)
 *   orc_gpu_dense_floatx_{i32,f32,f64}    float-X per-thread semantics of the
 *                          reference GPU kernels IDs 0-4 (gpu_exact_dense.cu:71-399,
 *                          wrappers :401-699), evaluated thread by thread on the CPU
 *   orc_ryser_mod          exact Ryser modulo a prime (exact oracle building block)
 *   orc_ryser_dd           double-double Ryser (accurate oracle, n <= ~30)
 *   orc_lane_range_dd      double-double sum over the subsets of a range of
 *                          lanes of the GPU walk's aligned split (natural order)
 *   orc_ryser_u128         exact Ryser of an integer matrix in 128-bit
 *                          wrap-around arithmetic (exact when |perm| < 2^127)
 *
 * Build: see oracle/Makefile (-O2 -ffp-contract=off so that every
 * double/float operation rounds exactly once, as the reference's
 * source-level arithmetic does).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

typedef long long i64;

static i64 min_i64(i64 a, i64 b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
/* Nijenhuis-Wilf setup as done by the double-X GPU wrappers
 * (gpu_exact_dense.cu:647-662, identical in :473-488, :531-546, :589-604):
 *   rs_j  = sum_k a[j][k]               (double accumulation, k ascending)
 *   x_j   = a[j][n-1] - rs_j/2
 *   p0    = prod_j x_j                  (double, j ascending)
 *   mat_t = transpose(a)                                                  */
void orc_nw_setup_f64(const double* mat, int n, double* x, double* p0, double* mat_t) {
  double p = 1.0;
  for (int j = 0; j < n; j++) {
    double rs = 0.0;
    for (int k = 0; k < n; k++) rs += mat[j * n + k];
    x[j] = mat[j * n + (n - 1)] - rs / 2;
    p *= x[j];
  }
  *p0 = p;
  if (mat_t)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) mat_t[i * n + j] = mat[j * n + i];
}

/* ------------------------------------------------------------------------ */
/* Serial double-X Ryser, algo.h:1031-1088.  One Gray walk over
 * i = 1 .. 2^(n-1)-1 with k = ctz(i), s = +1 iff bit k of the new gray code
 * is set, x += s*col_k, p += (-1)^i prod(x).  Returns (4(n&1)-2) p.       */
double orc_perman64(const double* mat, int n) {
  double x[64], *mat_t = (double*)malloc(sizeof(double) * n * n), p;
  orc_nw_setup_f64(mat, n, x, &p, mat_t);
  unsigned long long gray = 0, last = (1ULL << (n - 1)) - 1ULL;
  for (unsigned long long i = 1; i <= last; i++) {
    int k = __builtin_ctzll(i);
    gray ^= 1ULL << k;
    double s = ((1ULL << k) & gray) ? 1.0 : -1.0;
    double prod = 1.0;
    for (int j = 0; j < n; j++) {
      x[j] += s * mat_t[k * n + j];
      prod *= x[j];
    }
    p += ((i & 1ULL) ? -1.0 : 1.0) * prod;
  }
  free(mat_t);
  return (4 * (n & 1) - 2) * p;
}

/* ------------------------------------------------------------------------ */
/* Closed-form start of a Gray range at index i0 (gpu_exact_dense.cu:27-38):
 * gray = (i0-1) ^ ((i0-1)>>1); add every column k < n-1 whose bit is set.   */
static void gray_start_dense(const double* mat_t, double* x, int n, i64 i0, i64* gray_out) {
  i64 g = (i0 - 1) ^ ((i0 - 1) >> 1);
  for (int k = 0; k < n - 1; k++)
    if ((g >> k) & 1LL)
      for (int j = 0; j < n; j++) x[j] += mat_t[k * n + j];
  *gray_out = g;
}

/* Double-X range evaluator, gpu_exact_dense.cu:6-69.  The range
 * [start,end) is cut into `threads` chunks of (end-start)/threads+1 Gray
 * steps, each walked independently from its closed-form start; the chunk
 * sums are added in chunk order (the reference adds them under an
 * `omp atomic`, in completion order).  Returns the raw sum (no p0, no
 * final scale), exactly like the reference helper.                         */
double orc_cpu_perman64(const double* mat_t, const double* x_in, int n, i64 start, i64 end,
                        int threads) {
  i64 chunk = (end - start) / threads + 1;
  double p = 0.0;
  for (int tid = 0; tid < threads; tid++) {
    double x[64];
    memcpy(x, x_in, sizeof(double) * n);
    i64 my_start = start + tid * chunk;
    i64 my_end = min_i64(start + (tid + 1) * chunk, end);
    i64 i = my_start, gray;
    gray_start_dense(mat_t, x, n, i, &gray);
    int sign = (i & 1LL) ? -1 : 1;
    double my_p = 0.0;
    while (i < my_end) {
      int k = __builtin_ctzll((unsigned long long)i);
      gray ^= 1LL << k;
      int s = ((1LL << k) & gray) ? 1 : -1;
      double prod = 1.0;
      for (int j = 0; j < n; j++) {
        x[j] += s * mat_t[k * n + j];
        prod *= x[j];
      }
      my_p += sign * prod;
      sign = -sign;
      i++;
    }
    p += my_p;
  }
  return p;
}

/* ------------------------------------------------------------------------ */
/* Sparse double-X range evaluator, gpu_exact_sparse.cu:6-87: CCS column
 * updates, product kept incrementally by divide/multiply with an exact-zero
 * counter.  Raw sum, chunks added in order.                                */
double orc_cpu_perman64_sparse(const int* cptrs, const int* rows, const double* cvals,
                               const double* x_in, int n, i64 start, i64 end, int threads) {
  i64 chunk = (end - start) / threads + 1;
  double p = 0.0;
  for (int tid = 0; tid < threads; tid++) {
    double x[64];
    memcpy(x, x_in, sizeof(double) * n);
    i64 my_start = start + tid * chunk;
    i64 my_end = min_i64(start + (tid + 1) * chunk, end);
    i64 i = my_start;
    i64 gray = (i - 1) ^ ((i - 1) >> 1);
    for (int k = 0; k < n - 1; k++)
      if ((gray >> k) & 1LL)
        for (int e = cptrs[k]; e < cptrs[k + 1]; e++) x[rows[e]] += cvals[e];
    double prod = 1.0;
    int zeros = 0;
    for (int j = 0; j < n; j++) {
      if (x[j] == 0) zeros++;
      else prod *= x[j];
    }
    int sign = (i & 1LL) ? -1 : 1;
    double my_p = 0.0;
    while (i < my_end) {
      int k = __builtin_ctzll((unsigned long long)i);
      gray ^= 1LL << k;
      int s = ((1LL << k) & gray) ? 1 : -1;
      for (int e = cptrs[k]; e < cptrs[k + 1]; e++) {
        int r = rows[e];
        if (x[r] == 0) {
          zeros--;
          x[r] += s * cvals[e];
          prod *= x[r];
        } else {
          prod /= x[r];
          x[r] += s * cvals[e];
          if (x[r] == 0) zeros++;
          else prod *= x[r];
        }
      }
      if (zeros == 0) my_p += sign * prod;
      sign = -sign;
      i++;
    }
    p += my_p;
  }
  return p;
}

/* ------------------------------------------------------------------------ */
/* Float-X OpenMP baseline, algo.h:662-746, for T in {int, float, double}.
 * x and the row sum are FLOAT; element arithmetic follows C's usual
 * conversions exactly as the reference's expressions do:
 *   rs (float) += a (T)       -> float op for int/float, double op rounded
 *                                to float for double
 *   x  = a[j][n-1] - rs/2     -> rounded to float
 *   x (float) += s*a (T)      -> same conversion rules
 *   prod (double) *= x        -> double
 * The static split is chunk = 2^(n-1)/threads + 1 starting at 1; the
 * per-thread sums are added to p0 in thread order (the reference uses an
 * `omp critical`, i.e. completion order).  `start_cap`/`end_cap` restrict
 * the walk to a sub-slice [start_cap, end_cap) of the Gray range (used to
 * time a bounded slice); pass 1 and 2^(n-1) for the full computation.
 * When `parallel` is nonzero the threads really run under OpenMP.         */
#define DEFINE_PARALLEL_PERMAN64(SUFFIX, T)                                                    \
  double orc_parallel_perman64_slice_##SUFFIX(const T* mat, int n, int threads, i64 start,    \
                                              i64 end, int parallel) {                         \
    float x[64];                                                                               \
    double p = 1.0;                                                                            \
    for (int j = 0; j < n; j++) {                                                              \
      float rs = 0.0f;                                                                         \
      for (int k = 0; k < n; k++) rs += mat[j * n + k];                                        \
      x[j] = mat[j * n + (n - 1)] - rs / 2;                                                    \
      p *= x[j];                                                                               \
    }                                                                                          \
    T* mat_t = (T*)malloc(sizeof(T) * n * n);                                                  \
    for (int i = 0; i < n; i++)                                                                \
      for (int j = 0; j < n; j++) mat_t[i * n + j] = mat[j * n + i];                           \
    i64 chunk = (end - start) / threads + 1;                                                   \
    double* part = (double*)calloc(threads, sizeof(double));                                   \
    _Pragma("omp parallel for num_threads(threads) schedule(static, 1) if(parallel)")         \
    for (int tid = 0; tid < threads; tid++) {                                                  \
      float my_x[64];                                                                          \
      memcpy(my_x, x, sizeof(float) * n);                                                      \
      i64 my_start = start + tid * chunk;                                                      \
      i64 my_end = min_i64(start + (tid + 1) * chunk, end);                                    \
      i64 i = my_start;                                                                        \
      i64 gray = (i - 1) ^ ((i - 1) >> 1);                                                     \
      for (int k = 0; k < n - 1; k++)                                                          \
        if ((gray >> k) & 1LL)                                                                 \
          for (int j = 0; j < n; j++) my_x[j] += mat_t[k * n + j];                             \
      int sign = (i & 1LL) ? -1 : 1;                                                           \
      double my_p = 0.0;                                                                       \
      while (i < my_end) {                                                                     \
        int k = __builtin_ctzll((unsigned long long)i);                                        \
        gray ^= 1LL << k;                                                                      \
        int s = ((1LL << k) & gray) ? 1 : -1;                                                  \
        double prod = 1.0;                                                                     \
        for (int j = 0; j < n; j++) {                                                          \
          my_x[j] += s * mat_t[k * n + j];                                                     \
          prod *= my_x[j];                                                                     \
        }                                                                                      \
        my_p += sign * prod;                                                                   \
        sign = -sign;                                                                          \
        i++;                                                                                   \
      }                                                                                        \
      part[tid] = my_p;                                                                        \
    }                                                                                          \
    for (int tid = 0; tid < threads; tid++) p += part[tid];                                    \
    free(part);                                                                                \
    free(mat_t);                                                                               \
    return (4 * (n & 1) - 2) * p;                                                              \
  }                                                                                            \
  double orc_parallel_perman64_##SUFFIX(const T* mat, int n, int threads) {                    \
    return orc_parallel_perman64_slice_##SUFFIX(mat, n, threads, 1, 1LL << (n - 1), 0);        \
  }

DEFINE_PARALLEL_PERMAN64(i32, int)
DEFINE_PARALLEL_PERMAN64(f32, float)
DEFINE_PARALLEL_PERMAN64(f64, double)

/* ------------------------------------------------------------------------ */
/* Float-X sparse OpenMP baseline, algo.h:568-660.  x-init from the dense
 * matrix (float row sum over every entry), CCS updates with the
 * divide/multiply product and exact-zero counter.                         */
#define DEFINE_PARALLEL_PERMAN64_SPARSE(SUFFIX, T)                                             \
  double orc_parallel_perman64_sparse_##SUFFIX(const T* mat, const int* cptrs, const int* rows, \
                                               const T* cvals, int n, int threads) {           \
    float x[64];                                                                               \
    double p = 1.0;                                                                            \
    for (int j = 0; j < n; j++) {                                                              \
      float rs = 0.0f;                                                                         \
      for (int k = 0; k < n; k++) rs += mat[j * n + k];                                        \
      x[j] = mat[j * n + (n - 1)] - rs / 2;                                                    \
      p *= x[j];                                                                               \
    }                                                                                          \
    i64 start = 1, end = 1LL << (n - 1);                                                       \
    i64 chunk = end / threads + 1;                                                             \
    for (int tid = 0; tid < threads; tid++) {                                                  \
      float my_x[64];                                                                          \
      memcpy(my_x, x, sizeof(float) * n);                                                      \
      i64 my_start = start + tid * chunk;                                                      \
      i64 my_end = min_i64(start + (tid + 1) * chunk, end);                                    \
      i64 i = my_start;                                                                        \
      i64 gray = (i - 1) ^ ((i - 1) >> 1);                                                     \
      for (int k = 0; k < n - 1; k++)                                                          \
        if ((gray >> k) & 1LL)                                                                 \
          for (int e = cptrs[k]; e < cptrs[k + 1]; e++) my_x[rows[e]] += cvals[e];             \
      double prod = 1.0;                                                                       \
      int zeros = 0;                                                                           \
      for (int j = 0; j < n; j++) {                                                            \
        if (my_x[j] == 0) zeros++;                                                             \
        else prod *= my_x[j];                                                                  \
      }                                                                                        \
      int sign = (i & 1LL) ? -1 : 1;                                                           \
      double my_p = 0.0;                                                                       \
      while (i < my_end) {                                                                     \
        int k = __builtin_ctzll((unsigned long long)i);                                        \
        gray ^= 1LL << k;                                                                      \
        int s = ((1LL << k) & gray) ? 1 : -1;                                                  \
        for (int e = cptrs[k]; e < cptrs[k + 1]; e++) {                                        \
          int r = rows[e];                                                                     \
          if (my_x[r] == 0) {                                                                  \
            zeros--;                                                                           \
            my_x[r] += s * cvals[e];                                                           \
            prod *= my_x[r];                                                                   \
          } else {                                                                             \
            prod /= my_x[r];                                                                   \
            my_x[r] += s * cvals[e];                                                           \
            if (my_x[r] == 0) zeros++;                                                         \
            else prod *= my_x[r];                                                              \
          }                                                                                    \
        }                                                                                      \
        if (zeros == 0) my_p += sign * prod;                                                   \
        sign = -sign;                                                                          \
        i++;                                                                                   \
      }                                                                                        \
      p += my_p;                                                                               \
    }                                                                                          \
    return (4 * (n & 1) - 2) * p;                                                              \
  }

DEFINE_PARALLEL_PERMAN64_SPARSE(i32, int)
DEFINE_PARALLEL_PERMAN64_SPARSE(f32, float)
DEFINE_PARALLEL_PERMAN64_SPARSE(f64, double)

/* ------------------------------------------------------------------------ */
/* Float-X semantics of the reference GPU kernels, one simulated thread at
 * a time.  `id` selects the kernel:
 *   0 kernel_xglobal   (gpu_exact_dense.cu:71-127; wrapper :401-464)
 *   1 kernel_xlocal    (:129-194; wrapper :466-522)
 *   2 kernel_xshared   (:196-260; wrapper :524-580)
 *   3 kernel_xshared_coalescing (:262-327; wrapper :582-638)
 *   4 kernel_xshared_coalescing_mshared (:329-399; wrapper :640-699)
 * IDs 1-3 share one arithmetic (only the storage of X differs); ID 4 uses
 * the chunk (end-start)/T+1 instead of end/T+1; ID 0 starts from a FLOAT
 * host x-vector (:403-415).  Thread partials are summed in thread order on
 * top of p0, as the wrappers do (:455-457, :514-516, :691-693).           */
#define DEFINE_GPU_DENSE_FLOATX(SUFFIX, T)                                                     \
  double orc_gpu_dense_floatx_##SUFFIX(const T* mat, int n, int grid_dim, int block_dim,       \
                                       int id) {                                               \
    double xd[64];                                                                             \
    float xf[64];                                                                              \
    double p = 1.0;                                                                            \
    for (int j = 0; j < n; j++) {                                                              \
      double rs = 0.0;                                                                         \
      for (int k = 0; k < n; k++) rs += mat[j * n + k];                                        \
      if (id == 0) {                                                                           \
        xf[j] = mat[j * n + (n - 1)] - rs / 2;                                                 \
        p *= xf[j];                                                                            \
      } else {                                                                                 \
        xd[j] = mat[j * n + (n - 1)] - rs / 2;                                                 \
        p *= xd[j];                                                                            \
        xf[j] = (float)xd[j];                                                                  \
      }                                                                                        \
    }                                                                                          \
    T* mat_t = (T*)malloc(sizeof(T) * n * n);                                                  \
    for (int i = 0; i < n; i++)                                                                \
      for (int j = 0; j < n; j++) mat_t[i * n + j] = mat[j * n + i];                           \
    i64 nthreads = (i64)grid_dim * block_dim;                                                  \
    i64 start = 1, end = 1LL << (n - 1);                                                       \
    i64 chunk = (id == 4) ? (end - start) / nthreads + 1 : end / nthreads + 1;                 \
    for (i64 tid = 0; tid < nthreads; tid++) {                                                 \
      i64 my_start = start + tid * chunk;                                                      \
      if (my_start >= end) break; /* later threads contribute exactly 0.0 */                   \
      i64 my_end = min_i64(start + (tid + 1) * chunk, end);                                    \
      float my_x[64];                                                                          \
      memcpy(my_x, xf, sizeof(float) * n);                                                     \
      i64 i = my_start;                                                                        \
      i64 gray = (i - 1) ^ ((i - 1) >> 1);                                                     \
      for (int k = 0; k < n - 1; k++)                                                          \
        if ((gray >> k) & 1LL)                                                                 \
          for (int j = 0; j < n; j++) my_x[j] += mat_t[k * n + j];                             \
      int sign = (i & 1LL) ? -1 : 1;                                                           \
      double my_p = 0.0;                                                                       \
      while (i < my_end) {                                                                     \
        int k = __builtin_ctzll((unsigned long long)i);                                        \
        gray ^= 1LL << k;                                                                      \
        int s = ((1LL << k) & gray) ? 1 : -1;                                                  \
        double prod = 1.0;                                                                     \
        for (int j = 0; j < n; j++) {                                                          \
          my_x[j] += s * mat_t[k * n + j];                                                     \
          prod *= my_x[j];                                                                     \
        }                                                                                      \
        my_p += sign * prod;                                                                   \
        sign = -sign;                                                                          \
        i++;                                                                                   \
      }                                                                                        \
      p += my_p;                                                                               \
    }                                                                                          \
    free(mat_t);                                                                               \
    return (4 * (n & 1) - 2) * p;                                                              \
  }

DEFINE_GPU_DENSE_FLOATX(i32, int)
DEFINE_GPU_DENSE_FLOATX(f32, float)
DEFINE_GPU_DENSE_FLOATX(f64, double)

/* ------------------------------------------------------------------------ */
/* Exact Ryser modulo an odd prime p < 2^62, Nijenhuis-Wilf form:
 *   perm = (4(n&1)-2) * sum_{i} (-1)^i prod_j x_j(i)   (mod p)
 * with x_j(0) = a[j][n-1] - rowsum_j * inv(2).  `a` holds the entries
 * already reduced mod p (row-major).  The exact oracle runs this for
 * enough primes and recombines by CRT.                                     */
static inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p) {
  return (uint64_t)(((unsigned __int128)a * b) % p);
}
static inline uint64_t addmod(uint64_t a, uint64_t b, uint64_t p) {
  uint64_t s = a + b;
  return s >= p ? s - p : s;
}
static inline uint64_t submod(uint64_t a, uint64_t b, uint64_t p) {
  return a >= b ? a - b : a + p - b;
}

uint64_t orc_ryser_mod(const uint64_t* a, int n, uint64_t p) {
  uint64_t inv2 = (p + 1) / 2, x[64], colt[64 * 64];
  for (int j = 0; j < n; j++) {
    uint64_t rs = 0;
    for (int k = 0; k < n; k++) rs = addmod(rs, a[j * n + k], p);
    x[j] = submod(a[j * n + n - 1], mulmod(rs, inv2, p), p);
  }
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) colt[k * n + j] = a[j * n + k];
  uint64_t acc_even = 1, acc_odd = 0;
  for (int j = 0; j < n; j++) acc_even = mulmod(acc_even, x[j], p);
  unsigned long long gray = 0, last = (1ULL << (n - 1)) - 1ULL;
  for (unsigned long long i = 1; i <= last; i++) {
    int k = __builtin_ctzll(i);
    gray ^= 1ULL << k;
    uint64_t prod = 1;
    if (gray & (1ULL << k)) {
      for (int j = 0; j < n; j++) {
        x[j] = addmod(x[j], colt[k * n + j], p);
        prod = mulmod(prod, x[j], p);
      }
    } else {
      for (int j = 0; j < n; j++) {
        x[j] = submod(x[j], colt[k * n + j], p);
        prod = mulmod(prod, x[j], p);
      }
    }
    if (i & 1ULL) acc_odd = addmod(acc_odd, prod, p);
    else acc_even = addmod(acc_even, prod, p);
  }
  uint64_t s = submod(acc_even, acc_odd, p);
  s = addmod(s, s, p); /* times 2 */
  if (!(n & 1)) s = submod(0, s, p);
  return s;
}

/* ------------------------------------------------------------------------ */
/* Double-double Ryser (accurate oracle).  X, products and the sum are kept
 * as unevaluated pairs (hi, lo) with error-free TwoSum / TwoProd (fma), so
 * each operation has relative error about 2^-104.  The Gray range is split
 * over OpenMP threads; each chunk restarts from its closed-form x.  Result
 * returned as hi + lo.                                                     */
typedef struct { double hi, lo; } dd;
static inline dd dd_two_sum(double a, double b) {
  double s = a + b, bb = s - a;
  dd r = {s, (a - (s - bb)) + (b - bb)};
  return r;
}
static inline dd dd_quick_two_sum(double a, double b) {
  double s = a + b;
  dd r = {s, b - (s - a)};
  return r;
}
/* IEEE-style double-double addition (TwoSum on both words, renormalise). */
static inline dd dd_add(dd a, dd b) {
  dd s = dd_two_sum(a.hi, b.hi), t = dd_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = dd_quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return dd_quick_two_sum(s.hi, s.lo);
}
static inline dd dd_mul(dd a, dd b) {
  double p = a.hi * b.hi, e = fma(a.hi, b.hi, -p);
  e += a.hi * b.lo + a.lo * b.hi;
  return dd_quick_two_sum(p, e);
}
static inline dd dd_from(double a) { dd r = {a, 0.0}; return r; }
static inline dd dd_neg(dd a) { dd r = {-a.hi, -a.lo}; return r; }

void orc_ryser_dd(const double* mat, int n, int threads, double* out_hi, double* out_lo) {
  dd x0[64];
  dd colt[64 * 64];
  for (int j = 0; j < n; j++) {
    dd rs = dd_from(0.0);
    for (int k = 0; k < n; k++) rs = dd_add(rs, dd_from(mat[j * n + k]));
    dd half = {rs.hi * 0.5, rs.lo * 0.5};
    x0[j] = dd_add(dd_from(mat[j * n + n - 1]), dd_neg(half));
  }
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) colt[k * n + j] = dd_from(mat[j * n + k]);
  i64 total = 1LL << (n - 1); /* i = 0 .. total-1, i = 0 is the p0 term */
  if (threads < 1) threads = 1;
  i64 chunk = total / threads + 1;
  dd* part = (dd*)calloc(threads, sizeof(dd));
#pragma omp parallel for num_threads(threads) schedule(static, 1)
  for (int tid = 0; tid < threads; tid++) {
    i64 st = (i64)tid * chunk, en = min_i64(st + chunk, total);
    dd acc = dd_from(0.0);
    if (st < en) {
      dd x[64];
      memcpy(x, x0, sizeof(dd) * n);
      i64 g = st ^ (st >> 1);
      for (int k = 0; k < n - 1; k++)
        if ((g >> k) & 1LL)
          for (int j = 0; j < n; j++) x[j] = dd_add(x[j], colt[k * n + j]);
      for (i64 i = st; i < en; i++) {
        if (i > st) {
          int k = __builtin_ctzll((unsigned long long)i);
          g ^= 1LL << k;
          int plus = (g >> k) & 1LL;
          for (int j = 0; j < n; j++)
            x[j] = dd_add(x[j], plus ? colt[k * n + j] : dd_neg(colt[k * n + j]));
        }
        dd prod = x[0];
        for (int j = 1; j < n; j++) prod = dd_mul(prod, x[j]);
        acc = dd_add(acc, (i & 1LL) ? dd_neg(prod) : prod);
      }
    }
    part[tid] = acc;
  }
  dd p = dd_from(0.0);
  for (int tid = 0; tid < threads; tid++) p = dd_add(p, part[tid]);
  free(part);
  double sc = (double)(4 * (n & 1) - 2);
  *out_hi = p.hi * sc;
  *out_lo = p.lo * sc;
}

/* ------------------------------------------------------------------------ */
/* Exact Ryser of an integer matrix in the ring Z/2^128 (unsigned __int128
 * arithmetic wraps modulo 2^128), using Ryser's original form over all 2^n
 * column subsets S:  perm = (-1)^n sum_S (-1)^|S| prod_i (sum_{j in S} a_ij),
 * so every quantity is an integer (the Nijenhuis-Wilf form's half-integers
 * would need a division by 2^n).  Every operation is a ring operation, so
 * the result is the permanent modulo 2^128: the exact permanent whenever
 * |perm| < 2^127, which the caller checks with prod_i sum_j |a_ij|.  Row
 * sums are exact in int64 (|a_ij| < 2^31, n <= 62).  The subset range is
 * split over OpenMP threads, each chunk starting from its closed-form row
 * sums; the chunk sums are added modulo 2^128 (order-independent).
 * Result: the two 64-bit halves of the wrapped sum.                       */
void orc_ryser_u128(const i64* a, int n, int threads, uint64_t* out_lo, uint64_t* out_hi) {
  typedef unsigned __int128 u128;
  const i64 total = (i64)1 << n;
  if (threads < 1) threads = 1;
  i64 chunk = total / threads + 1;
  u128* part = (u128*)calloc(threads, sizeof(u128));
#pragma omp parallel for num_threads(threads) schedule(static, 1)
  for (int tid = 0; tid < threads; tid++) {
    i64 st = (i64)tid * chunk, en = min_i64(st + chunk, total);
    u128 acc = 0;
    if (st < en) {
      i64 rs[64];
      i64 g = st ^ (st >> 1);
      for (int i = 0; i < n; i++) {
        rs[i] = 0;
        for (int j = 0; j < n; j++)
          if ((g >> j) & 1LL) rs[i] += a[i * n + j];
      }
      for (i64 t = st; t < en; t++) {
        if (t > st) {
          int k = __builtin_ctzll((unsigned long long)t);
          g ^= 1LL << k;
          if ((g >> k) & 1LL)
            for (int i = 0; i < n; i++) rs[i] += a[i * n + k];
          else
            for (int i = 0; i < n; i++) rs[i] -= a[i * n + k];
        }
        u128 prod = 1;
        for (int i = 0; i < n; i++) prod *= (u128)(__int128)rs[i];
        /* (-1)^|S| with |S| = popcount(gray) */
        if (__builtin_popcountll((unsigned long long)g) & 1) acc -= prod;
        else acc += prod;
      }
    }
    part[tid] = acc;
  }
  u128 sum = 0;
  for (int tid = 0; tid < threads; tid++) sum += part[tid];
  free(part);
  if (n & 1) sum = (u128)0 - sum; /* (-1)^n */
  *out_lo = (uint64_t)sum;
  *out_hi = (uint64_t)(sum >> 64);
}

/* ------------------------------------------------------------------------ */
/* Double-double sum over the subsets the GPU walk assigns to lanes
 * [lane_begin, lane_end) of its aligned split, in natural column order: lane
 * l fixes columns w..n-2 by the bits of l and walks columns 0..w-1 in Gray
 * order; the sum is sum (-1)^|S| prod_j x_j(S) with x(S) = x0 + sum_{k in S}
 * a_k and x0_j = a[j][n-1] - rowsum_j / 2 in double-double from the matrix
 * (no scale factor).  It restates the Nijenhuis-Wilf sum of the reference's
 * range evaluator (gpu_exact_dense.cu:6-69) over the lane decomposition of
 * the GPU kernels, with X, products and sum in double-double.            */
void orc_lane_range_dd(const double* mat, int n, int w, i64 lane_begin, i64 lane_end,
                       int threads, double* out_hi, double* out_lo) {
  dd x0[64];
  dd colt[64 * 64];
  for (int j = 0; j < n; j++) {
    dd rs = dd_from(0.0);
    for (int k = 0; k < n; k++) rs = dd_add(rs, dd_from(mat[j * n + k]));
    dd half = {rs.hi * 0.5, rs.lo * 0.5};
    x0[j] = dd_add(dd_from(mat[j * n + n - 1]), dd_neg(half));
  }
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++) colt[k * n + j] = dd_from(mat[j * n + k]);
  const i64 lanes = lane_end - lane_begin;
  if (threads < 1) threads = 1;
  dd* part = (dd*)calloc(lanes > 0 ? lanes : 1, sizeof(dd));
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (i64 li = 0; li < lanes; li++) {
    const i64 l = lane_begin + li;
    dd x[64];
    memcpy(x, x0, sizeof(dd) * n);
    int parity = 0;
    for (int b = 0; w + b < n - 1; b++)
      if ((l >> b) & 1LL) {
        parity ^= 1;
        for (int j = 0; j < n; j++) x[j] = dd_add(x[j], colt[(w + b) * n + j]);
      }
    dd acc = dd_from(0.0);
    i64 g = 0;
    for (i64 t = 0; t < ((i64)1 << w); t++) {
      if (t > 0) {
        int k = __builtin_ctzll((unsigned long long)t);
        g ^= 1LL << k;
        int plus = (g >> k) & 1LL;
        for (int j = 0; j < n; j++)
          x[j] = dd_add(x[j], plus ? colt[k * n + j] : dd_neg(colt[k * n + j]));
      }
      dd prod = x[0];
      for (int j = 1; j < n; j++) prod = dd_mul(prod, x[j]);
      acc = dd_add(acc, ((t & 1LL) ^ parity) ? dd_neg(prod) : prod);
    }
    part[li] = acc;
  }
  dd p = dd_from(0.0);
  for (i64 li = 0; li < lanes; li++) p = dd_add(p, part[li]);
  free(part);
  *out_hi = p.hi;
  *out_lo = p.lo;
}
