/*
 * perman_hip.h -- C ABI of the MI355X (gfx950) exact-permanent library
 * libperman_hip.so.  Plain pointers and sizes only; every entry point is
 * synchronous unless stated otherwise and thread-safe (calls on one device
 * are serialised).
 *
 * Reference: berkyaglioglu/matrix-permanent-calculator.  Each drop-in entry
 * point names the reference function whose contract it replaces.
 *
 * Element types: _i32 (int), _f32 (float), _f64 (double); `mat` is the dense
 * row-major n x n matrix, `nov` its order (the reference's name).  Every
 * element type is converted exactly to double; the X vector, the row
 * products and the outer sum are FP64 (the outer sum double-double).
 *
 * Error behaviour: the reference has no error channel.  Here the
 * double-returning drop-ins return NaN on failure and set the thread's last
 * status (perm_last_status / perm_last_error).  `grid_dim` and `block_dim`
 * are accepted for signature compatibility; the launch geometry is chosen by
 * the library (values are otherwise ignored).
 */
#ifndef PERMAN_HIP_H
#define PERMAN_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PERM_OK 0
#define PERM_ERR_ARG 1      /* bad argument (null pointer, n out of range, ...) */
#define PERM_ERR_SIZE 2     /* n above the largest compiled size */
#define PERM_ERR_HIP 3      /* a HIP runtime call failed */
#define PERM_ERR_NODEV 4    /* no usable GPU */

#define PERM_DTYPE_I32 0
#define PERM_DTYPE_F32 1
#define PERM_DTYPE_F64 2

/* Result of one permanent.  value = scale * (sum_hi + sum_lo) where
 * scale = 4(n&1)-2 and sum is the double-double Ryser sum; abs_sum is
 * |scale| * sum of |terms| (the condition of the alternating sum) and
 * err_bound the library's stated worst-case error for `value` (every rounding
 * aligned); err_est the expected error when roundings are independent,
 * 3 (n + sqrt(n)) u abs_sum / sqrt(2^(n-1)).  mode: 0 = FP64 products
 * (the fast kernel), 1 = double-double products (used for integer-valued
 * matrices whose FP64 result could not be guaranteed to round to the exact
 * integer). */
typedef struct perm_result {
  double value;
  double sum_hi;
  double sum_lo;
  double abs_sum;
  double err_bound;
  double kernel_ms;     /* device time of the Ryser kernel(s), HIP events */
  double subsets;       /* Gray-code subsets evaluated (2^(n-1)) */
  int n;
  int status;
  int mode;
  int integral;         /* 1 if every entry is an integer (X is then exact) */
  double err_est;       /* statistical error estimate (independent roundings) */
  int nu;               /* FP64 kernel: rows touched by Gray column 0 (pair variant), or n */
} perm_result;

/* ---- library state ------------------------------------------------------ */
int perm_last_status(void);
const char* perm_last_error(void);
/* Device used by subsequent calls of this thread (default 0; the reference
 * hard-codes cudaSetDevice(1), gpu_exact_dense.cu:490). */
int perm_set_device(int device);
/* Largest n accepted by the dense/sparse GPU paths. */
int perm_max_n(void);

/* ---- dense drop-ins: gpu_perman64_<variant>(T* mat, int nov, int grid_dim,
 * int block_dim), gpu_exact_dense.cu:402 (ID 0), :467 (ID 1), :525 (ID 2),
 * :583 (ID 3), :641 (ID 4). ------------------------------------------------ */
double perm_xglobal_dense_i32(const int* mat, int nov, int grid_dim, int block_dim);
double perm_xglobal_dense_f32(const float* mat, int nov, int grid_dim, int block_dim);
double perm_xglobal_dense_f64(const double* mat, int nov, int grid_dim, int block_dim);
double perm_xlocal_dense_i32(const int* mat, int nov, int grid_dim, int block_dim);
double perm_xlocal_dense_f32(const float* mat, int nov, int grid_dim, int block_dim);
double perm_xlocal_dense_f64(const double* mat, int nov, int grid_dim, int block_dim);
double perm_xshared_dense_i32(const int* mat, int nov, int grid_dim, int block_dim);
double perm_xshared_dense_f32(const float* mat, int nov, int grid_dim, int block_dim);
double perm_xshared_dense_f64(const double* mat, int nov, int grid_dim, int block_dim);
double perm_xshared_coalescing_dense_i32(const int* mat, int nov, int grid_dim, int block_dim);
double perm_xshared_coalescing_dense_f32(const float* mat, int nov, int grid_dim, int block_dim);
double perm_xshared_coalescing_dense_f64(const double* mat, int nov, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_dense_i32(const int* mat, int nov, int grid_dim,
                                                 int block_dim);
double perm_xshared_coalescing_mshared_dense_f32(const float* mat, int nov, int grid_dim,
                                                 int block_dim);
double perm_xshared_coalescing_mshared_dense_f64(const double* mat, int nov, int grid_dim,
                                                 int block_dim);

/* ---- sparse drop-ins: gpu_perman64_<variant>_sparse(T* mat, int* cptrs,
 * int* rows, T* cvals, int nov, int grid_dim, int block_dim),
 * gpu_exact_sparse.cu:674 (ID 1), :734 (ID 2), :794 (ID 3), :854 (ID 4).
 * As in the reference, x is initialised from the dense `mat` and the column
 * updates come from the CCS arrays (cptrs[nov+1], rows/cvals[cptrs[nov]]). */
double perm_xlocal_sparse_i32(const int* mat, const int* cptrs, const int* rows,
                              const int* cvals, int nov, int grid_dim, int block_dim);
double perm_xlocal_sparse_f32(const float* mat, const int* cptrs, const int* rows,
                              const float* cvals, int nov, int grid_dim, int block_dim);
double perm_xlocal_sparse_f64(const double* mat, const int* cptrs, const int* rows,
                              const double* cvals, int nov, int grid_dim, int block_dim);
double perm_xshared_sparse_i32(const int* mat, const int* cptrs, const int* rows,
                               const int* cvals, int nov, int grid_dim, int block_dim);
double perm_xshared_sparse_f32(const float* mat, const int* cptrs, const int* rows,
                               const float* cvals, int nov, int grid_dim, int block_dim);
double perm_xshared_sparse_f64(const double* mat, const int* cptrs, const int* rows,
                               const double* cvals, int nov, int grid_dim, int block_dim);
double perm_xshared_coalescing_sparse_i32(const int* mat, const int* cptrs, const int* rows,
                                          const int* cvals, int nov, int grid_dim,
                                          int block_dim);
double perm_xshared_coalescing_sparse_f32(const float* mat, const int* cptrs, const int* rows,
                                          const float* cvals, int nov, int grid_dim,
                                          int block_dim);
double perm_xshared_coalescing_sparse_f64(const double* mat, const int* cptrs, const int* rows,
                                          const double* cvals, int nov, int grid_dim,
                                          int block_dim);
double perm_xshared_coalescing_mshared_sparse_i32(const int* mat, const int* cptrs,
                                                  const int* rows, const int* cvals, int nov,
                                                  int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_sparse_f32(const float* mat, const int* cptrs,
                                                  const int* rows, const float* cvals, int nov,
                                                  int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_sparse_f64(const double* mat, const int* cptrs,
                                                  const int* rows, const double* cvals, int nov,
                                                  int grid_dim, int block_dim);

/* ---- full-result forms of the same paths (extension) ---------------------
 * algo_id: the reference's dense ID 0-4 / sparse ID 1-4 (main.cu:34-58,
 * :108-127).  Returns a PERM_* status and fills *out. */
int perm_dense(int dtype, const void* mat, int nov, int algo_id, perm_result* out);
int perm_sparse(int dtype, const void* mat, const int* cptrs, const int* rows, const void* cvals,
                int nov, int algo_id, perm_result* out);

/* ---- resident plans (extension) ------------------------------------------
 * A plan uploads one matrix's column store once; perm_plan_run enqueues the
 * Ryser kernel + grid reduction on the plan's stream and waits for it, so a
 * caller can time device work with the inputs already in HBM. */
typedef struct perm_plan perm_plan;
int perm_plan_dense(int dtype, const void* mat, int nov, perm_plan** plan);
int perm_plan_sparse(int dtype, const void* mat, const int* cptrs, const int* rows,
                     const void* cvals, int nov, perm_plan** plan);
int perm_plan_run(perm_plan* plan, perm_result* out);
/* Run one kernel mode explicitly: 0 FP64 (production), 1 double-double
 * products and sums, 2 double-double X as well (no X drift; verification
 * mode, several times slower).  perm_plan_run picks 0, then 1 when an
 * integer matrix needs it to round to the exact integer. */
int perm_plan_run_mode(perm_plan* plan, int mode, perm_result* out);
void perm_plan_destroy(perm_plan* plan);

/* ---- CPU path of `perman -c` (parallel_perman64, algo.h:663; and
 * parallel_perman64_sparse, algo.h:569).  Native OpenMP Ryser with FP64 X
 * and a double-double outer sum. -------------------------------------------- */
double perm_parallel_perman64_i32(const int* mat, int nov, int threads);
double perm_parallel_perman64_f32(const float* mat, int nov, int threads);
double perm_parallel_perman64_f64(const double* mat, int nov, int threads);
double perm_parallel_perman64_sparse_i32(const int* mat, const int* cptrs, const int* rows,
                                         const int* cvals, int nov, int threads);
double perm_parallel_perman64_sparse_f32(const float* mat, const int* cptrs, const int* rows,
                                         const float* cvals, int nov, int threads);
double perm_parallel_perman64_sparse_f64(const double* mat, const int* cptrs, const int* rows,
                                         const double* cvals, int nov, int threads);

#ifdef __cplusplus
}
#endif

#endif /* PERMAN_HIP_H */
