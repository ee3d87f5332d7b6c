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
#define PERM_PENDING 5      /* perm_server_poll: the request is not finished yet */

#define PERM_DTYPE_I32 0
#define PERM_DTYPE_F32 1
#define PERM_DTYPE_F64 2

/* Result of one permanent.  value = scale * (sum_hi + sum_lo) where
 * scale = 4(n&1)-2 and sum is the double-double Ryser sum; abs_sum is
 * |scale| * sum of |terms| (the condition of the alternating sum).
 *
 * err_bound: worst-case bound (every rounding aligned, first order in the
 * unit roundoff u = 2^-53) of the product, accumulation and start-vector
 * roundings.  For an integer matrix X is exact and this bounds the whole
 * error.  For a non-integer matrix in mode 0 or 1 it does NOT include the
 * drift of X along the Gray walk (each x_j accumulates the roundings of its
 * column additions; how they move a term depends on how close x_j comes to
 * zero, so no useful worst-case bound exists); err_est covers it.
 * err_est: three standard deviations of the error when roundings behave as
 * independent random errors, from moments the kernel measures (sums over
 * runs of consecutive terms of the squared absolute and signed run sums);
 * calibrated against the double-double and exact oracles.
 * mode: 0 = FP64 products (the fast kernel), 1 = double-double products
 * (integer-valued matrices: exact X, so the result rounds to the exact
 * integer), 2 = double-double X and products (no X drift).  Results are held
 * to 1e-10 relative: an FP64 result with err_est > 1e-10 |value| is computed
 * again in mode 1 (integer matrices) or mode 2 (others), automatically, and
 * `certified` says whether the final result meets the target (exact, or
 * err_est <= 1e-10 |value|).  On the single-matrix path a non-integer FP64
 * result whose modelled err_est misses the target first has its error
 * measured: 256 of its workgroups, sampled in proportion to their modelled
 * variance, are walked again in mode 2, and err_est becomes three standard
 * deviations of the error estimated from them (verified_units); the full
 * mode 2 pass runs only if that does not certify it.  A matrix whose nonzero pattern has no perfect
 * matching has permanent exactly 0: it is returned as 0 (zero_pattern = 1,
 * errors 0) without running a walk. */
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
  int nu;               /* FP64 kernel: rows touched by Gray column 0 (pair/nested variant), or n */
  double main_ms;       /* single-matrix path: device time of the main walk kernel */
  double main_subsets;  /* subsets that kernel evaluated */
  int nu1;              /* FP64 kernel: rows touched by Gray columns 0-1 (nested variant), or n */
  int nu2;              /* FP64 kernel: rows touched by Gray columns 0-2 (nested variant), or n */
  int certified;        /* 1: exact, or err_est <= 1e-10 |value| */
  int zero_pattern;     /* 1: no perfect matching in the nonzero pattern; value is exactly 0 */
  double clock_ghz;     /* single-matrix path: mean shader clock of the first pass's walk,
                           from its workgroups' cycle and 100 MHz real-time counters (0: none) */
  int verified_units;   /* single-matrix path: workgroups of the FP64 walk recomputed in mode 2
                           to measure its error (err_est then comes from them); 0: none */
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

/* ---- dense path without device allocation (extension) --------------------
 * No counterpart in the reference (its wrappers cudaMalloc and cudaFree per
 * call, gpu_exact_dense.cu:669-689).  For callers that cannot allocate device
 * memory -- tight serving loops, harnesses that forbid hipMalloc: the same
 * computation and bits as perm_dense, with every device buffer in one static
 * device array of 4 MiB per GPU (held until the process ends), so nov <= 35;
 * larger nov fail with PERM_ERR_SIZE.  Calls no hipMalloc, hipFree or other
 * allocating HIP routine (the first call on a GPU creates that GPU's stream
 * and events).  Calls on one GPU share its array: they are serialised by a
 * lock inside the library; calls from several processes each have their own
 * copy. */
int perm_dense_noalloc(int dtype, const void* mat, int nov, int algo_id, perm_result* out);

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
/* The plan's walk split: 2^high_bits lanes of 2^walk_bits subsets each, in
 * `blocks` workgroups of 256 lanes; lane l fixes the Gray columns
 * [walk_bits, n-1) (in the plan's walk order) by the bits of l. */
int perm_plan_geometry(const perm_plan* plan, int* walk_bits, int* high_bits,
                       unsigned long long* blocks);
/* The sum over the workgroups [block_begin, block_end) only (lanes
 * [256 block_begin, 256 block_end)) in kernel mode `mode`: the unit one
 * device of a multi-GPU run computes (the ranges of all devices add up to
 * the permanent), and a piece of the walk small enough to check against a
 * CPU evaluation of the same subsets.  out->sum_hi/sum_lo: the range's
 * signed sum; value = (4(n&1)-2) * sum; certified is 0 (the target applies
 * to whole results). */
int perm_plan_run_blocks(perm_plan* plan, int mode, unsigned long long block_begin,
                         unsigned long long block_end, perm_result* out);
void perm_plan_destroy(perm_plan* plan);

/* ---- batches of mixed n (extension) --------------------------------------
 * Many matrices (n <= 40) resident on the device, computed by persistent
 * kernels (one per variant: n and walk shape, concurrent on several streams);
 * the first perm_batch_run after a change launches them directly, the
 * second captures them in one hipGraph, and later runs replay it.  Integer matrices
 * whose FP64 error could reach 1/2 then get a double-double pass over the
 * same items (the batch kernels' double-double variant).  A matrix is cut into 64-lane
 * items of about 2^item_log2 subsets that depend only on its own n, so its
 * result is identical alone and in any mix.  item_log2 (default 17),
 * streams (default 8) and max_grid (workgroups per variant kernel, default
 * 1024) are the work-split parameters (<= 0 keeps a value). */
typedef struct perm_batch perm_batch;
int perm_batch_create(perm_batch** batch);
int perm_batch_set(perm_batch* batch, int item_log2, int streams, int max_grid);
int perm_batch_add_dense(perm_batch* batch, int dtype, const void* mat, int nov, int* index);
int perm_batch_add_sparse(perm_batch* batch, int dtype, const void* mat, const int* cptrs,
                          const int* rows, const void* cvals, int nov, int* index);
/* Adds `count` dense matrices at once (mats[i]: dtypes[i], order novs[i]),
 * preparing them on `threads` host threads; on failure nothing is added and
 * the first failing matrix is reported.  *first_index: index of the first. */
int perm_batch_add_dense_many(perm_batch* batch, int count, const int* dtypes,
                              const void* const* mats, const int* novs, int threads,
                              int* first_index);
int perm_batch_size(const perm_batch* batch);
/* One batch end to end: adds `count` dense matrices to an EMPTY batch and
 * runs it, pipelined -- the matrices are prepared on `threads` host threads
 * in decreasing order n in three phases (cut at 15 % and 45 % of the work),
 * and each phase's kernels run while the next phase is prepared and
 * uploaded.  results: one
 * per matrix in the caller's order, the same values as perm_batch_add_dense_many
 * + perm_batch_run (a matrix's items depend only on its n); the batch is then
 * built, and later perm_batch_run calls replay it (results still in the
 * caller's order). */
int perm_batch_compute_dense(perm_batch* batch, int count, const int* dtypes,
                             const void* const* mats, const int* novs, int threads,
                             perm_result* results, double* device_ms);
/* results: one per matrix, in insertion order; device_ms: graph time plus
 * the double-double passes, if any. */
int perm_batch_run(perm_batch* batch, perm_result* results, double* device_ms);
void perm_batch_destroy(perm_batch* batch);
/* Work split of a batch and the autotuner that picks it: subsets per item
 * (2^item_log2: Gray-code chunk of 2^(item_log2-6) subsets per lane),
 * streams carrying the per-variant kernels, workgroups per variant kernel.
 * perm_batch_autotune times candidate splits on the batch itself (a
 * coordinate search, each candidate rebuilt and replayed `reps` times) and
 * keeps the fastest; results then follow that split. */
typedef struct perm_batch_tuning {
  int item_log2, streams, max_grid;
  double device_ms;  /* mean device time of the chosen split */
  int trials;        /* candidate splits timed */
} perm_batch_tuning;
int perm_batch_autotune(perm_batch* batch, int reps, perm_batch_tuning* best);
int perm_batch_get(const perm_batch* batch, perm_batch_tuning* out);

/* ---- sparse skipper, ID 7: gpu_perman64_xshared_coalescing_mshared_skipper(
 * T* mat, int* rptrs, int* cols, int* cptrs, int* rows, T* cvals, int nov,
 * int grid_dim, int block_dim), gpu_exact_sparse.cu:1124.  Skips, per wave,
 * aligned runs of Gray steps whose terms are all zero (a row of x that is
 * zero and untouched by the run's columns); needs exact zeros, so it runs
 * for integer matrices and n >= 16 -- other inputs take the sparse path
 * (same sum).  x0 from the dense matrix, column updates from the CCS;
 * the CRS (rptrs/cols) is accepted and not needed.  perm_skip's result:
 * main_subsets = subsets actually walked (the rest were skipped). */
int perm_skip(int dtype, const void* mat, const int* rptrs, const int* cols, const int* cptrs,
              const int* rows, const void* cvals, int nov, perm_result* out);
double perm_xshared_coalescing_mshared_skipper_i32(const int* mat, const int* rptrs,
                                                    const int* cols, const int* cptrs,
                                                    const int* rows, const int* cvals, int nov,
                                                    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_skipper_f32(const float* mat, const int* rptrs,
                                                    const int* cols, const int* cptrs,
                                                    const int* rows, const float* cvals, int nov,
                                                    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_skipper_f64(const double* mat, const int* rptrs,
                                                    const int* cols, const int* cptrs,
                                                    const int* rows, const double* cvals,
                                                    int nov, int grid_dim, int block_dim);

/* ---- multi-GPU and hybrid IDs on one GPU ----------------------------------
 * Dense ID 5 gpu_perman64_xshared_coalescing_mshared_multigpu (gpu_exact_dense.cu:702),
 * ID 6 ..._multigpucpu_chunks (:777), ID 66 ..._multigpu_manual_distribution
 * (:914); sparse ID 5 ..._multigpu_sparse (gpu_exact_sparse.cu:917), ID 6
 * ..._multigpucpu_chunks_sparse (:996), ID 8 ..._multigpucpu_chunks_skipper
 * (:1193), ID 66 ..._multigpu_sparse_manual_distribution (:1328).
 * This is a single-GPU build: gpu_num is accepted and one GPU computes the
 * whole Gray range.  The chunked forms (ID 6) cut the FP64 walk into
 * 2^(n-29) chunks (2^(n-28) with cpu != 0, as the reference) handed out by a
 * shared counter to the GPU and, with cpu != 0, to a CPU worker using
 * `threads` OpenMP threads; the partial sums are combined by the same grid
 * reduction as the one-launch path.
 * Without the CPU worker the value equals the one-launch value bitwise; with
 * it, the chunks the CPU took differ in rounding only.  ID 8 runs the skipper
 * on the GPU alone.  perm_chunked is the full-result form (cptrs NULL:
 * dense); *cpu_chunks receives the number of chunks the CPU computed. */
int perm_chunked(int dtype, const void* mat, const int* cptrs, const int* rows,
                 const void* cvals, int nov, int use_cpu, int threads, perm_result* out,
                 int* cpu_chunks);
double perm_xshared_coalescing_mshared_multigpu_dense_i32(const int* mat, int nov, int gpu_num,
    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_manual_distribution_dense_i32(
    const int* mat, int nov, int gpu_num, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpucpu_chunks_dense_i32(
    const int* mat, int nov, int gpu_num, int cpu, int threads, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_sparse_i32(
    const int* mat, const int* cptrs, const int* rows, const int* cvals, int nov, int gpu_num,
    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_sparse_manual_distribution_i32(
    const int* mat, const int* cptrs, const int* rows, const int* cvals, int nov, int gpu_num,
    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpucpu_chunks_sparse_i32(
    const int* mat, const int* cptrs, const int* rows, const int* cvals, int nov, int gpu_num,
    int cpu, int threads, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpucpu_chunks_skipper_i32(
    const int* mat, const int* rptrs, const int* cols, const int* cptrs, const int* rows,
    const int* cvals, int nov, int gpu_num, int cpu, int threads, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_dense_f32(const float* mat, int nov, int gpu_num,
    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_manual_distribution_dense_f32(
    const float* mat, int nov, int gpu_num, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpucpu_chunks_dense_f32(
    const float* mat, int nov, int gpu_num, int cpu, int threads, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_sparse_f32(
    const float* mat, const int* cptrs, const int* rows, const float* cvals, int nov, int gpu_num,
    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_sparse_manual_distribution_f32(
    const float* mat, const int* cptrs, const int* rows, const float* cvals, int nov, int gpu_num,
    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpucpu_chunks_sparse_f32(
    const float* mat, const int* cptrs, const int* rows, const float* cvals, int nov, int gpu_num,
    int cpu, int threads, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpucpu_chunks_skipper_f32(
    const float* mat, const int* rptrs, const int* cols, const int* cptrs, const int* rows,
    const float* cvals, int nov, int gpu_num, int cpu, int threads, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_dense_f64(const double* mat, int nov, int gpu_num,
    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_manual_distribution_dense_f64(
    const double* mat, int nov, int gpu_num, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpucpu_chunks_dense_f64(
    const double* mat, int nov, int gpu_num, int cpu, int threads, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_sparse_f64(
    const double* mat, const int* cptrs, const int* rows, const double* cvals, int nov, int gpu_num,
    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpu_sparse_manual_distribution_f64(
    const double* mat, const int* cptrs, const int* rows, const double* cvals, int nov, int gpu_num,
    int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpucpu_chunks_sparse_f64(
    const double* mat, const int* cptrs, const int* rows, const double* cvals, int nov, int gpu_num,
    int cpu, int threads, int grid_dim, int block_dim);
double perm_xshared_coalescing_mshared_multigpucpu_chunks_skipper_f64(
    const double* mat, const int* rptrs, const int* cols, const int* cptrs, const int* rows,
    const double* cvals, int nov, int gpu_num, int cpu, int threads, int grid_dim, int block_dim);

/* ---- the reference's Python/MATLAB connector (py_connector.c:63-190) ------
 * perm_read_calculate_return(filename, algorithm, nt, x, y, z) replaces
 * read_calculate_return (py_connector.c:63) and
 * perm_matlab_calculate_return(mat, algorithm, nt, x, y, z, nov, nnz) replaces
 * matlab_calculate_return (py_connector.c:152); perm_connect replaces
 * connect (:59).  Algorithms 4-8 (the exact ones: 4/6/7 sparse, 5/8 dense)
 * run on the GPU; 0-3 (approximations) return NaN with PERM_ERR_ARG; nt, x,
 * y, z are unused.  Unlike the reference's connector, the file's values are
 * used (it reads every file as a 0/1 pattern) and the result is not
 * truncated to int. */
void perm_connect(void);
double perm_read_calculate_return(const char* filename, int algorithm, int nt, int x, int y,
                                  int z);
double perm_matlab_calculate_return(const int* mat, int algorithm, int nt, int x, int y, int z,
                                    int nov, int nnz);

/* ---- server: continuous batching of requests (extension) ---------------
 * No counterpart in the reference (its entry points compute one permanent
 * per call; its nearest structure is the chunk queue of its multi-GPU
 * wrappers, gpu_exact_dense.cu:776-904).  Requests (n <= 40, dense or
 * sparse) are submitted from any thread and return a ticket at once: the
 * submission only copies the matrix; a pool of host threads prepares it
 * (start vector, column store, walk shape, padded to a served order that is
 * a multiple of 4) and a scheduler thread writes it into a pinned mirror of
 * a device arena and publishes it to ONE persistent kernel, launched as a
 * captured hipGraph, that polls a ring of pass descriptors in pinned host
 * memory and copies each new pass's range of the mirror into the arena
 * itself (no host-issued copies).  The kernel's waves claim 64-lane items
 * in publication order; the wave that finishes a request's last item sums
 * its item records on the device in item order and appends the sum to a
 * completion queue in host memory: a result is bitwise the same alone or in
 * any mix.  Argument errors found while preparing (a
 * malformed CCS) are reported by the wait/poll of that request.
 * Admission is bounded by the resident requests, the arena and the queued
 * work.  perm_server_wait blocks until the request is done and frees its
 * ticket; perm_server_poll returns PERM_PENDING until then.  In a served
 * result, kernel_ms holds the request's latency (submit to result, also
 * returned in *latency_ms).  Every result meets the accuracy rule of
 * perm_dense (integer results exact; others certified to 1e-10 by the
 * model, a sampled measurement or a double-double pass). */
typedef struct perm_server perm_server;
typedef struct perm_server_config {
  int item_log2;          /* subsets per item, log2 (default 16) */
  int grid;               /* workgroups of the persistent kernel (default 1024: 4 waves per SIMD) */
  int max_items;          /* most items per pass (default 4096, at most 16384: more lane groups per item above) */
  double idle_us;         /* a launch of the persistent kernel ends after this long without work (2000) */
  int max_resident;       /* requests resident on the device (default 8192, at most 32768) */
  double max_queued_work; /* admitted work cap, FP64 lane-instructions (4e13) */
  int prep_threads;       /* host threads preparing submitted matrices (default 4) */
  double arena_mb;        /* device arena for matrices and item records, MiB (default 256) */
} perm_server_config;
typedef struct perm_server_stats {
  long long submitted, admitted, completed;
  long long launches;     /* launches (graph replays) of the persistent kernel */
  long long items;        /* items published */
  long long pending;      /* prepared requests not yet admitted */
  long long uploads;      /* publications (tail advances) of staged passes */
  long long stops;        /* launches stopped so that arena space could be reused */
  long long laps;         /* times the arena's block ring wrapped to its start */
  double admit_ms;        /* scheduler time admitting and staging requests */
  double publish_ms;      /* scheduler time publishing passes and launching */
  double retire_ms;       /* scheduler time turning completed passes into results */
  double stop_ms;         /* scheduler time waiting for stopped launches */
  double gpu_ms;          /* summed duration of the kernel's launches */
  long long dd_passes;    /* passes in a double-double mode: escalations and sampled error measurements */
  double prep_ms;         /* preparation time summed over the prep threads */
  double submit_ms;       /* time in the submission calls (copying matrices, queueing) */
} perm_server_stats;
void perm_server_default_config(perm_server_config* cfg);
int perm_server_create(const perm_server_config* cfg, perm_server** server);
int perm_server_submit_dense(perm_server* server, int dtype, const void* mat, int nov,
                             unsigned long long* id);
int perm_server_submit_sparse(perm_server* server, int dtype, const void* mat, const int* cptrs,
                              const int* rows, const void* cvals, int nov,
                              unsigned long long* id);
/* Submits `count` matrices at once (one lock, one wake-up): ids[i] receives
 * the ticket of mats[i] (dtypes[i], order novs[i]); a request takes the
 * sparse path when cptrs is not NULL and cptrs[i] is not NULL (with rows[i],
 * cvals[i]).  Returns the first argument error; requests before it are
 * submitted. */
int perm_server_submit_many(perm_server* server, int count, const int* dtypes,
                            const void* const* mats, const int* const* cptrs,
                            const int* const* rows, const void* const* cvals, const int* novs,
                            unsigned long long* ids);
int perm_server_wait(perm_server* server, unsigned long long id, perm_result* out,
                     double* latency_ms);
/* Waits for `count` requests in order; out[i] / latency_ms[i] (latency_ms
 * may be NULL) as perm_server_wait.  Every id must be known; returns the
 * first failing request's status after collecting all of them. */
int perm_server_wait_many(perm_server* server, const unsigned long long* ids, int count,
                          perm_result* out, double* latency_ms);
int perm_server_poll(perm_server* server, unsigned long long id, perm_result* out,
                     double* latency_ms);
int perm_server_stats_get(perm_server* server, perm_server_stats* out);
/* Finishes every submitted request, then frees the server (uncollected
 * results are dropped). */
void perm_server_destroy(perm_server* server);
/* Open-loop driver (load testing): submits request i (arguments as
 * perm_server_submit_many) at arrival_s[i] seconds after the call from a
 * thread of its own, and collects every result in order on the calling
 * thread: out[i], latency_ms[i] (from the arrival time to the result, so
 * a submitter that falls behind shows; may be NULL), *wall_s the time from
 * the call until the last result (may be NULL).  Requests already due when
 * the thread gets to them are submitted in one call. */
int perm_server_drive(perm_server* server, int count, const int* dtypes,
                      const void* const* mats, const int* const* cptrs, const int* const* rows,
                      const void* const* cvals, const int* novs, const double* arrival_s,
                      perm_result* out, double* latency_ms, double* wall_s);
/* Autotuner of the served work split: closed-loop trials of the sample
 * (arguments as perm_server_submit_many) over item_log2 {12..20} and the
 * kernel's grid {256..2048} by coordinate descent from *base (NULL:
 * defaults); *best receives the fastest configuration and
 * *best_req_per_s its throughput.  Each trial creates its own server. */
int perm_server_autotune(const perm_server_config* base, int count, const int* dtypes,
                         const void* const* mats, const int* const* cptrs, const int* const* rows,
                         const void* const* cvals, const int* novs, perm_server_config* best,
                         double* best_req_per_s);

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

/* ---- CPU skip variants of `perman -c -s -p2/-p3` (parallel_skip_perman64_w,
 * algo.h:749; parallel_skip_perman64_w_balanced, algo.h:886).  The Gray walk
 * jumps over runs of subsets in which some row sum x_j is exactly zero (x_j
 * stays zero until a column touching row j flips); both names run the same
 * dynamically scheduled walk over aligned blocks, combined in block order.
 * The CRS (rptrs/cols) must list the same entries as the CCS; rvals is
 * accepted for the signature.  perm_cpu_skip is the full-result form:
 * main_subsets = the terms actually evaluated, kernel_ms = host wall time. */
int perm_cpu_skip(int dtype, const int* rptrs, const int* cols, const void* rvals,
                  const int* cptrs, const int* rows, const void* cvals, int nov, int threads,
                  perm_result* out);
double perm_parallel_skip_perman64_w_i32(const int* rptrs, const int* cols, const int* rvals,
                                         const int* cptrs, const int* rows, const int* cvals,
                                         int nov, int threads);
double perm_parallel_skip_perman64_w_f32(const int* rptrs, const int* cols, const float* rvals,
                                         const int* cptrs, const int* rows, const float* cvals,
                                         int nov, int threads);
double perm_parallel_skip_perman64_w_f64(const int* rptrs, const int* cols, const double* rvals,
                                         const int* cptrs, const int* rows, const double* cvals,
                                         int nov, int threads);
double perm_parallel_skip_perman64_w_balanced_i32(const int* rptrs, const int* cols,
                                                  const int* rvals, const int* cptrs,
                                                  const int* rows, const int* cvals, int nov,
                                                  int threads);
double perm_parallel_skip_perman64_w_balanced_f32(const int* rptrs, const int* cols,
                                                  const float* rvals, const int* cptrs,
                                                  const int* rows, const float* cvals, int nov,
                                                  int threads);
double perm_parallel_skip_perman64_w_balanced_f64(const int* rptrs, const int* cols,
                                                  const double* rvals, const int* cptrs,
                                                  const int* rows, const double* cvals, int nov,
                                                  int threads);

#ifdef __cplusplus
}
#endif

#endif /* PERMAN_HIP_H */
