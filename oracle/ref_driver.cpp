// ref_driver.cpp -- C-ABI driver around the reference's own CPU algorithms.
//
// Test/benchmark infrastructure only.  It is compiled by `make -C oracle ref`
// against the reference checkout in place (-I<reference>), never copied, and
// the resulting oracle/_ref/libref_algo.so is loaded only by tests/ and by
// bench.py's cpu_baseline leg.  The functions below just instantiate the
// reference templates and forward to them:
//   parallel_perman64<T>          algo.h:662-746  (float-X OpenMP baseline, `-c`)
//   parallel_perman64_sparse<T>   algo.h:568-660
//   perman64<T>                   algo.h:1031-1088 (serial double-X)
//   matrix2compressed<T>          util.h:522-551  (CCS/CRS builder)
#include "algo.h"

extern "C" {

double ref_parallel_perman64_f64(double* mat, int nov, int threads) {
  return parallel_perman64(mat, nov, threads);
}
double ref_parallel_perman64_f32(float* mat, int nov, int threads) {
  return parallel_perman64(mat, nov, threads);
}
double ref_parallel_perman64_i32(int* mat, int nov, int threads) {
  return parallel_perman64(mat, nov, threads);
}
double ref_parallel_perman64_sparse_f64(double* mat, int* cptrs, int* rows, double* cvals,
                                        int nov, int threads) {
  return parallel_perman64_sparse(mat, cptrs, rows, cvals, nov, threads);
}
double ref_perman64_f64(double* mat, int nov) { return perman64(mat, nov); }

// CCS/CRS of a dense row-major matrix; arrays are caller-allocated with room
// for n+1 pointers and nnz entries.  Returns the number of stored entries.
int ref_matrix2compressed_f64(double* mat, int nov, int nnz, int* cptrs_out, int* rows_out,
                              double* cvals_out, int* rptrs_out, int* cols_out,
                              double* rvals_out) {
  int *cptrs, *rows, *rptrs, *cols;
  double *cvals, *rvals;
  matrix2compressed(mat, cptrs, rows, cvals, rptrs, cols, rvals, nov, nnz);
  int stored = cptrs[nov];
  for (int i = 0; i <= nov; i++) {
    cptrs_out[i] = cptrs[i];
    rptrs_out[i] = rptrs[i];
  }
  for (int e = 0; e < stored; e++) {
    rows_out[e] = rows[e];
    cvals_out[e] = cvals[e];
    cols_out[e] = cols[e];
    rvals_out[e] = rvals[e];
  }
  delete[] cptrs;
  delete[] rows;
  delete[] cvals;
  delete[] rptrs;
  delete[] cols;
  delete[] rvals;
  return stored;
}

}  // extern "C"
