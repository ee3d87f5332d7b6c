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
//   ReadMatrix<T>                 util.h:342-357  (text reader; the header
//                                 line is parsed as main.cu:484-488 does)
//   matrix2compressed_sortOrder<T> util.h:553-619 (-r1)
//   matrix2compressed_skipOrder<T> util.h:621-684 (-r2)
#include <fstream>
#include <sstream>
#include <string>

#include "algo.h"

namespace {

// Read a matrix file as the reference's main does: header `nov nnz type`,
// then ReadMatrix into a zero-filled buffer of nov*nov entries (the
// reference allocates with new[] and does not clear it).  Returns nov or -1.
template <class T>
int read_file(const char* path, int generic, T* mat, int cap, int* nnz_out) {
  std::ifstream inFile(path);
  std::string line;
  if (!std::getline(inFile, line)) return -1;
  std::istringstream iss(line);
  int nov, nnz;
  std::string type;
  if (!(iss >> nov >> nnz >> type) || nov * nov > cap) return -1;
  for (int i = 0; i < nov * nov; i++) mat[i] = 0;
  ReadMatrix(mat, inFile, nov, generic != 0);
  *nnz_out = nnz;
  return nov;
}

// One of the reference's CCS/CRS builders on mat (rewritten in place by the
// ordering ones) with the arrays copied out; nnz is the capacity the
// builders allocate (the file header's count).  Returns the stored count.
template <class T, class F>
int compressed(F builder, T* mat, int nov, int nnz, int* cptrs_out, int* rows_out, T* cvals_out,
               int* rptrs_out, int* cols_out, T* rvals_out) {
  int *cptrs, *rows, *rptrs, *cols;
  T *cvals, *rvals;
  builder(mat, cptrs, rows, cvals, rptrs, cols, rvals, nov, nnz);
  const int stored = cptrs[nov];
  for (int i = 0; i <= nov; i++) {
    cptrs_out[i] = cptrs[i];
    rptrs_out[i] = rptrs[i];
  }
  for (int e = 0; e < stored; e++) {
    rows_out[e] = rows[e];
    cvals_out[e] = cvals[e];
  }
  for (int e = 0; e < rptrs[nov]; e++) {
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

}  // namespace

extern "C" {

int ref_read_matrix_f64(const char* path, int generic, double* mat, int cap, int* nnz) {
  return read_file(path, generic, mat, cap, nnz);
}
int ref_read_matrix_f32(const char* path, int generic, float* mat, int cap, int* nnz) {
  return read_file(path, generic, mat, cap, nnz);
}
int ref_read_matrix_i32(const char* path, int generic, int* mat, int cap, int* nnz) {
  return read_file(path, generic, mat, cap, nnz);
}

// order: 0 matrix2compressed, 1 sortOrder (-r1), 2 skipOrder (-r2)
#define REF_ORDERED(SFX, T)                                                                \
  int ref_compressed_order_##SFX(int order, T* mat, int nov, int nnz, int* cptrs, int* rows, \
                                 T* cvals, int* rptrs, int* cols, T* rvals) {               \
    if (order == 1)                                                                        \
      return compressed<T>(matrix2compressed_sortOrder<T>, mat, nov, nnz, cptrs, rows,      \
                           cvals, rptrs, cols, rvals);                                     \
    if (order == 2)                                                                        \
      return compressed<T>(matrix2compressed_skipOrder<T>, mat, nov, nnz, cptrs, rows,      \
                           cvals, rptrs, cols, rvals);                                     \
    return compressed<T>(matrix2compressed<T>, mat, nov, nnz, cptrs, rows, cvals, rptrs,    \
                         cols, rvals);                                                     \
  }
REF_ORDERED(f64, double)
REF_ORDERED(f32, float)
REF_ORDERED(i32, int)


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
