// perman_ffi.cpp -- the foreign-function entry points of the reference's
// Python/MATLAB connector (py_connector.c:63-190), served by the GPU path.
//
// The reference's connector reads a matrix file (or takes a MATLAB int
// matrix), orders it for the chosen algorithm and calls one of its CPU
// permanent routines (decide_and_call, py_connector.c:17-58):
//   0-3 Rasmussen / scaling approximations, 4 parallel_perman64_sparse,
//   5 parallel_perman64, 6/7 parallel_skip_perman64_w[_balanced],
//   8 perman64 (serial double-X Ryser).
// Here the exact ones (4-8) run on the GPU: 5 and 8 through the dense path,
// 4, 6 and 7 through the sparse (-s) path.  The approximations are not part
// of this library: they return NaN with PERM_ERR_ARG.  `nt` (CPU threads)
// and `x`, `y`, `z` (approximation parameters) are accepted and unused.
//
// Two behaviours of the reference's connector are NOT reproduced (both are
// slips there): it reads every file as a 0/1 pattern (ReadMatrix with
// generic = 0, py_connector.c:75 and util.h:352-356), and it stores the
// result in an int (py_connector.c:20), truncating it.  Here the file's
// values are used as read and the permanent is returned as a double.  A
// caller that wants the 0/1 pattern passes a 0/1 matrix (for MATLAB,
// `double(M ~= 0)` converted to int32).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "../../include/perman_hip.h"
#include "../host/matrix_io.hpp"
#include "perman_internal.hpp"

namespace {

double nan_value() { return std::numeric_limits<double>::quiet_NaN(); }

bool exact_algorithm(int algorithm) { return algorithm >= 4 && algorithm <= 8; }
bool sparse_algorithm(int algorithm) {
  return algorithm == 4 || algorithm == 6 || algorithm == 7;
}

template <class T>
double run_matrix(const std::vector<T>& mat, int n, int dtype, int algorithm) {
  perm_result r;
  int st;
  if (sparse_algorithm(algorithm)) {
    const permio::Compressed<T> c = permio::compress(mat, n);
    st = perm_sparse(dtype, mat.data(), c.cptrs.data(), c.rows.data(), c.cvals.data(), n, 4, &r);
  } else {
    st = perm_dense(dtype, mat.data(), n, 4, &r);
  }
  return st == PERM_OK ? r.value : nan_value();
}

int dtype_of(const std::string& type) {
  if (type == "int") return PERM_DTYPE_I32;
  if (type == "float") return PERM_DTYPE_F32;
  if (type == "double") return PERM_DTYPE_F64;
  return -1;
}

}  // namespace

extern "C" {

void perm_connect(void) { std::printf("libperman_hip connected\n"); }

double perm_read_calculate_return(const char* filename, int algorithm, int nt, int x, int y,
                                  int z) {
  (void)nt, (void)x, (void)y, (void)z;
  if (!filename) return (perm::fail(PERM_ERR_ARG, "null file name"), nan_value());
  if (!exact_algorithm(algorithm))
    return (perm::fail(PERM_ERR_ARG, "algorithm " + std::to_string(algorithm) +
                                         ": only the exact algorithms 4-8 are available"),
            nan_value());
  std::ifstream in(filename);
  permio::Header h;
  if (!in || !permio::read_header(in, h))
    return (perm::fail(PERM_ERR_ARG, std::string("cannot read ") + filename), nan_value());
  const int dtype = dtype_of(h.type);
  if (dtype < 0 || h.n < 1)
    return (perm::fail(PERM_ERR_ARG, "corrupted header in " + std::string(filename)),
            nan_value());
  if (dtype == PERM_DTYPE_I32) {
    std::vector<int> m;
    permio::read_entries(in, h.n, false, m);
    return run_matrix(m, h.n, dtype, algorithm);
  }
  if (dtype == PERM_DTYPE_F32) {
    std::vector<float> m;
    permio::read_entries(in, h.n, false, m);
    return run_matrix(m, h.n, dtype, algorithm);
  }
  std::vector<double> m;
  permio::read_entries(in, h.n, false, m);
  return run_matrix(m, h.n, dtype, algorithm);
}

double perm_matlab_calculate_return(const int* mat, int algorithm, int nt, int x, int y, int z,
                                    int nov, int nnz) {
  (void)nt, (void)x, (void)y, (void)z, (void)nnz;
  if (!mat || nov < 1) return (perm::fail(PERM_ERR_ARG, "null matrix or nov < 1"), nan_value());
  if (!exact_algorithm(algorithm))
    return (perm::fail(PERM_ERR_ARG, "algorithm " + std::to_string(algorithm) +
                                         ": only the exact algorithms 4-8 are available"),
            nan_value());
  std::vector<int> m(mat, mat + (size_t)nov * nov);
  return run_matrix(m, nov, PERM_DTYPE_I32, algorithm);
}

}  // extern "C"
