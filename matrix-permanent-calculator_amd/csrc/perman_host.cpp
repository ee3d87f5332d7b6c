// perman_host.cpp -- host-side state and matrix preparation shared by the
// single-matrix, batch and serving entry points: per-thread status and
// device, conversion of the caller's matrix, Nijenhuis-Wilf start vector,
// zero-skipping walk order and signed column store.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/perman_hip.h"
#include "perman_internal.hpp"
#include "ryser_host.hpp"
#include "ryser_kernels.hpp"

namespace perm {

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(PERM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

thread_local int t_status = PERM_OK;
thread_local std::string t_error;
thread_local int t_device = 0;

int fail(int code, const std::string& msg) {
  t_status = code;
  t_error = msg;
  return code;
}
int ok() {
  t_status = PERM_OK;
  t_error.clear();
  return PERM_OK;
}

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v && *v ? std::atoi(v) : dflt;
}

int current_device() { return t_device; }
void set_device(int d) { t_device = d; }
int last_status() { return t_status; }
const char* last_error() { return t_error.c_str(); }

// Convert the caller's matrix to double (exact for int, float, double).
bool to_double(int dtype, const void* mat, size_t count, std::vector<double>& out) {
  out.resize(count);
  switch (dtype) {
    case PERM_DTYPE_I32:
      for (size_t i = 0; i < count; ++i) out[i] = (double)((const int*)mat)[i];
      return true;
    case PERM_DTYPE_F32:
      for (size_t i = 0; i < count; ++i) out[i] = (double)((const float*)mat)[i];
      return true;
    case PERM_DTYPE_F64:
      std::memcpy(out.data(), mat, count * sizeof(double));
      return true;
  }
  return false;
}

double elem(int dtype, const void* arr, size_t i) {
  switch (dtype) {
    case PERM_DTYPE_I32: return (double)((const int*)arr)[i];
    case PERM_DTYPE_F32: return (double)((const float*)arr)[i];
    default: return ((const double*)arr)[i];
  }
}

int check_n(int n) {
  if (n < 1) return fail(PERM_ERR_ARG, "nov must be >= 1");
  if (n > perm::kMaxDenseN)
    return fail(PERM_ERR_SIZE, "nov=" + std::to_string(n) + " above the compiled maximum " +
                                   std::to_string(perm::kMaxDenseN));
  return PERM_OK;
}

int select_device() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(PERM_ERR_NODEV, "no HIP device visible");
  if (t_device < 0 || t_device >= ndev)
    return fail(PERM_ERR_NODEV, "device " + std::to_string(t_device) + " not present");
  HIP_TRY(hipSetDevice(t_device));
  return PERM_OK;
}

// Host preparation.  Produces the start vector x0 and the Gray columns in
// column-major order (cols[k*n + j] = a[j][k] for the walked columns
// k < n-1), then chooses the walk order (pair_order) and builds the signed
// column store.

// Zero-skipping order for the pair kernels: the walked column with the
// fewest nonzeros goes to Gray position 0 and the rows it touches move to
// the front, padded to a multiple of 4 (nu).  Row and column permutations
// leave the permanent unchanged; PERMAN_PAIR=0 keeps the natural order.
void pair_order(int n, std::vector<double>& x0, std::vector<double>& cols, int& nu) {
  nu = n;
  if (n < perm::kMinPairN || env_int("PERMAN_PAIR", 1) == 0) return;
  const int B = n - 1;
  int best = 0, best_nnz = n + 1;
  for (int k = 0; k < B; ++k) {
    int c = 0;
    for (int j = 0; j < n; ++j) c += cols[(size_t)k * n + j] != 0.0;
    if (c < best_nnz) {
      best_nnz = c;
      best = k;
    }
  }
  const int padded = std::max(4, (best_nnz + 3) / 4 * 4);
  if (padded > n - 4) return;
  std::vector<int> rp, cp;
  for (int j = 0; j < n; ++j)
    if (cols[(size_t)best * n + j] != 0.0) rp.push_back(j);
  for (int j = 0; j < n; ++j)
    if (cols[(size_t)best * n + j] == 0.0) rp.push_back(j);
  cp.push_back(best);
  for (int k = 0; k < B; ++k)
    if (k != best) cp.push_back(k);
  std::vector<double> x2(n), c2(cols.size());
  for (int j = 0; j < n; ++j) x2[j] = x0[rp[j]];
  for (int k = 0; k < B; ++k)
    for (int j = 0; j < n; ++j) c2[(size_t)k * n + j] = cols[(size_t)cp[k] * n + rp[j]];
  x0.swap(x2);
  cols.swap(c2);
  nu = padded;
}

bool integral_matrix(const std::vector<double>& a) {
  for (double v : a)
    if (!(std::fabs(v) < 2147483648.0) || v != std::nearbyint(v)) return false;
  return true;
}

void finish_prepare(int n, std::vector<double>& cols, Prepared& p) {
  pair_order(n, p.x0, cols, p.nu);
  perm::build_column_store(n, [&](int k, int j) { return cols[(size_t)k * n + j]; }, p.cs);
}

int prepare_dense(int dtype, const void* mat, int n, Prepared& p) {
  if (!mat) return fail(PERM_ERR_ARG, "null pointer");
  if (int s = check_n(n)) return s;
  std::vector<double> a;
  if (!to_double(dtype, mat, (size_t)n * n, a)) return fail(PERM_ERR_ARG, "bad dtype");
  p.integral = integral_matrix(a);
  p.x0.assign(n, 0.0);
  perm::nw_start_vector(a.data(), n, p.x0.data());
  std::vector<double> cols((size_t)(n - 1) * n);
  for (int k = 0; k < n - 1; ++k)
    for (int j = 0; j < n; ++j) cols[(size_t)k * n + j] = a[(size_t)j * n + k];
  finish_prepare(n, cols, p);
  return PERM_OK;
}

// Sparse: x0 from the dense matrix, column updates from the CCS arrays
// (reference: gpu_exact_sparse.cu:854-871 and the kernels' CCS walk).
// cols[k*n + j] = CCS value of (j, k) for the n-1 walked columns.
int gather_sparse(int dtype, const void* mat, const int* cptrs, const int* rows,
                  const void* cvals, int n, std::vector<double>& x0, std::vector<double>& cols,
                  bool& integral) {
  if (!mat || !cptrs) return fail(PERM_ERR_ARG, "null pointer");
  if (int s = check_n(n)) return s;
  const int nnz = cptrs[n];
  if (nnz < 0 || cptrs[0] != 0) return fail(PERM_ERR_ARG, "malformed cptrs");
  if (nnz > 0 && (!rows || !cvals)) return fail(PERM_ERR_ARG, "null rows/cvals");
  for (int k = 0; k < n; ++k)
    if (cptrs[k + 1] < cptrs[k]) return fail(PERM_ERR_ARG, "cptrs not monotone");
  for (int e = 0; e < nnz; ++e)
    if (rows[e] < 0 || rows[e] >= n) return fail(PERM_ERR_ARG, "row index out of range");
  std::vector<double> a;
  if (!to_double(dtype, mat, (size_t)n * n, a)) return fail(PERM_ERR_ARG, "bad dtype");
  x0.assign(n, 0.0);
  perm::nw_start_vector(a.data(), n, x0.data());
  cols.assign((size_t)n * n, 0.0);
  for (int k = 0; k < n; ++k)
    for (int e = cptrs[k]; e < cptrs[k + 1]; ++e)
      cols[(size_t)k * n + rows[e]] += elem(dtype, cvals, (size_t)e);
  integral = integral_matrix(a) && integral_matrix(cols);
  cols.resize((size_t)(n - 1) * n);  // the last column lives in x0
  return PERM_OK;
}

int prepare_sparse(int dtype, const void* mat, const int* cptrs, const int* rows,
                   const void* cvals, int n, Prepared& p) {
  std::vector<double> cols;
  if (int s = gather_sparse(dtype, mat, cptrs, rows, cvals, n, p.x0, cols, p.integral)) return s;
  finish_prepare(n, cols, p);
  return PERM_OK;
}

}  // namespace perm
