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
// k < n-1), then chooses the walk order (nest_order) and builds the signed
// column store.

// Walk order for the pair and nested kernels (ryser_dense_impl.hpp).  Gray
// columns 0, 1, 2 become the columns whose row supports admit the compiled
// shape (nest_menu_shape) of least work (nest_cost): candidates for column 0
// are the 8 sparsest columns, for columns 1 and 2 the 12 sparsest.  Rows are
// ordered [column 0's rows | column 1's new rows | column 2's new rows |
// the rest] (ascending index inside each part), so column l touches only
// rows [0, K_l).  Row and column permutations leave the permanent unchanged.
// `menu`: the shapes the kernels that will run the matrix provide.
// PERMAN_PAIR=0 keeps the natural order; PERMAN_NEST=0 only pair shapes.
void nest_order(int n, perm::NestMenu menu, std::vector<double>& x0, std::vector<double>& cols,
                int& nu, int& nu1, int& nu2) {
  nu = nu1 = nu2 = n;
  if (n < perm::kMinPairN || n > perm::kMaxNestN || env_int("PERMAN_PAIR", 1) == 0) return;
  if (env_int("PERMAN_NEST", 1) == 0) menu.mask &= 1u;
  const bool nested = (menu.mask & ~1u) != 0 && n >= perm::kMinNestN;
  const int B = n - 1;
  std::vector<uint64_t> sup(B, 0);
  for (int k = 0; k < B; ++k)
    for (int j = 0; j < n; ++j)
      if (cols[(size_t)k * n + j] != 0.0) sup[k] |= 1ull << j;
  std::vector<int> by_count(B);
  for (int k = 0; k < B; ++k) by_count[k] = k;
  std::stable_sort(by_count.begin(), by_count.end(), [&](int a, int b) {
    return __builtin_popcountll(sup[a]) < __builtin_popcountll(sup[b]);
  });
  // pair shapes only: the cost grows with column 0's support, so the
  // sparsest column (first in by_count) is the choice
  const int n0 = std::min(B, nested ? 8 : 1), n12 = nested ? std::min(B, 12) : 0;
  int best_cost = perm::nest_cost(n, {n, n, n});
  perm::NestShape best{n, n, n};
  int bc[3] = {-1, -1, -1};
  // cheapest compiled shape covering supports of sizes (c0, c01, c012)
  auto cover = [&](int c0, int c01, int c012, perm::NestShape& out) {
    int cost = -1;
    for (int k0 = std::max(4, (c0 + 3) / 4 * 4); k0 <= n - 4; k0 += 4)
      for (int j = 0; j < perm::kNestShapesPerK0; ++j) {
        if (!perm::nest_menu_has(menu, n, k0, j)) continue;
        const perm::NestShape s = perm::nest_menu_shape(n, k0, j);
        if (s.k1 < c01 || s.k2 < c012) continue;
        const int c = perm::nest_cost(n, s);
        if (cost < 0 || c < cost) {
          cost = c;
          out = s;
        }
      }
    return cost;
  };
  for (int i0 = 0; i0 < n0; ++i0) {
    const int a = by_count[i0];
    const int c0 = __builtin_popcountll(sup[a]);
    perm::NestShape s;
    int cost = cover(c0, n, n, s);  // pair shape: columns 1, 2 unrestricted
    if (cost >= 0 && cost < best_cost) {
      best_cost = cost;
      best = s;
      bc[0] = a;
      bc[1] = bc[2] = -1;
    }
    for (int i1 = 0; i1 < n12; ++i1) {
      const int b = by_count[i1];
      if (b == a) continue;
      const uint64_t u01 = sup[a] | sup[b];
      // cover() can only grow with the third union's size, so no column 2
      // beats best_cost unless the smallest union does: skip the scan then
      // (the choice is the one the full scan makes, at a tenth of the cost)
      int min012 = n + 1;
      for (int i2 = 0; i2 < n12; ++i2) {
        const int c = by_count[i2];
        if (c != a && c != b) min012 = std::min(min012, __builtin_popcountll(u01 | sup[c]));
      }
      if (min012 > n) continue;
      perm::NestShape lb_shape;
      const int lb = cover(c0, __builtin_popcountll(u01), min012, lb_shape);
      if (lb < 0 || lb >= best_cost) continue;
      for (int i2 = 0; i2 < n12; ++i2) {
        const int c = by_count[i2];
        if (c == a || c == b) continue;
        const uint64_t u012 = u01 | sup[c];
        cost = cover(c0, __builtin_popcountll(u01), __builtin_popcountll(u012), s);
        if (cost >= 0 && cost < best_cost) {
          best_cost = cost;
          best = s;
          bc[0] = a;
          bc[1] = b;
          bc[2] = c;
        }
      }
    }
  }
  if (bc[0] < 0) return;
  // rows: column 0's, then column 1's new ones, then column 2's, then the rest
  std::vector<int> rp, cp;
  uint64_t taken = 0;
  for (int l = 0; l < 3; ++l) {
    if (bc[l] < 0) continue;
    for (int j = 0; j < n; ++j)
      if (((sup[bc[l]] >> j) & 1) && !((taken >> j) & 1)) {
        rp.push_back(j);
        taken |= 1ull << j;
      }
  }
  for (int j = 0; j < n; ++j)
    if (!((taken >> j) & 1)) rp.push_back(j);
  for (int l = 0; l < 3; ++l)
    if (bc[l] >= 0) cp.push_back(bc[l]);
  for (int k = 0; k < B; ++k)
    if (k != bc[0] && k != bc[1] && k != bc[2]) cp.push_back(k);
  std::vector<double> x2(x0.size()), c2(cols.size());
  for (int j = 0; j < n; ++j) x2[j] = x0[rp[j]];
  for (int j = 0; j + n < (int)x0.size(); ++j) x2[n + j] = x0[n + rp[j]];  // low parts
  for (int k = 0; k < B; ++k)
    for (int j = 0; j < n; ++j) c2[(size_t)k * n + j] = cols[(size_t)cp[k] * n + rp[j]];
  x0.swap(x2);
  cols.swap(c2);
  nu = best.k0;
  nu1 = best.k1;
  nu2 = best.k2;
}

// Kuhn's augmenting path from row `row` over columns not yet visited.
static bool augment(int row, const uint64_t* rowsup, uint64_t& visited, int* col_row) {
  uint64_t cand = rowsup[row] & ~visited;
  while (cand) {
    const int c = __builtin_ctzll(cand);
    cand &= cand - 1;
    visited |= 1ull << c;
    if (col_row[c] < 0 || augment(col_row[c], rowsup, visited, col_row)) {
      col_row[c] = row;
      return true;
    }
  }
  return false;
}

bool has_perfect_matching(int n, const uint64_t* rowsup) {
  int col_row[64];
  for (int c = 0; c < n; ++c) col_row[c] = -1;
  for (int r = 0; r < n; ++r) {
    uint64_t visited = 0;
    if (!augment(r, rowsup, visited, col_row)) return false;
  }
  return true;
}

void fill_zero_result(int n, bool integral, perm_result& o) {
  std::memset(&o, 0, sizeof(o));
  o.n = n;
  o.value = 0.0;
  o.subsets = std::ldexp(1.0, n - 1);
  o.status = PERM_OK;
  o.mode = 0;
  o.integral = integral ? 1 : 0;
  o.nu = o.nu1 = o.nu2 = n;
  o.certified = 1;
  o.zero_pattern = 1;
}

// Nonzero pattern of a dense row-major matrix (rowsup[j] bit k: a[j][k] != 0).
static bool pattern_has_matching(int n, const std::vector<double>& a) {
  uint64_t rowsup[64];
  for (int j = 0; j < n; ++j) {
    rowsup[j] = 0;
    for (int k = 0; k < n; ++k)
      if (a[(size_t)j * n + k] != 0.0) rowsup[j] |= 1ull << k;
  }
  return has_perfect_matching(n, rowsup);
}

bool is_certified(const perm_result& o) {
  if (o.zero_pattern) return true;
  if (o.err_bound == 0.0 && o.err_est == 0.0) return true;  // exact arithmetic
  return o.value != 0.0 && o.err_est <= kTargetRel * std::fabs(o.value);
}

int escalation_mode(const perm_result& o) {
  if (o.zero_pattern) return -1;
  if (o.mode == 0 && o.integral && needs_dd_pass(o)) return 1;
  if (is_certified(o)) return -1;
  if (o.mode == 0) return o.integral ? 1 : 2;
  if (o.mode == 1 && !o.integral) return 2;
  return -1;
}

void pps_sample(const std::vector<double>& z, uint32_t K, std::vector<uint32_t>& ids,
                std::vector<double>& weight, uint32_t& certain) {
  ids.clear();
  weight.clear();
  certain = 0;
  const uint64_t units = z.size();
  if (units == 0 || K == 0) return;
  K = (uint32_t)std::min<uint64_t>(K, units);
  // One pass over the units in blocks of 1024: the block sums (four
  // interleaved partial sums) and the K largest units (z descending, index
  // ascending among equal z), kept in a heap whose top is the least of them.
  constexpr uint64_t kBlock = 1024;
  const uint64_t nblocks = (units + kBlock - 1) / kBlock;
  std::vector<double> bsum(nblocks);
  auto better = [&](uint32_t a, uint32_t b) { return z[a] > z[b] || (z[a] == z[b] && a < b); };
  std::vector<uint32_t> top;
  top.reserve(K);
  for (uint64_t i = 0; i < K; ++i) {
    top.push_back((uint32_t)i);
    std::push_heap(top.begin(), top.end(), better);
  }
  double least = z[top.front()];  // a later unit enters only with a larger z
  double Z = 0.0;
  for (uint64_t b = 0; b < nblocks; ++b) {
    const uint64_t i0 = b * kBlock, i1 = std::min(units, i0 + kBlock);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t i = i0; i < i1; ++i) {
      const double v = z[i];
      acc[i & 3] += v;
      if (i >= K && v > least) {
        std::pop_heap(top.begin(), top.end(), better);
        top.back() = (uint32_t)i;
        std::push_heap(top.begin(), top.end(), better);
        least = z[top.front()];
      }
    }
    bsum[b] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    Z += bsum[b];
  }
  if (!(Z > 0.0)) return;
  if (K == units) {  // every unit, each with certainty
    for (uint64_t i = 0; i < units; ++i) {
      ids.push_back((uint32_t)i);
      weight.push_back(1.0);
    }
    certain = (uint32_t)units;
    return;
  }
  std::sort(top.begin(), top.end(), better);
  // certain units: while the largest remaining one holds a whole sampling
  // interval of the rest
  double z_rest = Z;
  uint32_t k_rest = K;
  for (uint32_t c = 0; c < K && k_rest > 0 && z[top[c]] * k_rest >= z_rest; ++c) {
    ids.push_back(top[c]);
    weight.push_back(1.0);
    z_rest -= z[top[c]];
    --k_rest;
  }
  certain = (uint32_t)ids.size();
  if (k_rest == 0 || !(z_rest > 0.0)) return;
  std::vector<uint32_t> taken(ids.begin(), ids.end());
  std::sort(taken.begin(), taken.end());
  // systematic sampling of the rest in index order (a fixed mid-interval
  // start, so the choice is deterministic); a block without a certain unit
  // that the next sampling point lies beyond is passed by its sum
  const double step = z_rest / k_rest;
  double next = 0.5 * step, cum = 0.0;
  size_t t = 0;
  for (uint64_t b = 0; b < nblocks && ids.size() < K; ++b) {
    const uint64_t i0 = b * kBlock, i1 = std::min(units, i0 + kBlock);
    const bool has_taken = t < taken.size() && taken[t] < i1;
    if (!has_taken && cum + bsum[b] < next) {
      cum += bsum[b];
      continue;
    }
    for (uint64_t i = i0; i < i1 && ids.size() < K; ++i) {
      if (t < taken.size() && taken[t] == i) {
        ++t;
        continue;
      }
      if (!(z[i] > 0.0)) continue;
      cum += z[i];
      if (cum > next) {
        ids.push_back((uint32_t)i);
        weight.push_back(step / z[i]);
        while (next < cum) next += step;
      }
    }
  }
}

double unit_weight(double sq_sum, double sq_run, int walk_bits, int high_bits) {
  return sq_sum + (std::ldexp(1.0, walk_bits - 1) + (double)high_bits) * sq_run;
}

double sampled_error(const std::vector<double>& e, const std::vector<double>& weight,
                     uint32_t certain, int n, double value) {
  double V = 0.0;
  for (size_t k = 0; k < e.size(); ++k) V += weight[k] * e[k] * e[k];
  const size_t sampled = e.size() - std::min<size_t>(certain, e.size());
  const double margin = 1.0 + (sampled > 0 ? 3.0 * std::sqrt(2.0 / (double)sampled) : 0.0);
  const double scale = std::fabs((double)(4 * (n & 1) - 2));
  return 3.0 * scale * std::sqrt(V * margin) +
         2.0 * std::sqrt((double)n) * std::ldexp(std::fabs(value), -53);
}

bool integral_matrix(const std::vector<double>& a) {
  for (double v : a)
    if (!(std::fabs(v) < 2147483648.0) || v != std::nearbyint(v)) return false;
  return true;
}

void finish_prepare(int n, const NestMenu& menu, std::vector<double>& cols, Prepared& p) {
  nest_order(n, menu, p.x0, cols, p.nu, p.nu1, p.nu2);
  if (p.integral) p.product_bits = perm::integer_product_bits(p.x0.data(), cols.data(), n);
  perm::build_column_store(n, [&](int k, int j) { return cols[(size_t)k * n + j]; }, p.cs);
}

int prepare_dense(int dtype, const void* mat, int n, Prepared& p, const NestMenu& menu) {
  if (!mat) return fail(PERM_ERR_ARG, "null pointer");
  if (int s = check_n(n)) return s;
  std::vector<double> a;
  if (!to_double(dtype, mat, (size_t)n * n, a)) return fail(PERM_ERR_ARG, "bad dtype");
  p.integral = integral_matrix(a);
  p.zero = !pattern_has_matching(n, a);
  p.x0.assign(2 * (size_t)n, 0.0);
  perm::nw_start_vector(a.data(), n, p.x0.data(), p.x0.data() + n);
  std::vector<double> cols((size_t)(n - 1) * n);
  for (int k = 0; k < n - 1; ++k)
    for (int j = 0; j < n; ++j) cols[(size_t)k * n + j] = a[(size_t)j * n + k];
  finish_prepare(n, menu, cols, p);
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
  x0.assign(2 * (size_t)n, 0.0);  // start vector, then its low parts
  perm::nw_start_vector(a.data(), n, x0.data(), x0.data() + n);
  cols.assign((size_t)n * n, 0.0);
  for (int k = 0; k < n; ++k)
    for (int e = cptrs[k]; e < cptrs[k + 1]; ++e)
      cols[(size_t)k * n + rows[e]] += elem(dtype, cvals, (size_t)e);
  integral = integral_matrix(a) && integral_matrix(cols);
  cols.resize((size_t)(n - 1) * n);  // the last column lives in x0
  return PERM_OK;
}

int prepare_sparse(int dtype, const void* mat, const int* cptrs, const int* rows,
                   const void* cvals, int n, Prepared& p, const NestMenu& menu) {
  std::vector<double> cols;
  if (int s = gather_sparse(dtype, mat, cptrs, rows, cvals, n, p.x0, cols, p.integral)) return s;
  p.zero = sparse_pattern_zero(dtype, mat, cols, n);
  finish_prepare(n, menu, cols, p);
  return PERM_OK;
}

// The sparse sum is the permanent of the dense matrix when the CCS lists
// exactly its nonzero entries of the walked columns (every nonzero entry
// positive, as the reference's CCS keeps entries > 0 only); then the
// matching check of the dense pattern applies.  Otherwise (entries <= 0
// dropped by the CCS) no shortcut is taken.
bool sparse_pattern_zero(int dtype, const void* mat, const std::vector<double>& cols, int n) {
  std::vector<double> a;
  if (!to_double(dtype, mat, (size_t)n * n, a)) return false;
  for (int k = 0; k < n - 1; ++k)
    for (int j = 0; j < n; ++j)
      if (cols[(size_t)k * n + j] != a[(size_t)j * n + k]) return false;
  return !pattern_has_matching(n, a);
}

}  // namespace perm
