// perman_capi.cpp -- the C ABI of libperman_hip.so (declared in
// include/perman_hip.h): drop-in entry points for the reference's dense and
// sparse exact GPU wrappers, full-result and resident-plan forms.
//
// Host flow of one permanent (reference wrapper gpu_exact_dense.cu:640-699):
// x-init + column store on the host, upload, Ryser kernel, grid reduction,
// download of one 32-byte record, scale by 4(n&1)-2.  Unlike the reference,
// device buffers come from a per-device cache (no cudaMalloc/cudaFree per
// call), the launch geometry is chosen here, and every HIP status is
// checked.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/perman_hip.h"
#include "ryser_host.hpp"
#include "ryser_kernels.hpp"

namespace {

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

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(PERM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v && *v ? std::atoi(v) : dflt;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Device memory owned by one plan (or borrowed from the per-device cache).
struct DeviceBuffers {
  void* base = nullptr;
  size_t bytes = 0;
};

}  // namespace

// A resident plan: one matrix's column store and start vector on the device.
struct perm_plan {
  int device = 0;
  int n = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DeviceBuffers mem;
  perm::DenseLaunch L{};
  int unroll_bits = 0;
  int nu = 0;             // active rows of Gray column 0 for the FP64 kernel (n: none)
  bool integral = false;  // every entry an integer of magnitude < 2^31
  bool borrowed = false;  // stream/events/memory belong to the device cache
};

namespace {

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

// Per-device cache used by the one-shot entry points (perm_dense,
// perm_sparse and the drop-ins): one stream, two events and a grow-only
// device buffer, reused call after call under a mutex.
struct DeviceCache {
  std::mutex mu;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void* base = nullptr;
  size_t bytes = 0;
};

DeviceCache& device_cache(int dev) {
  static std::mutex table_mu;
  static std::vector<DeviceCache*> table;
  std::lock_guard<std::mutex> g(table_mu);
  if ((int)table.size() <= dev) table.resize(dev + 1, nullptr);
  if (!table[dev]) table[dev] = new DeviceCache();  // lives until process exit
  return *table[dev];
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

// Fill the launch descriptor of a plan for order n and return the device
// bytes it needs.
size_t layout_plan(perm_plan* P, int n, size_t col_doubles) {
  P->n = n;
  P->unroll_bits = perm::dense_unroll_bits(n);
  const perm::Split sp = perm::choose_split(n, P->unroll_bits, env_int("PERMAN_WALK_BITS", 0));
  P->L.n = n;
  P->L.walk_bits = sp.walk_bits;
  P->L.high_bits = sp.high_bits;
  P->L.lanes = 1ull << sp.high_bits;
  P->L.blocks = (P->L.lanes + 255) / 256;
  const size_t col_bytes = align_up(col_doubles * sizeof(double), 256);
  const size_t x_bytes = align_up((size_t)n * sizeof(double), 256);
  const size_t part_bytes = align_up(P->L.blocks * sizeof(perm::Partial), 256);
  const size_t res_bytes = align_up(sizeof(perm::Partial), 256);
  return col_bytes + x_bytes + part_bytes + res_bytes;
}

void bind_plan_memory(perm_plan* P, void* base, size_t col_doubles) {
  const size_t col_bytes = align_up(col_doubles * sizeof(double), 256);
  const size_t x_bytes = align_up((size_t)P->n * sizeof(double), 256);
  const size_t part_bytes = align_up(P->L.blocks * sizeof(perm::Partial), 256);
  char* b = (char*)base;
  P->L.colpm = (const double*)b;
  P->L.x0 = (const double*)(b + col_bytes);
  P->L.partials = (perm::Partial*)(b + col_bytes + x_bytes);
  P->L.result = (perm::Partial*)(b + col_bytes + x_bytes + part_bytes);
}

int upload_plan(perm_plan* P, const std::vector<double>& colstore, const std::vector<double>& x0) {
  HIP_TRY(hipMemcpyAsync((void*)P->L.colpm, colstore.data(), colstore.size() * sizeof(double),
                         hipMemcpyHostToDevice, P->stream));
  HIP_TRY(hipMemcpyAsync((void*)P->L.x0, x0.data(), x0.size() * sizeof(double),
                         hipMemcpyHostToDevice, P->stream));
  return PERM_OK;
}

// Create an owning plan (its own stream, events and device memory).
int make_plan(int n, const std::vector<double>& colstore, const std::vector<double>& x0, int nu,
              bool integral, perm_plan** out) {
  if (int s = select_device()) return s;
  auto* P = new perm_plan();
  P->device = t_device;
  P->nu = nu;
  P->integral = integral;
  P->mem.bytes = layout_plan(P, n, colstore.size());
  auto cleanup = [&](int code, const std::string& msg) {
    perm_plan_destroy(P);
    return fail(code, msg);
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking)) != hipSuccess)
    return cleanup(PERM_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  if ((e = hipEventCreate(&P->ev0)) != hipSuccess || (e = hipEventCreate(&P->ev1)) != hipSuccess)
    return cleanup(PERM_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
  if ((e = hipMalloc(&P->mem.base, P->mem.bytes)) != hipSuccess)
    return cleanup(PERM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  bind_plan_memory(P, P->mem.base, colstore.size());
  if (upload_plan(P, colstore, x0) != PERM_OK ||
      (e = hipStreamSynchronize(P->stream)) != hipSuccess) {
    std::string msg = t_error.empty() ? hipGetErrorString(e) : t_error;
    return cleanup(PERM_ERR_HIP, "upload: " + msg);
  }
  *out = P;
  return ok();
}

// Host preparation.  Produces the start vector x0 and the Gray columns in
// column-major order (cols[k*n + j] = a[j][k] for the walked columns
// k < n-1), then chooses the walk order (pair_order) and builds the signed
// column store.
struct Prepared {
  std::vector<double> x0, cs;
  int nu = 0;
  bool integral = false;
};

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
int prepare_sparse(int dtype, const void* mat, const int* cptrs, const int* rows,
                   const void* cvals, int n, Prepared& p) {
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
  p.x0.assign(n, 0.0);
  perm::nw_start_vector(a.data(), n, p.x0.data());
  std::vector<double> cols((size_t)n * n, 0.0);
  for (int k = 0; k < n; ++k)
    for (int e = cptrs[k]; e < cptrs[k + 1]; ++e)
      cols[(size_t)k * n + rows[e]] += elem(dtype, cvals, (size_t)e);
  p.integral = integral_matrix(a) && integral_matrix(cols);
  cols.resize((size_t)(n - 1) * n);  // the last column lives in x0
  finish_prepare(n, cols, p);
  return PERM_OK;
}

// Launch one pass (FP64 or double-double products) and read the result.
int run_pass(perm_plan* P, int mode, perm::Partial& r, float& ms) {
  P->L.mode = mode;
  P->L.nu = mode == 0 ? P->nu : P->n;
  HIP_TRY(hipEventRecord(P->ev0, P->stream));
  HIP_TRY(perm::launch_dense(P->L, P->stream));
  HIP_TRY(hipEventRecord(P->ev1, P->stream));
  HIP_TRY(hipMemcpyAsync(&r, P->L.result, sizeof(r), hipMemcpyDeviceToHost, P->stream));
  HIP_TRY(hipStreamSynchronize(P->stream));
  HIP_TRY(hipEventElapsedTime(&ms, P->ev0, P->ev1));
  return PERM_OK;
}

void fill_result(const perm_plan* P, const perm::Partial& r, int mode, perm_result* out) {
  const int n = P->n;
  const double scale = (double)(4 * (n & 1) - 2);
  out->n = n;
  out->sum_hi = r.hi;
  out->sum_lo = r.lo;
  out->value = scale * (r.hi + r.lo);
  out->abs_sum = std::fabs(scale) * r.abs_sum;
  out->err_bound = perm::error_statement(out->abs_sum, n, P->L.walk_bits, P->unroll_bits,
                                         P->integral, mode);
  out->subsets = std::ldexp(1.0, n - 1);
  out->err_est = perm::error_estimate(out->abs_sum, n, mode);
  out->status = PERM_OK;
  out->mode = mode;
  out->nu = mode == 0 ? P->nu : P->n;
  out->integral = P->integral ? 1 : 0;
}

// mode < 0 (automatic): FP64 pass; for an integer matrix whose result is
// below 2^53 but whose stated error could reach 1/4, a second pass with
// double-double products makes the value round to the exact integer.
// mode 0/1/2 runs exactly that kernel mode.
int plan_run_impl(perm_plan* P, int mode, perm_result* out) {
  if (!P || !out) return fail(PERM_ERR_ARG, "null pointer");
  if (mode > 2) return fail(PERM_ERR_ARG, "unknown mode");
  HIP_TRY(hipSetDevice(P->device));
  perm::Partial r;
  float ms = 0.f;
  const int first = mode < 0 ? 0 : mode;
  if (int s = run_pass(P, first, r, ms)) return s;
  fill_result(P, r, first, out);
  double total_ms = ms;
  if (mode < 0 && P->integral && std::fabs(out->value) < 9007199254740992.0 &&
      out->err_bound >= 0.25) {
    if (int s = run_pass(P, 1, r, ms)) return s;
    fill_result(P, r, 1, out);
    total_ms += ms;
  }
  out->kernel_ms = total_ms;
  return ok();
}

// One-shot run through the per-device cache.
int run_oneshot(int n, const Prepared& prep, perm_result* out) {
  const std::vector<double>& cs = prep.cs;
  const std::vector<double>& x0 = prep.x0;
  if (int s = select_device()) return s;
  DeviceCache& C = device_cache(t_device);
  std::lock_guard<std::mutex> guard(C.mu);
  if (!C.stream) {
    HIP_TRY(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&C.ev0));
    HIP_TRY(hipEventCreate(&C.ev1));
  }
  perm_plan P;
  P.device = t_device;
  P.borrowed = true;
  P.stream = C.stream;
  P.ev0 = C.ev0;
  P.ev1 = C.ev1;
  P.integral = prep.integral;
  P.nu = prep.nu;
  const size_t need = layout_plan(&P, n, cs.size());
  if (need > C.bytes) {
    if (C.base) HIP_TRY(hipFree(C.base));
    C.base = nullptr;
    C.bytes = 0;
    HIP_TRY(hipMalloc(&C.base, need));
    C.bytes = need;
  }
  bind_plan_memory(&P, C.base, cs.size());
  if (int s = upload_plan(&P, cs, x0)) return s;
  return plan_run_impl(&P, -1, out);
}

double nan_result() { return std::numeric_limits<double>::quiet_NaN(); }

double dense_dropin(int dtype, const void* mat, int n) {
  perm_result r;
  if (perm_dense(dtype, mat, n, 4, &r) != PERM_OK) return nan_result();
  return r.value;
}

double sparse_dropin(int dtype, const void* mat, const int* cptrs, const int* rows,
                     const void* cvals, int n) {
  perm_result r;
  if (perm_sparse(dtype, mat, cptrs, rows, cvals, n, 4, &r) != PERM_OK) return nan_result();
  return r.value;
}

}  // namespace

extern "C" {

int perm_last_status(void) { return t_status; }
const char* perm_last_error(void) { return t_error.c_str(); }
int perm_set_device(int device) {
  if (device < 0) return fail(PERM_ERR_ARG, "negative device");
  t_device = device;
  return ok();
}
int perm_max_n(void) { return perm::kMaxDenseN; }

int perm_plan_dense(int dtype, const void* mat, int nov, perm_plan** plan) {
  if (!plan) return fail(PERM_ERR_ARG, "null plan");
  Prepared p;
  if (int s = prepare_dense(dtype, mat, nov, p)) return s;
  return make_plan(nov, p.cs, p.x0, p.nu, p.integral, plan);
}
int perm_plan_sparse(int dtype, const void* mat, const int* cptrs, const int* rows,
                     const void* cvals, int nov, perm_plan** plan) {
  if (!plan) return fail(PERM_ERR_ARG, "null plan");
  Prepared p;
  if (int s = prepare_sparse(dtype, mat, cptrs, rows, cvals, nov, p)) return s;
  return make_plan(nov, p.cs, p.x0, p.nu, p.integral, plan);
}
int perm_plan_run(perm_plan* plan, perm_result* out) { return plan_run_impl(plan, -1, out); }
int perm_plan_run_mode(perm_plan* plan, int mode, perm_result* out) {
  return plan_run_impl(plan, mode, out);
}
void perm_plan_destroy(perm_plan* P) {
  if (!P || P->borrowed) return;
  (void)hipSetDevice(P->device);
  if (P->stream) (void)hipStreamSynchronize(P->stream);
  if (P->mem.base) (void)hipFree(P->mem.base);
  if (P->ev0) (void)hipEventDestroy(P->ev0);
  if (P->ev1) (void)hipEventDestroy(P->ev1);
  if (P->stream) (void)hipStreamDestroy(P->stream);
  delete P;
}

int perm_dense(int dtype, const void* mat, int nov, int algo_id, perm_result* out) {
  if (!out) return fail(PERM_ERR_ARG, "null result");
  if (algo_id < 0 || algo_id > 4) return fail(PERM_ERR_ARG, "unknown dense algorithm ID");
  Prepared p;
  if (int s = prepare_dense(dtype, mat, nov, p)) return s;
  return run_oneshot(nov, p, out);
}

int perm_sparse(int dtype, const void* mat, const int* cptrs, const int* rows, const void* cvals,
                int nov, int algo_id, perm_result* out) {
  if (!out) return fail(PERM_ERR_ARG, "null result");
  if (algo_id < 1 || algo_id > 4) return fail(PERM_ERR_ARG, "unknown sparse algorithm ID");
  Prepared p;
  if (int s = prepare_sparse(dtype, mat, cptrs, rows, cvals, nov, p)) return s;
  return run_oneshot(nov, p, out);
}

#define PERM_DENSE_DROPIN(NAME, SFX, T, DT)                                   \
  double perm_##NAME##_dense_##SFX(const T* mat, int nov, int, int) {         \
    return dense_dropin(DT, mat, nov);                                        \
  }
#define PERM_SPARSE_DROPIN(NAME, SFX, T, DT)                                  \
  double perm_##NAME##_sparse_##SFX(const T* mat, const int* cptrs,           \
                                    const int* rows, const T* cvals, int nov, \
                                    int, int) {                               \
    return sparse_dropin(DT, mat, cptrs, rows, cvals, nov);                   \
  }
#define PERM_ALL_TYPES(M, NAME)              \
  M(NAME, i32, int, PERM_DTYPE_I32)          \
  M(NAME, f32, float, PERM_DTYPE_F32)        \
  M(NAME, f64, double, PERM_DTYPE_F64)

PERM_ALL_TYPES(PERM_DENSE_DROPIN, xglobal)
PERM_ALL_TYPES(PERM_DENSE_DROPIN, xlocal)
PERM_ALL_TYPES(PERM_DENSE_DROPIN, xshared)
PERM_ALL_TYPES(PERM_DENSE_DROPIN, xshared_coalescing)
PERM_ALL_TYPES(PERM_DENSE_DROPIN, xshared_coalescing_mshared)
PERM_ALL_TYPES(PERM_SPARSE_DROPIN, xlocal)
PERM_ALL_TYPES(PERM_SPARSE_DROPIN, xshared)
PERM_ALL_TYPES(PERM_SPARSE_DROPIN, xshared_coalescing)
PERM_ALL_TYPES(PERM_SPARSE_DROPIN, xshared_coalescing_mshared)

}  // extern "C"
