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
#include "perman_internal.hpp"
#include "ryser_host.hpp"
#include "ryser_kernels.hpp"

namespace {

using perm::fail;
using perm::ok;
using perm::env_int;
using perm::select_device;
using perm::Prepared;

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(PERM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

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
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evm = nullptr;  // evm: after the main pass
  DeviceBuffers mem;
  perm::DenseLaunch L{};
  int unroll_bits = 0;
  int walk_bits = 0, high_bits = 0;  // base split of the n-1 Gray columns
  perm::PassGeometry cut_main{}, cut_side{};  // FP64 passes under the cut
  int nu = 0;             // active rows of Gray column 0 for the FP64 kernel (n: none)
  int ext_k = -1;         // extreme-subset cut of the FP64 walk (-1: none)
  bool integral = false;  // every entry an integer of magnitude < 2^31
  bool borrowed = false;  // stream/events/memory belong to the device cache
};

namespace {

// Per-device cache used by the one-shot entry points (perm_dense,
// perm_sparse and the drop-ins): one stream, two events and a grow-only
// device buffer, reused call after call under a mutex.
struct DeviceCache {
  std::mutex mu;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evm = nullptr;
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

// Device layout of a plan: column store, x0, x_all (hi, lo), partial
// records (sized for the larger of the plain and the cut configuration),
// result record.
struct PlanLayout {
  size_t col, x, xall, part, res;
  size_t total() const { return col + x + xall + part + res; }
};

uint64_t blocks_of(uint64_t lanes) { return (lanes + 255) / 256; }

uint64_t partial_records(const perm_plan* P) {
  const uint64_t plain = blocks_of(1ull << P->high_bits);
  if (P->ext_k < 0) return plain;
  const uint64_t cut =
      blocks_of(P->cut_main.lanes) + blocks_of(P->cut_side.lanes) + P->L.ext_blocks;
  return std::max(plain, cut);
}

PlanLayout plan_layout(const perm_plan* P, size_t col_doubles) {
  PlanLayout l;
  l.col = align_up(col_doubles * sizeof(double), 256);
  l.x = align_up((size_t)P->n * sizeof(double), 256);
  l.xall = align_up(2 * (size_t)P->n * sizeof(double), 256);
  l.part = align_up(partial_records(P) * sizeof(perm::Partial), 256);
  l.res = align_up(sizeof(perm::Partial), 256);
  return l;
}

// Fill the geometry of a plan for order n and return the device bytes it
// needs.
size_t layout_plan(perm_plan* P, int n, size_t col_doubles) {
  P->n = n;
  P->unroll_bits = perm::dense_unroll_bits(n);
  const perm::Split sp = perm::choose_split(n, P->unroll_bits, env_int("PERMAN_WALK_BITS", 0));
  P->walk_bits = sp.walk_bits;
  P->high_bits = sp.high_bits;
  P->L.n = n;
  P->ext_k = perm::choose_ext_k(n, perm::kExtMaxK, env_int("PERMAN_EXT_K", -2));
  P->L.ext_per_side = P->ext_k >= 0 ? perm::extreme_per_side(n, P->ext_k) : 0;
  P->L.ext_blocks = blocks_of(2 * P->L.ext_per_side);
  if (P->ext_k >= 0)
    perm::cut_passes(n, sp.walk_bits, sp.high_bits, P->ext_k, P->unroll_bits, P->cut_main,
                     P->cut_side);
  return plan_layout(P, col_doubles).total();
}

void bind_plan_memory(perm_plan* P, void* base, size_t col_doubles) {
  const PlanLayout l = plan_layout(P, col_doubles);
  char* b = (char*)base;
  P->L.colpm = (const double*)b;
  P->L.x0 = (const double*)(b + l.col);
  P->L.xall = (const double*)(b + l.col + l.x);
  P->L.partials = (perm::Partial*)(b + l.col + l.x + l.xall);
  P->L.result = (perm::Partial*)(b + l.col + l.x + l.xall + l.part);
}

perm::LanePass make_pass(const perm::PassGeometry& g, bool checked, perm::Partial* partials) {
  perm::LanePass p;
  p.lanes = g.lanes;
  p.blocks = blocks_of(g.lanes);
  p.class_mask = g.class_mask;
  p.walk_bits = g.walk_bits;
  p.high_bits = g.high_bits;
  p.free_bits = g.free_bits;
  p.checked = checked ? 1 : 0;
  p.partials = partials;
  return p;
}

// Passes of one run in `mode`: with the extreme-subset cut (FP64 mode and
// ext_k >= 0) the main and side passes and the extreme-subset kernel, else
// one pass over all 2^h lanes.  Partial records: [main | side | extreme].
void configure_passes(perm_plan* P, int mode) {
  perm::DenseLaunch& L = P->L;
  L.mode = mode;
  L.nu = mode == 0 ? P->nu : P->n;
  L.ext_k = mode == 0 ? P->ext_k : -1;
  if (L.ext_k >= 0) {
    L.main = make_pass(P->cut_main, false, L.partials);
    L.side = make_pass(P->cut_side, true, L.partials + L.main.blocks);
    L.ext_partials = L.side.partials + L.side.blocks;
    L.total_blocks = L.main.blocks + L.side.blocks + L.ext_blocks;
  } else {
    const perm::PassGeometry all{1ull << P->high_bits, ~0ull, P->walk_bits, P->high_bits, 0};
    L.main = make_pass(all, false, L.partials);
    L.side = make_pass(perm::PassGeometry{0, 0, 0, 0, 0}, false, nullptr);
    L.ext_partials = nullptr;
    L.total_blocks = L.main.blocks;
  }
}

int upload_plan(perm_plan* P, const std::vector<double>& colstore, const std::vector<double>& x0) {
  std::vector<double> xall;
  perm::extreme_total_vector(P->n, colstore, x0, xall);
  HIP_TRY(hipMemcpyAsync((void*)P->L.colpm, colstore.data(), colstore.size() * sizeof(double),
                         hipMemcpyHostToDevice, P->stream));
  HIP_TRY(hipMemcpyAsync((void*)P->L.x0, x0.data(), x0.size() * sizeof(double),
                         hipMemcpyHostToDevice, P->stream));
  HIP_TRY(hipMemcpyAsync((void*)P->L.xall, xall.data(), xall.size() * sizeof(double),
                         hipMemcpyHostToDevice, P->stream));
  // the host vectors die on return: wait for the copies
  HIP_TRY(hipStreamSynchronize(P->stream));
  return PERM_OK;
}

// Create an owning plan (its own stream, events and device memory).
int make_plan(int n, const std::vector<double>& colstore, const std::vector<double>& x0, int nu,
              bool integral, perm_plan** out) {
  if (int s = select_device()) return s;
  auto* P = new perm_plan();
  P->device = perm::current_device();
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
  if ((e = hipEventCreate(&P->ev0)) != hipSuccess || (e = hipEventCreate(&P->ev1)) != hipSuccess ||
      (e = hipEventCreate(&P->evm)) != hipSuccess)
    return cleanup(PERM_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
  if ((e = hipMalloc(&P->mem.base, P->mem.bytes)) != hipSuccess)
    return cleanup(PERM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  bind_plan_memory(P, P->mem.base, colstore.size());
  if (upload_plan(P, colstore, x0) != PERM_OK ||
      (e = hipStreamSynchronize(P->stream)) != hipSuccess) {
    std::string msg = *perm::last_error() ? std::string(perm::last_error()) : hipGetErrorString(e);
    return cleanup(PERM_ERR_HIP, "upload: " + msg);
  }
  *out = P;
  return ok();
}

// Launch one run (FP64 or double-double products) and read the result;
// ms = device time of all its kernels, main_ms = of the main pass.
int run_pass(perm_plan* P, int mode, perm::Partial& r, float& ms, float& main_ms) {
  configure_passes(P, mode);
  HIP_TRY(hipEventRecord(P->ev0, P->stream));
  HIP_TRY(perm::launch_dense(P->L, P->stream, P->evm));
  HIP_TRY(hipEventRecord(P->ev1, P->stream));
  HIP_TRY(hipMemcpyAsync(&r, P->L.result, sizeof(r), hipMemcpyDeviceToHost, P->stream));
  HIP_TRY(hipStreamSynchronize(P->stream));
  HIP_TRY(hipEventElapsedTime(&ms, P->ev0, P->ev1));
  HIP_TRY(hipEventElapsedTime(&main_ms, P->ev0, P->evm));
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
  out->err_bound = perm::error_statement(out->abs_sum, out->value, n, P->walk_bits,
                                         P->unroll_bits, P->integral, mode);
  out->subsets = std::ldexp(1.0, n - 1);
  out->err_est =
      perm::error_estimate(scale * scale * r.sq_sum, out->value, n, mode, P->integral);
  out->ext_k = mode == 0 ? P->ext_k : -1;
  perm::round_integral(P->integral, out->value, out->err_bound, out->err_est);
  out->status = PERM_OK;
  out->mode = mode;
  out->nu = mode == 0 ? P->nu : P->n;
  out->integral = P->integral ? 1 : 0;
}

// mode < 0 (automatic): FP64 pass; for an integer matrix whose result is
// below 2^53 but whose stated error could reach 1/2 (so that rounding could
// miss the exact integer), a second pass with double-double products.
// mode 0/1/2 runs exactly that kernel mode.
int plan_run_impl(perm_plan* P, int mode, perm_result* out) {
  if (!P || !out) return fail(PERM_ERR_ARG, "null pointer");
  if (mode > 2) return fail(PERM_ERR_ARG, "unknown mode");
  HIP_TRY(hipSetDevice(P->device));
  perm::Partial r;
  float ms = 0.f, main_ms = 0.f;
  const int first = mode < 0 ? 0 : mode;
  if (int s = run_pass(P, first, r, ms, main_ms)) return s;
  fill_result(P, r, first, out);
  double total_ms = ms;
  if (mode < 0 && perm::needs_dd_pass(*out)) {
    if (int s = run_pass(P, 1, r, ms, main_ms)) return s;
    fill_result(P, r, 1, out);
    total_ms += ms;
  }
  out->kernel_ms = total_ms;
  out->main_ms = main_ms;
  out->main_subsets = (double)P->L.main.lanes * std::ldexp(1.0, P->L.main.walk_bits);
  return ok();
}

// One-shot run through the per-device cache (mode as in plan_run_impl).
int run_oneshot(int n, const Prepared& prep, int mode, perm_result* out) {
  const std::vector<double>& cs = prep.cs;
  const std::vector<double>& x0 = prep.x0;
  if (int s = select_device()) return s;
  DeviceCache& C = device_cache(perm::current_device());
  std::lock_guard<std::mutex> guard(C.mu);
  if (!C.stream) {
    HIP_TRY(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&C.ev0));
    HIP_TRY(hipEventCreate(&C.ev1));
    HIP_TRY(hipEventCreate(&C.evm));
  }
  perm_plan P;
  P.device = perm::current_device();
  P.borrowed = true;
  P.stream = C.stream;
  P.ev0 = C.ev0;
  P.ev1 = C.ev1;
  P.evm = C.evm;
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
  return plan_run_impl(&P, mode, out);
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

namespace perm {
int run_prepared(int n, const Prepared& p, int mode, perm_result* out) {
  return run_oneshot(n, p, mode, out);
}
}  // namespace perm

extern "C" {

int perm_last_status(void) { return perm::last_status(); }
const char* perm_last_error(void) { return perm::last_error(); }
int perm_set_device(int device) {
  if (device < 0) return fail(PERM_ERR_ARG, "negative device");
  perm::set_device(device);
  return ok();
}
int perm_max_n(void) { return perm::kMaxDenseN; }

int perm_plan_dense(int dtype, const void* mat, int nov, perm_plan** plan) {
  if (!plan) return fail(PERM_ERR_ARG, "null plan");
  Prepared p;
  if (int s = perm::prepare_dense(dtype, mat, nov, p)) return s;
  return make_plan(nov, p.cs, p.x0, p.nu, p.integral, plan);
}
int perm_plan_sparse(int dtype, const void* mat, const int* cptrs, const int* rows,
                     const void* cvals, int nov, perm_plan** plan) {
  if (!plan) return fail(PERM_ERR_ARG, "null plan");
  Prepared p;
  if (int s = perm::prepare_sparse(dtype, mat, cptrs, rows, cvals, nov, p)) return s;
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
  if (P->evm) (void)hipEventDestroy(P->evm);
  if (P->stream) (void)hipStreamDestroy(P->stream);
  delete P;
}

int perm_dense(int dtype, const void* mat, int nov, int algo_id, perm_result* out) {
  if (!out) return fail(PERM_ERR_ARG, "null result");
  if (algo_id < 0 || algo_id > 4) return fail(PERM_ERR_ARG, "unknown dense algorithm ID");
  Prepared p;
  if (int s = perm::prepare_dense(dtype, mat, nov, p)) return s;
  return run_oneshot(nov, p, -1, out);
}

int perm_sparse(int dtype, const void* mat, const int* cptrs, const int* rows, const void* cvals,
                int nov, int algo_id, perm_result* out) {
  if (!out) return fail(PERM_ERR_ARG, "null result");
  if (algo_id < 1 || algo_id > 4) return fail(PERM_ERR_ARG, "unknown sparse algorithm ID");
  Prepared p;
  if (int s = perm::prepare_sparse(dtype, mat, cptrs, rows, cvals, nov, p)) return s;
  return run_oneshot(nov, p, -1, out);
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
