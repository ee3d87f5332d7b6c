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
#include <array>
#include <atomic>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
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
  int walk_bits = 0, high_bits = 0;  // split of the n-1 Gray columns
  int nu = 0;             // rows touched by Gray column 0 for the FP64 kernel (n: none)
  int nu1 = 0, nu2 = 0;   // rows touched by Gray columns 0-1 and 0-2 (n: not nested)
  bool integral = false;  // every entry an integer of magnitude < 2^31
  double product_bits = 1e9;  // integer matrices: log2 bound of the walk's product numerators
  bool zero = false;      // no perfect matching: the permanent is exactly 0
  bool borrowed = false;  // stream/events/memory belong to the device cache
  uint32_t* ids = nullptr;  // device: room for kVerifyUnits workgroup indices
  // Workgroup records (and their variance weights): inside `mem` for a
  // borrowed plan; for an owning plan a separate arena of rec_cap records,
  // allocated when a run first needs it (a block-range run only needs
  // records for its range).
  void* rec_base = nullptr;
  uint64_t rec_cap = 0;
  double* zdev = nullptr;   // device: one variance weight per workgroup record
  perm::Partial* vrec = nullptr;  // device: 2 kVerifyUnits records (FP64, then mode 2)
  std::vector<double> zhost;      // host copy of zdev (kept between runs)
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
  char* stage = nullptr;  // pinned upload staging: [column store | pad | x0]
  size_t stage_bytes = 0;
};

DeviceCache& device_cache(int dev) {
  static std::mutex table_mu;
  static std::vector<DeviceCache*> table;
  std::lock_guard<std::mutex> g(table_mu);
  if ((int)table.size() <= dev) table.resize(dev + 1, nullptr);
  if (!table[dev]) table[dev] = new DeviceCache();  // lives until process exit
  return *table[dev];
}

// Device layout of a plan: column store, x0, result record, sample ids and
// records, then the workgroup records and their weights (the last two are a
// separate arena in owning plans).
struct PlanLayout {
  size_t col, x, res, ids, vrec, part, z;
  size_t fixed() const { return col + x + res + ids + vrec; }
  size_t total() const { return fixed() + part + z; }
};

using perm::kVerifyUnits;

uint64_t blocks_of(uint64_t lanes) { return (lanes + 255) / 256; }

PlanLayout plan_layout(const perm_plan* P, size_t col_doubles) {
  PlanLayout l;
  l.col = align_up(col_doubles * sizeof(double), 256);
  l.x = align_up(2 * (size_t)P->n * sizeof(double), 256);  // start vector and its low parts
  l.part = align_up(blocks_of(1ull << P->high_bits) * sizeof(perm::Partial), 256);
  l.res = align_up(sizeof(perm::Partial), 256);
  l.ids = align_up(kVerifyUnits * sizeof(uint32_t), 256);
  l.z = align_up(blocks_of(1ull << P->high_bits) * sizeof(double), 256);
  l.vrec = align_up(2 * kVerifyUnits * sizeof(perm::Partial), 256);
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
  return plan_layout(P, col_doubles).total();
}

// Bind the fixed regions at `base`; with `records`, also the workgroup
// records of all 2^h lanes and their weights right after them.
void bind_plan_memory(perm_plan* P, void* base, size_t col_doubles, bool records = true) {
  const PlanLayout l = plan_layout(P, col_doubles);
  char* b = (char*)base;
  P->L.colpm = (const double*)b;
  P->L.x0 = (const double*)(b + l.col);
  P->L.result = (perm::Partial*)(b + l.col + l.x);
  P->ids = (uint32_t*)(b + l.col + l.x + l.res);
  P->vrec = (perm::Partial*)(b + l.col + l.x + l.res + l.ids);
  if (records) {
    P->L.partials = (perm::Partial*)(b + l.fixed());
    P->zdev = (double*)(b + l.fixed() + l.part);
  }
}

// Room for `count` workgroup records (and weights) in an owning plan's
// record arena (grow-only); a borrowed plan always has all of them.
int ensure_records(perm_plan* P, uint64_t count) {
  if (P->borrowed || count <= P->rec_cap) return PERM_OK;
  if (P->rec_base) {
    HIP_TRY(hipStreamSynchronize(P->stream));
    HIP_TRY(hipFree(P->rec_base));
    P->rec_base = nullptr;
    P->rec_cap = 0;
    P->L.partials = nullptr;
    P->zdev = nullptr;
  }
  const size_t part = align_up(count * sizeof(perm::Partial), 256);
  HIP_TRY(hipMalloc(&P->rec_base, part + count * sizeof(double)));
  P->rec_cap = count;
  P->L.partials = (perm::Partial*)P->rec_base;
  P->zdev = (double*)((char*)P->rec_base + part);
  return PERM_OK;
}

// The FP64 kernel's walk shape from the host preparation (0: not restricted).
void set_shape(perm_plan* P, int n, const Prepared& prep) {
  P->product_bits = prep.product_bits;
  P->zero = prep.zero;
  P->nu = prep.nu > 0 ? prep.nu : n;
  P->nu1 = prep.nu1 > 0 ? prep.nu1 : n;
  P->nu2 = prep.nu2 > 0 ? prep.nu2 : n;
}

// The walk of one run in `mode`: all 2^h lanes.
void configure_passes(perm_plan* P, int mode) {
  perm::DenseLaunch& L = P->L;
  L.mode = mode;
  // FP64 and double-double products run the prepared walk shape; the
  // double-double X mode the plain walk
  L.nu = mode <= 1 ? P->nu : P->n;
  L.nu1 = mode <= 1 ? P->nu1 : P->n;
  L.nu2 = mode <= 1 ? P->nu2 : P->n;
  L.main.lanes = 1ull << P->high_bits;
  L.main.blocks = blocks_of(L.main.lanes);
  L.main.walk_bits = P->walk_bits;
  L.main.high_bits = P->high_bits;
  L.main.partials = L.partials;
  L.main.block_begin = 0;
  L.main.block_end = 0;
  L.total_blocks = L.main.blocks;
}

int upload_plan(perm_plan* P, const std::vector<double>& colstore, const std::vector<double>& x0) {
  HIP_TRY(hipMemcpyAsync((void*)P->L.colpm, colstore.data(), colstore.size() * sizeof(double),
                         hipMemcpyHostToDevice, P->stream));
  HIP_TRY(hipMemcpyAsync((void*)P->L.x0, x0.data(), x0.size() * sizeof(double),
                         hipMemcpyHostToDevice, P->stream));
  // the host vectors die on return: wait for the copies
  HIP_TRY(hipStreamSynchronize(P->stream));
  return PERM_OK;
}

// Create an owning plan (its own stream, events and device memory).
int make_plan(int n, const Prepared& prep, perm_plan** out) {
  const std::vector<double>& colstore = prep.cs;
  const std::vector<double>& x0 = prep.x0;
  const bool integral = prep.integral;
  if (int s = select_device()) return s;
  auto* P = new perm_plan();
  P->device = perm::current_device();
  set_shape(P, n, prep);
  P->integral = integral;
  layout_plan(P, n, colstore.size());
  P->mem.bytes = plan_layout(P, colstore.size()).fixed();  // records: ensure_records
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
  bind_plan_memory(P, P->mem.base, colstore.size(), false);
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
  // one workgroup (n <= 19): its record is the sum, the walk the only kernel
  const bool single = P->L.total_blocks <= 1;
  HIP_TRY(hipEventRecord(P->ev0, P->stream));
  HIP_TRY(perm::launch_dense(P->L, P->stream, single ? nullptr : P->evm));
  HIP_TRY(hipEventRecord(P->ev1, P->stream));
  HIP_TRY(hipMemcpyAsync(&r, single ? P->L.partials : P->L.result, sizeof(r),
                         hipMemcpyDeviceToHost, P->stream));
  HIP_TRY(hipStreamSynchronize(P->stream));
  HIP_TRY(hipEventElapsedTime(&ms, P->ev0, P->ev1));
  if (single) main_ms = ms;
  else HIP_TRY(hipEventElapsedTime(&main_ms, P->ev0, P->evm));
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
  out->err_est = perm::error_estimate(scale * scale * r.sq_sum, scale * scale * r.sq_run,
                                      out->value, n, P->walk_bits, P->high_bits, mode,
                                      P->integral);
  if (perm::integer_sum_exact(P->integral, P->product_bits, out->abs_sum, scale, n))
    out->err_bound = out->err_est = 0.0;  // exact arithmetic throughout
  perm::round_integral(P->integral, out->value, out->err_bound, out->err_est);
  out->status = PERM_OK;
  out->mode = mode;
  out->nu = mode <= 1 ? P->nu : P->n;
  out->nu1 = mode <= 1 ? P->nu1 : P->n;
  out->nu2 = mode <= 1 ? P->nu2 : P->n;
  out->integral = P->integral ? 1 : 0;
  out->zero_pattern = 0;
  out->certified = perm::is_certified(*out) ? 1 : 0;
}

// Sampled verification of an FP64 result of a non-integer matrix whose
// model estimate (perm::error_estimate) misses the accuracy target: the
// error is measured instead of modelled.  The FP64 pass left one record
// per workgroup; the error of the sum is the sum of the workgroups' errors,
// independent from one workgroup to the next (each has its own roundings;
// the shared x0 rounding is stated separately).  kVerifyUnits workgroups
// are chosen with probability proportional to their modelled variance
// z_i = Q_i + (2^(w-1) + h) Qr_i (systematic sampling; a workgroup whose
// share exceeds one sampling interval is taken with certainty) and walked
// again in mode 2 (double-double X and products: no drift), which gives
// their FP64 errors E_i exactly.  V = sum over certain workgroups of E_i^2
// + (Z_rest / K_rest) sum over sampled ones of E_i^2 / z_i estimates
// sum_i E_i^2 without bias; err_est = 3 |scale| sqrt(V (1 + 3 sqrt(2 /
// K_rest))) (a margin for the sampling spread) + 2 sqrt(n) u |value|.
// The value itself is unchanged.  units_out: workgroups recomputed; ms_out:
// device time.  Returns a PERM status; out->err_est / certified updated.
int verify_sampled(perm_plan* P, perm_result* out, uint32_t& units_out, float& ms_out) {
  units_out = 0;
  ms_out = 0.f;
  const uint64_t blocks = P->L.total_blocks;
  // variance weights of the FP64 pass's workgroup records, formed on the
  // device (8 bytes per record come back instead of the record)
  const double run_weight = std::ldexp(1.0, P->walk_bits - 1) + (double)P->high_bits;
  HIP_TRY(perm::launch_unit_weights(P->L.partials, blocks, run_weight, P->zdev, P->stream));
  P->zhost.resize(blocks);
  HIP_TRY(hipMemcpyAsync(P->zhost.data(), P->zdev, blocks * sizeof(double),
                         hipMemcpyDeviceToHost, P->stream));
  HIP_TRY(hipStreamSynchronize(P->stream));
  std::vector<uint32_t> ids;
  std::vector<double> weight;
  uint32_t certain = 0;
  perm::pps_sample(P->zhost, perm::kVerifyUnits, ids, weight, certain);
  const uint32_t count = (uint32_t)ids.size();
  if (count == 0) return PERM_OK;  // nothing to measure
  // the chosen FP64 records go to vrec[0, count), their mode 2 walks to
  // vrec[kVerifyUnits, +count)
  HIP_TRY(hipMemcpyAsync(P->ids, ids.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice,
                         P->stream));
  HIP_TRY(perm::launch_gather_partials(P->L.partials, P->ids, count, P->vrec, P->stream));
  perm::DenseLaunch L = P->L;
  L.mode = 2;
  L.nu = L.nu1 = L.nu2 = P->n;
  L.main.block_ids = P->ids;
  L.main.block_count = count;
  L.main.partials = P->vrec + kVerifyUnits;
  HIP_TRY(hipEventRecord(P->ev0, P->stream));
  HIP_TRY(perm::launch_dense_pass_only(L, P->stream));
  HIP_TRY(hipEventRecord(P->ev1, P->stream));
  std::vector<perm::Partial> rec(2 * (size_t)kVerifyUnits);
  HIP_TRY(hipMemcpyAsync(rec.data(), P->vrec, rec.size() * sizeof(perm::Partial),
                         hipMemcpyDeviceToHost, P->stream));
  HIP_TRY(hipStreamSynchronize(P->stream));
  HIP_TRY(hipEventElapsedTime(&ms_out, P->ev0, P->ev1));
  std::vector<double> e(count);
  for (uint32_t k = 0; k < count; ++k) {
    const perm::Partial& f = rec[k];
    const perm::Partial& r = rec[kVerifyUnits + k];
    e[k] = (f.hi - r.hi) + (f.lo - r.lo);
  }
  out->err_est = perm::sampled_error(e, weight, certain, P->n, out->value);
  out->certified = perm::is_certified(*out) ? 1 : 0;
  units_out = count;
  return PERM_OK;
}

// mode < 0 (automatic): a matrix without a perfect matching is exactly 0
// (no pass); otherwise the FP64 pass, then, while the result misses the
// accuracy target (perm::escalation_mode), a pass in the more accurate mode:
// double-double products for an integer matrix (exact X; below 2^53 the
// result then rounds to the exact integer), double-double X for any other
// matrix.  mode 0/1/2 runs exactly that kernel mode.  kernel_ms sums the
// passes; main_ms is the main walk of the first pass.
int plan_run_impl(perm_plan* P, int mode, perm_result* out) {
  if (!P || !out) return fail(PERM_ERR_ARG, "null pointer");
  if (mode > 2) return fail(PERM_ERR_ARG, "unknown mode");
  if (mode < 0 && P->zero) {
    perm::fill_zero_result(P->n, P->integral, *out);
    return ok();
  }
  HIP_TRY(hipSetDevice(P->device));
  if (int s = ensure_records(P, blocks_of(1ull << P->high_bits))) return s;
  perm::Partial r;
  float ms = 0.f, main_ms = 0.f, first_main_ms = 0.f;
  const int first = mode < 0 ? 0 : mode;
  if (int s = run_pass(P, first, r, ms, first_main_ms)) return s;
  fill_result(P, r, first, out);
  const double clock_ghz = 0.1 * r.pad / (double)std::max<uint64_t>(1, P->L.total_blocks);
  double total_ms = ms;
  const double first_subsets = (double)P->L.main.lanes * std::ldexp(1.0, P->L.main.walk_bits);
  uint32_t verified = 0;
  if (mode < 0) {
    // a non-integer result over many workgroups: measure its error on a
    // sample first (verify_sampled); the full double-double X pass only if
    // that does not certify it
    if (first == 0 && !P->integral && perm::escalation_mode(*out) == 2 &&
        P->L.total_blocks >= perm::kVerifyMinUnits) {
      float vms = 0.f;
      if (int s = verify_sampled(P, out, verified, vms)) return s;
      total_ms += vms;
    }
    // at most two more passes: the mode strictly increases
    for (int next = perm::escalation_mode(*out); next >= 0; next = perm::escalation_mode(*out)) {
      if (int s = run_pass(P, next, r, ms, main_ms)) return s;
      fill_result(P, r, next, out);
      total_ms += ms;
    }
  }
  out->verified_units = (int)verified;
  out->kernel_ms = total_ms;
  out->main_ms = first_main_ms;
  out->main_subsets = first_subsets;
  out->clock_ghz = clock_ghz;
  return ok();
}

// Bind a borrowed plan to the per-device cache and upload a prepared
// matrix (caller holds C.mu).
int bind_oneshot(int n, const Prepared& prep, DeviceCache& C, perm_plan& P) {
  const std::vector<double>& cs = prep.cs;
  if (!C.stream) {
    HIP_TRY(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&C.ev0));
    HIP_TRY(hipEventCreate(&C.ev1));
    HIP_TRY(hipEventCreate(&C.evm));
  }
  P.device = perm::current_device();
  P.borrowed = true;
  P.stream = C.stream;
  P.ev0 = C.ev0;
  P.ev1 = C.ev1;
  P.evm = C.evm;
  P.integral = prep.integral;
  set_shape(&P, n, prep);
  const size_t need = layout_plan(&P, n, cs.size());
  if (need > C.bytes) {
    if (C.base) HIP_TRY(hipFree(C.base));
    C.base = nullptr;
    C.bytes = 0;
    HIP_TRY(hipMalloc(&C.base, need));
    C.bytes = need;
  }
  bind_plan_memory(&P, C.base, cs.size());
  // one asynchronous copy from pinned staging in the device layout; the
  // stream orders the kernels after it, and the staging is reused only by
  // the next call under C.mu, after this call's final synchronisation
  const size_t col_bytes = (const char*)P.L.x0 - (const char*)P.L.colpm;
  const size_t up = col_bytes + prep.x0.size() * sizeof(double);
  if (up > C.stage_bytes) {
    if (C.stage) HIP_TRY(hipHostFree(C.stage));
    C.stage = nullptr;
    C.stage_bytes = 0;
    HIP_TRY(hipHostMalloc((void**)&C.stage, up, hipHostMallocDefault));
    C.stage_bytes = up;
  }
  HIP_TRY(hipStreamSynchronize(P.stream));  // idle unless an earlier call failed midway
  std::memcpy(C.stage, cs.data(), cs.size() * sizeof(double));
  std::memcpy(C.stage + col_bytes, prep.x0.data(), prep.x0.size() * sizeof(double));
  HIP_TRY(hipMemcpyAsync((void*)P.L.colpm, C.stage, up, hipMemcpyHostToDevice, P.stream));
  return PERM_OK;
}

// One-shot run through the per-device cache (mode as in plan_run_impl).
int run_oneshot(int n, const Prepared& prep, int mode, perm_result* out) {
  if (mode < 0 && prep.zero) {  // exactly 0: nothing to upload or run
    perm::fill_zero_result(n, prep.integral, *out);
    return ok();
  }
  if (int s = select_device()) return s;
  DeviceCache& C = device_cache(perm::current_device());
  std::lock_guard<std::mutex> guard(C.mu);
  perm_plan P;
  if (int s = bind_oneshot(n, prep, C, P)) return s;
  return plan_run_impl(&P, mode, out);
}

// The reference's chunked GPU+CPU scheduler (multigpucpu_chunks,
// gpu_exact_dense.cu:777-911) on one GPU.  The FP64 walk is cut into
// chunks of workgroups (2^(n-29) chunks, 2^(n-28) with the CPU worker, as
// the reference); a shared counter hands them out.  The calling thread is
// the GPU worker: it takes runs of chunks large enough to fill the GPU and
// launches the walk over their block range.  With `use_cpu` a second
// thread takes one chunk at a time and computes its lanes on the host with
// `threads` OpenMP threads.  When the counter is exhausted the CPU chunks'
// sums are written over their blocks' partial records and the device grid
// reduction sums all records, in the one-launch order.
// Without the CPU worker the result equals the one-launch result bitwise;
// with it, chunks the CPU took differ in rounding only.
int run_chunked(int n, const Prepared& prep, bool use_cpu, int threads, perm_result* out,
                int* cpu_chunks_out) {
  if (cpu_chunks_out) *cpu_chunks_out = 0;
  if (prep.zero) {
    perm::fill_zero_result(n, prep.integral, *out);
    return ok();
  }
  if (int s = select_device()) return s;
  DeviceCache& C = device_cache(perm::current_device());
  std::lock_guard<std::mutex> guard(C.mu);
  perm_plan P;
  if (int s = bind_oneshot(n, prep, C, P)) return s;
  configure_passes(&P, 0);
  perm::DenseLaunch& L = P.L;
  const uint64_t mb = L.main.blocks;
  const int shift = std::max(0, n - (use_cpu ? 28 : 29));
  const uint64_t chunks =
      std::max<uint64_t>(1, std::min<uint64_t>(mb, 1ull << std::min(shift, 20)));
  const uint64_t chunk_blocks = (mb + chunks - 1) / chunks;
  const uint64_t n_chunks = mb == 0 ? 0 : (mb + chunk_blocks - 1) / chunk_blocks;
  const uint64_t gpu_take = std::max<uint64_t>(1, 1024 / chunk_blocks);  // ~2^18 lanes a launch
  std::atomic<uint64_t> next{0};
  std::vector<std::array<double, 5>> cpu_part;
  std::vector<uint64_t> cpu_ids;
  std::thread cpu;
  if (use_cpu && n_chunks > 0) {
    cpu = std::thread([&] {
      for (;;) {
        const uint64_t c = next.fetch_add(1);
        if (c >= n_chunks) break;
        const uint64_t l0 = c * chunk_blocks * 256;
        const uint64_t l1 = std::min<uint64_t>((c + 1) * chunk_blocks * 256, L.main.lanes);
        std::array<double, 5> r;
        perm::cpu_lanes(prep.cs, prep.x0, n, L.main.walk_bits, L.main.high_bits, l0, l1,
                        threads, r.data());
        cpu_part.push_back(r);
        cpu_ids.push_back(c);
      }
    });
  }
  hipError_t e = hipEventRecord(P.ev0, P.stream);
  for (;;) {
    if (e != hipSuccess) break;
    const uint64_t c0 = next.fetch_add(gpu_take);
    if (c0 >= n_chunks) break;
    const uint64_t c1 = std::min(n_chunks, c0 + gpu_take);
    perm::DenseLaunch Lc = L;
    Lc.main.block_begin = c0 * chunk_blocks;
    Lc.main.block_end = std::min(mb, c1 * chunk_blocks);
    e = perm::launch_dense_pass_only(Lc, P.stream);
  }
  if (cpu.joinable()) cpu.join();
  if (e == hipSuccess) e = hipEventRecord(P.ev1, P.stream);
  // the CPU's chunks: its sum in the chunk's first record, zeros in the rest,
  // then the device reduction over all records, in the one-launch order
  std::vector<perm::Partial> staged(cpu_ids.size() * chunk_blocks,
                                    perm::Partial{0.0, 0.0, 0.0, 0.0, 0.0, 0.0});
  for (size_t i = 0; i < cpu_ids.size() && e == hipSuccess; ++i) {
    const uint64_t b0 = cpu_ids[i] * chunk_blocks, b1 = std::min(mb, b0 + chunk_blocks);
    perm::Partial* h = staged.data() + i * chunk_blocks;
    h[0] = {cpu_part[i][0], cpu_part[i][1], cpu_part[i][2], cpu_part[i][3], cpu_part[i][4], 0.0};
    e = hipMemcpyAsync(L.partials + b0, h, (b1 - b0) * sizeof(perm::Partial),
                       hipMemcpyHostToDevice, P.stream);
  }
  perm::Partial r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (e == hipSuccess) e = perm::launch_reduce(L.partials, L.total_blocks, L.result, P.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&r, L.result, sizeof(r), hipMemcpyDeviceToHost, P.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(P.stream);
  if (e != hipSuccess)
    return fail(PERM_ERR_HIP, std::string("chunked run: ") + hipGetErrorString(e));
  fill_result(&P, r, 0, out);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, P.ev0, P.ev1);
  out->kernel_ms = ms;
  if (cpu_chunks_out) *cpu_chunks_out = (int)cpu_ids.size();
  // a result that misses the accuracy target is computed again in the more
  // accurate mode (one launch), as plan_run_impl does
  const int esc = perm::escalation_mode(*out);
  if (esc >= 0) {
    const double chunk_ms = out->kernel_ms;
    if (int s = plan_run_impl(&P, esc, out)) return s;
    for (int m = perm::escalation_mode(*out); m >= 0; m = perm::escalation_mode(*out)) {
      const double prev = out->kernel_ms;
      if (int s = plan_run_impl(&P, m, out)) return s;
      out->kernel_ms += prev;
    }
    out->kernel_ms += chunk_ms;
  }
  return ok();
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

// ---- perm_dense_noalloc: the dense path with no device allocation ----------
// All device memory of one call -- column store, start vector, one partial
// record per workgroup and the result record -- lives in `scratch`, a
// file-scope array every GPU holds one copy of until the process ends.  In
// host code the name `scratch` is only a stand-in for the symbol: its
// device address on each GPU comes from hipGetSymbolAddress (kept in
// NoallocDevice::ptr) and only that address is ever passed to a kernel or
// a copy.  4 MiB fit the walk of every n <= 35 (2^24 lanes, 2^16 partial
// records of 48 bytes); larger n are rejected with PERM_ERR_SIZE.
constexpr size_t kScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

constexpr int kMaxNoallocDevices = 64;
// Per-GPU state of perm_dense_noalloc, set up by its first call on that GPU
// (with the GPU current): the device address of `scratch`, a stream and the
// timing events.  `mu` serialises the calls on one GPU -- they share its
// copy of the array, whose contents the last call leaves behind.
struct NoallocDevice {
  std::mutex mu;
  bool ready = false;
  char* ptr = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evm = nullptr;
};
NoallocDevice g_noalloc[kMaxNoallocDevices];

int noalloc_setup(NoallocDevice& D) {
  void* p = nullptr;
  HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(PERM_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
  HIP_TRY(hipStreamCreateWithFlags(&D.stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&D.ev0));
  HIP_TRY(hipEventCreate(&D.ev1));
  HIP_TRY(hipEventCreate(&D.evm));
  D.ptr = (char*)p;
  D.ready = true;
  return PERM_OK;
}

}  // namespace

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
  return make_plan(nov, p, plan);
}
int perm_plan_sparse(int dtype, const void* mat, const int* cptrs, const int* rows,
                     const void* cvals, int nov, perm_plan** plan) {
  if (!plan) return fail(PERM_ERR_ARG, "null plan");
  Prepared p;
  if (int s = perm::prepare_sparse(dtype, mat, cptrs, rows, cvals, nov, p)) return s;
  return make_plan(nov, p, plan);
}
int perm_plan_geometry(const perm_plan* P, int* walk_bits, int* high_bits,
                       unsigned long long* blocks) {
  if (!P) return fail(PERM_ERR_ARG, "null plan");
  if (walk_bits) *walk_bits = P->walk_bits;
  if (high_bits) *high_bits = P->high_bits;
  if (blocks) *blocks = blocks_of(1ull << P->high_bits);
  return ok();
}

int perm_plan_run_blocks(perm_plan* P, int mode, unsigned long long block_begin,
                         unsigned long long block_end, perm_result* out) {
  if (!P || !out) return fail(PERM_ERR_ARG, "null pointer");
  if (mode < 0 || mode > 2) return fail(PERM_ERR_ARG, "mode 0, 1 or 2");
  const uint64_t blocks = blocks_of(1ull << P->high_bits);
  if (block_begin >= block_end || block_end > blocks)
    return fail(PERM_ERR_ARG, "block range outside [0, " + std::to_string(blocks) + ")");
  HIP_TRY(hipSetDevice(P->device));
  if (int s = ensure_records(P, block_end - block_begin)) return s;
  configure_passes(P, mode);
  perm::DenseLaunch L = P->L;
  L.main.block_begin = block_begin;
  L.main.block_end = block_end;
  L.main.record_begin = block_begin;  // the range's records from partials[0]
  HIP_TRY(hipEventRecord(P->ev0, P->stream));
  HIP_TRY(perm::launch_dense_pass_only(L, P->stream));
  HIP_TRY(hipEventRecord(P->ev1, P->stream));
  HIP_TRY(perm::launch_reduce(L.partials, block_end - block_begin, L.result, P->stream));
  perm::Partial r;
  HIP_TRY(hipMemcpyAsync(&r, L.result, sizeof(r), hipMemcpyDeviceToHost, P->stream));
  HIP_TRY(hipStreamSynchronize(P->stream));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, P->ev0, P->ev1));
  fill_result(P, r, mode, out);
  out->kernel_ms = out->main_ms = ms;
  out->main_subsets = (double)(block_end - block_begin) * 256.0 * std::ldexp(1.0, P->walk_bits);
  out->certified = 0;  // a part of the sum: the accuracy target applies to whole results
  return ok();
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
  if (P->rec_base) (void)hipFree(P->rec_base);
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

// perm_dense with every device buffer in `scratch` (see above): no
// hipMalloc/hipFree or other allocating HIP call; the first call on a GPU
// also creates that GPU's stream and events.  Same walk split, kernels and
// reduction as perm_dense, so the same bits.
int perm_dense_noalloc(int dtype, const void* mat, int nov, int algo_id, perm_result* out) {
  if (!out) return fail(PERM_ERR_ARG, "null result");
  if (algo_id < 0 || algo_id > 4) return fail(PERM_ERR_ARG, "unknown dense algorithm ID");
  Prepared p;
  if (int s = perm::prepare_dense(dtype, mat, nov, p)) return s;
  if (p.zero) {  // exactly 0: nothing to upload or run
    perm::fill_zero_result(nov, p.integral, *out);
    return ok();
  }
  if (int s = select_device()) return s;
  const int dev = perm::current_device();
  if (dev >= kMaxNoallocDevices) return fail(PERM_ERR_NODEV, "device index above 63");
  NoallocDevice& D = g_noalloc[dev];
  std::lock_guard<std::mutex> guard(D.mu);
  if (!D.ready)
    if (int s = noalloc_setup(D)) return s;
  perm_plan P;
  P.device = dev;
  P.borrowed = true;
  P.stream = D.stream;
  P.ev0 = D.ev0;
  P.ev1 = D.ev1;
  P.evm = D.evm;
  P.integral = p.integral;
  set_shape(&P, nov, p);
  const size_t need = layout_plan(&P, nov, p.cs.size());
  if (need > kScratchBytes)
    return fail(PERM_ERR_SIZE, "nov=" + std::to_string(nov) + " needs " + std::to_string(need) +
                                   " bytes of device scratch; perm_dense_noalloc has " +
                                   std::to_string(kScratchBytes) + " (n <= 35)");
  bind_plan_memory(&P, D.ptr, p.cs.size());
  // pageable sources: plan_run_impl waits for the stream before returning,
  // and p outlives the call
  HIP_TRY(hipMemcpyAsync((void*)P.L.colpm, p.cs.data(), p.cs.size() * sizeof(double),
                         hipMemcpyHostToDevice, P.stream));
  HIP_TRY(hipMemcpyAsync((void*)P.L.x0, p.x0.data(), p.x0.size() * sizeof(double),
                         hipMemcpyHostToDevice, P.stream));
  return plan_run_impl(&P, -1, out);
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

// Multi-GPU and hybrid IDs on one GPU (gpu_exact_dense.cu:702, :777, :914;
// gpu_exact_sparse.cu:917, :996, :1193, :1328): gpu_num is accepted and one
// GPU computes the whole range; the chunked forms run the chunk scheduler
// (run_chunked) with the CPU worker when `cpu` is set.
int perm_chunked(int dtype, const void* mat, const int* cptrs, const int* rows,
                 const void* cvals, int nov, int use_cpu, int threads, perm_result* out,
                 int* cpu_chunks) {
  if (!out) return fail(PERM_ERR_ARG, "null result");
  Prepared p;
  const int s = cptrs ? perm::prepare_sparse(dtype, mat, cptrs, rows, cvals, nov, p)
                      : perm::prepare_dense(dtype, mat, nov, p);
  if (s) return s;
  return run_chunked(nov, p, use_cpu != 0, threads, out, cpu_chunks);
}

#define PERM_HYBRID_DENSE(SFX, T, DT)                                                       \
  double perm_xshared_coalescing_mshared_multigpu_dense_##SFX(const T* mat, int nov, int,   \
                                                               int, int) {                  \
    return dense_dropin(DT, mat, nov);                                                      \
  }                                                                                         \
  double perm_xshared_coalescing_mshared_multigpu_manual_distribution_dense_##SFX(          \
      const T* mat, int nov, int, int, int) {                                               \
    return dense_dropin(DT, mat, nov);                                                      \
  }                                                                                         \
  double perm_xshared_coalescing_mshared_multigpucpu_chunks_dense_##SFX(                    \
      const T* mat, int nov, int, int cpu, int threads, int, int) {                         \
    perm_result r;                                                                          \
    if (perm_chunked(DT, mat, nullptr, nullptr, nullptr, nov, cpu, threads, &r, nullptr))   \
      return nan_result();                                                                  \
    return r.value;                                                                         \
  }
#define PERM_HYBRID_SPARSE(SFX, T, DT)                                                      \
  double perm_xshared_coalescing_mshared_multigpu_sparse_##SFX(                             \
      const T* mat, const int* cptrs, const int* rows, const T* cvals, int nov, int, int,   \
      int) {                                                                                \
    return sparse_dropin(DT, mat, cptrs, rows, cvals, nov);                                 \
  }                                                                                         \
  double perm_xshared_coalescing_mshared_multigpu_sparse_manual_distribution_##SFX(         \
      const T* mat, const int* cptrs, const int* rows, const T* cvals, int nov, int, int,   \
      int) {                                                                                \
    return sparse_dropin(DT, mat, cptrs, rows, cvals, nov);                                 \
  }                                                                                         \
  double perm_xshared_coalescing_mshared_multigpucpu_chunks_sparse_##SFX(                   \
      const T* mat, const int* cptrs, const int* rows, const T* cvals, int nov, int,        \
      int cpu, int threads, int, int) {                                                     \
    perm_result r;                                                                          \
    if (perm_chunked(DT, mat, cptrs, rows, cvals, nov, cpu, threads, &r, nullptr))          \
      return nan_result();                                                                  \
    return r.value;                                                                         \
  }                                                                                         \
  double perm_xshared_coalescing_mshared_multigpucpu_chunks_skipper_##SFX(                  \
      const T* mat, const int* rptrs, const int* cols, const int* cptrs, const int* rows,   \
      const T* cvals, int nov, int, int, int, int, int) {                                   \
    perm_result r;                                                                          \
    if (perm_skip(DT, mat, rptrs, cols, cptrs, rows, cvals, nov, &r)) return nan_result();  \
    return r.value;                                                                         \
  }
PERM_HYBRID_DENSE(i32, int, PERM_DTYPE_I32)
PERM_HYBRID_DENSE(f32, float, PERM_DTYPE_F32)
PERM_HYBRID_DENSE(f64, double, PERM_DTYPE_F64)
PERM_HYBRID_SPARSE(i32, int, PERM_DTYPE_I32)
PERM_HYBRID_SPARSE(f32, float, PERM_DTYPE_F32)
PERM_HYBRID_SPARSE(f64, double, PERM_DTYPE_F64)

}  // extern "C"
