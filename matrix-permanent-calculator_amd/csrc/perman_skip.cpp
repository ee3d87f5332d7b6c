// perman_skip.cpp -- the sparse skipper (reference sparse ID 7,
// gpu_perman64_xshared_coalescing_mshared_skipper, gpu_exact_sparse.cu:1123-1190)
// on the zero-run skipping kernel (ryser_skip_impl.hpp).
//
// Inputs follow the reference's skipper: the dense matrix for x0 and the
// CCS (entries > 0) for the column updates; its CRS (rptrs/cols) is accepted
// and not needed.  Skipping relies on exact zeros of x, so it runs for
// integer matrices (x exact); other matrices, and n below the smallest
// skipping kernel, take the regular sparse path, which computes the same
// sum.
//
// Host preparation: walk columns ordered greedily (next, the column adding
// the fewest rows not yet touched; ties by fewer nonzeros, then index) so the
// low columns touch few rows; rows ordered by the first walk column that
// touches them (lev, ascending; rows untouched by the walked part get lev =
// walk_bits).  Both orders leave the permanent unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/perman_hip.h"
#include "perman_internal.hpp"
#include "ryser_host.hpp"
#include "ryser_kernels.hpp"

namespace {

#define HIP_TRYK(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return perm::fail(PERM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

struct SkipPrepared {
  std::vector<double> x0, cs;
  std::vector<int> lev;
  int walk_bits = 0, high_bits = 0;
  double product_bits = 1e9;  // integer_product_bits of the matrix
};

// Greedy column order and row levels; builds the permuted column store.
void prepare_order(int n, const std::vector<double>& x0, const std::vector<double>& cols,
                   int walk_bits, SkipPrepared& sp) {
  const int B = n - 1;
  std::vector<int> order;
  std::vector<char> used(B, 0), touched(n, 0);
  for (int step = 0; step < B; ++step) {
    int best = -1, best_new = 0, best_nnz = 0;
    for (int k = 0; k < B; ++k) {
      if (used[k]) continue;
      int nw = 0, nz = 0;
      for (int j = 0; j < n; ++j)
        if (cols[(size_t)k * n + j] != 0.0) {
          ++nz;
          nw += touched[j] ? 0 : 1;
        }
      if (best < 0 || nw < best_new || (nw == best_new && nz < best_nnz)) {
        best = k;
        best_new = nw;
        best_nnz = nz;
      }
    }
    used[best] = 1;
    order.push_back(best);
    for (int j = 0; j < n; ++j)
      if (cols[(size_t)best * n + j] != 0.0) touched[j] = 1;
  }
  std::vector<int> lev(n, walk_bits);
  for (int j = 0; j < n; ++j)
    for (int c = 0; c < walk_bits; ++c)
      if (cols[(size_t)order[c] * n + j] != 0.0) {
        lev[j] = c;
        break;
      }
  std::vector<int> rp(n);
  std::iota(rp.begin(), rp.end(), 0);
  std::stable_sort(rp.begin(), rp.end(), [&](int a, int b) { return lev[a] < lev[b]; });
  sp.x0.resize(n);
  sp.lev.resize(n);
  for (int r = 0; r < n; ++r) {
    sp.x0[r] = x0[rp[r]];
    sp.lev[r] = lev[rp[r]];
  }
  perm::build_column_store(
      n, [&](int k, int r) { return cols[(size_t)order[k] * n + rp[r]]; }, sp.cs);
}

// Sum of two records (double-double sum, plain moments), in this order.
perm::Partial add_records(const perm::Partial& u, const perm::Partial& v) {
  const double s = u.hi + v.hi;
  const double bb = s - u.hi;
  const double err = (u.hi - (s - bb)) + (v.hi - bb);
  const double lo = err + u.lo + v.lo;
  const double hi = s + lo;
  return {hi, lo - (hi - s), u.abs_sum + v.abs_sum, u.sq_sum + v.sq_sum, u.sq_run + v.sq_run, 0.0};
}

// Walked groups a wave may spend before the rest of its range goes to
// continuation waves (PERMAN_SKIP_BUDGET; 0: none), from order
// kSkipContMinN on, and the log2 of the pieces a continuation range is cut
// into, relative to the walk (PERMAN_SKIP_SPLIT_BITS).  Measured on grid
// graphs (profiles/ryser_skip_grid_graphs.md): 8x12 (n = 48) 341 -> 233 ms;
// below n = 44 the second launch and the host's planning between the two
// cost more than the balance gains (8x10: 4.6 -> 9.3 ms), so the walk runs
// in one launch there.
constexpr int kSkipBudget = 64;
constexpr int kSkipContMinN = 44;
constexpr int kContSplitBits = 8;

// One run on the device (mode 0 FP64, 1 double-double products).  The
// skipping walk leaves most waves with little to walk and a few with most
// of it (on the 8x10 grid graph the busiest wave walks ~30x the mean); so,
// for large orders, every wave stops after a budget of walked groups and
// the remaining ranges of the waves that stopped are cut into aligned
// pieces (2^-8 of the walk at most), each walked by a wave of a second
// launch (its x in closed form at the piece's first step).  Both launches'
// records are summed in a fixed order: the result does not depend on
// scheduling.
int run_skip(int n, const SkipPrepared& sp, int mode, int check_bits, perm::Partial& r,
             double& walked_frac, float& ms) {
  const uint64_t lanes = 1ull << sp.high_bits;
  const uint64_t blocks = (lanes + 255) / 256;
  const uint64_t waves = blocks * 4;
  const size_t cs_b = align256(sp.cs.size() * sizeof(double));
  const size_t x_b = align256((size_t)n * sizeof(double));
  const size_t lev_b = align256((size_t)(n + 3) / 4 * sizeof(uint32_t));
  const size_t part_b = align256(blocks * sizeof(perm::Partial));
  const size_t walk_b = align256(waves * sizeof(uint32_t));
  const size_t res_b = align256(2 * sizeof(perm::Partial));
  const size_t cont_b = align256(waves * sizeof(uint32_t));
  char* d = nullptr;
  char* d2 = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    if (st) (void)hipStreamSynchronize(st);
    if (d) (void)hipFree(d);
    if (d2) (void)hipFree(d2);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  };
  struct Guard {
    decltype(cleanup)& f;
    ~Guard() { f(); }
  } guard{cleanup};
  HIP_TRYK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  HIP_TRYK(hipEventCreate(&e0));
  HIP_TRYK(hipEventCreate(&e1));
  HIP_TRYK(hipMalloc((void**)&d, cs_b + x_b + lev_b + part_b + walk_b + res_b + cont_b));
  const int budget_env = perm::env_int("PERMAN_SKIP_BUDGET", n >= kSkipContMinN ? kSkipBudget : 0);
  const uint32_t budget = (uint32_t)std::max(0, budget_env);
  perm::SkipLaunch L{};
  L.colpm = (const double*)d;
  L.x0 = (const double*)(d + cs_b);
  L.row_level = (const uint32_t*)(d + cs_b + x_b);
  L.partials = (perm::Partial*)(d + cs_b + x_b + lev_b);
  L.walked = (uint32_t*)(d + cs_b + x_b + lev_b + part_b);
  L.result = (perm::Partial*)(d + cs_b + x_b + lev_b + part_b + walk_b);
  uint32_t* d_cont = (uint32_t*)(d + cs_b + x_b + lev_b + part_b + walk_b + res_b);
  L.lanes = lanes;
  L.blocks = blocks;
  L.n = n;
  L.walk_bits = sp.walk_bits;
  L.high_bits = sp.high_bits;
  L.check_bits = check_bits;
  L.mode = mode;
  // bit-reversed lane ids: the lanes of a wave differ in the last high
  // columns; measured 1.5-1.9x faster on grid graphs 8x9, 8x10, 6x12
  L.lane_reverse = perm::env_int("PERMAN_SKIP_LANE_REVERSE", 1);
  L.budget = budget;
  L.cont = budget ? d_cont : nullptr;
  HIP_TRYK(hipMemcpyAsync((void*)L.colpm, sp.cs.data(), sp.cs.size() * sizeof(double),
                          hipMemcpyHostToDevice, st));
  HIP_TRYK(hipMemcpyAsync((void*)L.x0, sp.x0.data(), n * sizeof(double), hipMemcpyHostToDevice,
                          st));
  std::vector<uint32_t> packed((n + 3) / 4, 0u);  // levels <= walk_bits < 32: one byte each
  for (int r = 0; r < n; ++r) packed[r / 4] |= (uint32_t)sp.lev[r] << (8 * (r % 4));
  HIP_TRYK(hipMemcpyAsync((void*)L.row_level, packed.data(), packed.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice, st));
  HIP_TRYK(hipEventRecord(e0, st));
  HIP_TRYK(perm::launch_skip(L, st));
  HIP_TRYK(hipEventRecord(e1, st));  // (recorded again after the continuations, if any)
  std::vector<uint32_t> walked(waves);
  std::vector<uint32_t> cont(budget ? waves : 0);
  perm::Partial r1{};
  HIP_TRYK(hipMemcpyAsync(&r1, L.result, sizeof(r1), hipMemcpyDeviceToHost, st));
  HIP_TRYK(hipMemcpyAsync(walked.data(), L.walked, walked.size() * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, st));
  if (budget)
    HIP_TRYK(hipMemcpyAsync(cont.data(), d_cont, cont.size() * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, st));
  HIP_TRYK(hipStreamSynchronize(st));
  r = r1;
  // the continuations: each stopped wave's remaining range [stop, 2^w) in
  // aligned pieces of at most 2^(w - kContSplitBits) steps
  std::vector<uint32_t> work;
  const int split_bits = perm::env_int("PERMAN_SKIP_SPLIT_BITS", kContSplitBits);
  const int piece_bits = std::max(check_bits, sp.walk_bits - std::max(0, split_bits));
  const uint64_t steps = 1ull << sp.walk_bits;
  for (uint64_t w = 0; w < cont.size(); ++w) {
    uint64_t t = cont[w];
    if (t == 0) continue;
    while (t < steps) {
      const int b = std::min(__builtin_ctzll(t), piece_bits);
      work.insert(work.end(), {(uint32_t)w, (uint32_t)t, (uint32_t)b, 0u});
      t += 1ull << b;
    }
  }
  if (!work.empty()) {
    const uint64_t n_work = work.size() / 4;
    const uint64_t blocks2 = (n_work + 3) / 4;
    const size_t work_b = align256(work.size() * sizeof(uint32_t));
    const size_t part2_b = align256(blocks2 * sizeof(perm::Partial));
    const size_t walk2_b = align256(blocks2 * 4 * sizeof(uint32_t));
    HIP_TRYK(hipMalloc((void**)&d2, work_b + part2_b + walk2_b));
    perm::SkipLaunch L2 = L;
    L2.work = (const uint32_t*)d2;
    L2.n_work = (uint32_t)n_work;
    L2.blocks = blocks2;
    L2.partials = (perm::Partial*)(d2 + work_b);
    L2.walked = (uint32_t*)(d2 + work_b + part2_b);
    L2.result = L.result + 1;
    L2.budget = 0;
    L2.cont = nullptr;
    HIP_TRYK(hipMemcpyAsync((void*)L2.work, work.data(), work.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, st));
    HIP_TRYK(perm::launch_skip(L2, st));
    HIP_TRYK(hipEventRecord(e1, st));  // the time includes the host's planning between launches
    std::vector<uint32_t> walked2(blocks2 * 4);
    perm::Partial r2{};
    HIP_TRYK(hipMemcpyAsync(&r2, L2.result, sizeof(r2), hipMemcpyDeviceToHost, st));
    HIP_TRYK(hipMemcpyAsync(walked2.data(), L2.walked, walked2.size() * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, st));
    HIP_TRYK(hipStreamSynchronize(st));
    r = add_records(r1, r2);
    walked.insert(walked.end(), walked2.begin(), walked2.end());
  }
  HIP_TRYK(hipEventElapsedTime(&ms, e0, e1));
  // walked[] counts groups of 2^R steps per wave (64 lanes); lanes past
  // `lanes` in the last block walk with their wave
  const int R = perm::dense_unroll_bits(n);
  double groups = 0.0;
  for (uint32_t w : walked) groups += w;
  walked_frac = groups * std::ldexp(1.0, R) / ((double)waves * std::ldexp(1.0, sp.walk_bits));
  return PERM_OK;
}

void fill(int n, const SkipPrepared& sp, const perm::Partial& r, int mode, perm_result* out) {
  std::memset(out, 0, sizeof(*out));
  const double scale = (double)(4 * (n & 1) - 2);
  out->n = n;
  out->sum_hi = r.hi;
  out->sum_lo = r.lo;
  out->value = scale * (r.hi + r.lo);
  out->abs_sum = std::fabs(scale) * r.abs_sum;
  out->err_bound = perm::error_statement(out->abs_sum, out->value, n, sp.walk_bits,
                                         perm::dense_unroll_bits(n), true, mode);
  out->err_est = perm::error_estimate(scale * scale * r.sq_sum, 0.0, out->value, n, sp.walk_bits,
                                      sp.high_bits, mode, true);
  out->subsets = std::ldexp(1.0, n - 1);
  out->mode = mode;
  out->integral = 1;
  out->nu = out->nu1 = out->nu2 = n;
  out->status = PERM_OK;
  if (perm::integer_sum_exact(true, sp.product_bits, out->abs_sum, scale, n))
    out->err_bound = out->err_est = 0.0;  // exact arithmetic throughout
  perm::round_integral(true, out->value, out->err_bound, out->err_est);
  out->certified = perm::is_certified(*out) ? 1 : 0;
}

}  // namespace

extern "C" {

int perm_skip(int dtype, const void* mat, const int* rptrs, const int* cols, const int* cptrs,
              const int* rows, const void* cvals, int nov, perm_result* out) {
  (void)rptrs, (void)cols;
  if (!out) return perm::fail(PERM_ERR_ARG, "null result");
  std::vector<double> x0, cc;
  bool integral = false;
  if (int s = perm::gather_sparse(dtype, mat, cptrs, rows, cvals, nov, x0, cc, integral))
    return s;
  if (!integral || nov < perm::kMinSkipN)
    return perm_sparse(dtype, mat, cptrs, rows, cvals, nov, 4, out);
  if (perm::sparse_pattern_zero(dtype, mat, cc, nov)) {  // exactly 0: nothing to walk
    perm::fill_zero_result(nov, true, *out);
    return perm::ok();
  }
  if (int s = perm::select_device()) return s;
  SkipPrepared sp;
  const int R = perm::dense_unroll_bits(nov);
  const perm::Split split =
      perm::choose_skip_split(nov, R, perm::env_int("PERMAN_WALK_BITS", 0));
  sp.walk_bits = split.walk_bits;
  sp.high_bits = split.high_bits;
  prepare_order(nov, x0, cc, sp.walk_bits, sp);
  sp.product_bits = perm::integer_product_bits(x0.data(), cc.data(), nov);
  // check every 2^3 steps (one unrolled group): measured on grid graphs
  // 8x9, 8x10, 6x12 against 2^4..2^6
  int check_bits = perm::env_int("PERMAN_SKIP_CHECK_BITS", 3);
  check_bits = std::max(R, std::min(check_bits, sp.walk_bits));
  perm::Partial r;
  double walked = 0.0;
  float ms = 0.f, ms2 = 0.f;
  if (int s = run_skip(nov, sp, 0, check_bits, r, walked, ms)) return s;
  fill(nov, sp, r, 0, out);
  // integer matrix: X is exact, so the only escalation is double-double
  // products (exact rounding below 2^53, error ~2^-104 A above)
  if (perm::escalation_mode(*out) == 1) {
    double walked2 = 0.0;
    if (int s = run_skip(nov, sp, 1, check_bits, r, walked2, ms2)) return s;
    fill(nov, sp, r, 1, out);
  }
  out->kernel_ms = (double)ms + (double)ms2;
  out->main_ms = ms;
  out->main_subsets = walked * out->subsets;  // subsets actually walked (rest skipped)
  return perm::ok();
}

#define PERM_SKIP_DROPIN(SFX, T, DT)                                                        \
  double perm_xshared_coalescing_mshared_skipper_##SFX(const T* mat, const int* rptrs,      \
                                                        const int* cols, const int* cptrs,  \
                                                        const int* rows, const T* cvals,    \
                                                        int nov, int, int) {                \
    perm_result r;                                                                          \
    if (perm_skip(DT, mat, rptrs, cols, cptrs, rows, cvals, nov, &r) != PERM_OK)            \
      return std::nan("");                                                                  \
    return r.value;                                                                         \
  }
PERM_SKIP_DROPIN(i32, int, PERM_DTYPE_I32)
PERM_SKIP_DROPIN(f32, float, PERM_DTYPE_F32)
PERM_SKIP_DROPIN(f64, double, PERM_DTYPE_F64)

}  // extern "C"
