// ryser_batch.hip -- persistent kernels for many permanents of mixed n.
//
// Work is a flat list of items; an item is one wave of 64 lanes over one
// matrix: lanes [lane_base, lane_base+64) of that matrix's aligned split
// (the same lane walk as the single-matrix kernel, ryser_dense_impl.hpp).
// Items are grouped by the matrix's variant (n and walk shape, see
// ryser_batch_impl.hpp); each variant present in
// a batch is drained by one persistent kernel instance (one wave per item,
// taken from the variant's atomic counter), the instances running
// concurrently on several streams.  One kernel per variant keeps each body's
// register allocation that of the single-matrix kernel (a single kernel
// switching over all variants spills SGPRs in the hot loops).
//
// Each item writes its double-double partial; a reduction kernel sums each
// matrix's partials in item order, so a matrix's result depends only on its
// own items: bitwise the same whether it is computed alone or in any mix,
// and whatever wave ran which item.  Counters are cleared by a kernel (no
// memset node), so a batch can be captured in a hipGraph and replayed.
#include "ryser_batch_impl.hpp"

namespace perm {

// First variant id of each order n (the table is sorted by n); entry
// kMaxBatchN + 1 is kNumVariants.
constexpr std::array<int, kMaxBatchN + 2> make_variant_start() {
  std::array<int, kMaxBatchN + 2> st{};
  int i = 0;
  for (int n = 0; n <= kMaxBatchN + 1; ++n) {
    while (i < kNumVariants && kVariants[i].n < n) ++i;
    st[n] = i;
  }
  return st;
}
constexpr std::array<int, kMaxBatchN + 2> kVariantStart = make_variant_start();

// Host lookup: variant id of walk shape s and kernel mode for order n, or -1.
int batch_variant(int n, NestShape s, int mode) {
  if (n < 1 || n > kMaxBatchN) return -1;
  for (int i = kVariantStart[n]; i < kVariantStart[n + 1]; ++i)
    if (kVariants[i].k0 == s.k0 && kVariants[i].k1 == s.k1 && kVariants[i].k2 == s.k2 &&
        kVariants[i].mode == mode)
      return i;
  return -1;
}

// Cheapest variant shape of order n covering `want` (every column's rows
// inside the variant's prefixes; extra rows add zeros); (n, n, n) if none.
NestShape batch_shape(int n, NestShape want) {
  NestShape best{n, n, n};
  if (n < 1 || n > kMaxBatchN) return best;
  for (int i = kVariantStart[n]; i < kVariantStart[n + 1]; ++i) {
    const Variant& v = kVariants[i];
    if (v.mode != kFast || v.k0 < want.k0 || v.k1 < want.k1 || v.k2 < want.k2)
      continue;
    if (nest_cost(n, {v.k0, v.k1, v.k2}) < nest_cost(n, best)) best = {v.k0, v.k1, v.k2};
  }
  return best;
}

__global__ void batch_reset_kernel(uint32_t* counters, uint32_t count) {
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) counters[i] = 0u;
}

hipError_t launch_batch_reset(uint32_t* counters, uint32_t count, hipStream_t stream) {
  batch_reset_kernel<<<1, 256, 0, stream>>>(counters, count);
  return hipGetLastError();
}

hipError_t launch_batch_variant(int vid, const BatchMatrix* mats, const BatchItem* items,
                                uint32_t begin, uint32_t end, uint32_t* counter,
                                Partial* partials, uint32_t grid, hipStream_t stream) {
  if (vid < 0 || vid >= kNumVariants || end <= begin) return hipErrorInvalidValue;
  using RangeFn = const void* (*)(int);
  static const RangeFn kRangeFns[kBatchRanges] = {
      batch_kernel_range<0>, batch_kernel_range<1>, batch_kernel_range<2>,
      batch_kernel_range<3>, batch_kernel_range<4>, batch_kernel_range<5>,
      batch_kernel_range<6>, batch_kernel_range<7>};
  int r = 0;
  while (r + 1 < kBatchRanges && vid >= batch_range_lo(r + 1)) ++r;
  const void* fn = kRangeFns[r](vid);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {(void*)&mats, (void*)&items, (void*)&begin, (void*)&end, (void*)&counter,
                  (void*)&partials};
  return hipLaunchKernel(fn, dim3(grid), dim3(256), args, 0, stream);
}

// One wave per matrix: its items' partials [first_item, first_item+n_item_count)
// summed in a fixed order.
__global__ __launch_bounds__(64) void batch_reduce_kernel(const BatchMatrix* __restrict__ mats,
                                                          const Partial* __restrict__ partials,
                                                          Partial* __restrict__ results) {
  const BatchMatrix m = mats[blockIdx.x];
  DD s = {0.0, 0.0};
  Moments mo = {0.0, 0.0, 0.0};
  for (uint32_t i = threadIdx.x; i < m.item_count; i += kWave) {
    const Partial p = partials[m.first_item + i];
    s = dd_add(s, DD{p.hi, p.lo});
    mo.a += p.abs_sum;
    mo.q += p.sq_sum;
    mo.r += p.sq_run;
  }
  wave_reduce(s, mo);
  if (threadIdx.x == 0) results[blockIdx.x] = make_partial(s, mo);
}

hipError_t launch_batch_reduce(const BatchMatrix* mats, const Partial* partials, Partial* results,
                               uint32_t n_mats, hipStream_t stream) {
  batch_reduce_kernel<<<n_mats, kWave, 0, stream>>>(mats, partials, results);
  return hipGetLastError();
}

}  // namespace perm
