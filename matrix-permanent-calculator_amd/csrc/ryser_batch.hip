// ryser_batch.hip -- persistent kernels for many permanents of mixed n.
//
// Work is a flat list of items; an item is one wave of 64 lanes over one
// matrix: lanes [lane_base, lane_base+64) of that matrix's aligned split
// (the same lane walk as the single-matrix kernel, ryser_dense_impl.hpp).
// Items are grouped by the matrix's variant (n, nu); each variant present in
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
#include <array>

#include "ryser_batch.hpp"
#include "ryser_dense_impl.hpp"

namespace perm {

struct Variant {
  int n, nu;
};

constexpr int kBatchNus[] = {4, 8, 12, 16, 20};

constexpr bool batch_pair_ok(int n, int nu) { return n >= kMinPairN && nu <= n - 4; }

constexpr int count_variants() {
  int c = 0;
  for (int n = 1; n <= kMaxBatchN; ++n) {
    ++c;
    for (int nu : kBatchNus) c += batch_pair_ok(n, nu) ? 1 : 0;
  }
  return c;
}
constexpr int kNumVariants = count_variants();

constexpr std::array<Variant, kNumVariants> make_variants() {
  std::array<Variant, kNumVariants> v{};
  int c = 0;
  for (int n = 1; n <= kMaxBatchN; ++n) {
    v[c++] = {n, n};
    for (int nu : kBatchNus)
      if (batch_pair_ok(n, nu)) v[c++] = {n, nu};
  }
  return v;
}
constexpr std::array<Variant, kNumVariants> kVariants = make_variants();

// Host lookup: variant id of (n, nu), or -1.
int batch_variant(int n, int nu) {
  for (int i = 0; i < kNumVariants; ++i)
    if (kVariants[i].n == n && kVariants[i].nu == nu) return i;
  return -1;
}

// Largest supported pair width <= want for order n (n itself when none).
int batch_nu(int n, int want) {
  int best = n;
  for (int nu : kBatchNus)
    if (batch_pair_ok(n, nu) && nu >= want && nu < best) best = nu;
  return best;
}

// Persistent kernel of one variant (n, nu): each wave takes the next item of
// the variant's range [begin, end) from the variant's counter and walks its
// 64 lanes.  Items and matrices are read with scalar loads (wave-uniform).
template <int N, int NU>
__global__ __launch_bounds__(256) void ryser_batch_variant_kernel(
    const BatchMatrix* __restrict__ mats, const BatchItem* __restrict__ items, uint32_t begin,
    uint32_t end, uint32_t* __restrict__ counter, Partial* __restrict__ partials) {
  constexpr int R = unroll_bits_for<N>();
  for (;;) {
    // Ticket: every lane adds (lane == 0) -- one increment per wave, no
    // divergent branch around the atomic -- and lane 0's old value is the
    // wave's item.  (With `if (lane == 0) atomicAdd` the loop exit was
    // compiled as a per-lane exec-mask loop that never let lanes 1..63 out.)
    const uint32_t inc = (threadIdx.x & (kWave - 1)) == 0 ? 1u : 0u;
    const uint32_t it = __builtin_amdgcn_readfirstlane(atomicAdd(counter, inc)) + begin;
    if (it >= end) break;  // every wave leaves once the range is drained
    const BatchItem d = items[it];
    const BatchMatrix m = mats[d.matrix];
    const uint64_t lane = d.lane_base + (threadIdx.x & (kWave - 1));
    DD s;
    double a, q;
    lane_walk<N, R, kFast, Nest<NU, N, N>>(m.colpm, m.x0, lane, m.walk_bits, m.high_bits, s, a, q);
    if (lane >= m.lanes) {
      s = {0.0, 0.0};
      a = 0.0;
      q = 0.0;
    }
    wave_reduce(s, a, q);
    if ((threadIdx.x & (kWave - 1)) == 0) partials[it] = {s.hi, s.lo, a, q};
  }
}

using VariantFn = void (*)(const BatchMatrix*, const BatchItem*, uint32_t, uint32_t, uint32_t*,
                           Partial*);

template <int... I>
constexpr std::array<const void*, sizeof...(I)> variant_kernels(std::integer_sequence<int, I...>) {
  return {{(const void*)&ryser_batch_variant_kernel<kVariants[I].n, kVariants[I].nu>...}};
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
  static const auto table = variant_kernels(std::make_integer_sequence<int, kNumVariants>{});
  if (vid < 0 || vid >= kNumVariants || end <= begin) return hipErrorInvalidValue;
  void* args[] = {(void*)&mats, (void*)&items, (void*)&begin, (void*)&end, (void*)&counter,
                  (void*)&partials};
  return hipLaunchKernel(table[vid], dim3(grid), dim3(256), args, 0, stream);
}

// One wave per matrix: its items' partials [first_item, first_item+n_item_count)
// summed in a fixed order.
__global__ __launch_bounds__(64) void batch_reduce_kernel(const BatchMatrix* __restrict__ mats,
                                                          const Partial* __restrict__ partials,
                                                          Partial* __restrict__ results) {
  const BatchMatrix m = mats[blockIdx.x];
  DD s = {0.0, 0.0};
  double a = 0.0, q = 0.0;
  for (uint32_t i = threadIdx.x; i < m.item_count; i += kWave) {
    const Partial p = partials[m.first_item + i];
    s = dd_add(s, DD{p.hi, p.lo});
    a += p.abs_sum;
    q += p.sq_sum;
  }
  wave_reduce(s, a, q);
  if (threadIdx.x == 0) results[blockIdx.x] = {s.hi, s.lo, a, q};
}

hipError_t launch_batch_reduce(const BatchMatrix* mats, const Partial* partials, Partial* results,
                               uint32_t n_mats, hipStream_t stream) {
  batch_reduce_kernel<<<n_mats, kWave, 0, stream>>>(mats, partials, results);
  return hipGetLastError();
}

}  // namespace perm
