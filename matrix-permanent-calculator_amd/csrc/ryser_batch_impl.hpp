// ryser_batch_impl.hpp -- the persistent batch kernel and its variant table
// (shared by the translation units that instantiate ranges of variants).
#pragma once
#include <array>
#include <utility>

#include "ryser_batch.hpp"
#include "ryser_dense_impl.hpp"

namespace perm {

// A batch kernel variant: order n, walk shape (k0, k1, k2) and kernel mode
// (kFast; kDDProd for the exact pass of integer matrices; kDDAll for the
// drift-free pass of other matrices whose FP64 result misses the accuracy
// target).
struct Variant {
  int n, k0, k1, k2, mode;
};

// Variants: for every n <= kMaxBatchN the plain walk in all three modes,
// then the kMenuBatch shapes (every pair width; for n >= kMinNestN also the
// two three-level shapes per width) in the FP64 and double-double product
// modes (double-double X runs the plain walk only).
constexpr int count_variants() {
  int c = 0;
  for (int n = 1; n <= kMaxBatchN; ++n) {
    c += 3;
    for (int k0 = 4; k0 <= n - 4; k0 += 4)
      for (int j = 0; j < kNestShapesPerK0; ++j)
        c += nest_menu_has(kMenuBatch, n, k0, j) ? 2 : 0;
  }
  return c;
}
constexpr int kNumVariants = count_variants();

constexpr std::array<Variant, kNumVariants> make_variants() {
  std::array<Variant, kNumVariants> v{};
  int c = 0;
  for (int n = 1; n <= kMaxBatchN; ++n) {
    v[c++] = {n, n, n, n, kFast};
    v[c++] = {n, n, n, n, kDDProd};
    v[c++] = {n, n, n, n, kDDAll};
    for (int k0 = 4; k0 <= n - 4; k0 += 4)
      for (int j = 0; j < kNestShapesPerK0; ++j)
        if (nest_menu_has(kMenuBatch, n, k0, j)) {
          const NestShape s = nest_menu_shape(n, k0, j);
          v[c++] = {n, s.k0, s.k1, s.k2, kFast};
          v[c++] = {n, s.k0, s.k1, s.k2, kDDProd};
        }
  }
  return v;
}
constexpr std::array<Variant, kNumVariants> kVariants = make_variants();

// Persistent kernel of one variant: each wave takes the next item of the
// variant's range [begin, end) from the variant's counter and walks its 64
// lanes.  Items and matrices are read with scalar loads (wave-uniform).
template <int N, int MODE, int K0, int K1, int K2>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MODE == kFast ? kFastWavesPerEU : 1)))
void ryser_batch_variant_kernel(
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
    Moments mo;
    lane_walk<N, R, MODE, Nest<K0, K1, K2>>(m.colpm, m.x0, lane, m.walk_bits, m.high_bits, s, mo);
    if (lane >= m.lanes) {
      s = {0.0, 0.0};
      mo = {0.0, 0.0, 0.0};
    }
    wave_reduce(s, mo);
    if ((threadIdx.x & (kWave - 1)) == 0) partials[it] = make_partial(s, mo);
  }
}

template <int LO, int... I>
constexpr std::array<const void*, sizeof...(I)> variant_kernels(std::integer_sequence<int, I...>) {
  return {{(const void*)&ryser_batch_variant_kernel<kVariants[LO + I].n, kVariants[LO + I].mode,
                                                    kVariants[LO + I].k0, kVariants[LO + I].k1,
                                                    kVariants[LO + I].k2>...}};
}

// Kernel entry of variant vid for vid in range IDX of kBatchRanges (one
// translation unit per range keeps the build parallel); null outside it.
constexpr int kBatchRanges = 8;
template <int IDX>
const void* batch_kernel_range(int vid);
constexpr int batch_range_lo(int idx) { return (int)((long)kNumVariants * idx / kBatchRanges); }

}  // namespace perm

#define PERM_BATCH_RANGE(IDX)                                                              \
  namespace perm {                                                                         \
  template <>                                                                              \
  const void* batch_kernel_range<IDX>(int vid) {                                           \
    constexpr int lo = batch_range_lo(IDX), hi = batch_range_lo(IDX + 1);                  \
    static const auto table = variant_kernels<lo>(std::make_integer_sequence<int, hi - lo>{}); \
    return vid >= lo && vid < hi ? table[vid - lo] : nullptr;                              \
  }                                                                                        \
  }
