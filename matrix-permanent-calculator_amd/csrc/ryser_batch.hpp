// ryser_batch.hpp -- descriptors of the persistent mixed-size batch kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ryser_device.hpp"
#include "ryser_kernels.hpp"

namespace perm {

// Largest n the batch kernel handles (larger matrices go through the
// single-matrix kernels).
constexpr int kMaxBatchN = 40;
// Lanes per item: one wave.
constexpr int kItemLanes = 64;

// One matrix of a batch (device copy, 48 bytes).
struct BatchMatrix {
  const double* colpm;  // column store (see ryser_dense_impl.hpp)
  const double* x0;     // start vector (n doubles), then its low parts (n more)
  uint64_t lanes;       // 2^high_bits
  int32_t walk_bits;    // subsets per lane = 2^walk_bits
  int32_t high_bits;
  int32_t variant;      // index of (n, walk shape) in the kernel's variant table
  uint32_t first_item;  // this matrix's items are [first_item, first_item + item_count)
  uint32_t item_count;
  uint32_t pad;
};

// One work item: one wave over lanes [lane_base, lane_base + 64).
struct BatchItem {
  uint32_t matrix;
  uint32_t pad;
  uint64_t lane_base;
};

// Variant id of walk shape s and kernel mode (0 FP64, 1 double-double
// products, 2 double-double X: plain shape only) for order n (-1 if none),
// and the cheapest FP64 variant shape covering `want` ((n, n, n) if none).
int batch_variant(int n, NestShape s, int mode = 0);
NestShape batch_shape(int n, NestShape want);

// Zero `count` work counters.
hipError_t launch_batch_reset(uint32_t* counters, uint32_t count, hipStream_t stream);
// Persistent kernel of variant `vid` over items [begin, end); `counter`
// must be zero; partial of item i goes to partials[i].
hipError_t launch_batch_variant(int vid, const BatchMatrix* mats, const BatchItem* items,
                                uint32_t begin, uint32_t end, uint32_t* counter,
                                Partial* partials, uint32_t grid, hipStream_t stream);
// results[m] = fixed-order sum of matrix m's item partials.
hipError_t launch_batch_reduce(const BatchMatrix* mats, const Partial* partials, Partial* results,
                               uint32_t n_mats, hipStream_t stream);

}  // namespace perm
