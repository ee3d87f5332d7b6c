// ryser_kernels.hpp -- host-visible launch descriptors of the Ryser kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ryser_device.hpp"

namespace perm {

// Largest n the dense kernels are instantiated for (x lives in 2n VGPRs).
constexpr int kMaxDenseN = 48;
// Smallest n with pair (zero-skipping) variants; below it the whole
// permanent takes microseconds.
constexpr int kMinPairN = 16;

// Largest extreme-subset size k summed by the double-double kernel.
constexpr int kExtMaxK = 12;

// One pass of the Ryser walk over a set of lanes.  A lane fixes the columns
// [walk_bits, walk_bits + high_bits) by its pattern and walks the Gray code
// of the low walk_bits columns.  class_mask == ~0: pattern = lane id, all
// 2^high_bits patterns.  Otherwise the pattern's top high_bits - free_bits
// bits run over the patterns whose popcount c has bit c set in class_mask
// (in order of c, colex within c), and its low free_bits bits are the low
// bits of the lane id.
struct LanePass {
  uint64_t lanes;       // lanes of the pass
  uint64_t blocks;      // workgroups of 256 lanes
  uint64_t class_mask;
  int walk_bits;
  int high_bits;        // walk_bits + high_bits = n - 1
  int free_bits;
  int checked;          // drop the extreme subsets met by the walk
  Partial* partials;    // device: `blocks` records
};

// One dense launch over all 2^(n-1) subsets of one matrix.
struct DenseLaunch {
  const double* colpm;  // device: 2(n-1)+1 blocks of n doubles, [2k]=+a_k, [2k+1]=-a_k, zeros
  const double* x0;     // device: n doubles, Nijenhuis-Wilf start vector
  const double* xall;   // device: x0 + all walk columns as hi[n], lo[n]
  Partial* partials;    // device: base of all partial records (reduced in order)
  Partial* result;      // device: 1 record (grid sum)
  uint64_t total_blocks;  // partial records written by the passes below
  int n;
  int mode;             // 0 FP64, 1 double-double products, 2 double-double X too
  int nu;               // active rows of Gray column 0 (multiple of 4, < n) or n
  int ext_k;            // FP64 mode: |S| <= ext_k and >= n-1-ext_k summed apart (-1: none)
  LanePass main;        // lanes whose walk meets no extreme subset (all lanes if ext_k < 0)
  LanePass side;        // lanes whose walk may meet one (ext_k >= 0; lanes may be 0)
  uint64_t ext_per_side;  // sum_{i <= ext_k} C(n-1, i)
  uint64_t ext_blocks;    // workgroups of 256 threads over 2 ext_per_side subsets
  Partial* ext_partials;  // device: `ext_blocks` records
};

// Enqueue the passes of the dense kernel, in FP64 mode with ext_k >= 0 the
// extreme-subset kernel, and the grid reduction of all partials on `stream`.
// `main_done`, if not null, is recorded on `stream` after the main pass.
hipError_t launch_dense(const DenseLaunch& L, hipStream_t stream, hipEvent_t main_done = nullptr);
hipError_t launch_extreme(const DenseLaunch& L, hipStream_t stream);
// Unrolled inner steps per group (log2) used for a given n.
int dense_unroll_bits(int n);

}  // namespace perm
