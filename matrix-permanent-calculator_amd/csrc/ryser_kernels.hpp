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

// One dense launch over all 2^(n-1) subsets of one matrix.
struct DenseLaunch {
  const double* colpm;  // device: 2(n-1) blocks of n doubles, [2k]=+a_k, [2k+1]=-a_k
  const double* x0;     // device: n doubles, Nijenhuis-Wilf start vector
  Partial* partials;    // device: `blocks` records
  Partial* result;      // device: 1 record (grid sum)
  uint64_t lanes;       // 2^high_bits
  uint64_t blocks;      // workgroups of 256 lanes
  int n;
  int walk_bits;        // subsets walked per lane = 2^walk_bits
  int high_bits;        // walk_bits + high_bits = n - 1
  int mode;             // 0 FP64, 1 double-double products, 2 double-double X too
  int nu;               // active rows of Gray column 0 (multiple of 4, < n) or n
};

// Enqueue the dense kernel and the grid reduction on `stream`.
hipError_t launch_dense(const DenseLaunch& L, hipStream_t stream);
// Unrolled inner steps per group (log2) used for a given n.
int dense_unroll_bits(int n);

}  // namespace perm
