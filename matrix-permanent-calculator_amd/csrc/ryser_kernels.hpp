// ryser_kernels.hpp -- host-visible launch descriptors of the Ryser kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nest_shape.hpp"
#include "ryser_device.hpp"

namespace perm {

// The Ryser walk over lanes [0, lanes): lane l fixes the columns
// [walk_bits, walk_bits + high_bits) by the bits of l and walks the Gray code
// of the low walk_bits columns.  A launch may cover a range of workgroups.
struct LanePass {
  uint64_t lanes;       // 2^high_bits
  uint64_t blocks;      // workgroups of 256 lanes
  int walk_bits;
  int high_bits;        // walk_bits + high_bits = n - 1
  Partial* partials;    // device: `blocks` records
  uint64_t block_begin = 0, block_end = 0;  // workgroups to launch (end 0: all)
  uint64_t record_begin = 0;  // the record of workgroup b goes to partials[b - record_begin]
  // Instead of a range: the listed workgroups (device array of block_count
  // indices); the record of the i-th listed workgroup goes to partials[i].
  const uint32_t* block_ids = nullptr;
  uint32_t block_count = 0;
};

// One dense launch over all 2^(n-1) subsets of one matrix.
struct DenseLaunch {
  const double* colpm;  // device: 2(n-1)+1 blocks of n doubles, [2k]=+a_k, [2k+1]=-a_k, zeros
  const double* x0;     // device: 2n doubles, Nijenhuis-Wilf start vector, then its low parts
  Partial* partials;    // device: base of all partial records (reduced in order)
  Partial* result;      // device: 1 record (grid sum)
  uint64_t total_blocks;  // partial records written by the passes below
  int n;
  int mode;             // 0 FP64, 1 double-double products, 2 double-double X too
  int nu;               // rows touched by Gray column 0 (multiple of 4, < n) or n
  int nu1 = 0, nu2 = 0; // rows touched by Gray columns 0-1 and 0-2 (nested shape)
  LanePass main;        // the lanes of the walk
};

// One skipping walk (integer matrices; see ryser_skip_impl.hpp).
struct SkipLaunch {
  const double* colpm;        // device: column store in walk order
  const double* x0;           // device: start vector
  const uint32_t* row_level;  // device: byte r of word r/4 = first walk column with an
                              // entry in row r (walk_bits if none); rows ordered by it
  Partial* partials;          // device: `blocks` records
  uint32_t* walked;           // device: 4 * blocks counters (groups walked per wave)
  Partial* result;            // device: 1 record
  uint64_t lanes, blocks;
  int n, walk_bits, high_bits, check_bits, mode;  // mode 0 FP64, 1 double-double products
  int lane_reverse = 0;  // 1: lane l fixes the high columns by the bit-reversal of l
  // Without a work list: every wave of the plain split walks its range and,
  // with budget > 0, stops after `budget` walked groups of 2^R steps,
  // cont[wave] = the step it stopped at (0: finished).  With a work list
  // (n_work entries of 4 words: wave, first step, log2 steps, 0; blocks =
  // ceil(n_work / 4)): wave j walks entry j to its end.
  const uint32_t* work = nullptr;
  uint32_t n_work = 0;
  uint32_t budget = 0;
  uint32_t* cont = nullptr;
};
// Smallest n with a skipping kernel (below it the plain kernel takes microseconds).
constexpr int kMinSkipN = 16;
// Enqueue the skipping kernel and the grid reduction on `stream`.
hipError_t launch_skip(const SkipLaunch& L, hipStream_t stream);

// Enqueue the passes of the dense kernel and the grid reduction of all
// partials on `stream` (with a single workgroup record, no reduction: the
// sum is partials[0]).
// `main_done`, if not null, is recorded on `stream` after the main pass.
hipError_t launch_dense(const DenseLaunch& L, hipStream_t stream, hipEvent_t main_done = nullptr);
// Only the walk of L (its block range), no reduction (the caller sums the
// partial records).
hipError_t launch_dense_pass_only(const DenseLaunch& L, hipStream_t stream);
// The grid reduction alone: the DD sum of in[0, count) into *out (the same
// order as launch_dense's).
hipError_t launch_reduce(const Partial* in, uint64_t count, Partial* out, hipStream_t stream);
// z[i] = rec[i].sq_sum + run_weight * rec[i].sq_run for i < units: the
// modelled variance weights of a walk's workgroup records.
hipError_t launch_unit_weights(const Partial* rec, uint64_t units, double run_weight, double* z,
                               hipStream_t stream);
// dst[k] = src[ids[k]] for k < count (ids checked by the caller).
hipError_t launch_gather_partials(const Partial* src, const uint32_t* ids, uint32_t count,
                                  Partial* dst, hipStream_t stream);
// Unrolled inner steps per group (log2) used for a given n.
int dense_unroll_bits(int n);

}  // namespace perm
