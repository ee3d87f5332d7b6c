// perman_internal.hpp -- host functions shared by the C ABI translation units
// (single-matrix, batch, serving).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "nest_shape.hpp"

struct perm_result;

namespace perm {

// Per-thread status (perm_last_status / perm_last_error) and device.
int fail(int code, const std::string& msg);
int ok();
int last_status();
const char* last_error();
int current_device();
void set_device(int device);
// hipSetDevice(current device) after checking it exists.
int select_device();
int env_int(const char* name, int dflt);
int check_n(int n);

// A matrix prepared for the Ryser kernels: start vector, signed column
// store (walk order applied), the walk shape (Gray column 0 touches rows
// [0, nu), columns 0-1 rows [0, nu1), columns 0-2 rows [0, nu2); n or 0
// where not restricted), whether every entry is an integer.
struct Prepared {
  // x0: start vector (n doubles) followed by its double-double low parts
  // (n more, read by the double-double X mode): 2n doubles, uploaded as one
  std::vector<double> x0, cs;
  int nu = 0, nu1 = 0, nu2 = 0;
  bool integral = false;
  double product_bits = 1e9;  // integer_product_bits (integer matrices)
  // no perfect matching in the matrix's nonzero pattern: every term of the
  // permanent vanishes, the result is exactly 0 and no walk is run
  bool zero = false;
};

// Perfect matching in the bipartite graph of a 0/1 pattern: rowsup[j] has
// bit k set when entry (j, k) is nonzero (n <= 64).  Kuhn's augmenting
// paths, O(n^3) at most.
bool has_perfect_matching(int n, const uint64_t* rowsup);
// Sparse inputs (cols as gather_sparse returns them): true when the CCS
// lists exactly the dense matrix's entries and its pattern has no perfect
// matching, so the sparse sum is exactly 0.
bool sparse_pattern_zero(int dtype, const void* mat, const std::vector<double>& cols, int n);
// The result of a permanent whose pattern has no perfect matching: exactly 0.
void fill_zero_result(int n, bool integral, perm_result& out);
// menu: the walk shapes the kernels that will run the matrix are compiled
// for (kMenuSingle: single-matrix paths; kMenuBatch: batch and server).
int prepare_dense(int dtype, const void* mat, int n, Prepared& p,
                  const NestMenu& menu = kMenuSingle);
int prepare_sparse(int dtype, const void* mat, const int* cptrs, const int* rows,
                   const void* cvals, int n, Prepared& p, const NestMenu& menu = kMenuSingle);
// The sparse path's inputs before any reordering: x0 from the dense matrix
// (2n doubles: start vector, then its low parts),
// cols[k*n + j] = CCS value of (j, k) for the walked columns k < n-1.
int gather_sparse(int dtype, const void* mat, const int* cptrs, const int* rows,
                  const void* cvals, int n, std::vector<double>& x0, std::vector<double>& cols,
                  bool& integral);

}  // namespace perm

struct perm_result;

namespace perm {

// One prepared matrix through the single-matrix path on the current device
// (per-device cached buffers).  mode -1: automatic (FP64, then double-double
// products for integer matrices whose FP64 error could reach 1/2); 0, 1, 2:
// that kernel mode only.
int run_prepared(int n, const Prepared& p, int mode, perm_result* out);

struct Partial;
// Result of a matrix computed as 64-lane items (batch and server paths):
// value, error statement and estimate from the fixed-order sum of its item
// partials; integer results with error below 1/2 rounded (round_integral).
void fill_item_result(int n, int walk_bits, bool integral, double product_bits, NestShape shape,
                      int mode, const Partial& r, perm_result& o);
// An integer matrix whose FP64 result could be off by 1/2 or more while
// below 2^53: run again with double-double products (mode 1).
bool needs_dd_pass(const perm_result& o);
// A result meets the accuracy target: exact (stated error 0), or its
// estimated error is at most kTargetRel |value|.
bool is_certified(const perm_result& o);
// The mode to run next for a result computed in o.mode, or -1 if none:
// integer matrices go to double-double products (mode 1: exact X, exact
// integer rounding below 2^53), other matrices whose estimated error
// exceeds kTargetRel |value| to double-double X (mode 2: no X drift).
int escalation_mode(const perm_result& o);

// Lanes [l0, l1) of the GPU walk computed on the host (the hybrid chunk
// scheduler's CPU worker); out = {hi, lo, sum |term|, sq_sum, sq_run}.
void cpu_lanes(const std::vector<double>& cs, const std::vector<double>& x0, int n,
               int walk_bits, int high_bits, uint64_t l0, uint64_t l1, int threads,
               double out[5]);

// Sampled error measurement (the FP64 result of a non-integer matrix whose
// modelled estimate misses the target).  The units of a walk (workgroups,
// batch items) have independent errors; pps_sample draws up to K of them
// with probability proportional to their modelled variance z_i (systematic
// sampling; a unit whose share exceeds one sampling interval is taken with
// certainty) and returns their indices, weights (1 / inclusion probability)
// and how many were certain.
void pps_sample(const std::vector<double>& z, uint32_t K, std::vector<uint32_t>& ids,
                std::vector<double>& weight, uint32_t& certain);
// Modelled variance weight of one unit's moments (sq_sum, sq_run) for a walk
// of 2^walk_bits steps after high_bits start additions.
double unit_weight(double sq_sum, double sq_run, int walk_bits, int high_bits);
// err_est from the units' measured errors e_k (FP64 sum minus the mode 2 sum
// of the same unit): 3 |scale| sqrt(V (1 + 3 sqrt(2 / sampled))) + 2 sqrt(n)
// u |value|, V = sum_k weight_k e_k^2 (unbiased for the sum of all units'
// squared errors).
double sampled_error(const std::vector<double>& e, const std::vector<double>& weight,
                     uint32_t certain, int n, double value);
// Units recomputed per matrix, and the fewest units for which sampling is
// used (below that the full double-double X pass is short anyway).
constexpr uint32_t kVerifyUnits = 256;
constexpr uint64_t kVerifyMinUnits = 512;

// Split of an n-matrix into 64-lane batch items of about 2^item_log2
// subsets (walk_bits + high_bits = n - 1); depends on n and item_log2 only.
void batch_split(int n, int item_log2, int& walk_bits, int& high_bits);

}  // namespace perm
