// nest_shape.hpp -- sizes and walk shapes of the Ryser kernels (host and
// device; no HIP headers, so host-only code can include it).
#pragma once

namespace perm {

// Largest n the dense, sparse and skipping kernels are instantiated for (x
// lives in 2n VGPRs), as the reference's kernels and CPU paths accept
// (n <= 63: 64-bit subset masks).
constexpr int kMaxDenseN = 63;
// Largest n with pair and nested walk shapes; above it the plain walk only
// (2^48 subsets and more: the compiled shape menu would multiply the kernels
// for orders a walk of every subset cannot finish anyway).
constexpr int kMaxNestN = 48;

// Smallest n with pair (zero-skipping) variants; below it the whole
// permanent takes microseconds.
constexpr int kMinPairN = 16;
// Smallest n with nested variants (tails of Gray columns 1 and 2 shared too).
constexpr int kMinNestN = 24;

// Walk shape: Gray column l < 3 touches rows [0, k_l) only (k0 <= k1 <= k2).
struct NestShape {
  int k0, k1, k2;
};
// Compiled shapes per multiple-of-4 K0 in [4, n-4]:
//   0: (K0, n, n)  pair: only column 0 restricted (every n >= kMinPairN);
//   1: (K0, K0+4, n)   2: (K0, K0+8, n)
//   3: (K0, K0+4, K0+8)   4: (K0, K0+8, K0+12)   (n >= kMinNestN)
// A restricted level keeps at least 4 rows outside it.  {0, 0, 0}: not
// compiled.  Any shape that covers a matrix's supports is correct (extra
// rows add zeros); the host picks the covering shape of least work.
constexpr int kNestShapesPerK0 = 5;
constexpr NestShape nest_menu_shape(int n, int k0, int j) {
  if (n < kMinPairN || n > kMaxNestN || k0 < 4 || k0 > n - 4 || k0 % 4 != 0 || j < 0 ||
      j >= kNestShapesPerK0)
    return {0, 0, 0};
  if (j == 0) return {k0, n, n};
  if (n < kMinNestN) return {0, 0, 0};
  const int k1 = (j == 1 || j == 3) ? k0 + 4 : k0 + 8;
  const int k2 = j <= 2 ? n : (j == 3 ? k0 + 8 : k0 + 12);
  if (k1 > n - 4 || (j >= 3 && k2 > n - 4)) return {0, 0, 0};
  return {k0, k1, k2};
}
// A subset of the compiled shapes: K0 <= max_k0 and shape indices j with
// bit j of `mask` set.
struct NestMenu {
  int max_k0;
  unsigned mask;
};
constexpr NestMenu kMenuSingle = {kMaxDenseN, 0x1fu};  // single-matrix kernels: all
constexpr NestMenu kMenuBatch = {kMaxDenseN, 0x19u};  // batch kernels: shapes j = 0, 3, 4
constexpr bool nest_menu_has(const NestMenu& m, int n, int k0, int j) {
  return k0 <= m.max_k0 && ((m.mask >> j) & 1u) && nest_menu_shape(n, k0, j).k0 != 0;
}

// FP64 instructions per 8 subsets of a shape (adds + multiplies + 2 FMAs
// per subset); the host minimises it.
constexpr int nest_cost(int n, NestShape s) {
  return 2 * (4 * s.k0 + 2 * s.k1 + s.k2 + n) + 16;
}


}  // namespace perm
