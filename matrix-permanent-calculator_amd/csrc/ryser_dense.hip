// ryser_dense.hip -- dense exact permanent (Nijenhuis-Wilf form of Ryser's
// formula, Gray-code subset order) for gfx950.
//
// Reference semantics: gpu_exact_dense.cu:129-399 (kernels) and :466-699
// (wrappers) of berkyaglioglu/matrix-permanent-calculator.  The sum computed
// is the same -- sum over all subsets S of the first n-1 columns of
// (-1)^|S| prod_j x_j(S), x(S) = x0 + sum_{k in S} a_k, scaled by
// (4(n&1)-2) -- but the work split is designed for CDNA4:
//
//  * Every lane owns an ALIGNED block of 2^w subsets: its "high" columns
//    [w, n-1) are fixed by the lane id (x built in closed form), and it walks
//    the low columns [0, w) in standard Gray order.  All 64 lanes of a wave
//    therefore flip the SAME column with the SAME sign at every step: the
//    column index k is wave-uniform.
//  * Because k is uniform, column k is fetched with scalar loads
//    (s_load_dwordx16 through the scalar cache) into SGPRs and feeds
//    v_add_f64 directly as the SGPR operand.  No LDS or vector-memory
//    bandwidth is spent on the matrix; each lane keeps x in VGPRs.
//  * The sign of the X update is folded into the column choice: the column
//    store holds +a_k and -a_k, so an update is a single v_add_f64.
//  * The innermost 2^R steps are unrolled with compile-time column indices.
//  * Per step: n adds, n-2 multiplies and one FMA that multiplies in the
//    last x and accumulates the signed term, plus one FMA for sum |term|
//    (the error estimate): 2n VALU FP64 instructions per subset.
//  * Lane sums are folded into double-double every 2^(R+5) steps; lanes,
//    waves and workgroups are reduced in double-double in a fixed order.
#include "ryser_device.hpp"
#include "ryser_kernels.hpp"

namespace perm {

// Pointer into the column store in the constant address space: loads through
// it are known not to alias any store of the kernel, so uniform loads are
// selected as scalar (SMEM) loads.
typedef const double __attribute__((address_space(4)))* ColPtr;

// Load column block `blk` (N doubles, wave-uniform address) into SGPRs.
// The pointer is laundered through an empty asm so that the compiler
// neither merges repeated loads of one column across steps (which would pin
// N SGPR pairs and spill) nor hoists them out of the loop.
template <int N>
__device__ __forceinline__ void load_col(double (&cv)[N], ColPtr base, int blk) {
  ColPtr c = base + (size_t)blk * N;
  asm volatile("" : "+s"(c));
#pragma unroll
  for (int j = 0; j < N; ++j) cv[j] = c[j];
}

// Kernel modes.
//  kFast: FP64 X, FP64 products (the production kernel).
//  kDDProd: FP64 X, double-double products and lane sums -- exact products
//    for integer matrices (whose X is exact), used to round 0/1 results to
//    the exact integer.
//  kDDAll: X itself kept as a double-double (every column addition through
//    TwoSum), double-double products and sums: no X drift; a verification
//    mode for large n where no CPU oracle finishes.
enum Mode : int { kFast = 0, kDDProd = 1, kDDAll = 2 };

#ifndef PERM_FAST_WAVES_PER_EU
#define PERM_FAST_WAVES_PER_EU 1
#endif
// Minimum waves per SIMD the fast kernel is compiled for (caps its VGPRs at
// 512/waves).  1 = no cap: unconstrained, n <= 40 already fits in 128 VGPRs
// (4 waves/SIMD) without spills; measured, 3 and 4 waves run at the same
// VALU issue rate.
constexpr int kFastWavesPerEU = PERM_FAST_WAVES_PER_EU;

// Per-lane accumulator.  kFast: plain double `p` folded into double-double
// every 32 groups.  Other modes: the signed sum is kept as an unevaluated
// pair (p, plo) updated with TwoSum on every term.
struct LaneAcc {
  double p = 0.0, plo = 0.0, a = 0.0;
};
// p += sign * prod(x);  a += |prod(x)|.
// kFast: the product of x[0..N-2] runs as four interleaved multiply chains
// (independent work for the scheduler), N-2 multiplies; the last factor and
// the signed accumulation are one FMA.
// kDDProd / kDDAll: the product is a double-double (TwoProd through FMA:
// hi*x = hi' + e exactly), two interleaved chains joined by a double-double
// multiply; the term enters (p, plo) through TwoSum.  In kDDAll each factor
// is the pair (x[j], xl[j]).
template <int N, bool NEG, int MODE>
__device__ __forceinline__ void accumulate_term(const double (&x)[N], const double (&xl)[N],
                                                LaneAcc& acc) {
  if constexpr (N == 1) {
    const double v = MODE == kDDAll ? x[0] + xl[0] : x[0];
    acc.p += NEG ? -v : v;
    acc.a += __builtin_fabs(v);
  } else if constexpr (MODE == kFast) {
    constexpr int M = N - 1;  // factors before the last one
    double c[4];
#pragma unroll
    for (int i = 0; i < 4 && i < M; ++i) c[i] = x[i];
#pragma unroll
    for (int j = 4; j < M; ++j) c[j & 3] *= x[j];
    double pp;
    if constexpr (M == 1) pp = c[0];
    else if constexpr (M == 2) pp = c[0] * c[1];
    else if constexpr (M == 3) pp = (c[0] * c[1]) * c[2];
    else pp = (c[0] * c[1]) * (c[2] * c[3]);
    const double last = x[N - 1];
    acc.p = __builtin_fma(NEG ? -pp : pp, last, acc.p);
    acc.a = __builtin_fma(__builtin_fabs(pp), __builtin_fabs(last), acc.a);
    // Pin the |term| accumulation to this step: left alone, the compiler
    // defers it to the end of the group and keeps 2 registers per step live.
    asm volatile("" : "+v"(acc.a));
  } else {
    // chain 0 over even j, chain 1 over odd j
    double h0 = x[0], l0 = MODE == kDDAll ? xl[0] : 0.0;
    double h1 = x[1], l1 = MODE == kDDAll ? xl[1] : 0.0;
#pragma unroll
    for (int j = 2; j < N; ++j) {
      double& h = (j & 1) ? h1 : h0;
      double& l = (j & 1) ? l1 : l0;
      const double ph = h * x[j];
      double e = __builtin_fma(h, x[j], -ph);
      if constexpr (MODE == kDDAll) e = __builtin_fma(h, xl[j], e);
      l = __builtin_fma(l, x[j], e);
      h = ph;
    }
    // (h0, l0) * (h1, l1)
    double th = h0 * h1;
    double tl = __builtin_fma(h0, h1, -th);
    tl = __builtin_fma(h0, l1, __builtin_fma(l0, h1, tl));
    if (NEG) {
      th = -th;
      tl = -tl;
    }
    // (p, plo) += (th, tl)
    const double s = acc.p + th;
    const double bb = s - acc.p;
    const double err = (acc.p - (s - bb)) + (th - bb);
    acc.p = s;
    acc.plo += err + tl;
    acc.a += __builtin_fabs(th);
  }
}

// x += column.  kDDAll: (x, xl) += c through TwoSum, the rounding error of
// every addition is carried in xl (exact, never renormalised).
template <int N, int MODE>
__device__ __forceinline__ void add_column(double (&x)[N], double (&xl)[N], const double (&cv)[N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) {
    if constexpr (MODE == kDDAll) {
      const double s = x[j] + cv[j];
      const double bb = s - x[j];
      xl[j] += (x[j] - (s - bb)) + (cv[j] - bb);
      x[j] = s;
    } else {
      x[j] += cv[j];
    }
  }
}

// Step v of group g (step index t = g*2^R + v).  Phase A adds the column
// that was prefetched into SGPRs by the previous step; phase B issues the
// scalar loads of the next step's column and, while they are in flight,
// forms the product and accumulates the term.  sched_barrier keeps the two
// phases apart so the loads always overlap a product.
template <int N, int R, int MODE, int V>
__device__ __forceinline__ void group_steps(double (&x)[N], double (&xl)[N], double (&cv)[N],
                                            ColPtr cols, uint32_t g, uint32_t groups,
                                            int zero_blk, LaneAcc& acc) {
  if constexpr (V < (1 << R)) {
    add_column<N, MODE>(x, xl, cv);
    __builtin_amdgcn_sched_barrier(0);
    int blk;
    if constexpr (V + 1 < (1 << R)) {
      // Next step inside the group: column k = ctz(v+1); the new gray bit k
      // is set iff bit k+1 of the step index is clear.  For k+1 < R that
      // bit is inside the group, for k+1 == R it is bit 0 of g.
      constexpr int k = __builtin_ctz(V + 1);
      int neg;
      if constexpr (k + 1 < R) neg = ((V + 1) >> (k + 1)) & 1;
      else neg = (int)(g & 1u);
      blk = 2 * k + neg;
    } else {
      // Next step opens group g+1: column R + ctz(g+1).  Past the last group
      // the (unused) prefetch reads the zero column to stay in bounds.
      const uint32_t gn = g + 1;
      const int c = __builtin_ctz(gn);
      const int neg = (int)((gn >> (c + 1)) & 1u);
      blk = gn < groups ? 2 * (R + c) + neg : zero_blk;
    }
    load_col<N>(cv, cols, blk);
    accumulate_term<N, (V & 1) != 0, MODE>(x, xl, acc);
    __builtin_amdgcn_sched_barrier(0);
    group_steps<N, R, MODE, V + 1>(x, xl, cv, cols, g, groups, zero_blk, acc);
  }
}

// colpm: 2(n-1)+1 blocks of N doubles: block 2k = +a_k, 2k+1 = -a_k for the
// Gray columns k < n-1, block 2(n-1) = zeros.
template <int N, int R, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == kFast ? kFastWavesPerEU : 1)))
void ryser_dense_kernel(const double* __restrict__ colpm,
                                                          const double* __restrict__ x0,
                                                          Partial* __restrict__ partials,
                                                          uint64_t lanes, int n_walk_bits,
                                                          int n_high_bits) {
  __shared__ Partial scratch[256 / kWave];
  const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const ColPtr cols = (ColPtr)colpm;
  const int zero_blk = 2 * (N - 1);

  double x[N], xl[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    x[j] = x0[j];
    xl[j] = 0.0;
  }

  // Closed-form start: add the high columns selected by the lane id
  // (adding sel*c with sel in {0, 1} is exact selection, no branch).
  int parity = 0;
  for (int b = 0; b < n_high_bits; ++b) {
    ColPtr c = cols + (size_t)(2 * (n_walk_bits + b)) * N;
    const int bit = (int)((lane >> b) & 1);
    parity ^= bit;
    const double sel = bit ? 1.0 : 0.0;
    double cv[N];
#pragma unroll
    for (int j = 0; j < N; ++j) cv[j] = sel * c[j];
    add_column<N, MODE>(x, xl, cv);
  }

  DD s = {0.0, 0.0};
  LaneAcc acc;
  const uint32_t groups = 1u << (n_walk_bits - R);
  double cv[N];
  load_col<N>(cv, cols, zero_blk);  // step 0 of the walk changes nothing
  for (uint32_t g = 0; g < groups; ++g) {
    group_steps<N, R, MODE, 0>(x, xl, cv, cols, g, groups, zero_blk, acc);
    if constexpr (MODE == kFast) {
      if ((g & 31u) == 31u) {  // fold into double-double every 32 groups
        s = dd_add_d(s, acc.p);
        acc.p = 0.0;
      }
    }
  }
  if constexpr (MODE == kFast) s = dd_add_d(s, acc.p);
  else s = two_sum(acc.p, acc.plo);
  double a = acc.a;
  if (parity) {
    s.hi = -s.hi;
    s.lo = -s.lo;
  }
  if (lane >= lanes) {
    s = {0.0, 0.0};
    a = 0.0;
  }
  block_reduce_store(s, a, scratch, partials + blockIdx.x);
}

// Grid-level reduction of `count` partials into out[0] (one workgroup of
// 256 threads, fixed order).
__global__ __launch_bounds__(256) void reduce_partials_kernel(const Partial* __restrict__ in,
                                                              uint64_t count,
                                                              Partial* __restrict__ out) {
  __shared__ Partial scratch[256 / kWave];
  DD s = {0.0, 0.0};
  double a = 0.0;
  for (uint64_t i = threadIdx.x; i < count; i += blockDim.x) {
    s = dd_add(s, DD{in[i].hi, in[i].lo});
    a += in[i].abs_sum;
  }
  block_reduce_store(s, a, scratch, out);
}

template <int N>
static hipError_t launch_dense_n(const DenseLaunch& L, hipStream_t stream) {
  constexpr int R = N - 1 < 3 ? N - 1 : 3;
  if (L.walk_bits < R) return hipErrorInvalidValue;
  const dim3 grid((unsigned)L.blocks), block(256);
  if (L.mode == kDDProd)
    ryser_dense_kernel<N, R, kDDProd><<<grid, block, 0, stream>>>(
        L.colpm, L.x0, L.partials, L.lanes, L.walk_bits, L.high_bits);
  else if (L.mode == kDDAll)
    ryser_dense_kernel<N, R, kDDAll><<<grid, block, 0, stream>>>(
        L.colpm, L.x0, L.partials, L.lanes, L.walk_bits, L.high_bits);
  else
    ryser_dense_kernel<N, R, kFast><<<grid, block, 0, stream>>>(
        L.colpm, L.x0, L.partials, L.lanes, L.walk_bits, L.high_bits);
  return hipGetLastError();
}

int dense_unroll_bits(int n) { return n - 1 < 3 ? n - 1 : 3; }

using DenseFn = hipError_t (*)(const DenseLaunch&, hipStream_t);

template <int... Ns>
struct DenseTable {
  static constexpr DenseFn fns[sizeof...(Ns)] = {&launch_dense_n<Ns>...};
};

template <int... Is>
static constexpr auto make_table(std::integer_sequence<int, Is...>) {
  return DenseTable<(Is + 1)...>{};
}

hipError_t launch_dense(const DenseLaunch& L, hipStream_t stream) {
  static constexpr auto table = make_table(std::make_integer_sequence<int, kMaxDenseN>{});
  if (L.n < 1 || L.n > kMaxDenseN) return hipErrorInvalidValue;
  hipError_t e = table.fns[L.n - 1](L, stream);
  if (e != hipSuccess) return e;
  reduce_partials_kernel<<<1, 256, 0, stream>>>(L.partials, L.blocks, L.result);
  return hipGetLastError();
}

}  // namespace perm
