// ryser_extreme.hip -- the extreme subsets of the Ryser sum in double-double.
//
// In the Nijenhuis-Wilf form x_j(S) = x0_j + sum_{k in S} a_jk runs from
// x0_j = a_j,n-1 - r_j/2 (S empty, r_j the row sum) to x_j = r_j/2 (S = all
// n-1 walk columns).  For nonnegative matrices |x_j| is largest at both
// ends, so the subsets with |S| <= k or |S| >= n-1-k carry the
// largest terms of the alternating sum -- at n = 40 a few thousand of them
// are 10^4 times the permanent.  Their FP64 rounding (and the X drift of the
// lanes that reach them) is the dominant error of the FP64 walk.  The FP64
// walk therefore leaves them out and this kernel sums them with every X
// formed exactly in double-double (x0 + the |S| columns, or the total
// x_all - the n-1-|S| missing columns) and the product in double-double:
// relative error ~ n 2^-104 per term.
//
// One thread per extreme subset: thread r < M (M = sum_{i<=k} C(n-1, i))
// takes the r-th small subset C in size-then-colex order and computes the
// term of S = C; thread M + r computes the term of S = complement(C).  The
// x0 used is the same double vector the FP64 walk starts from, so the two
// parts sum the same function of x0 exactly over disjoint subsets.
#include "ryser_kernels.hpp"

namespace perm {

static_assert(kBinomRows >= kMaxDenseN, "binomial table too small");

__device__ __forceinline__ DD dd_mul(DD a, DD b) {
  const double p = a.hi * b.hi;
  double e = __builtin_fma(a.hi, b.hi, -p);
  e = __builtin_fma(a.hi, b.lo, e);
  e = __builtin_fma(a.lo, b.hi, e);
  return quick_two_sum(p, e);
}

__global__ __launch_bounds__(256) void ryser_extreme_kernel(const double* __restrict__ colpm,
                                                            const double* __restrict__ x0,
                                                            const double* __restrict__ xall,
                                                            Partial* __restrict__ partials,
                                                            int n, int k, uint64_t per_side) {
  __shared__ Partial scratch[256 / kWave];
  const int B = n - 1;
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  DD s = {0.0, 0.0};
  double a = 0.0;
  if (r < 2 * per_side) {
    const bool high = r >= per_side;
    uint64_t rank = high ? r - per_side : r;
    // size of the small subset C: sizes 0..k hold C(B, size) subsets each
    int size = 0;
    for (; size < k; ++size) {
      const uint64_t cnt = kBinom.v[B][size];
      if (rank < cnt) break;
      rank -= cnt;
    }
    // colex unranking: c_p is the largest c with C(c, p) <= rank
    int cols[kExtMaxK];
    int c = B - 1;
#pragma unroll
    for (int p = kExtMaxK; p >= 1; --p) {
      cols[p - 1] = 0;
      if (p <= size) {
        while (kBinom.v[c][p] > rank) --c;
        cols[p - 1] = c;
        rank -= kBinom.v[c][p];
        --c;
      }
    }
    // S = C: x0 + columns of C; S = complement of C: x_all - columns of C
    // (store block 2c holds +a_c, block 2c+1 holds -a_c).
    const int neg = high ? 1 : 0;
    DD prod = {1.0, 0.0};
    for (int j = 0; j < n; ++j) {
      DD x = high ? DD{xall[j], xall[n + j]} : DD{x0[j], 0.0};
#pragma unroll
      for (int p = 0; p < kExtMaxK; ++p)
        if (p < size) x = dd_add_d(x, colpm[(size_t)(2 * cols[p] + neg) * n + j]);
      prod = dd_mul(prod, x);
    }
    const int s_size = high ? B - size : size;
    if (s_size & 1) prod = {-prod.hi, -prod.lo};
    s = prod;
    a = __builtin_fabs(prod.hi);
  }
  block_reduce_store(s, a, 0.0, scratch, partials + blockIdx.x);
}

hipError_t launch_extreme(const DenseLaunch& L, hipStream_t stream) {
  if (L.ext_k < 0 || L.ext_k > kExtMaxK || 2 * L.ext_k >= L.n - 1 || L.ext_blocks == 0)
    return hipErrorInvalidValue;
  ryser_extreme_kernel<<<dim3((unsigned)L.ext_blocks), dim3(256), 0, stream>>>(
      L.colpm, L.x0, L.xall, L.ext_partials, L.n, L.ext_k, L.ext_per_side);
  return hipGetLastError();
}

}  // namespace perm
