// ryser_dense.hip -- dispatcher of the dense Ryser kernel family (the kernels
// live in ryser_dense_impl.hpp and are instantiated for ranges of n in
// ryser_dense_n*.hip, compiled in parallel) and the grid reduction.
#include "ryser_dense_impl.hpp"

namespace perm {

hipError_t launch_dense_range_a(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_b(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_c(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_d(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_e(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_f(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_g(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_h(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_i(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_j(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_k(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_l(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_m(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);
hipError_t launch_dense_range_o(const DenseLaunch& L, hipStream_t s, hipEvent_t ev);

// Grid-level reduction of `count` partials into out[0] (one workgroup of
// 256 threads, fixed order).
__global__ __launch_bounds__(256) void reduce_partials_kernel(const Partial* __restrict__ in,
                                                              uint64_t count,
                                                              Partial* __restrict__ out) {
  __shared__ Partial scratch[256 / kWave];
  __shared__ double clk[256];
  DD s = {0.0, 0.0};
  Moments m = {0.0, 0.0, 0.0};
  double c = 0.0;
  for (uint64_t i = threadIdx.x; i < count; i += blockDim.x) {
    s = dd_add(s, DD{in[i].hi, in[i].lo});
    m.a += in[i].abs_sum;
    m.q += in[i].sq_sum;
    m.r += in[i].sq_run;
    c += in[i].pad;
  }
  clk[threadIdx.x] = c;
  block_reduce_store(s, m, scratch, out, 0.0, clk);  // (its first __syncthreads orders clk)
}

int dense_unroll_bits(int n) { return n - 1 < 3 ? n - 1 : 3; }

// z[i] = sq_sum + run_weight * sq_run of record i (grid-stride).
__global__ __launch_bounds__(256) void unit_weights_kernel(const Partial* __restrict__ rec,
                                                           uint64_t units, double run_weight,
                                                           double* __restrict__ z) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < units;
       i += (uint64_t)gridDim.x * blockDim.x)
    z[i] = rec[i].sq_sum + run_weight * rec[i].sq_run;
}

// dst[k] = src[ids[k]], k < count.
__global__ __launch_bounds__(256) void gather_partials_kernel(const Partial* __restrict__ src,
                                                              const uint32_t* __restrict__ ids,
                                                              uint32_t count,
                                                              Partial* __restrict__ dst) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) dst[k] = src[ids[k]];
}

hipError_t launch_unit_weights(const Partial* rec, uint64_t units, double run_weight, double* z,
                               hipStream_t stream) {
  if (units == 0) return hipSuccess;
  const uint64_t want = (units + 255) / 256;
  const unsigned grid = (unsigned)(want < 4096 ? want : 4096);
  unit_weights_kernel<<<grid, 256, 0, stream>>>(rec, units, run_weight, z);
  return hipGetLastError();
}

hipError_t launch_gather_partials(const Partial* src, const uint32_t* ids, uint32_t count,
                                  Partial* dst, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  gather_partials_kernel<<<(count + 255) / 256, 256, 0, stream>>>(src, ids, count, dst);
  return hipGetLastError();
}

hipError_t launch_skip_range_a(const SkipLaunch& L, hipStream_t s);
hipError_t launch_skip_range_b(const SkipLaunch& L, hipStream_t s);
hipError_t launch_skip_range_c(const SkipLaunch& L, hipStream_t s);

hipError_t launch_skip(const SkipLaunch& L, hipStream_t stream) {
  if (L.n < kMinSkipN || L.n > kMaxDenseN) return hipErrorInvalidValue;
  hipError_t e = L.n <= 32   ? launch_skip_range_a(L, stream)
                 : L.n <= 48 ? launch_skip_range_b(L, stream)
                             : launch_skip_range_c(L, stream);
  if (e != hipSuccess) return e;
  reduce_partials_kernel<<<1, 256, 0, stream>>>(L.partials, L.blocks, L.result);
  return hipGetLastError();
}

static hipError_t launch_passes(const DenseLaunch& L, hipStream_t stream, hipEvent_t main_done) {
  if (L.n < 1 || L.n > kMaxDenseN) return hipErrorInvalidValue;
  using RangeFn = hipError_t (*)(const DenseLaunch&, hipStream_t, hipEvent_t);
  static const struct {
    int hi;
    RangeFn fn;
  } kRanges[] = {{20, launch_dense_range_a}, {26, launch_dense_range_b},
                 {30, launch_dense_range_c}, {33, launch_dense_range_d},
                 {36, launch_dense_range_e}, {38, launch_dense_range_f},
                 {40, launch_dense_range_g}, {42, launch_dense_range_h},
                 {44, launch_dense_range_i}, {46, launch_dense_range_j},
                 {48, launch_dense_range_k}, {53, launch_dense_range_l},
                 {58, launch_dense_range_m}, {63, launch_dense_range_o}};
  for (const auto& r : kRanges)
    if (L.n <= r.hi) return r.fn(L, stream, main_done);
  return hipErrorInvalidValue;
}

hipError_t launch_dense_pass_only(const DenseLaunch& L, hipStream_t stream) {
  return launch_passes(L, stream, nullptr);
}

hipError_t launch_reduce(const Partial* in, uint64_t count, Partial* out, hipStream_t stream) {
  reduce_partials_kernel<<<1, 256, 0, stream>>>(in, count, out);
  return hipGetLastError();
}

hipError_t launch_dense(const DenseLaunch& L, hipStream_t stream, hipEvent_t main_done) {
  hipError_t e = launch_passes(L, stream, main_done);
  if (e != hipSuccess || L.total_blocks <= 1) return e;  // one record: it is the sum
  reduce_partials_kernel<<<1, 256, 0, stream>>>(L.partials, L.total_blocks, L.result);
  return hipGetLastError();
}

}  // namespace perm
