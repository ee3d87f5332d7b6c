// copy_probe.hip -- a plain device copy of a known size, used to calibrate
// the HBM traffic counters (FETCH_SIZE / WRITE_SIZE) that the profiles of the
// Ryser kernels are scaled by: the copy reads and writes exactly `bytes`.
//
//   copy_probe [MiB]   (default 256)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst,
                            size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

#define CHECK(x)                                                                   \
  do {                                                                             \
    hipError_t e = (x);                                                            \
    if (e != hipSuccess) {                                                         \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));                  \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 256;
  const size_t bytes = mib << 20, n = bytes / sizeof(double4);
  void *src, *dst;
  CHECK(hipMalloc(&src, bytes));
  CHECK(hipMalloc(&dst, bytes));
  CHECK(hipMemset(src, 1, bytes));
  CHECK(hipDeviceSynchronize());
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  CHECK(hipEventRecord(a));
  copy_kernel<<<4096, 256>>>((const double4*)src, (double4*)dst, n);
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(b));
  CHECK(hipEventSynchronize(b));
  float ms = 0.f;
  CHECK(hipEventElapsedTime(&ms, a, b));
  std::printf("copy_probe: %zu MiB read + %zu MiB written in %.3f ms (%.1f GB/s)\n", mib, mib, ms,
              2.0 * bytes / (ms * 1e-3) / 1e9);
  CHECK(hipFree(src));
  CHECK(hipFree(dst));
  return 0;
}
