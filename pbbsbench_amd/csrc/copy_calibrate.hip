// HBM counter calibration: one grid-stride copy of exactly 1 GiB (16-byte
// elements), timed with hipEvents.  Profiled with `rocprofv3 --pmc
// FETCH_SIZE` and `--pmc WRITE_SIZE`, it gives the factor between each
// counter and the bytes a kernel actually moves (1 GiB read, 1 GiB written).
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_1gib(const uint4* __restrict__ s, uint4* __restrict__ d, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    d[i] = s[i];
}

#define CK(x)                                                         \
  do {                                                                \
    if ((x) != hipSuccess) {                                          \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(x)); \
      return 1;                                                       \
    }                                                                 \
  } while (0)

int main() {
  const long long bytes = 1ll << 30, n = bytes / 16;
  uint4 *s = nullptr, *d = nullptr;
  CK(hipMalloc(&s, bytes));
  CK(hipMalloc(&d, bytes));
  CK(hipMemset(s, 1, bytes));
  CK(hipMemset(d, 0, bytes));
  CK(hipDeviceSynchronize());
  hipLaunchKernelGGL(copy_1gib, dim3(4096), dim3(256), 0, 0, s, d, n);
  CK(hipGetLastError());
  CK(hipDeviceSynchronize());
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  CK(hipEventRecord(a));
  hipLaunchKernelGGL(copy_1gib, dim3(4096), dim3(256), 0, 0, s, d, n);
  CK(hipEventRecord(b));
  CK(hipEventSynchronize(b));
  float ms = 0;
  CK(hipEventElapsedTime(&ms, a, b));
  std::printf("copy 1 GiB: %.3f ms = %.1f GB/s (read + write)\n", ms, 2.0 * bytes / ms / 1e6);
  CK(hipFree(s));
  CK(hipFree(d));
  return 0;
}
