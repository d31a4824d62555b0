// Device-atomic calibration: throughput of scattered 4-byte loads (plain,
// and device-coherent: relaxed agent-scope atomic loads) and of scattered
// atomicMax (with and without a returned value) over working sets from
// 64 KiB to 64 MiB, and the round-trip latency of one wavefront's dependent
// chain of each.  The BFS claim step is one atomicMax per frontier edge into
// an unvisited vertex, so these rates bound the levels whose time is claims.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(x)                                                            \
  do {                                                                   \
    const hipError_t e_ = (x);                                           \
    if (e_ != hipSuccess) {                                              \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_)); \
      return 1;                                                          \
    }                                                                    \
  } while (0)

// A 32-bit mixing function (xorshift-multiply rounds) for scattered indices.
__device__ __forceinline__ unsigned mix32(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

enum Op { kLoad = 0, kMaxRet = 1, kMaxNoRet = 2, kLoadDev = 3 };

__device__ __forceinline__ int load_dev(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every thread makes `iters` accesses at scattered indices in [0, n).
template <int kOp>
__global__ void scattered(int* a, unsigned n, int iters, int* sink) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  int acc = 0;
  for (int i = 0; i < iters; ++i) {
    const unsigned idx = mix32(t * 131u + (unsigned)i * 7919u) % n;
    if (kOp == kLoad) acc += a[idx];
    if (kOp == kMaxRet) acc += atomicMax(&a[idx], (int)t);
    if (kOp == kMaxNoRet) atomicMax(&a[idx], (int)t);
    if (kOp == kLoadDev) acc += load_dev(&a[idx]);
  }
  if (acc == 0x7fffffff) sink[0] = acc;   // keeps the loads
}

// One wavefront, `steps` dependent accesses per lane: each index comes from
// the value the previous access returned (the array holds a permutation-like
// next-index table, so every address stays in [0, n)).
template <int kOp>
__global__ void chain(int* a, unsigned n, int steps, int* sink) {
  unsigned idx = mix32(threadIdx.x) % n;
  for (int i = 0; i < steps; ++i) {
    const int v = kOp == kLoad      ? a[idx]
                  : kOp == kLoadDev ? load_dev(&a[idx])
                                    : atomicMax(&a[idx], -1);   // -1 never changes a value
    idx = (unsigned)v % n;
  }
  sink[1 + threadIdx.x] = (int)idx;
}

__global__ void fill_next(int* a, unsigned n) {
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    a[i] = (int)(mix32(i * 2654435761u + 17u) % n);
}

int main() {
  const unsigned nmax = 16u << 20;   // 64 MiB of int
  int *a = nullptr, *sink = nullptr;
  CK(hipMalloc(&a, nmax * sizeof(int)));
  CK(hipMalloc(&sink, 128 * sizeof(int)));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const int grid = 4096, block = 256, iters = 16;
  const char* names[] = {"load", "atomicMax (returned)", "atomicMax (no return)",
                         "load, agent scope"};
  std::printf("scattered accesses, %d x %d threads x %d each\n", grid, block, iters);
  std::printf("| working set | op | us | G accesses/s |\n|---|---|---|---|\n");
  for (unsigned n : {16u << 10, 256u << 10, 1u << 20, 4u << 20, 16u << 20}) {
    for (int op = 0; op < 4; ++op) {
      CK(hipMemset(a, 0, n * sizeof(int)));
      float ms = 0.f;
      for (int rep = 0; rep < 2; ++rep) {   // the first run warms up
        CK(hipEventRecord(e0));
        if (op == kLoad) hipLaunchKernelGGL(scattered<kLoad>, dim3(grid), dim3(block), 0, 0, a, n, iters, sink);
        if (op == kMaxRet) hipLaunchKernelGGL(scattered<kMaxRet>, dim3(grid), dim3(block), 0, 0, a, n, iters, sink);
        if (op == kMaxNoRet) hipLaunchKernelGGL(scattered<kMaxNoRet>, dim3(grid), dim3(block), 0, 0, a, n, iters, sink);
        if (op == kLoadDev) hipLaunchKernelGGL(scattered<kLoadDev>, dim3(grid), dim3(block), 0, 0, a, n, iters, sink);
        CK(hipGetLastError());
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms, e0, e1));
      }
      std::printf("| %u KiB | %s | %.1f | %.2f |\n", n / 256, names[op], ms * 1e3,
                  (double)grid * block * iters / (ms * 1e-3) / 1e9);
    }
  }
  const int steps = 2000;
  std::printf("\none wavefront, %d dependent accesses per lane\n", steps);
  std::printf("| working set | op | ns per access |\n|---|---|---|\n");
  for (unsigned n : {16u << 10, 16u << 20}) {
    hipLaunchKernelGGL(fill_next, dim3(1024), dim3(256), 0, 0, a, n);
    CK(hipGetLastError());
    for (int op : {kLoad, kLoadDev, kMaxRet}) {
      float ms = 0.f;
      for (int rep = 0; rep < 2; ++rep) {
        CK(hipEventRecord(e0));
        if (op == kLoad) hipLaunchKernelGGL(chain<kLoad>, dim3(1), dim3(64), 0, 0, a, n, steps, sink);
        else if (op == kLoadDev) hipLaunchKernelGGL(chain<kLoadDev>, dim3(1), dim3(64), 0, 0, a, n, steps, sink);
        else hipLaunchKernelGGL(chain<kMaxRet>, dim3(1), dim3(64), 0, 0, a, n, steps, sink);
        CK(hipGetLastError());
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms, e0, e1));
      }
      std::printf("| %u KiB | %s | %.0f |\n", n / 256, names[op], ms * 1e6 / steps);
    }
  }
  CK(hipFree(a));
  CK(hipFree(sink));
  return 0;
}
