// The seeded generators of the benchmark inputs, built directly in HBM.
//
// Same steps as the host construction (graphgen.cpp, after
// testData/graphData/rMatGraph.C:33-79, gridGraph.C:53-72 and
// common/graphUtils.h:54-94), so the CSR is bit-identical to the host one:
//   1. one 64-bit key (u << vb) | v per directed edge, both directions,
//      self-loops marked ~0 (the same counter-based draws as the host);
//   2. an optional vertex relabelling (the host's seeded permutation);
//   3. a radix sort on the low 2*vb bits, then duplicates removed; the ~0
//      markers sort last and are dropped;
//   4. offsets and adjacency scattered from the sorted keys;
//   5. pbbs_graph_upload of the device arrays (the usual validation).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include "../../include/pbbs_bfs.h"
#include "bfs_device.hpp"
#include "engine.hpp"

using pbbs::fail;

namespace pbbs {
hipStream_t ctx_stream(pbbs_ctx* c);
int ctx_device(const pbbs_ctx* c);
int ctx_num_cus(const pbbs_ctx* c);
const void* graph_dev_off(const pbbs_graph* g);
const int32_t* graph_dev_adj(const pbbs_graph* g);
bool graph_off64(const pbbs_graph* g);
bool graph_symmetric(const pbbs_graph* g);
int64_t graph_n(const pbbs_graph* g);
int64_t graph_m(const pbbs_graph* g);
int graph_device(const pbbs_graph* g);
}  // namespace pbbs

namespace {

// SplitMix64 finaliser (Steele, Lea, Flood 2014); the host generator uses
// the same function and constants.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
uint64_t mix64_host(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// (x + 0.5) / 2^32: exact in double, so the quadrant comparisons match the
// host's bit for bit.
__device__ __forceinline__ double unit32(uint32_t x) { return (x + 0.5) * (1.0 / 4294967296.0); }

__global__ void rmat_keys_kernel(long long draws, int scale, double a, double ab, double abc,
                                 uint64_t hseed, unsigned long long* __restrict__ keys) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < draws; i += stride) {
    uint64_t st = mix64(hseed ^ ((uint64_t)i * 0xD1B54A32D192ED03ull));
    uint64_t u = 0, v = 0;
    for (int l = 0; l < scale; l += 2) {
      st = mix64(st);
      for (int h = 0; h < 2 && l + h < scale; ++h) {
        const double r = unit32((uint32_t)(h ? (st >> 32) : st));
        const uint64_t ub = r >= ab, vbit = (r >= a && r < ab) || r >= abc;
        u = (u << 1) | ub;
        v = (v << 1) | vbit;
      }
    }
    const bool loop = u == v;
    keys[2 * i] = loop ? ~0ull : (u << scale) | v;
    keys[2 * i + 1] = loop ? ~0ull : (v << scale) | u;
  }
}

__global__ void grid_keys_kernel(long long D, int vb, unsigned long long* __restrict__ keys) {
  const long long n = D * D * D;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += stride) {
    const long long i = l / (D * D), j = (l / D) % D, k = l % D;
    const long long nb[3] = {((i + 1) % D) * D * D + j * D + k, i * D * D + ((j + 1) % D) * D + k,
                             i * D * D + j * D + (k + 1) % D};
    for (int e = 0; e < 3; ++e) {
      const bool loop = nb[e] == l;
      keys[6 * l + 2 * e] = loop ? ~0ull : ((uint64_t)l << vb) | (uint64_t)nb[e];
      keys[6 * l + 2 * e + 1] = loop ? ~0ull : ((uint64_t)nb[e] << vb) | (uint64_t)l;
    }
  }
}

__global__ void permute_keys_kernel(unsigned long long* __restrict__ keys, long long N, int vb,
                                    const int32_t* __restrict__ perm) {
  const uint64_t vmask = (uint64_t(1) << vb) - 1;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
    const uint64_t k = keys[i];
    if (k == ~0ull) continue;
    keys[i] = ((uint64_t)perm[k >> vb] << vb) | (uint64_t)perm[k & vmask];
  }
}

// keys[0..m) sorted and distinct: adj[i] = low bits; offsets[u] = the first
// i of row u (rows with no key take the next row's start).
template <typename OffT>
__global__ void csr_scatter_kernel(const unsigned long long* __restrict__ keys, long long m,
                                   long long n, int vb, OffT* __restrict__ off,
                                   int32_t* __restrict__ adj) {
  const uint64_t vmask = (uint64_t(1) << vb) - 1;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i <= m; i += stride) {
    const long long u = i < m ? (long long)(keys[i] >> vb) : n;
    const long long prev = i == 0 ? -1 : (long long)(keys[i - 1] >> vb);
    for (long long x = prev + 1; x <= u; ++x) off[x] = (OffT)i;
    if (i < m) adj[i] = (int32_t)(keys[i] & vmask);
  }
}

int bits_for(int64_t n) {
  int b = 0;
  while ((int64_t(1) << b) < n) ++b;
  return std::max(b, 1);
}

// Device buffers freed on every exit path.
struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() {
    for (void* x : p) (void)hipFree(x);
  }
  template <typename T>
  hipError_t alloc(T** out, size_t bytes) {
    void* x = nullptr;
    const hipError_t e = hipMalloc(&x, std::max<size_t>(bytes, 1));
    if (e == hipSuccess) p.push_back(x);
    *out = static_cast<T*>(x);
    return e;
  }
};

#define GEN_TRY(expr)                                                                 \
  do {                                                                                \
    const hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                             \
      return fail(e_ == hipErrorOutOfMemory ? PBBS_ERR_NOMEM : PBBS_ERR_HIP,          \
                  "%s: %s", #expr, hipGetErrorString(e_));                            \
  } while (0)

// Steps 2-5 for keys[0..N) of a graph of n vertices.
int finish_graph(pbbs_ctx* c, int64_t n, int vb, unsigned long long* keys, int64_t N,
                 uint64_t permute_seed, uint32_t flags, DevBufs& bufs, pbbs_graph** out) {
  hipStream_t st = pbbs::ctx_stream(c);
  const int grid = std::max(1, pbbs::ctx_num_cus(c) * 8);
  if (permute_seed) {
    std::vector<int32_t> perm(n);
    pbbs::make_permutation(n, permute_seed, perm.data());
    int32_t* dperm = nullptr;
    GEN_TRY(bufs.alloc(&dperm, n * sizeof(int32_t)));
    GEN_TRY(hipMemcpyAsync(dperm, perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(permute_keys_kernel, dim3(grid), dim3(256), 0, st, keys, (long long)N, vb,
                       dperm);
    GEN_TRY(hipGetLastError());
    GEN_TRY(hipStreamSynchronize(st));   // perm (host) must outlive the copy
  }
  unsigned long long *sorted = nullptr, *uniq = nullptr;
  int64_t* count = nullptr;
  GEN_TRY(bufs.alloc(&sorted, N * sizeof(unsigned long long)));
  GEN_TRY(bufs.alloc(&uniq, N * sizeof(unsigned long long)));
  GEN_TRY(bufs.alloc(&count, sizeof(int64_t)));
  size_t b1 = 0, b2 = 0;
  GEN_TRY(rocprim::radix_sort_keys(nullptr, b1, keys, sorted, (size_t)N, 0, 2 * vb, st));
  GEN_TRY(rocprim::unique(nullptr, b2, sorted, uniq, count, (size_t)N,
                          rocprim::equal_to<unsigned long long>(), st));
  void* tmp = nullptr;
  GEN_TRY(bufs.alloc(&tmp, std::max(b1, b2)));
  GEN_TRY(rocprim::radix_sort_keys(tmp, b1, keys, sorted, (size_t)N, 0, 2 * vb, st));
  GEN_TRY(rocprim::unique(tmp, b2, sorted, uniq, count, (size_t)N,
                          rocprim::equal_to<unsigned long long>(), st));
  int64_t m = 0;
  GEN_TRY(hipMemcpyAsync(&m, count, sizeof m, hipMemcpyDeviceToHost, st));
  GEN_TRY(hipStreamSynchronize(st));
  if (m > 0) {   // the self-loop marker sorts last
    unsigned long long last = 0;
    GEN_TRY(hipMemcpy(&last, uniq + (m - 1), sizeof last, hipMemcpyDeviceToHost));
    if (last == ~0ull) --m;
  }
  if (m > (int64_t)UINT32_MAX) return fail(PBBS_ERR_LIMIT, "graph has %lld >= 2^32 edges", (long long)m);
  const bool wide = (flags & PBBS_GRAPH_OFFSETS_I64) != 0;
  void* off = nullptr;
  int32_t* adj = nullptr;
  GEN_TRY(bufs.alloc(&off, (n + 1) * (wide ? 8 : 4)));
  GEN_TRY(bufs.alloc(&adj, m * sizeof(int32_t)));
  if (wide)
    hipLaunchKernelGGL(csr_scatter_kernel<int64_t>, dim3(grid), dim3(256), 0, st, uniq,
                       (long long)m, (long long)n, vb, static_cast<int64_t*>(off), adj);
  else
    hipLaunchKernelGGL(csr_scatter_kernel<uint32_t>, dim3(grid), dim3(256), 0, st, uniq,
                       (long long)m, (long long)n, vb, static_cast<uint32_t*>(off), adj);
  GEN_TRY(hipGetLastError());
  GEN_TRY(hipStreamSynchronize(st));
  // symmetric by construction; the upload still checks the offsets, ids and
  // row order
  return pbbs_graph_upload(c, n, m, off, adj,
                           (wide ? PBBS_GRAPH_OFFSETS_I64 : 0u) | PBBS_GRAPH_DEVICE_INPUT |
                               PBBS_GRAPH_ASSUME_SYMMETRIC,
                           out);
}

// Edge arrays from a resident graph: per row, the entries above the row id.
template <typename OffT>
__global__ void upper_count_kernel(const OffT* __restrict__ off, const int32_t* __restrict__ adj,
                                   long long n, unsigned long long* __restrict__ cnt) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    unsigned long long c = 0;
    for (OffT e = off[v]; e < off[v + 1]; ++e) c += adj[e] > v;
    cnt[v] = c;
  }
}

// Row-order pairs packed (u << 32 | v), u < v, at each row's prefix position.
template <typename OffT>
__global__ void upper_write_kernel(const OffT* __restrict__ off, const int32_t* __restrict__ adj,
                                   long long n, const unsigned long long* __restrict__ pre,
                                   unsigned long long* __restrict__ pairs) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    unsigned long long w = pre[v];
    for (OffT e = off[v]; e < off[v + 1]; ++e)
      if (adj[e] > v) pairs[w++] = ((unsigned long long)v << 32) | (uint32_t)adj[e];
  }
}

__global__ void shuffle_keys_kernel(long long m, uint64_t hseed, unsigned long long* __restrict__ k) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
    k[i] = mix64(hseed ^ ((uint64_t)i * 0xD1B54A32D192ED03ull));
}

__global__ void unpack_pairs_kernel(long long m, const unsigned long long* __restrict__ pairs,
                                    int32_t* __restrict__ u, int32_t* __restrict__ v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    const unsigned long long p = pairs[i];
    u[i] = (int32_t)(p >> 32);
    v[i] = (int32_t)(uint32_t)p;
  }
}

}  // namespace

extern "C" {

int pbbs_graph_edges(pbbs_ctx* c, const pbbs_graph* g, uint64_t shuffle_seed, int32_t* u,
                     int32_t* v, int64_t cap, int64_t* m_out) {
  pbbs::clear_error();
  if (!c || !g || !m_out) return fail(PBBS_ERR_INVALID, "NULL argument");
  if (!pbbs::graph_symmetric(g)) return fail(PBBS_ERR_INVALID, "the graph is not symmetric");
  if (pbbs::graph_device(g) != pbbs::ctx_device(c))
    return fail(PBBS_ERR_INVALID, "graph lives on another device");
  GEN_TRY(hipSetDevice(pbbs::ctx_device(c)));
  hipStream_t st = pbbs::ctx_stream(c);
  const int grid = std::max(1, pbbs::ctx_num_cus(c) * 8);
  const int64_t n = pbbs::graph_n(g);
  DevBufs bufs;
  unsigned long long *cnt = nullptr, *pre = nullptr;
  GEN_TRY(bufs.alloc(&cnt, (n + 1) * sizeof(unsigned long long)));
  GEN_TRY(bufs.alloc(&pre, (n + 1) * sizeof(unsigned long long)));
  const bool w = pbbs::graph_off64(g);
  const void* off = pbbs::graph_dev_off(g);
  const int32_t* adj = pbbs::graph_dev_adj(g);
  if (w)
    hipLaunchKernelGGL(upper_count_kernel<int64_t>, dim3(grid), dim3(256), 0, st,
                       static_cast<const int64_t*>(off), adj, (long long)n, cnt);
  else
    hipLaunchKernelGGL(upper_count_kernel<uint32_t>, dim3(grid), dim3(256), 0, st,
                       static_cast<const uint32_t*>(off), adj, (long long)n, cnt);
  GEN_TRY(hipGetLastError());
  GEN_TRY(hipMemsetAsync(cnt + n, 0, sizeof(unsigned long long), st));
  size_t b1 = 0;
  GEN_TRY(rocprim::exclusive_scan(nullptr, b1, cnt, pre, 0ull, (size_t)(n + 1),
                                  rocprim::plus<unsigned long long>(), st));
  void* tmp = nullptr;
  GEN_TRY(bufs.alloc(&tmp, b1));
  GEN_TRY(rocprim::exclusive_scan(tmp, b1, cnt, pre, 0ull, (size_t)(n + 1),
                                  rocprim::plus<unsigned long long>(), st));
  unsigned long long m = 0;
  GEN_TRY(hipMemcpyAsync(&m, pre + n, sizeof m, hipMemcpyDeviceToHost, st));
  GEN_TRY(hipStreamSynchronize(st));
  *m_out = (int64_t)m;
  if (!u && !v && cap == 0) return PBBS_OK;   // the count only
  if ((int64_t)m > cap) return fail(PBBS_ERR_LIMIT, "%lld edges, room for %lld", (long long)m,
                                    (long long)cap);
  if (m == 0) return PBBS_OK;
  if (!u || !v) return fail(PBBS_ERR_INVALID, "NULL edge buffer");
  unsigned long long* pairs = nullptr;
  GEN_TRY(bufs.alloc(&pairs, m * sizeof(unsigned long long)));
  if (w)
    hipLaunchKernelGGL(upper_write_kernel<int64_t>, dim3(grid), dim3(256), 0, st,
                       static_cast<const int64_t*>(off), adj, (long long)n, pre, pairs);
  else
    hipLaunchKernelGGL(upper_write_kernel<uint32_t>, dim3(grid), dim3(256), 0, st,
                       static_cast<const uint32_t*>(off), adj, (long long)n, pre, pairs);
  GEN_TRY(hipGetLastError());
  const unsigned long long* src = pairs;
  if (shuffle_seed) {
    unsigned long long *keys = nullptr, *skeys = nullptr, *spairs = nullptr;
    GEN_TRY(bufs.alloc(&keys, m * sizeof(unsigned long long)));
    GEN_TRY(bufs.alloc(&skeys, m * sizeof(unsigned long long)));
    GEN_TRY(bufs.alloc(&spairs, m * sizeof(unsigned long long)));
    hipLaunchKernelGGL(shuffle_keys_kernel, dim3(grid), dim3(256), 0, st, (long long)m,
                       mix64_host(shuffle_seed), keys);
    GEN_TRY(hipGetLastError());
    size_t b2 = 0;
    GEN_TRY(rocprim::radix_sort_pairs(nullptr, b2, keys, skeys, pairs, spairs, (size_t)m, 0, 64,
                                      st));
    void* tmp2 = nullptr;
    GEN_TRY(bufs.alloc(&tmp2, b2));
    GEN_TRY(rocprim::radix_sort_pairs(tmp2, b2, keys, skeys, pairs, spairs, (size_t)m, 0, 64,
                                      st));
    src = spairs;
  }
  hipLaunchKernelGGL(unpack_pairs_kernel, dim3(grid), dim3(256), 0, st, (long long)m, src, u, v);
  GEN_TRY(hipGetLastError());
  GEN_TRY(hipStreamSynchronize(st));
  return PBBS_OK;
}

int pbbs_graph_gen_rmat(pbbs_ctx* c, int32_t scale, int64_t draws, double a, double b, double cq,
                        uint64_t seed, uint64_t permute_seed, uint32_t flags, pbbs_graph** out) {
  pbbs::clear_error();
  if (!c || !out) return fail(PBBS_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (scale < 1 || scale > 30) return fail(PBBS_ERR_LIMIT, "scale %d outside [1, 30]", scale);
  if (draws < 0) return fail(PBBS_ERR_INVALID, "negative edge count");
  if (!(a >= 0 && b >= 0 && cq >= 0 && a + b + cq <= 1.0))
    return fail(PBBS_ERR_INVALID, "quadrant probabilities must be >= 0 with a+b+c <= 1");
  if (flags & ~(uint32_t)PBBS_GRAPH_OFFSETS_I64)
    return fail(PBBS_ERR_INVALID, "only PBBS_GRAPH_OFFSETS_I64 is a valid flag here");
  GEN_TRY(hipSetDevice(pbbs::ctx_device(c)));
  const int64_t n = int64_t(1) << scale;
  DevBufs bufs;
  unsigned long long* keys = nullptr;
  GEN_TRY(bufs.alloc(&keys, draws * 2 * sizeof(unsigned long long)));
  hipStream_t st = pbbs::ctx_stream(c);
  if (draws > 0) {
    const double ab = a + b, abc = a + b + cq;   // computed as the host generator does
    hipLaunchKernelGGL(rmat_keys_kernel, dim3(std::max(1, pbbs::ctx_num_cus(c) * 8)), dim3(256), 0,
                       st, (long long)draws, scale, a, ab, abc, mix64_host(seed), keys);
    GEN_TRY(hipGetLastError());
  }
  return finish_graph(c, n, scale, keys, draws * 2, permute_seed, flags, bufs, out);
}

int pbbs_graph_gen_grid3d(pbbs_ctx* c, int32_t d, uint32_t flags, pbbs_graph** out) {
  pbbs::clear_error();
  if (!c || !out) return fail(PBBS_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (d < 1 || (int64_t)d * d * d > INT32_MAX) return fail(PBBS_ERR_LIMIT, "bad grid side %d", d);
  if (flags & ~(uint32_t)PBBS_GRAPH_OFFSETS_I64)
    return fail(PBBS_ERR_INVALID, "only PBBS_GRAPH_OFFSETS_I64 is a valid flag here");
  GEN_TRY(hipSetDevice(pbbs::ctx_device(c)));
  const int64_t D = d, n = D * D * D;
  const int vb = bits_for(n);
  DevBufs bufs;
  unsigned long long* keys = nullptr;
  GEN_TRY(bufs.alloc(&keys, n * 6 * sizeof(unsigned long long)));
  hipLaunchKernelGGL(grid_keys_kernel, dim3(std::max(1, pbbs::ctx_num_cus(c) * 8)), dim3(256), 0,
                     pbbs::ctx_stream(c), (long long)D, vb, keys);
  GEN_TRY(hipGetLastError());
  return finish_graph(c, n, vb, keys, n * 6, 0, flags, bufs, out);
}

}  // extern "C"
