// BFS output checker on the GPU (bench/BFSCheck.C:39-90, with the source as
// a parameter, plus the max-parent rule of deterministicBFS):
//   1. parents outside [-1, n) fail at the first such vertex;
//   2. level_number (BFSCheck.C:39-56): (count, ptr) pointer doubling to a
//      fixed point, unreached -> n; no fixed point within 40 doublings means
//      a cycle in the parents;
//   3. per vertex v (one wavefront per row): the source must be its own
//      parent; any other v needs level[v] == min level over N(v) + 1, or
//      both n (BFSCheck.C:59-90); the max rule wants parent[v] = the largest
//      u in N(v) with level[u] == level[v] - 1 (-1 if unreached); given
//      depths must equal the levels.
// The max rule reads N(v) as v's in-neighbours, which holds on symmetric
// graphs only; on any other graph it is skipped and max_rule_bad is -1.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "../../include/pbbs_bfs.h"
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

constexpr int kWave = 64;

// counters[0] bad (reference rule), [1] max-rule bad, [2] depth bad,
// [3] reached, [4] max level, [5] first bad vertex (min), [6] range bad,
// [7] changed flag of a doubling round
enum { kBad, kMaxBad, kDepthBad, kReached, kMaxLevel, kFirstBad, kRangeBad, kChanged, kCounters };

__global__ void range_kernel(const int32_t* __restrict__ parent, long long n,
                             long long* __restrict__ ctr) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    const int32_t p = parent[v];
    if (p < -1 || p >= n) {
      atomicAdd((unsigned long long*)&ctr[kRangeBad], 1ull);
      atomicMin(&ctr[kFirstBad], v);
    }
  }
}

__global__ void level_init_kernel(const int32_t* __restrict__ parent, long long n,
                                  long long* __restrict__ cnt, int32_t* __restrict__ ptr) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t p = parent[i];
    if (p == -1 || p == i) {
      cnt[i] = 0;
      ptr[i] = (int32_t)i;
    } else {
      cnt[i] = 1;
      ptr[i] = p;
    }
  }
}

__global__ void level_round_kernel(const long long* __restrict__ cnt,
                                   const int32_t* __restrict__ ptr, long long n,
                                   long long* __restrict__ cnt2, int32_t* __restrict__ ptr2,
                                   long long* __restrict__ ctr) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  bool changed = false;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t p = ptr[i];
    const long long c = cnt[p] + cnt[i];
    const int32_t q = ptr[p];
    cnt2[i] = c;
    ptr2[i] = q;
    changed |= c != cnt[i] || q != p;
  }
  if (__ballot(changed) && (threadIdx.x % kWave) == 0) ctr[kChanged] = 1;
}

__global__ void level_final_kernel(const int32_t* __restrict__ parent,
                                   const long long* __restrict__ cnt,
                                   const int32_t* __restrict__ ptr, long long n,
                                   int32_t* __restrict__ level) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    level[i] = parent[ptr[i]] == -1 ? (int32_t)n : (int32_t)cnt[i];
}

template <typename OffT>
__global__ void check_rows_kernel(const OffT* __restrict__ off, const int32_t* __restrict__ adj,
                                  long long n, int32_t s, const int32_t* __restrict__ parent,
                                  const int32_t* __restrict__ depth,
                                  const int32_t* __restrict__ level, long long* __restrict__ ctr) {
  const int lane = threadIdx.x % kWave;
  const long long W = (long long)gridDim.x * (blockDim.x / kWave);
  const long long w = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  long long bad = 0, maxbad = 0, depthbad = 0, reached = 0, first = LLONG_MAX;
  int maxlev = 0;
  for (long long v = w; v < n; v += W) {
    const int32_t lv = level[v];
    const OffT lo = off[v], hi = off[v + 1];
    int32_t mn = (int32_t)n, mx = -1;
    for (OffT e = lo + lane; e < hi; e += kWave) {
      const int32_t u = adj[e];
      const int32_t lu = level[u];
      mn = min(mn, lu);
      if (lu == lv - 1 && u > mx) mx = u;
    }
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      mn = min(mn, __shfl_xor(mn, d));
      mx = max(mx, __shfl_xor(mx, d));
    }
    if (lane == 0) {
      const int32_t p = parent[v];
      bool ok, okmax;
      if (v == s) {
        ok = p == s;
        okmax = ok;
      } else {
        ok = (lv == n && mn == n) || lv == mn + 1;
        okmax = p == (lv == n ? -1 : mx);
      }
      if (!ok) {
        ++bad;
        first = min(first, v);
      }
      maxbad += !okmax;
      if (depth && depth[v] != lv) ++depthbad;
      if (lv < n) {
        ++reached;
        maxlev = max(maxlev, (int)lv);
      }
    }
  }
  if (lane == 0) {
    if (bad) atomicAdd((unsigned long long*)&ctr[kBad], (unsigned long long)bad);
    if (maxbad) atomicAdd((unsigned long long*)&ctr[kMaxBad], (unsigned long long)maxbad);
    if (depthbad) atomicAdd((unsigned long long*)&ctr[kDepthBad], (unsigned long long)depthbad);
    if (reached) atomicAdd((unsigned long long*)&ctr[kReached], (unsigned long long)reached);
    if (maxlev) atomicMax(&ctr[kMaxLevel], (long long)maxlev);
    if (first != LLONG_MAX) atomicMin(&ctr[kFirstBad], first);
  }
}

// Maximal-independent-set check (maximalIndependentSet/bench/MISCheck.C:37-63),
// one wavefront per row: a member with a member neighbour is a bad edge; a
// non-member (any flag other than 1) with no member neighbour is a bad
// vertex.  The row scan stops at its first member neighbour unless v is a
// member (then the whole row is read: any member neighbour is the fault).
// ctr[0] bad edges (one per member with a member neighbour), [1] bad vertices,
// [2] the smallest vertex of either kind.
template <typename OffT>
__global__ void check_mis_kernel(const OffT* __restrict__ off, const int32_t* __restrict__ adj,
                                 long long n, const int32_t* __restrict__ flags,
                                 long long* __restrict__ ctr) {
  const int lane = threadIdx.x % kWave;
  const long long W = (long long)gridDim.x * (blockDim.x / kWave);
  const long long w = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  long long bad_edges = 0, bad_vertices = 0, first = LLONG_MAX;
  for (long long v = w; v < n; v += W) {
    const bool member = flags[v] == 1;
    const OffT lo = off[v], hi = off[v + 1];
    bool has = false;
    for (OffT base = lo; base < hi; base += kWave) {   // wave-uniform bounds
      const OffT e = base + (OffT)lane;
      const bool hit = e < hi && flags[adj[e]] == 1;
      if (__ballot(hit)) {
        has = true;
        break;
      }
    }
    if (lane == 0) {
      if (member && has) {
        ++bad_edges;
        first = min(first, v);
      } else if (!member && !has) {
        ++bad_vertices;
        first = min(first, v);
      }
    }
  }
  if (lane == 0) {
    if (bad_edges) atomicAdd((unsigned long long*)&ctr[0], (unsigned long long)bad_edges);
    if (bad_vertices) atomicAdd((unsigned long long*)&ctr[1], (unsigned long long)bad_vertices);
    if (first != LLONG_MAX) atomicMin(&ctr[2], first);
  }
}

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

#define CHK_TRY(expr)                                                                 \
  do {                                                                                \
    const hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                             \
      return fail(e_ == hipErrorOutOfMemory ? PBBS_ERR_NOMEM : PBBS_ERR_HIP,          \
                  "%s: %s", #expr, hipGetErrorString(e_));                            \
  } while (0)

}  // namespace

extern "C" int pbbs_check_bfs(pbbs_ctx* c, const pbbs_graph* g, int32_t source,
                              const int32_t* parent, const int32_t* depth, uint32_t flags,
                              pbbs_check_result* res) {
  pbbs::clear_error();
  if (!c || !g || !parent || !res) return fail(PBBS_ERR_INVALID, "NULL argument");
  if (flags & ~(uint32_t)PBBS_CHECK_DEVICE_INPUT)
    return fail(PBBS_ERR_INVALID, "unknown check flags 0x%x", flags);
  const int64_t n = pbbs::graph_n(g);
  if (source < 0 || source >= n) return fail(PBBS_ERR_INVALID, "source %d outside the graph", source);
  if (pbbs::graph_device(g) != pbbs::ctx_device(c))
    return fail(PBBS_ERR_INVALID, "graph lives on another device");
  CHK_TRY(hipSetDevice(pbbs::ctx_device(c)));
  hipStream_t st = pbbs::ctx_stream(c);
  const int grid = std::max(1, pbbs::ctx_num_cus(c) * 8);
  DevBufs bufs;
  const int32_t* dp = parent;
  const int32_t* dd = depth;
  if (!(flags & PBBS_CHECK_DEVICE_INPUT)) {
    int32_t *p2 = nullptr, *d2 = nullptr;
    CHK_TRY(bufs.alloc(&p2, n * sizeof(int32_t)));
    CHK_TRY(hipMemcpyAsync(p2, parent, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    dp = p2;
    if (depth) {
      CHK_TRY(bufs.alloc(&d2, n * sizeof(int32_t)));
      CHK_TRY(hipMemcpyAsync(d2, depth, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
      dd = d2;
    }
  }
  long long* ctr = nullptr;
  CHK_TRY(bufs.alloc(&ctr, kCounters * sizeof(long long)));
  long long init[kCounters] = {};
  init[kFirstBad] = LLONG_MAX;
  CHK_TRY(hipMemcpyAsync(ctr, init, sizeof init, hipMemcpyHostToDevice, st));
  long long h[kCounters];
  hipLaunchKernelGGL(range_kernel, dim3(grid), dim3(256), 0, st, dp, (long long)n, ctr);
  CHK_TRY(hipGetLastError());
  CHK_TRY(hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, st));
  CHK_TRY(hipStreamSynchronize(st));
  pbbs_check_result r{};
  r.first_bad = -1;
  if (h[kRangeBad]) {   // levels are undefined: report the first such vertex
    r.valid = 0;
    r.bad = h[kRangeBad];
    r.first_bad = h[kFirstBad];
    *res = r;
    return PBBS_OK;
  }
  long long *cnt = nullptr, *cnt2 = nullptr;
  int32_t *ptr = nullptr, *ptr2 = nullptr, *level = nullptr;
  CHK_TRY(bufs.alloc(&cnt, n * sizeof(long long)));
  CHK_TRY(bufs.alloc(&cnt2, n * sizeof(long long)));
  CHK_TRY(bufs.alloc(&ptr, n * sizeof(int32_t)));
  CHK_TRY(bufs.alloc(&ptr2, n * sizeof(int32_t)));
  CHK_TRY(bufs.alloc(&level, n * sizeof(int32_t)));
  hipLaunchKernelGGL(level_init_kernel, dim3(grid), dim3(256), 0, st, dp, (long long)n, cnt, ptr);
  CHK_TRY(hipGetLastError());
  bool converged = false;
  for (int round = 0; round < 40 && !converged; ++round) {
    const long long zero = 0;
    CHK_TRY(hipMemcpyAsync(ctr + kChanged, &zero, sizeof zero, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(level_round_kernel, dim3(grid), dim3(256), 0, st, cnt, ptr, (long long)n,
                       cnt2, ptr2, ctr);
    CHK_TRY(hipGetLastError());
    long long changed = 0;
    CHK_TRY(hipMemcpyAsync(&changed, ctr + kChanged, sizeof changed, hipMemcpyDeviceToHost, st));
    CHK_TRY(hipStreamSynchronize(st));
    std::swap(cnt, cnt2);
    std::swap(ptr, ptr2);
    converged = changed == 0;
  }
  if (!converged) {
    r.valid = 0;
    r.cycle = 1;
    *res = r;
    return PBBS_OK;
  }
  hipLaunchKernelGGL(level_final_kernel, dim3(grid), dim3(256), 0, st, dp, cnt, ptr, (long long)n,
                     level);
  CHK_TRY(hipGetLastError());
  if (pbbs::graph_off64(g))
    hipLaunchKernelGGL(check_rows_kernel<int64_t>, dim3(grid), dim3(256), 0, st,
                       static_cast<const int64_t*>(pbbs::graph_dev_off(g)), pbbs::graph_dev_adj(g),
                       (long long)n, source, dp, dd, level, ctr);
  else
    hipLaunchKernelGGL(check_rows_kernel<uint32_t>, dim3(grid), dim3(256), 0, st,
                       static_cast<const uint32_t*>(pbbs::graph_dev_off(g)), pbbs::graph_dev_adj(g),
                       (long long)n, source, dp, dd, level, ctr);
  CHK_TRY(hipGetLastError());
  CHK_TRY(hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, st));
  CHK_TRY(hipStreamSynchronize(st));
  r.valid = h[kBad] == 0;
  r.bad = h[kBad];
  r.first_bad = h[kBad] ? h[kFirstBad] : -1;
  r.max_rule_bad = pbbs::graph_symmetric(g) ? h[kMaxBad] : -1;
  r.depth_bad = h[kDepthBad];
  r.reached = h[kReached];
  r.max_level = (int32_t)h[kMaxLevel];
  *res = r;
  return PBBS_OK;
}

extern "C" int pbbs_check_mis(pbbs_ctx* c, const pbbs_graph* g, const int32_t* flags,
                              uint32_t check_flags, pbbs_mis_check_result* res) {
  pbbs::clear_error();
  if (!c || !g || !flags || !res) return fail(PBBS_ERR_INVALID, "NULL argument");
  if (check_flags & ~(uint32_t)PBBS_CHECK_DEVICE_INPUT)
    return fail(PBBS_ERR_INVALID, "unknown check flags 0x%x", check_flags);
  if (pbbs::graph_device(g) != pbbs::ctx_device(c))
    return fail(PBBS_ERR_INVALID, "graph lives on another device");
  const int64_t n = pbbs::graph_n(g);
  CHK_TRY(hipSetDevice(pbbs::ctx_device(c)));
  hipStream_t st = pbbs::ctx_stream(c);
  const int grid = std::max(1, pbbs::ctx_num_cus(c) * 8);
  DevBufs bufs;
  const int32_t* df = flags;
  if (!(check_flags & PBBS_CHECK_DEVICE_INPUT)) {
    int32_t* f2 = nullptr;
    CHK_TRY(bufs.alloc(&f2, n * sizeof(int32_t)));
    CHK_TRY(hipMemcpyAsync(f2, flags, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    df = f2;
  }
  long long* ctr = nullptr;
  CHK_TRY(bufs.alloc(&ctr, 3 * sizeof(long long)));
  const long long init[3] = {0, 0, LLONG_MAX};
  CHK_TRY(hipMemcpyAsync(ctr, init, sizeof init, hipMemcpyHostToDevice, st));
  if (pbbs::graph_off64(g))
    hipLaunchKernelGGL(check_mis_kernel<int64_t>, dim3(grid), dim3(256), 0, st,
                       static_cast<const int64_t*>(pbbs::graph_dev_off(g)), pbbs::graph_dev_adj(g),
                       (long long)n, df, ctr);
  else
    hipLaunchKernelGGL(check_mis_kernel<uint32_t>, dim3(grid), dim3(256), 0, st,
                       static_cast<const uint32_t*>(pbbs::graph_dev_off(g)), pbbs::graph_dev_adj(g),
                       (long long)n, df, ctr);
  CHK_TRY(hipGetLastError());
  long long h[3];
  CHK_TRY(hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, st));
  CHK_TRY(hipStreamSynchronize(st));
  pbbs_mis_check_result r{};
  r.bad_edges = h[0];
  r.bad_vertices = h[1];
  r.valid = h[0] == 0 && h[1] == 0;
  r.first_bad = r.valid ? -1 : h[2];
  *res = r;
  return PBBS_OK;
}
