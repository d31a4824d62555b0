// Spanning forest of an edge array: the forest PBBS's incrementalST returns
// (benchmarks/spanningForest/incrementalST/ST.C:35-71), which is the forest
// its speculative_for round schedule (common/speculative_for.h:30-100, with
// granularity 100) produces.  That schedule decides the answer: edges that
// share a round reserve and commit against the union-find state of the
// rounds before, so a later edge of the round can join the forest where a
// serial pass would have taken an earlier one.  The GPU therefore runs the
// same schedule, round by round, with every round's work spread over the
// device:
//   reserve  each iterate i of the round (the failed ones of the round before
//            first, in order, then the next unprocessed edges) finds the
//            roots u < v of its ends and write-mins i into R[v];
//   commit   an iterate still holding R[v] links v under u (parents[v] = u;
//            roots are always their tree's smallest vertex, so links go from
//            larger to smaller ids and never close a cycle);
//   pack     the iterates that reserved and lost, in order, form the front of
//            the next round; one workgroup scans the per-workgroup counts and
//            applies the schedule's size rule (halve above 20% failures,
//            double below 10%).
// The forest is R's reserved entries in vertex order (ST.C:67-69).  The
// schedule's control words live in device memory, so rounds are launched
// without the host reading anything between them.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include "bfs_device.hpp"

namespace pbbs {

namespace {

__device__ __forceinline__ unsigned lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Root of i (algorithm/union_find.h:36-52): -1 marks a root; the walk points
// each visited vertex at its grandparent.  Finds run only in the reserve
// kernel, where no link happens, so every pointer it reads or writes is an
// ancestor and a stale read is still one.
__device__ __forceinline__ int32_t st_find(int32_t* parents, int32_t i) {
  int32_t p = parents[i];
  if (p < 0) return i;
  int32_t gp = parents[p];
  while (gp >= 0) {
    parents[i] = gp;
    i = p;
    p = gp;
    gp = parents[p];
  }
  return p;
}

// Contiguous share of [0, size) for workgroup b of g, in whole workgroup tiles.
__device__ __forceinline__ void st_share(long long size, long long& lo, long long& hi) {
  const long long tiles = (size + kBlock - 1) / kBlock;
  const long long per = (tiles + gridDim.x - 1) / gridDim.x * kBlock;
  lo = (long long)blockIdx.x * per;
  hi = lo + per < size ? lo + per : size;
}

__global__ __launch_bounds__(kBlock) void st_init_kernel(StArgs a) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  for (long long v = tid; v < a.n; v += nth) {
    a.parents[v] = -1;
    a.R[v] = UINT_MAX;
  }
  if (tid == 0) {
    StCtl& c = *a.ctl;
    c.done = 0;
    c.nkeep = 0;
    c.cur = a.max_round / 4;
    c.size = c.cur < a.m ? c.cur : a.m;
    c.scatter_size = 0;
    c.rounds = 0;
    c.count = 0;
    c.error = 0;
    c.finished = a.m == 0;
    __threadfence_system();
    __hip_atomic_store(&a.host_status[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&a.host_status[0], c.finished, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__global__ __launch_bounds__(kBlock) void st_reserve_kernel(StArgs a) {
  const StCtl& c = *a.ctl;
  if (c.finished) return;
  const long long size = c.size, nk = c.nkeep, done = c.done;
  const uint32_t* hold = (c.rounds & 1) ? a.hold[1] : a.hold[0];
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  for (long long i = tid; i < size; i += nth) {
    const uint32_t it = i < nk ? hold[i] : (uint32_t)(done + i);
    const int32_t eu = a.eu[it], ev = a.ev[it];
    a.I[i] = it;
    if ((uint32_t)eu >= (uint32_t)a.n || (uint32_t)ev >= (uint32_t)a.n) {
      // an end outside [0, n): the run reports it and its forest is not used;
      // the edge takes no part (it is never read as a vertex)
      a.keep[i] = 0;
      atomicOr(&a.ctl->error, 1);
      continue;
    }
    int32_t u = st_find(a.parents, eu);
    int32_t v = st_find(a.parents, ev);
    if (u > v) {
      const int32_t t = u;
      u = v;
      v = t;
    }
    a.su[i] = u;
    a.sv[i] = v;
    a.keep[i] = u != v;
    if (u != v) atomicMin(&a.R[v], it);
  }
}

// Commit, then count the iterates that reserved and lost, per workgroup.
__global__ __launch_bounds__(kBlock) void st_commit_kernel(StArgs a) {
  __shared__ unsigned wsum[kWavesPerBlock];
  const StCtl& c = *a.ctl;
  if (c.finished) return;
  long long lo, hi;
  st_share(c.size, lo, hi);
  unsigned cnt = 0;
  for (long long i = lo + threadIdx.x; i < hi; i += kBlock) {
    if (!a.keep[i]) continue;
    const int32_t v = a.sv[i];
    if (a.R[v] == a.I[i]) {
      a.parents[v] = a.su[i];
      a.keep[i] = 0;
    } else {
      ++cnt;
    }
  }
  for (int d = 32; d >= 1; d >>= 1) cnt += __shfl_xor(cnt, d);
  if (__lane_id() == 0) wsum[threadIdx.x / kWave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) t += wsum[w];
    a.blk[blockIdx.x] = t;
  }
}

// One workgroup: exclusive scan of the per-workgroup counts (in place), then
// the schedule's bookkeeping (speculative_for.h:78-95).
__global__ __launch_bounds__(1024) void st_schedule_kernel(StArgs a, int groups) {
  __shared__ unsigned long long part[1024 / kWave];
  StCtl& c = *a.ctl;
  if (c.finished) return;
  const int t = threadIdx.x;
  // groups <= 1024 * kStScanPer: each thread sums a run of kStScanPer counts
  unsigned long long mine = 0;
  for (int j = 0; j < kStScanPer; ++j) {
    const int g = t * kStScanPer + j;
    if (g < groups) mine += a.blk[g];
  }
  unsigned long long x = mine;   // inclusive scan over the workgroup
  const unsigned lane = __lane_id();
  for (int d = 1; d < kWave; d <<= 1) {
    const unsigned long long y = __shfl_up(x, d);
    if (lane >= (unsigned)d) x += y;
  }
  if (lane == kWave - 1) part[t / kWave] = x;
  __syncthreads();
  if (t < kWave) {
    unsigned long long y = t < 1024 / kWave ? part[t] : 0;
    for (int d = 1; d < kWave; d <<= 1) {
      const unsigned long long z = __shfl_up(y, d);
      if (lane >= (unsigned)d) y += z;
    }
    if (t < 1024 / kWave) part[t] = y;
  }
  __syncthreads();
  unsigned long long run = x - mine + (t >= kWave ? part[t / kWave - 1] : 0);
  for (int j = 0; j < kStScanPer; ++j) {
    const int g = t * kStScanPer + j;
    if (g < groups) {
      const unsigned long long cg = a.blk[g];
      a.blk[g] = (unsigned)run;
      run += cg;
    }
  }
  __syncthreads();
  if (t == 1023) {
    const long long size = c.size, nkeep = (long long)run;
    c.nkeep = nkeep;
    c.done += size - nkeep;
    const float share = (float)nkeep / (float)size;
    if (share > .2) {
      long long lo = a.max_round / 64 + 1;
      if (nkeep > lo) lo = nkeep;
      c.cur = c.cur / 2 > lo ? c.cur / 2 : lo;
    } else if (share < .1) {
      c.cur = c.cur * 2 < a.max_round ? c.cur * 2 : a.max_round;
    }
    c.rounds += 1;
    c.size = c.cur < a.m - c.done ? c.cur : a.m - c.done;
    int fin = c.done >= a.m;
    if (!fin && c.rounds > a.max_tries) {   // speculative_for gives up (size 0 never moves)
      c.error |= 2;
      fin = 1;
    }
    c.finished = fin;
    c.scatter_size = fin ? 0 : size;   // the last round leaves nothing to carry
    __threadfence_system();
    __hip_atomic_store(&a.host_status[1], (int32_t)c.rounds, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
    if (fin) __hip_atomic_store(&a.host_status[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The failed iterates of the round, in order, to the next round's front.
// Runs after the schedule kernel, so `rounds` already names the next round
// and `finished` may be set by this very round: it reads `scatter_size`.
__global__ __launch_bounds__(kBlock) void st_pack_kernel(StArgs a) {
  __shared__ unsigned wcnt[kWavesPerBlock];
  const StCtl& c = *a.ctl;
  if (c.scatter_size == 0) return;
  long long lo, hi;
  st_share(c.scatter_size, lo, hi);
  uint32_t* out = (c.rounds & 1) ? a.hold[1] : a.hold[0];
  unsigned base = a.blk[blockIdx.x];
  const unsigned lane = __lane_id(), w = threadIdx.x / kWave;
  for (long long t0 = lo; t0 < hi; t0 += kBlock) {
    const long long i = t0 + threadIdx.x;
    const bool k = i < hi && a.keep[i];
    const unsigned long long m = __ballot(k);
    if (lane == 0) wcnt[w] = (unsigned)__popcll(m);
    __syncthreads();
    unsigned before = 0, total = 0;
    for (int j = 0; j < kWavesPerBlock; ++j) {
      before += j < (int)w ? wcnt[j] : 0;
      total += wcnt[j];
    }
    if (k) out[base + before + lanes_below(m)] = a.I[i];
    base += total;
    __syncthreads();
  }
}

// The ndST contract (spanningForest/ndST/ST.C): any spanning forest.  One
// pass over the edges: an edge finds the roots u < v of its ends and, if it
// is the first to claim hooks[u] (compare-and-swap from free), links u under
// v; a lost claim or a stale root means finding again.  Links go from the
// smaller root to the larger, so no cycle forms; each root is hooked once.
// Finds here run beside links, so they read parents at agent scope (not a
// stale L1 copy).  A claim lost to an edge whose link is not yet visible is
// retried; kAnyTries bounds that wait, and running out marks the run as
// failed (error bit 4) rather than spinning on.
constexpr int kAnyTries = 1 << 16;

__device__ __forceinline__ int32_t ld_parent(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int32_t st_find_any(int32_t* parents, int32_t i) {
  int32_t p = ld_parent(&parents[i]);
  if (p < 0) return i;
  int32_t gp = ld_parent(&parents[p]);
  while (gp >= 0) {
    __hip_atomic_store(&parents[i], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    i = p;
    p = gp;
    gp = ld_parent(&parents[p]);
  }
  return p;
}

__global__ __launch_bounds__(kBlock) void st_any_kernel(StArgs a) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  uint32_t* hooks = a.R;
  for (long long i = tid; i < a.m; i += nth) {
    int32_t u = a.eu[i], v = a.ev[i];
    if ((uint32_t)u >= (uint32_t)a.n || (uint32_t)v >= (uint32_t)a.n) {
      atomicOr(&a.ctl->error, 1);
      continue;
    }
    // The retry loop runs while any lane of the wavefront still retries, and
    // a winning lane links inside it.  With a per-lane exit the compiler may
    // move the link store to the loop's exit, which the wavefront reaches
    // only when every lane is done: a lane waiting for that link would then
    // never see it.
    bool active = true;
    int t = 0;
    while (__ballot(active)) {
      if (active) {
        u = st_find_any(a.parents, u);
        v = st_find_any(a.parents, v);
        if (u == v) {
          active = false;
        } else {
          if (u > v) {
            const int32_t x = u;
            u = v;
            v = x;
          }
          if (__hip_atomic_load(&hooks[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                  UINT_MAX &&
              atomicCAS(&hooks[u], UINT_MAX, (uint32_t)i) == UINT_MAX) {
            __hip_atomic_store(&a.parents[u], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            active = false;
          } else if (++t == kAnyTries) {
            atomicOr(&a.ctl->error, 4);
            active = false;
          }
        }
      }
    }
  }
}

// filter_map of R (ST.C:67-69): per-workgroup counts of reserved entries,
// one workgroup's scan (st_schedule_kernel's scan with `out_scan`), then the
// ordered writes.
__global__ __launch_bounds__(kBlock) void st_out_count_kernel(StArgs a) {
  __shared__ unsigned wsum[kWavesPerBlock];
  long long lo, hi;
  st_share(a.n, lo, hi);
  unsigned cnt = 0;
  for (long long v = lo + threadIdx.x; v < hi; v += kBlock) cnt += a.R[v] != UINT_MAX;
  for (int d = 32; d >= 1; d >>= 1) cnt += __shfl_xor(cnt, d);
  if (__lane_id() == 0) wsum[threadIdx.x / kWave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) t += wsum[w];
    a.blk[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void st_out_scan_kernel(StArgs a, int groups) {
  __shared__ unsigned long long part[1024 / kWave];
  const int t = threadIdx.x;
  unsigned long long mine = 0;
  for (int j = 0; j < kStScanPer; ++j) {
    const int g = t * kStScanPer + j;
    if (g < groups) mine += a.blk[g];
  }
  unsigned long long x = mine;
  const unsigned lane = __lane_id();
  for (int d = 1; d < kWave; d <<= 1) {
    const unsigned long long y = __shfl_up(x, d);
    if (lane >= (unsigned)d) x += y;
  }
  if (lane == kWave - 1) part[t / kWave] = x;
  __syncthreads();
  if (t < kWave) {
    unsigned long long y = t < 1024 / kWave ? part[t] : 0;
    for (int d = 1; d < kWave; d <<= 1) {
      const unsigned long long z = __shfl_up(y, d);
      if (lane >= (unsigned)d) y += z;
    }
    if (t < 1024 / kWave) part[t] = y;
  }
  __syncthreads();
  unsigned long long run = x - mine + (t >= kWave ? part[t / kWave - 1] : 0);
  for (int j = 0; j < kStScanPer; ++j) {
    const int g = t * kStScanPer + j;
    if (g < groups) {
      const unsigned long long cg = a.blk[g];
      a.blk[g] = (unsigned)run;
      run += cg;
    }
  }
  if (t == 1023) a.ctl->count = (long long)run;
}

__global__ __launch_bounds__(kBlock) void st_out_write_kernel(StArgs a, int32_t* forest) {
  __shared__ unsigned wcnt[kWavesPerBlock];
  long long lo, hi;
  st_share(a.n, lo, hi);
  unsigned base = a.blk[blockIdx.x];
  const unsigned lane = __lane_id(), w = threadIdx.x / kWave;
  for (long long t0 = lo; t0 < hi; t0 += kBlock) {
    const long long v = t0 + threadIdx.x;
    const uint32_t r = v < hi ? a.R[v] : UINT_MAX;
    const bool k = r != UINT_MAX;
    const unsigned long long m = __ballot(k);
    if (lane == 0) wcnt[w] = (unsigned)__popcll(m);
    __syncthreads();
    unsigned before = 0, total = 0;
    for (int j = 0; j < kWavesPerBlock; ++j) {
      before += j < (int)w ? wcnt[j] : 0;
      total += wcnt[j];
    }
    if (k) forest[base + before + lanes_below(m)] = (int32_t)r;
    base += total;
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_st_init(const StArgs& a, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(st_init_kernel, dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

// One round of the schedule: four launches, each a no-op once the run is over.
hipError_t launch_st_round(const StArgs& a, int grid, hipStream_t stream) {
  if (grid < 1 || grid > kStMaxGroups) return hipErrorInvalidValue;
  hipLaunchKernelGGL(st_reserve_kernel, dim3(grid), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(st_commit_kernel, dim3(grid), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(st_schedule_kernel, dim3(1), dim3(1024), 0, stream, a, grid);
  hipLaunchKernelGGL(st_pack_kernel, dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_st_any(const StArgs& a, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(st_any_kernel, dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_st_output(const StArgs& a, int32_t* forest, int grid, hipStream_t stream) {
  if (grid < 1 || grid > kStMaxGroups) return hipErrorInvalidValue;
  hipLaunchKernelGGL(st_out_count_kernel, dim3(grid), dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(st_out_scan_kernel, dim3(1), dim3(1024), 0, stream, a, grid);
  hipLaunchKernelGGL(st_out_write_kernel, dim3(grid), dim3(kBlock), 0, stream, a, forest);
  return hipGetLastError();
}

}  // namespace pbbs
