// Maximal independent set on the BFS's CSR: the set PBBS's incrementalMIS
// computes (benchmarks/maximalIndependentSet/incrementalMIS/MIS.C:35-66), the
// lexicographically first one.  Vertex v joins unless a lower-numbered
// neighbour joins; it is decided once every lower neighbour is.
//
// Rounds over a worklist of undecided vertices.  Each vertex keeps a resume
// position in its row: entries before it are neighbours already known to be
// out of the set or not lower than v.  A round walks each listed vertex's row
// from there: a lower neighbour in the set makes v out (2); an undecided one
// stops the walk (v waits for the next round); reaching the end of the lower
// neighbours (the first entry >= v on sorted rows, the row end otherwise)
// with all of them out puts v in (1).  Flags only go from 0 to a final value,
// so a round may read a flag another workgroup has just written, or an older
// copy of it: either way v decides only on final values, and the result does
// not depend on timing.  Undecided vertices are appended to the next round's
// list; a round that finds the list empty ends the run.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bfs_device.hpp"

namespace pbbs {

namespace {

__device__ __forceinline__ unsigned lanes_below_mask(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Round r reads count ctr[r % 3] and list[r & 1] (round 0: every vertex, in
// id order, without a list), appends to list[(r + 1) & 1] and ctr[(r + 1) % 3];
// workgroup 0 clears ctr[(r + 2) % 3], which held round r - 1's count.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void mis_round_kernel(MisArgs<OffT> a) {
  const int r = (int)a.wg_round[blockIdx.x];
  const unsigned long long count = a.ctr[r % 3];
  if (count == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && !*a.finished) {
      *a.finished = 1;
      __threadfence_system();
      __hip_atomic_store(&a.host_status[1], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&a.host_status[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.ctr[(r + 2) % 3] = 0;
    __hip_atomic_store(&a.host_status[1], r + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  const int32_t* in = a.list[r & 1];
  int32_t* out = a.list[(r + 1) & 1];
  unsigned long long* out_ctr = &a.ctr[(r + 1) % 3];
  const unsigned lane = __lane_id();
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  // every lane of a wavefront runs the same number of iterations, so the
  // ballot below sees the whole wavefront
  const unsigned long long base0 = (unsigned long long)blockIdx.x * blockDim.x + (threadIdx.x - lane);
  for (unsigned long long b = base0; b < count; b += stride) {
    const unsigned long long i = b + lane;
    bool wait = false;
    int32_t v = 0;
    if (i < count) {
      v = r == 0 ? (int32_t)i : in[i];
      const OffT lo = a.off[v], hi = a.off[v + 1];
      OffT p = lo + (OffT)a.pos[v];
      uint8_t f = 1;
      for (; p < hi; ++p) {
        const int32_t u = a.adj[p];
        if (u >= v) {
          if (a.sorted) break;   // the rest of the row is not lower
          continue;
        }
        const uint8_t fu = a.flags[u];
        if (fu == 1) {
          f = 2;
          break;
        }
        if (fu == 0) {
          f = 0;
          break;
        }
      }
      if (f == 0) {
        a.pos[v] = (uint32_t)(p - lo);
        wait = true;
      } else {
        a.flags[v] = f;
      }
    }
    const unsigned long long wm = __ballot(wait);
    if (wm) {
      unsigned long long at = 0;
      if (lane == 0) at = atomicAdd(out_ctr, (unsigned long long)__popcll(wm));
      at = __shfl(at, 0);
      if (wait) out[at + lanes_below_mask(wm)] = v;
    }
  }
  if (threadIdx.x == 0) a.wg_round[blockIdx.x] = (uint32_t)(r + 1);
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void mis_init_kernel(MisArgs<OffT> a) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  for (long long v = tid; v < a.n; v += nth) {
    a.flags[v] = 0;
    a.pos[v] = 0;
  }
  for (long long i = tid; i < kMaxGrid; i += nth) a.wg_round[i] = 0;
  if (tid == 0) {
    a.ctr[0] = (unsigned long long)a.n;
    a.ctr[1] = 0;
    a.ctr[2] = 0;
    a.ctr[3] = 0;   // the set's size, counted by the output kernel
    *a.finished = 0;
    __threadfence_system();
    __hip_atomic_store(&a.host_status[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&a.host_status[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Output in the reference's format (MISTime.C:38-43: one int per vertex, 1 in
// the set, 2 not) and the set's size.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void mis_output_kernel(MisArgs<OffT> a, int32_t* flags_out,
                                                            unsigned long long* members) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  unsigned c = 0;
  for (long long v = tid; v < a.n; v += nth) {
    const uint8_t f = a.flags[v];
    flags_out[v] = f;
    c += f == 1;
  }
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d);
  if (__lane_id() == 0 && c) atomicAdd(members, (unsigned long long)c);
}

}  // namespace

template <typename OffT>
hipError_t launch_mis_init(const MisArgs<OffT>& a, int grid, hipStream_t stream) {
  hipLaunchKernelGGL((mis_init_kernel<OffT>), dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_mis_round(const MisArgs<OffT>& a, int grid, hipStream_t stream) {
  if (grid < 1 || grid > kMaxGrid) return hipErrorInvalidValue;
  hipLaunchKernelGGL((mis_round_kernel<OffT>), dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_mis_output(const MisArgs<OffT>& a, int32_t* flags_out,
                             unsigned long long* members, int grid, hipStream_t stream) {
  hipLaunchKernelGGL((mis_output_kernel<OffT>), dim3(grid), dim3(kBlock), 0, stream, a,
                     flags_out, members);
  return hipGetLastError();
}

#define PBBS_MIS_INSTANTIATE(OffT)                                                             \
  template hipError_t launch_mis_init<OffT>(const MisArgs<OffT>&, int, hipStream_t);           \
  template hipError_t launch_mis_round<OffT>(const MisArgs<OffT>&, int, hipStream_t);          \
  template hipError_t launch_mis_output<OffT>(const MisArgs<OffT>&, int32_t*,                  \
                                              unsigned long long*, int, hipStream_t);
PBBS_MIS_INSTANTIATE(uint32_t)
PBBS_MIS_INSTANTIATE(int64_t)
#undef PBBS_MIS_INSTANTIATE

}  // namespace pbbs
