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

#include <climits>

#include "bfs_device.hpp"

namespace pbbs {

namespace {

__device__ __forceinline__ unsigned lanes_below_mask(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// One look at vertex v from row position p: up to kStep entries are read
// together (their flag loads overlap; 4 on rows averaging 16 entries or more,
// 2 on sparser graphs, where a wider look mostly reads flags it does not need).  Returns 1 (in: every lower neighbour
// is out), 2 (out: a lower neighbour is in), 0 (wait at *p: an undecided lower
// neighbour, or more of the row to read; *p moved past settled entries), so
// that the caller can look again or defer v.
template <int kStep, typename OffT>
__device__ __forceinline__ int mis_look(const MisArgs<OffT>& a, int32_t v, OffT& p, OffT hi,
                                        bool& blocked) {
  int32_t u[kStep];
  uint8_t f[kStep];
#pragma unroll
  for (int j = 0; j < kStep; ++j) u[j] = (OffT)j < hi - p ? a.adj[p + j] : INT32_MAX;
#pragma unroll
  for (int j = 0; j < kStep; ++j) f[j] = u[j] < v ? a.flags[u[j]] : (uint8_t)2;
  bool any_in = false;
#pragma unroll
  for (int j = 0; j < kStep; ++j) any_in |= f[j] == 1;
  if (any_in) return 2;
  int first_open = kStep;
#pragma unroll
  for (int j = kStep - 1; j >= 0; --j)
    if (f[j] == 0) first_open = j;
  if (first_open < kStep) {
    p += (OffT)first_open;
    blocked = true;
    return 0;
  }
  // all kStep entries settled (out, not lower, or past the row end)
  if (hi - p <= (OffT)kStep) return 1;
  if (a.sorted && u[kStep - 1] >= v) return 1;   // the rest of the row is not lower
  p += (OffT)kStep;
  blocked = false;
  return 0;
}

static_assert(kMisStripes == kWave, "one region count per lane of the scan");

// Region s of the round's list holds entries [pre[s], pre[s + 1]) of its
// index space; `s` is a hint no larger than the answer (indices a lane takes
// only grow within a chunk).
__device__ __forceinline__ int region_from(const unsigned long long* pre, unsigned long long i,
                                           int s) {
  while (pre[s + 1] <= i) ++s;
  return s;
}

__device__ __forceinline__ int region_of(const unsigned long long* pre, unsigned long long i) {
  int lo = 0, hi = kMisStripes;   // pre[lo] <= i < pre[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (pre[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Round r reads the count set r % 3 and list[r & 1] (round 0: every vertex in
// id order, without a list), appends to list[(r + 1) & 1] and count set
// (r + 1) % 3; workgroup 0 clears set (r + 2) % 3, which held round r - 1's
// counts.
//
// Each wavefront takes chunks of consecutive list entries.  A lane holds one
// vertex and looks at it (mis_look) until it decides or meets an undecided
// lower neighbour; then it takes the chunk's next vertex, so a long row does
// not hold up the other 63 lanes.  Deferred vertices collect in LDS and go to
// the next list with one atomicAdd per chunk, on the count of the region that
// owns the chunk: chunk k of nchunks goes to region k * kMisStripes / nchunks,
// so the next list keeps the ids roughly ascending (lower ids first lets a
// round see more decisions made earlier in the same round), and a region
// receives at most ceil(nchunks / kMisStripes) chunks: at most
// count / kMisStripes + chunk entries, within `region`.
template <typename OffT, int kStep>
__global__ __launch_bounds__(kBlock) void mis_round_kernel(MisArgs<OffT> a) {
  __shared__ int32_t defer[kWavesPerBlock][kMisMaxChunk];
  __shared__ unsigned long long pre[kMisStripes + 1];
  const int r = (int)a.wg_round[blockIdx.x];
  const unsigned lane = __lane_id();
  unsigned long long count = (unsigned long long)a.n;
  if (r > 0) {
    if (threadIdx.x < kWave) {   // kMisStripes == kWave: one count per lane, scanned
      const unsigned long long* cin = a.ctr + (size_t)(r % 3) * kMisStripes * 16;
      unsigned long long c = cin[threadIdx.x * 16];
      // a count past the region (never expected; the host reports it) is cut
      // so that every read stays inside the list
      if (c > (unsigned long long)a.region) c = (unsigned long long)a.region;
      for (int d = 1; d < kWave; d <<= 1) {
        const unsigned long long t = __shfl_up(c, d);
        if (lane >= (unsigned)d) c += t;
      }
      pre[threadIdx.x + 1] = c;
      if (threadIdx.x == 0) pre[0] = 0;
    }
    __syncthreads();
    count = pre[kMisStripes];
  }
  if (count == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && !*a.finished) {
      *a.finished = 1;
      __threadfence_system();
      __hip_atomic_store(&a.host_status[1], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&a.host_status[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x < kMisStripes) {
    a.ctr[((size_t)((r + 2) % 3) * kMisStripes + threadIdx.x) * 16] = 0;
    if (threadIdx.x == 0)
      __hip_atomic_store(&a.host_status[1], r + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  // selects, not an index into the argument block (which would copy it to scratch)
  const int32_t* in = (r & 1) ? a.list[1] : a.list[0];
  int32_t* out = (r & 1) ? a.list[0] : a.list[1];
  unsigned long long* cout = a.ctr + (size_t)((r + 1) % 3) * kMisStripes * 16;
  const unsigned wib = threadIdx.x / kWave;
  int32_t* dbuf = defer[wib];
  const unsigned long long nwaves = (unsigned long long)gridDim.x * kWavesPerBlock;
  const unsigned long long wave = (unsigned long long)blockIdx.x * kWavesPerBlock + wib;
  // chunk: the list spread over every wavefront, in multiples of 64, 64..kMisMaxChunk
  unsigned long long chunk = (count + nwaves - 1) / nwaves;
  chunk = ((chunk + kWave - 1) / kWave) * kWave;
  chunk = chunk < kWave ? kWave : (chunk > kMisMaxChunk ? kMisMaxChunk : chunk);
  const unsigned long long nchunks = (count + chunk - 1) / chunk;
  for (unsigned long long k = wave; k < nchunks; k += nwaves) {
    const unsigned long long c0 = k * chunk;
    const unsigned long long c1 = c0 + chunk < count ? c0 + chunk : count;
    unsigned long long next = c0 + kWave;   // the chunk's first entry no lane holds yet
    unsigned ndefer = 0;                    // wave-uniform
    int s = r == 0 ? 0 : region_of(pre, c0);
    bool have = c0 + lane < c1;
    int32_t v = 0;
    OffT p = 0, lo = 0, hi = 0;
    if (have) {
      const unsigned long long i = c0 + lane;
      if (r == 0) {
        v = (int32_t)i;
      } else {
        s = region_from(pre, i, s);
        v = in[(unsigned long long)s * a.region + (i - pre[s])];
      }
      lo = a.off[v];
      hi = a.off[v + 1];
      p = r == 0 ? lo : lo + (OffT)a.pos[v];
    }
    while (__ballot(have)) {
      int res = -1;   // -1 looking on, 0 deferred, 1 / 2 decided
      if (have) {
        bool blocked = false;
        const int d = mis_look<kStep>(a, v, p, hi, blocked);
        if (d != 0) res = d;
        else if (blocked) res = 0;
      }
      if (res > 0) a.flags[v] = (uint8_t)res;
      const unsigned long long dm = __ballot(res == 0);
      if (res == 0) {
        a.pos[v] = (uint32_t)(p - lo);
        dbuf[ndefer + lanes_below_mask(dm)] = v;
      }
      ndefer += (unsigned)__popcll(dm);
      // lanes that finished take the chunk's next entries
      const bool take = res >= 0 || !have;
      const unsigned long long free_m = __ballot(take);
      if (take) {
        const unsigned long long i = next + lanes_below_mask(free_m);
        have = i < c1;
        if (have) {
          if (r == 0) {
            v = (int32_t)i;
          } else {
            s = region_from(pre, i, s);
            v = in[(unsigned long long)s * a.region + (i - pre[s])];
          }
          lo = a.off[v];
          hi = a.off[v + 1];
          p = r == 0 ? lo : lo + (OffT)a.pos[v];
        }
      }
      next += (unsigned long long)__popcll(free_m);
    }
    // Look again at the chunk's deferred vertices (in LDS; two passes on
    // graphs read 4 entries a look, one on sparser ones, by measurement):
    // by now their blockers have had a chunk's time to decide, and their rows
    // and flags were just read, so a second look mostly hits in cache where
    // a look in the next round would not.  Vertices still blocked are
    // compacted to the front of the buffer (writes land below the read
    // cursor: a pass never writes more entries than it has taken).
    constexpr int kRetry = kStep == 4 ? 2 : 1;
    for (int pass = 0; pass < kRetry && ndefer; ++pass) {
      __builtin_amdgcn_wave_barrier();
      const unsigned total = ndefer;
      unsigned nd2 = 0, nxt = kWave;
      bool hv = lane < total;
      int32_t w = 0;
      OffT q = 0, wlo = 0, whi = 0;
      if (hv) {
        w = dbuf[lane];
        wlo = a.off[w];
        whi = a.off[w + 1];
        q = wlo + (OffT)a.pos[w];
      }
      while (__ballot(hv)) {
        int res = -1;
        if (hv) {
          bool blocked = false;
          const int d = mis_look<kStep>(a, w, q, whi, blocked);
          if (d != 0) res = d;
          else if (blocked) res = 0;
        }
        if (res > 0) a.flags[w] = (uint8_t)res;
        const unsigned long long dm = __ballot(res == 0);
        if (res == 0) {
          a.pos[w] = (uint32_t)(q - wlo);
          dbuf[nd2 + lanes_below_mask(dm)] = w;
        }
        nd2 += (unsigned)__popcll(dm);
        __builtin_amdgcn_wave_barrier();
        const bool take = res >= 0 || !hv;
        const unsigned long long fm = __ballot(take);
        if (take) {
          const unsigned i = nxt + lanes_below_mask(fm);
          hv = i < total;
          if (hv) {
            w = dbuf[i];
            wlo = a.off[w];
            whi = a.off[w + 1];
            q = wlo + (OffT)a.pos[w];
          }
        }
        nxt += (unsigned)__popcll(fm);
      }
      ndefer = nd2;
    }
    __builtin_amdgcn_wave_barrier();
    if (ndefer) {
      const unsigned long long t = k * kMisStripes / nchunks;
      unsigned long long at = 0;
      if (lane == 0) at = atomicAdd(&cout[t * 16], (unsigned long long)ndefer);
      at = __shfl(at, 0);
      if (at + ndefer <= (unsigned long long)a.region) {
        int32_t* dst = out + t * (unsigned long long)a.region + at;
        for (unsigned i = lane; i < ndefer; i += kWave) dst[i] = dbuf[i];
      } else if (lane == 0) {   // cannot happen (see above); reported, never written
        __hip_atomic_store(&a.host_status[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
  if (threadIdx.x == 0) a.wg_round[blockIdx.x] = (uint32_t)(r + 1);
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void mis_init_kernel(MisArgs<OffT> a) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  for (long long v = tid; v < a.n; v += nth) a.flags[v] = 0;
  for (long long i = tid; i < kMaxGrid; i += nth) a.wg_round[i] = 0;
  for (long long i = tid; i < 3 * kMisStripes; i += nth) a.ctr[i * 16] = 0;
  if (tid == 0) {
    *a.finished = 0;
    __threadfence_system();
    __hip_atomic_store(&a.host_status[2], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&a.host_status[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&a.host_status[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Output in the reference's format (MISTime.C:38-43: one int per vertex, 1 in
// the set, 2 not), four vertices per thread, and each workgroup's count of
// set members in part[blockIdx.x].
template <typename OffT>
__global__ __launch_bounds__(kBlock) void mis_output_kernel(MisArgs<OffT> a, int32_t* flags_out) {
  __shared__ unsigned wsum[kWavesPerBlock];
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  const long long n4 = a.n / 4;
  const uchar4* f4 = reinterpret_cast<const uchar4*>(a.flags);
  int4* o4 = reinterpret_cast<int4*>(flags_out);
  const bool vec = (reinterpret_cast<uintptr_t>(flags_out) & 15) == 0;
  unsigned c = 0;
  if (vec) {
    for (long long i = tid; i < n4; i += nth) {
      const uchar4 f = f4[i];
      o4[i] = make_int4(f.x, f.y, f.z, f.w);
      c += (f.x == 1) + (f.y == 1) + (f.z == 1) + (f.w == 1);
    }
  }
  for (long long v = (vec ? 4 * n4 : 0) + tid; v < a.n; v += nth) {
    const uint8_t f = a.flags[v];
    flags_out[v] = f;
    c += f == 1;
  }
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d);
  if (__lane_id() == 0) wsum[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) t += wsum[w];
    a.part[blockIdx.x] = t;
  }
}

// The set's size: the sum of the output workgroups' counts (one workgroup).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void mis_count_kernel(MisArgs<OffT> a, int parts,
                                                           unsigned long long* members) {
  __shared__ unsigned long long wsum[kWavesPerBlock];
  unsigned long long c = 0;
  for (int i = threadIdx.x; i < parts; i += kBlock) c += a.part[i];
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d);
  if (__lane_id() == 0) wsum[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) t += wsum[w];
    *members = t;
  }
}

}  // namespace

template <typename OffT>
hipError_t launch_mis_init(const MisArgs<OffT>& a, int grid, hipStream_t stream) {
  hipLaunchKernelGGL((mis_init_kernel<OffT>), dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

// step: row entries per look, 2 or 4 (the caller picks it from the graph's
// average degree).
template <typename OffT>
hipError_t launch_mis_round(const MisArgs<OffT>& a, int step, int grid, hipStream_t stream) {
  if (grid < 1 || grid > kMaxGrid || (step != 2 && step != 4)) return hipErrorInvalidValue;
  if (step == 4)
    hipLaunchKernelGGL((mis_round_kernel<OffT, 4>), dim3(grid), dim3(kBlock), 0, stream, a);
  else
    hipLaunchKernelGGL((mis_round_kernel<OffT, 2>), dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_mis_output(const MisArgs<OffT>& a, int32_t* flags_out,
                             unsigned long long* members, int grid, hipStream_t stream) {
  if (grid < 1 || grid > kMaxGrid) return hipErrorInvalidValue;
  hipLaunchKernelGGL((mis_output_kernel<OffT>), dim3(grid), dim3(kBlock), 0, stream, a,
                     flags_out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((mis_count_kernel<OffT>), dim3(1), dim3(kBlock), 0, stream, a, grid,
                     members);
  return hipGetLastError();
}

#define PBBS_MIS_INSTANTIATE(OffT)                                                             \
  template hipError_t launch_mis_init<OffT>(const MisArgs<OffT>&, int, hipStream_t);           \
  template hipError_t launch_mis_round<OffT>(const MisArgs<OffT>&, int, int, hipStream_t);     \
  template hipError_t launch_mis_output<OffT>(const MisArgs<OffT>&, int32_t*,                  \
                                              unsigned long long*, int, hipStream_t);
PBBS_MIS_INSTANTIATE(uint32_t)
PBBS_MIS_INSTANTIATE(int64_t)
#undef PBBS_MIS_INSTANTIATE

}  // namespace pbbs
