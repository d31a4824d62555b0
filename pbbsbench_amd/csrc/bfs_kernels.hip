// gfx950 kernels of the deterministic breadth-first search.
//
// One BFS = one init launch + one level launch per BFS level.  A level
// kernel runs either top-down (sparse frontier queue) or bottom-up (dense
// frontier bitmap).  Every workgroup derives the level number from its own
// counter and the direction from the frontier counters, so no workgroup waits
// for another and the host never waits between levels.
//
// Output rule (deterministicBFS/BFS.C:59-74 with common/atomics.h:75-81):
// parent[v] = max{u in F_k : (u,v) in E} for v first reached at level k+1.
//   * top-down: every frontier edge whose target was not reached before this
//     level does atomicMax(&parent[v], u).  parent starts at -1, so the one
//     atomic that returns -1 is the first claim; that lane owns the append of
//     v to the next frontier.  The final value is the max whatever the order.
//     The "reached before this level" guard is depth[v] > k: depth only moves
//     n -> k+1 during level k and both values pass the guard, so a claimer
//     never hides v from a later claimer of the same level (the property the
//     reference gets from setting visited only in its winner pass, BFS.C:73).
//   * bottom-up: an unreached v scans its sorted row from the top and stops at
//     the first neighbour in F_k; that neighbour is the largest one in F_k.
//     Unsorted rows are scanned completely with a max.
//
// Same-address device atomics serialise (about 10 ns each), so appends to
// the next frontier are staged per wavefront in LDS and reserved once per
// workgroup, spread over kSubQueues counters.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <hipcub/hipcub.hpp>

#include "bfs_device.hpp"

namespace pbbs {

#ifdef PBBS_WAVE_TIMING
// Diagnostic build only: per-wavefront wall-clock stamps (start, end of the
// first pass, end) of the first 16 levels, for 16384 wavefronts.
constexpr int kTimingLevels = 16;
constexpr int kTimingWaves = 16384;
__device__ unsigned long long g_wave_times[kTimingLevels * kTimingWaves * 3];
__device__ __forceinline__ void stamp(int k, int slot) {
  const long long gw = (long long)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (k < kTimingLevels && gw < kTimingWaves && __lane_id() == 0)
    g_wave_times[((long long)k * kTimingWaves + gw) * 3 + slot] = wall_clock64();
}
#define PBBS_STAMP(k, slot) stamp((k), (slot))
#else
#define PBBS_STAMP(k, slot) ((void)0)
#endif

// ---------------------------------------------------------------- wave utils
__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ unsigned lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ unsigned wave_inclusive_sum(unsigned x) {
  const unsigned lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const unsigned y = __shfl_up(x, d);
    if (lane >= (unsigned)d) x += y;
  }
  return x;
}

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) x = max(x, __shfl_xor(x, d));
  return x;
}

// Orders this wavefront's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t shfl_off(uint32_t x, int src) { return __shfl(x, src); }
__device__ __forceinline__ int64_t shfl_off(int64_t x, int src) {
  return (int64_t)__shfl((long long)x, src);
}

__device__ __forceinline__ bool in_set(const unsigned* bits, int u) {
  return (bits[u >> 5] >> (u & 31)) & 1u;
}

__device__ __forceinline__ long long global_wave() {
  return (long long)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
}
__device__ __forceinline__ long long num_waves() {
  return (long long)gridDim.x * kWavesPerBlock;
}

// ------------------------------------------------------------- level view
// The pointers one level uses, resolved once from the rotating buffers so
// that no array of the argument block is indexed dynamically.
template <typename OffT>
struct LevelView {
  const OffT* __restrict__ off;
  const int32_t* __restrict__ adj;
  int32_t* parent;
  int32_t* depth;
  unsigned long long* visited;
  const unsigned* fb_cur;           // F_k
  unsigned long long* fb_next;      // F_{k+1}
  const int32_t* q_cur;             // sub-queue regions of F_k
  const uint32_t* qe_cur;
  int32_t* q_next;                  // this workgroup's sub-queue region of F_{k+1}
  uint32_t* qe_next;
  unsigned long long* ctr_next;     // its counter
  uint32_t* d0_next;
  const unsigned long long* ctr_cur;  // F_k's counters (all sub-queues)
  const uint32_t* bu_bounds;        // weighted wavefront shares, or null for equal shares
  long long n;
  long long nwords;
  long long qcap;
  int k;
  int serial_depth;
};

// --------------------------------------------------------- frontier staging
// Per-wavefront LDS buffer of (vertex, degree offset within the buffer).
struct StageBuf {
  int32_t v[kWavesPerBlock][kWaveBuf];
  uint32_t eo[kWavesPerBlock][kWaveBuf];
  uint32_t cnt[kWavesPerBlock];
  uint32_t deg[kWavesPerBlock];
  unsigned long long base;
};

struct Stage {
  unsigned cnt = 0;   // wave-uniform
  unsigned deg = 0;   // wave-uniform
};

// Writes buffer entries [0, cnt) of wavefront w to the next frontier at
// (qbase, ebase) and empties it.
template <typename View>
__device__ __forceinline__ void stage_copy_out(const View& a, StageBuf& sb, int w,
                                               unsigned cnt, unsigned qbase, unsigned ebase) {
  for (unsigned i = lane_id(); i < cnt; i += kWave) {
    a.q_next[qbase + i] = sb.v[w][i];
    a.qe_next[qbase + i] = ebase + sb.eo[w][i];
  }
}

template <typename View>
__device__ __forceinline__ void stage_flush_wave(const View& a, StageBuf& sb, Stage& s) {
  const int w = threadIdx.x / kWave;
  unsigned long long base = 0;
  if (lane_id() == 0)
    base = atomicAdd(a.ctr_next, ((unsigned long long)s.deg << 32) | s.cnt);
  base = __shfl(base, 0);
  wave_lds_sync();
  stage_copy_out(a, sb, w, s.cnt, (unsigned)base, (unsigned)(base >> 32));
  wave_lds_sync();
  s.cnt = 0;
  s.deg = 0;
}

// Stages the lanes with `take` (vertex v of degree deg) for F_{k+1}.
template <typename View>
__device__ __forceinline__ void stage_push(const View& a, StageBuf& sb, Stage& s, bool take, int v,
                                           uint32_t deg) {
  const unsigned long long mask = __ballot(take);
  if (mask == 0) return;
  const unsigned c = (unsigned)__popcll(mask);
  const unsigned incl = wave_inclusive_sum(take ? deg : 0u);
  const unsigned total = __shfl(incl, kWave - 1);
  if (s.cnt + c > (unsigned)kWaveBuf) stage_flush_wave(a, sb, s);
  if (take) {
    const int w = threadIdx.x / kWave;
    const unsigned pos = s.cnt + lanes_below(mask);
    sb.v[w][pos] = v;
    sb.eo[w][pos] = s.deg + incl - deg;
  }
  s.cnt += c;
  s.deg += total;
}

// End of level: one reservation for the whole workgroup, then copy-out.
template <typename View>
__device__ __forceinline__ void stage_flush_block(const View& a, StageBuf& sb, const Stage& s) {
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    sb.cnt[w] = s.cnt;
    sb.deg[w] = s.deg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tc = 0, td = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
      tc += sb.cnt[i];
      td += sb.deg[i];
    }
    sb.base = tc ? atomicAdd(a.ctr_next, ((unsigned long long)td << 32) | tc) : 0ull;
  }
  __syncthreads();
  unsigned qb = (unsigned)sb.base, eb = (unsigned)(sb.base >> 32);
  for (int i = 0; i < w; ++i) {
    qb += sb.cnt[i];
    eb += sb.deg[i];
  }
  stage_copy_out(a, sb, w, s.cnt, qb, eb);
}

// ------------------------------------------------------------------ top-down
// Edges [eb, ee) of one sub-queue (nf entries, E edges; E < 2^32 since m is):
// find the entry holding edge eb with a 64-ary search over qe, then walk
// windows of 64 entries; lane l of a batch takes edge base+l and finds its
// source by a register binary search over the window's qe values.
template <typename OffT>
__device__ void td_range(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                         const int32_t* __restrict__ q, const uint32_t* __restrict__ qe,
                         unsigned nf, uint32_t E, uint32_t eb, uint32_t ee) {
  const unsigned lane = lane_id();
  const int k = a.k;
  // Largest i with qe[i] <= eb (qe[0] == 0 and qe is strictly increasing).
  unsigned lo = 0, hi = nf;
  while (hi - lo > (unsigned)kWave) {
    const unsigned step = (hi - lo + kWave - 1) / kWave;
    const unsigned idx = lo + lane * step;
    const bool p = idx < hi && qe[idx] <= eb;
    const unsigned cnt = (unsigned)__popcll(__ballot(p));
    lo += (cnt - 1) * step;
    hi = min(hi, lo + step);
  }
  unsigned i;
  {
    const unsigned idx = lo + lane;
    const bool p = idx < hi && qe[idx] <= eb;
    i = lo + (unsigned)__popcll(__ballot(p)) - 1;
  }

  uint32_t e = eb;
  while (e < ee) {
    const unsigned idx = i + lane;
    int32_t qv = 0;
    uint32_t qo = 0xffffffffu;
    OffT rs = 0;
    if (idx < nf) {
      qv = q[idx];
      qo = qe[idx];
      rs = a.off[qv];
    }
    const uint32_t wend = (i + kWave < nf) ? qe[i + kWave] : E;
    const uint32_t lim = min(wend, ee);
    for (uint32_t base = e; base < lim; base += kWave) {
      const uint32_t my = base + lane;  // < lim + 64 <= 2^32 - 1 + 64: compare in 64 bits
      const bool valid = (unsigned long long)base + lane < (unsigned long long)lim;
      int j = 0;
#pragma unroll
      for (int s = kWave / 2; s >= 1; s >>= 1) {
        const uint32_t qt = __shfl(qo, j + s);
        if (qt <= my) j += s;
      }
      const int u = __shfl(qv, j);
      const OffT rsj = shfl_off(rs, j);
      const uint32_t qoj = __shfl(qo, j);
      bool fresh = false;
      int v = 0;
      if (valid) {
        v = a.adj[rsj + (OffT)(my - qoj)];
        if (a.depth[v] > k) {
          const int old = atomicMax(&a.parent[v], u);
          if (old == -1) {
            fresh = true;
            a.depth[v] = k + 1;
          }
        }
      }
      if (__ballot(fresh)) {
        uint32_t deg = 0;
        if (fresh) {
          deg = (uint32_t)(a.off[v + 1] - a.off[v]);
          atomicOr(reinterpret_cast<unsigned*>(a.fb_next) + (v >> 5), 1u << (v & 31));
        }
        const unsigned long long z = __ballot(fresh && deg == 0);
        if (z && lane == 0) atomicAdd(a.d0_next, (unsigned)__popcll(z));
        stage_push(a, sb, stg, fresh && deg > 0, v, deg);
      }
    }
    e = lim;
    i += kWave;
  }
}

// The frontier's edge space is the concatenation of the sub-queues; each
// wavefront takes an equal slice of it.
template <typename OffT>
__device__ void top_down(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                         unsigned long long E) {
  const unsigned long long W = (unsigned long long)num_waves();
  const unsigned long long gw = (unsigned long long)global_wave();
  const unsigned long long eb = E * gw / W, ee = E * (gw + 1) / W;
  if (eb >= ee) return;
  unsigned long long P = 0;
  for (int s = 0; s < kSubQueues; ++s) {
    const unsigned long long c = a.ctr_cur[s];
    const unsigned long long Es = c >> 32;
    const unsigned long long lo = eb > P ? eb : P;
    const unsigned long long hi = ee < P + Es ? ee : P + Es;
    if (lo < hi)
      td_range(a, sb, stg, a.q_cur + s * a.qcap, a.qe_cur + s * a.qcap, (unsigned)c,
               (uint32_t)Es, (uint32_t)(lo - P), (uint32_t)(hi - P));
    P += Es;
  }
}

// ----------------------------------------------------------------- bottom-up
// Each wavefront owns a contiguous run of frontier words (64 vertices each).
// It folds F_k into the visited bitmap word by word and gathers the ids of
// the still-unreached vertices into an LDS queue (ascending), so that every
// probe pass runs 64 unreached vertices at once whatever the density.  A
// pass tests the top four entries of each row with independent loads; rows
// ascend, so the first hit from the top is the largest neighbour in F_k.
// Vertices left unresolved are parked and finished 64 at a time: up to
// `serial_depth` entries per lane, then a whole-wavefront scan of the rest of
// the row, 64 entries per step from the top.  Found vertices only set bits of
// F_{k+1} (merged per word before one atomicOr); the visited bitmap learns
// them when level k+1 folds F_{k+1} in.

// Tests entries [p-N, p) of row [lo, p) from the top with N independent
// loads; returns the first hit (or -1) and moves p down past them.
template <int N, typename OffT>
__device__ __forceinline__ int scan_top(const LevelView<OffT>& a, OffT lo, OffT& p) {
  const unsigned* __restrict__ fbc = a.fb_cur;
  const OffT have = p - lo;
  int c[N];
#pragma unroll
  for (int j = 0; j < N; ++j) c[j] = (OffT)j < have ? a.adj[p - 1 - j] : -1;
  bool h[N];
#pragma unroll
  for (int j = 0; j < N; ++j) h[j] = c[j] >= 0 && in_set(fbc, c[j]);
  int found = -1;
#pragma unroll
  for (int j = N - 1; j >= 0; --j)
    if (h[j]) found = c[j];
  p = have > (OffT)N ? p - (OffT)N : lo;
  return found;
}

// Whole-wavefront scan of the rows of the lanes in `left`, from the top, 64
// entries per step; the first hit (lowest lane) is the highest entry.
template <typename OffT>
__device__ __forceinline__ void scan_rows_together(const LevelView<OffT>& a,
                                                   unsigned long long left, OffT lo, OffT p,
                                                   int& found) {
  const unsigned lane = lane_id();
  const unsigned* __restrict__ fbc = a.fb_cur;
  while (left) {
    const int L = __builtin_ctzll(left);
    left &= left - 1;
    const OffT loL = shfl_off(lo, L);
    OffT pL = shfl_off(p, L);
    int fu = -1;
    while (pL > loL) {
      const long long pos = (long long)pL - 1 - (long long)lane;
      bool hit = false;
      int u = -1;
      if (pos >= (long long)loL) {
        u = a.adj[pos];
        hit = in_set(fbc, u);
      }
      const unsigned long long hm = __ballot(hit);
      if (hm) {
        fu = __shfl(u, __builtin_ctzll(hm));
        break;
      }
      pL = (pL - loL > (OffT)kWave) ? pL - (OffT)kWave : loL;
    }
    if (lane == (unsigned)L) found = fu;
  }
}

// Full-row max scan for unsorted rows (no early exit).
template <typename OffT>
__device__ __forceinline__ int scan_unsorted(const LevelView<OffT>& a, bool mine, OffT lo, OffT hi) {
  const unsigned lane = lane_id();
  const unsigned* __restrict__ fbc = a.fb_cur;
  int found = -1;
  const bool serial = mine && (hi - lo) <= (OffT)kWave;
  if (serial) {
    for (OffT j = lo; j < hi; ++j) {
      const int u = a.adj[j];
      if (u > found && in_set(fbc, u)) found = u;
    }
  }
  unsigned long long left = __ballot(mine && !serial);
  while (left) {
    const int L = __builtin_ctzll(left);
    left &= left - 1;
    const OffT loL = shfl_off(lo, L);
    const OffT hiL = shfl_off(hi, L);
    int best = -1;
    for (OffT base = loL; base < hiL; base += (OffT)kWave) {
      const OffT pos = base + (OffT)lane;
      if (pos < hiL) {
        const int u = a.adj[pos];
        if (u > best && in_set(fbc, u)) best = u;
      }
    }
    best = wave_max(best);
    if (lane == (unsigned)L) found = best;
  }
  return found;
}

template <typename OffT>
struct BuLds {
  int32_t q[kWavesPerBlock][2 * kWave];      // unreached vertices, ascending
  int32_t lv[kWavesPerBlock][2 * kWave];     // parked vertices
  OffT lp[kWavesPerBlock][2 * kWave];        // their scan positions
};

// Records the found lanes: parent/depth, their F_{k+1} bits (one atomicOr per
// distinct word; lanes hold ascending ids so a word's lanes are adjacent),
// and the frontier append.
template <typename OffT>
__device__ __forceinline__ void bu_commit(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                                          bool valid, bool nw, int v, int found, uint32_t deg) {
  const unsigned lane = lane_id();
  if (nw) {
    a.parent[v] = found;
    a.depth[v] = a.k + 1;
  }
  const unsigned long long fm = __ballot(nw);
  if (fm) {
    const int wi = valid ? (v >> 6) : -1 - (int)lane;  // idle lanes join no word
    unsigned long long bits = nw ? (1ull << (v & 63)) : 0ull;
    // OR of `bits` over the lanes at or above this one with the same word
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const unsigned long long o = __shfl_down(bits, d);
      const int ow = __shfl_down(wi, d);
      if (lane + d < (unsigned)kWave && ow == wi) bits |= o;
    }
    const int pw = __shfl_up(wi, 1);
    const bool head = lane == 0 || pw != wi;
    if (head && bits) atomicOr(&a.fb_next[wi], bits);
  }
  stage_push(a, sb, stg, nw, v, deg);
}

template <typename OffT>
__device__ void bu_finish_parked(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                                 BuLds<OffT>& L, unsigned& npark) {
  const unsigned lane = lane_id();
  const int w = threadIdx.x / kWave;
  wave_lds_sync();
  for (unsigned base = 0; base < npark; base += kWave) {
    const bool mine = base + lane < npark;
    int v = 0;
    OffT lo = 0, hi = 0, p = 0;
    if (mine) {
      v = L.lv[w][base + lane];
      p = L.lp[w][base + lane];
      lo = a.off[v];
      hi = a.off[v + 1];
    }
    int found = -1;
    // parked rows are the deep ones: sixteen loads in flight per pass
    for (int t = 4; t < a.serial_depth; t += 16) {
      const bool active = mine && found < 0 && p > lo;
      if (!__ballot(active)) break;
      if (active) found = scan_top<16>(a, lo, p);
    }
    scan_rows_together(a, __ballot(mine && found < 0 && p > lo), lo, p, found);
    bu_commit(a, sb, stg, mine, mine && found >= 0, v, found, (uint32_t)(hi - lo));
  }
  wave_lds_sync();
  npark = 0;
}

// One probe pass over queue entries [0, cnt) (cnt <= 64).
template <typename OffT, bool kSorted>
__device__ void bu_batch(const LevelView<OffT>& a, StageBuf& sb, Stage& stg, BuLds<OffT>& L,
                         unsigned cnt, unsigned& npark) {
  const unsigned lane = lane_id();
  const int w = threadIdx.x / kWave;
  wave_lds_sync();
  const bool mine = lane < cnt;
  int v = 0;
  OffT lo = 0, hi = 0;
  if (mine) {
    v = L.q[w][lane];
    lo = a.off[v];
    hi = a.off[v + 1];
  }
  int found = -1;
  if (kSorted) {
    OffT p = hi;
    if (mine && p > lo) found = scan_top<4>(a, lo, p);
    const bool park = mine && found < 0 && p > lo;
    const unsigned long long pm = __ballot(park);
    bu_commit(a, sb, stg, mine, mine && found >= 0, v, found, (uint32_t)(hi - lo));
    if (pm) {
      if (npark + (unsigned)__popcll(pm) > 2 * kWave) bu_finish_parked(a, sb, stg, L, npark);
      if (park) {
        const unsigned slot = npark + lanes_below(pm);
        L.lv[w][slot] = v;
        L.lp[w][slot] = p;
      }
      npark += (unsigned)__popcll(pm);
    }
  } else {
    found = scan_unsorted(a, mine, lo, hi);
    bu_commit(a, sb, stg, mine, mine && found >= 0, v, found, (uint32_t)(hi - lo));
  }
}

template <typename OffT, bool kSorted>
__device__ void bottom_up(const LevelView<OffT>& a, StageBuf& sb, Stage& stg, BuLds<OffT>& L) {
  const unsigned lane = lane_id();
  const int w = threadIdx.x / kWave;
  const unsigned long long* __restrict__ fbc64 =
      reinterpret_cast<const unsigned long long*>(a.fb_cur);
  const long long W = num_waves();
  const long long gw = global_wave();
  const long long b0 = a.bu_bounds ? (long long)a.bu_bounds[gw] : a.nwords * gw / W;
  const long long b1 = a.bu_bounds ? (long long)a.bu_bounds[gw + 1] : a.nwords * (gw + 1) / W;
  unsigned nq = 0, npark = 0;
  for (long long g0 = b0; g0 < b1; g0 += kWave) {
    const long long mb = g0 + lane;
    unsigned long long unv = 0;
    if (mb < b1) {
      const unsigned long long old = a.visited[mb];
      const unsigned long long vis = old | fbc64[mb];  // fold F_k in
      if (vis != old) a.visited[mb] = vis;
      unv = ~vis;
    }
    unsigned long long todo = __ballot(unv != 0);
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1;
      const unsigned long long word = __shfl(unv, j);
      const bool take = (word >> lane) & 1ull;
      const unsigned c = (unsigned)__popcll(word);
      if (nq + c > 2 * kWave) {
        // keep the queue below two batches: run the full batch first
        bu_batch<OffT, kSorted>(a, sb, stg, L, kWave, npark);
        wave_lds_sync();
        const unsigned rest = nq - kWave;
        int tmp = lane < rest ? L.q[w][kWave + lane] : 0;
        wave_lds_sync();
        if (lane < rest) L.q[w][lane] = tmp;
        nq = rest;
      }
      if (take) L.q[w][nq + lanes_below(word)] = (int)((g0 + j) * kWave + lane);
      nq += c;
      if (nq >= (unsigned)kWave) {
        bu_batch<OffT, kSorted>(a, sb, stg, L, kWave, npark);
        wave_lds_sync();
        const unsigned rest = nq - kWave;
        int tmp = lane < rest ? L.q[w][kWave + lane] : 0;
        wave_lds_sync();
        if (lane < rest) L.q[w][lane] = tmp;
        nq = rest;
      }
    }
  }
  if (nq) bu_batch<OffT, kSorted>(a, sb, stg, L, nq, npark);
  PBBS_STAMP(a.k, 1);
  if (kSorted && npark) bu_finish_parked(a, sb, stg, L, npark);
}

// -------------------------------------------------------------- level driver
template <typename OffT>
__device__ __forceinline__ int level_mode(const BfsArgs<OffT>& A, int prev, unsigned long long cnt,
                                          unsigned long long E) {
  if (!A.symmetric || A.td_to_bu_div <= 0) return kTopDown;
  if (prev == kTopDown)
    return (cnt + E > (unsigned long long)A.m / (unsigned)A.td_to_bu_div) ? kBottomUp : kTopDown;
  return ((long long)cnt > A.n / A.bu_to_td_div) ? kBottomUp : kTopDown;
}

#ifndef PBBS_LEVEL_MIN_BLOCKS
#define PBBS_LEVEL_MIN_BLOCKS 1
#endif
template <typename OffT, bool kSorted>
__global__ __launch_bounds__(kBlock, PBBS_LEVEL_MIN_BLOCKS) void bfs_level_kernel(const BfsArgs<OffT>* __restrict__ A) {
  __shared__ StageBuf sb;
  __shared__ BuLds<OffT> bl;
  BfsState* st = A->st;
  const int k = (int)A->wg_level[blockIdx.x];
  const int c0 = k % 3, c1 = (k + 1) % 3, c2 = (k + 2) % 3;
  unsigned long long cnt = 0, E = 0;
#pragma unroll
  for (int s = 0; s < kSubQueues; ++s) {
    const unsigned long long c = st->ctr[c0][s];
    cnt += (uint32_t)c;
    E += c >> 32;
  }
  if (cnt == 0) {
    // F_k has no vertex to expand: the BFS is over.  The first kernel to see
    // it closes the books; later (overshoot) launches return at once.
    if (blockIdx.x == 0 && threadIdx.x == 0 && !st->finished) {
      const uint32_t z = st->d0[c0];
      st->reached += z;
      if (z) st->max_depth = k;
      st->finished = 1;
      __threadfence_system();
      __hip_atomic_store(&A->host_status[1], k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&A->host_status[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return;
  }
  PBBS_STAMP(k, 0);
  const int prev = k == 0 ? kTopDown : st->modes[(k - 1) & 3];
  const int mode = level_mode(*A, prev, cnt, E);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->modes[k & 3] = mode;
    st->reached += cnt + st->d0[c0];
    st->edges_reached += E;
    st->max_depth = k;
    if (k < kLevelLogMax) {
      st->log[k].mode = mode;
      st->log[k].count = (uint32_t)cnt;
      st->log[k].edges = E;
    }
    if (mode == kTopDown) st->td_levels += 1; else st->bu_levels += 1;
    // F_{k+2}'s counters held F_{k-1}, which no kernel reads any more.
#pragma unroll
    for (int s = 0; s < kSubQueues; ++s) st->ctr[c2][s] = 0;
    st->d0[c2] = 0;
    __hip_atomic_store(&A->host_status[1], k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  {
    // F_{k+2}'s bitmap held F_{k-1}.  A top-down level also folds F_k into
    // the visited bitmap, which only bottom-up levels read.
    unsigned long long* z = A->fb[c2];
    const unsigned long long* f = A->fb[c0];
    unsigned long long* vis = A->visited;
    const bool fold = mode == kTopDown && A->symmetric && A->td_to_bu_div > 0;
    const long long nwords = A->nwords;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += stride) {
      z[w] = 0;
      if (fold) {
        const unsigned long long fw = f[w];
        if (fw) vis[w] |= fw;
      }
    }
  }
  const int sq = blockIdx.x % kSubQueues;
  LevelView<OffT> a;
  a.off = A->off;
  a.adj = A->adj;
  a.parent = A->parent;
  a.depth = A->depth;
  a.visited = A->visited;
  a.fb_cur = reinterpret_cast<const unsigned*>(A->fb[c0]);
  a.fb_next = A->fb[c1];
  a.qcap = A->qcap;
  a.q_cur = A->q[k & 1];
  a.qe_cur = A->qe[k & 1];
  a.q_next = A->q[(k + 1) & 1] + sq * a.qcap;
  a.qe_next = A->qe[(k + 1) & 1] + sq * a.qcap;
  a.ctr_next = &st->ctr[c1][sq];
  a.d0_next = &st->d0[c1];
  a.ctr_cur = st->ctr[c0];
  // The first bottom-up level after top-down meets almost every vertex
  // unreached; its per-vertex cost grows with the degree, so the shares are
  // cut at equal sums of 1 + log2(1 + degree).  Later bottom-up levels meet
  // few, low-degree vertices and split the words evenly.
  a.bu_bounds = (prev == kTopDown) ? A->bu_bounds : nullptr;
  a.n = A->n;
  a.nwords = A->nwords;
  a.k = k;
  a.serial_depth = A->bu_serial_depth;
  Stage stg;
  if (mode == kTopDown)
    top_down(a, sb, stg, E);
  else
    bottom_up<OffT, kSorted>(a, sb, stg, bl);
  PBBS_STAMP(k, 2);
  stage_flush_block(a, sb, stg);
  if (threadIdx.x == 0) A->wg_level[blockIdx.x] = (uint32_t)(k + 1);
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void bfs_init_kernel(BfsArgs<OffT> a, BfsArgs<OffT>* dst,
                                                          const unsigned long long* __restrict__ iso,
                                                          int32_t s) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  const int32_t n = (int32_t)a.n;
  for (long long i = tid; i < a.n; i += nth) {
    a.parent[i] = (i == s) ? s : -1;
    a.depth[i] = (i == s) ? 0 : n;
  }
  const long long sw = s >> 6;
  const unsigned long long sb = 1ull << (s & 63);
  const unsigned long long pad = (a.n & 63) ? (~0ull << (a.n & 63)) : 0ull;
  for (long long w = tid; w < a.nwords; w += nth) {
    unsigned long long vv = iso[w];
    if (w == a.nwords - 1) vv |= pad;
    if (w == sw) vv |= sb;
    a.visited[w] = vv;
    a.fb[0][w] = (w == sw) ? sb : 0ull;
    a.fb[1][w] = 0ull;
    a.fb[2][w] = 0ull;
  }
  for (long long i = tid; i < kMaxGrid; i += nth) a.wg_level[i] = 0;
  if (tid == 0) {
    *dst = a;
    const unsigned long long deg = (unsigned long long)(a.off[s + 1] - a.off[s]);
    BfsState* st = a.st;
    a.q[0][0] = s;
    a.qe[0][0] = 0;
    for (int j = 0; j < 3; ++j) {
      for (int q = 0; q < kSubQueues; ++q) st->ctr[j][q] = 0;
      st->d0[j] = 0;
    }
    if (deg) st->ctr[0][0] = (deg << 32) | 1ull;
    else st->d0[0] = 1;
    for (int j = 0; j < 4; ++j) st->modes[j] = kTopDown;
    st->finished = 0;
    st->td_levels = 0;
    st->bu_levels = 0;
    st->max_depth = 0;
    st->reached = 0;
    st->edges_reached = 0;
    __threadfence_system();
    __hip_atomic_store(&a.host_status[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&a.host_status[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---------------------------------------------------------------- batched BFS
// Up to 64 sources share one traversal.  Bit i of a per-vertex 64-bit mask
// stands for source i; F_k of all sources is the mask array fr[k % 4] and the
// vertices with a non-zero mask form the union frontier queue.  Each source's
// outputs follow the single-source rule (largest neighbour one level up), so
// a batch returns exactly what separate BFS calls return.
//
//   * push level k: every edge (u,v) of the union frontier ORs
//     fr_k[u] & ~reached(v) into fr_{k+1}[v]; the lane whose atomicOr saw 0
//     appends v.  The parents of those bits are found in level k+1's kernel
//     by scanning v's sorted row from the top against fr_k (symmetric graph:
//     the row holds v's in-neighbours), clearing each bit at its first hit.
//   * pull level k: every vertex with sources still unreached scans its row
//     from the top against fr_k, assigning parents as bits hit, until all of
//     its unreached bits are found or the row ends.
// "Reached" bits of v during level k are vis[v] | fr_k[v]: vis gets level-k
// bits during level k (resolve), so the union keeps the guard exact.
template <typename OffT>
struct MbfsView {
  const OffT* __restrict__ off;
  const int32_t* __restrict__ adj;
  int32_t* parent;
  int32_t* depth;
  unsigned long long* vis;
  const unsigned long long* fr_prev;   // F_{k-1}
  const unsigned long long* fr_cur;    // F_k
  unsigned long long* fr_next;         // F_{k+1}
  const int32_t* q_cur;
  const uint32_t* qe_cur;
  const unsigned long long* ctr_cur;
  int32_t* q_next;                     // this workgroup's sub-queue region
  uint32_t* qe_next;
  unsigned long long* ctr_next;
  long long n;
  long long nwords;
  long long qcap;
  unsigned long long srcmask;
  int k;
};

template <typename OffT>
__device__ __forceinline__ void assign_sources(const MbfsView<OffT>& a, int v, int u,
                                               unsigned long long bits, int d) {
  while (bits) {
    const long long i = __builtin_ctzll(bits);
    bits &= bits - 1;
    a.parent[i * a.n + v] = u;
    a.depth[i * a.n + v] = d;
  }
}

// Parents of the bits F_k[v] for the queue entries of F_k (found by a push).
template <typename OffT>
__device__ void mbfs_resolve(const MbfsView<OffT>& a, unsigned long long total) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < (long long)total;
       t += stride) {
    // entry t of the concatenated sub-queues
    long long idx = t;
    int sq = 0;
    for (; sq < kSubQueues; ++sq) {
      const long long c = (uint32_t)a.ctr_cur[sq];
      if (idx < c) break;
      idx -= c;
    }
    const int v = a.q_cur[sq * a.qcap + idx];
    unsigned long long rem = a.fr_cur[v];
    const unsigned long long bits = rem;
    OffT lo = a.off[v], p = a.off[v + 1];
    while (rem && p > lo) {
      const OffT have = p - lo;
      int c[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) c[j] = (OffT)j < have ? a.adj[p - 1 - j] : -1;
      unsigned long long mk[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) mk[j] = c[j] >= 0 ? a.fr_prev[c[j]] : 0ull;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const unsigned long long hit = mk[j] & rem;
        if (hit) {
          assign_sources(a, v, c[j], hit, a.k);
          rem &= ~hit;
        }
      }
      p = have > 4 ? p - 4 : lo;
    }
    atomicOr(&a.vis[v], bits);
  }
}

template <typename OffT>
__device__ void mbfs_push(const MbfsView<OffT>& a, StageBuf& sb, Stage& stg,
                          unsigned long long E) {
  const unsigned lane = lane_id();
  const unsigned long long W = (unsigned long long)num_waves();
  const unsigned long long gw = (unsigned long long)global_wave();
  const unsigned long long ebeg = E * gw / W, eend = E * (gw + 1) / W;
  if (ebeg >= eend) return;
  unsigned long long P = 0;
  for (int sq = 0; sq < kSubQueues; ++sq) {
    const unsigned long long cc = a.ctr_cur[sq];
    const unsigned long long Es = cc >> 32;
    const unsigned nf = (unsigned)cc;
    const unsigned long long lo_e = ebeg > P ? ebeg : P;
    const unsigned long long hi_e = eend < P + Es ? eend : P + Es;
    P += Es;
    if (lo_e >= hi_e) continue;
    const int32_t* __restrict__ q = a.q_cur + sq * a.qcap;
    const uint32_t* __restrict__ qe = a.qe_cur + sq * a.qcap;
    const uint32_t eb = (uint32_t)(lo_e - (P - Es)), ee = (uint32_t)(hi_e - (P - Es));
    unsigned lo = 0, hi = nf;
    while (hi - lo > (unsigned)kWave) {
      const unsigned step = (hi - lo + kWave - 1) / kWave;
      const unsigned idx = lo + lane * step;
      const bool pr = idx < hi && qe[idx] <= eb;
      const unsigned cnt = (unsigned)__popcll(__ballot(pr));
      lo += (cnt - 1) * step;
      hi = min(hi, lo + step);
    }
    unsigned i;
    {
      const unsigned idx = lo + lane;
      const bool pr = idx < hi && qe[idx] <= eb;
      i = lo + (unsigned)__popcll(__ballot(pr)) - 1;
    }
    uint32_t e = eb;
    while (e < ee) {
      const unsigned idx = i + lane;
      int32_t qv = 0;
      uint32_t qo = 0xffffffffu;
      OffT rs = 0;
      unsigned long long qm = 0;
      if (idx < nf) {
        qv = q[idx];
        qo = qe[idx];
        rs = a.off[qv];
        qm = a.fr_cur[qv];
      }
      const uint32_t wend = (i + kWave < nf) ? qe[i + kWave] : (uint32_t)Es;
      const uint32_t lim = min(wend, ee);
      for (uint32_t base = e; base < lim; base += kWave) {
        const uint32_t my = base + lane;
        const bool valid = (unsigned long long)base + lane < (unsigned long long)lim;
        int j = 0;
#pragma unroll
        for (int st = kWave / 2; st >= 1; st >>= 1) {
          const uint32_t qt = __shfl(qo, j + st);
          if (qt <= my) j += st;
        }
        const OffT rsj = shfl_off(rs, j);
        const uint32_t qoj = __shfl(qo, j);
        const unsigned long long um = __shfl(qm, j);
        bool fresh = false;
        int v = 0;
        if (valid) {
          v = a.adj[rsj + (OffT)(my - qoj)];
          const unsigned long long nb = um & ~(a.vis[v] | a.fr_cur[v]) & a.srcmask;
          if (nb) fresh = atomicOr(&a.fr_next[v], nb) == 0ull;
        }
        if (__ballot(fresh)) {
          uint32_t deg = 0;
          if (fresh) deg = (uint32_t)(a.off[v + 1] - a.off[v]);
          stage_push(a, sb, stg, fresh, v, deg);
        }
      }
      e = lim;
      i += kWave;
    }
  }
}

template <typename OffT>
__device__ void mbfs_pull(const MbfsView<OffT>& a, StageBuf& sb, Stage& stg) {
  const unsigned lane = lane_id();
  const long long W = num_waves();
  const long long gw = global_wave();
  const long long b0 = a.nwords * gw / W, b1 = a.nwords * (gw + 1) / W;
  for (long long b = b0; b < b1; ++b) {
    const long long v = b * kWave + lane;
    unsigned long long rem = 0;
    if (v < a.n) rem = ~(a.vis[v] | a.fr_cur[v]) & a.srcmask;
    if (!__ballot(rem != 0)) continue;
    unsigned long long found = 0;
    OffT lo = 0, hi = 0;
    if (rem) {
      lo = a.off[v];
      hi = a.off[v + 1];
      OffT p = hi;
      while (rem && p > lo) {
        const OffT have = p - lo;
        int c[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) c[j] = (OffT)j < have ? a.adj[p - 1 - j] : -1;
        unsigned long long mk[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) mk[j] = c[j] >= 0 ? a.fr_cur[c[j]] : 0ull;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const unsigned long long hit = mk[j] & rem;
          if (hit) {
            assign_sources(a, (int)v, c[j], hit, a.k + 1);
            found |= hit;
            rem &= ~hit;
          }
        }
        p = have > 4 ? p - 4 : lo;
      }
      if (found) {
        a.fr_next[v] = found;
        atomicOr(&a.vis[v], found);
      }
    }
    stage_push(a, sb, stg, found != 0, (int)v, (uint32_t)(hi - lo));
  }
}

template <typename OffT>
__device__ __forceinline__ int mbfs_mode(const MbfsArgs<OffT>& A, int prev, unsigned long long cnt,
                                         unsigned long long E) {
  if (A.td_to_bu_div <= 0) return kTopDown;
  if (prev == kTopDown)
    return (cnt + E > (unsigned long long)A.m / (unsigned)A.td_to_bu_div) ? kBottomUp : kTopDown;
  return ((long long)cnt > A.n / A.bu_to_td_div) ? kBottomUp : kTopDown;
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void mbfs_level_kernel(const MbfsArgs<OffT>* __restrict__ A) {
  __shared__ StageBuf sb;
  MbfsState* st = A->st;
  const int k = (int)A->wg_level[blockIdx.x];
  const int c0 = k % 3, c1 = (k + 1) % 3, c2 = (k + 2) % 3;
  unsigned long long cnt = 0, E = 0;
#pragma unroll
  for (int s = 0; s < kSubQueues; ++s) {
    const unsigned long long c = st->ctr[c0][s];
    cnt += (uint32_t)c;
    E += c >> 32;
  }
  if (cnt == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && !st->finished) {
      st->finished = 1;
      __threadfence_system();
      __hip_atomic_store(&A->host_status[1], k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&A->host_status[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return;
  }
  const int prev = k == 0 ? kTopDown : st->modes[(k - 1) & 3];
  const int mode = mbfs_mode(*A, prev, cnt, E);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->modes[k & 3] = mode;
    st->max_depth = k;
    if (mode == kTopDown) st->push_levels += 1; else st->pull_levels += 1;
#pragma unroll
    for (int s = 0; s < kSubQueues; ++s) st->ctr[c2][s] = 0;
    __hip_atomic_store(&A->host_status[1], k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  {
    // F_{k+2}'s masks held F_{k-2}: level k reads F_{k-1} and F_k only.
    unsigned long long* z = A->fr[(k + 2) & 3];
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < A->n; v += stride)
      z[v] = 0;
  }
  const int sq = blockIdx.x % kSubQueues;
  MbfsView<OffT> a;
  a.off = A->off;
  a.adj = A->adj;
  a.parent = A->parent;
  a.depth = A->depth;
  a.vis = A->vis;
  a.fr_prev = A->fr[(k + 3) & 3];
  a.fr_cur = A->fr[k & 3];
  a.fr_next = A->fr[(k + 1) & 3];
  a.qcap = A->qcap;
  a.q_cur = A->q[k & 1];
  a.qe_cur = A->qe[k & 1];
  a.ctr_cur = st->ctr[c0];
  a.q_next = A->q[(k + 1) & 1] + sq * a.qcap;
  a.qe_next = A->qe[(k + 1) & 1] + sq * a.qcap;
  a.ctr_next = &st->ctr[c1][sq];
  a.n = A->n;
  a.nwords = A->nwords;
  a.srcmask = A->srcmask;
  a.k = k;
  if (k > 0 && prev == kTopDown) mbfs_resolve(a, cnt);
  Stage stg;
  if (mode == kTopDown)
    mbfs_push(a, sb, stg, E);
  else
    mbfs_pull(a, sb, stg);
  stage_flush_block(a, sb, stg);
  if (threadIdx.x == 0) A->wg_level[blockIdx.x] = (uint32_t)(k + 1);
}

// Clears the masks and presets every output to "unreached".
template <typename OffT>
__global__ __launch_bounds__(kBlock) void mbfs_clear_kernel(MbfsArgs<OffT> a) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  for (long long v = tid; v < a.n; v += nth) {
    a.vis[v] = 0;
    a.fr[0][v] = 0;
    a.fr[1][v] = 0;
    a.fr[2][v] = 0;
    a.fr[3][v] = 0;
  }
  const long long total = (long long)a.nsrc * a.n;
  for (long long t = tid; t < total; t += nth) {
    a.parent[t] = -1;
    a.depth[t] = (int32_t)a.n;
  }
  for (long long i = tid; i < kMaxGrid; i += nth) a.wg_level[i] = 0;
}

// One workgroup: marks the sources, queues the distinct ones as F_0 and
// resets the control block.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void mbfs_seed_kernel(MbfsArgs<OffT> a, MbfsArgs<OffT>* dst) {
  __shared__ int first[kMaxSources];
  const int t = threadIdx.x;
  if (t < a.nsrc) {
    const int s = a.sources[t];
    const unsigned long long bit = 1ull << t;
    a.parent[(long long)t * a.n + s] = s;
    a.depth[(long long)t * a.n + s] = 0;
    atomicOr(&a.vis[s], bit);
    first[t] = atomicOr(&a.fr[0][s], bit) == 0ull;
  }
  __syncthreads();
  if (t == 0) {
    *dst = a;
    MbfsState* st = a.st;
    unsigned cnt = 0;
    unsigned long long E = 0;
    for (int i = 0; i < a.nsrc; ++i) {
      if (!first[i]) continue;
      const int s = a.sources[i];
      a.q[0][cnt] = s;
      a.qe[0][cnt] = (uint32_t)E;
      E += (unsigned long long)(a.off[s + 1] - a.off[s]);
      ++cnt;
    }
    for (int j = 0; j < 3; ++j)
      for (int q = 0; q < kSubQueues; ++q) st->ctr[j][q] = 0;
    st->ctr[0][0] = (E << 32) | cnt;
    for (int j = 0; j < 4; ++j) st->modes[j] = kTopDown;
    st->finished = 0;
    st->push_levels = 0;
    st->pull_levels = 0;
    st->max_depth = 0;
    st->reached_pairs = 0;
    st->edges_pairs = 0;
    __threadfence_system();
    __hip_atomic_store(&a.host_status[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&a.host_status[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Totals over (source, vertex) pairs reached: count and degree sum.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void mbfs_stats_kernel(MbfsArgs<OffT> a) {
  __shared__ unsigned long long sp[kWavesPerBlock], se[kWavesPerBlock];
  unsigned long long pairs = 0, edges = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < a.n; v += stride) {
    const unsigned long long c = (unsigned long long)__popcll(a.vis[v] & a.srcmask);
    pairs += c;
    edges += c * (unsigned long long)(a.off[v + 1] - a.off[v]);
  }
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    pairs += __shfl_xor(pairs, d);
    edges += __shfl_xor(edges, d);
  }
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    sp[w] = pairs;
    se[w] = edges;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long P = 0, Ed = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
      P += sp[i];
      Ed += se[i];
    }
    atomicAdd(&a.st->reached_pairs, P);
    atomicAdd(&a.st->edges_pairs, Ed);
  }
}

// ---------------------------------------------------------- graph metadata
// counters: [0] isolated vertices, [1] max degree, [2] bad offsets (decreasing
// or beyond m), [3] row-start descents (adj[off[v]-1] > adj[off[v]] for a
// non-empty row v), [4] all descents (adj[j] > adj[j+1]), [5] out-of-range
// ids, [6] edges whose reverse is missing.
//
// Pass 1 reads offsets only, so malformed offsets are caught before any
// adjacency access depends on them.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void graph_offsets_kernel(
    const OffT* __restrict__ off, long long n, long long m, long long nwords,
    unsigned long long* __restrict__ iso, unsigned long long* __restrict__ counters) {
  const unsigned lane = lane_id();
  for (long long b = global_wave(); b < nwords; b += num_waves()) {
    const long long v = b * kWave + lane;
    bool isolated = false, bad = false;
    unsigned long long deg = 0;
    if (v < n) {
      const OffT o0 = off[v], o1 = off[v + 1];
      bad = o1 < o0 || (long long)o1 > m;
      if (!bad) {
        deg = (unsigned long long)(o1 - o0);
        isolated = deg == 0;
      }
    }
    const unsigned long long im = __ballot(isolated);
    const unsigned long long bm = __ballot(bad);
    unsigned long long md = deg;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const unsigned long long o = __shfl_xor(md, d);
      md = o > md ? o : md;
    }
    if (lane == 0) {
      iso[b] = im;
      if (im) atomicAdd(&counters[0], (unsigned long long)__popcll(im));
      if (md) atomicMax(&counters[1], md);
      if (bm) atomicAdd(&counters[2], (unsigned long long)__popcll(bm));
    }
  }
}

// Pass 2 (offsets validated: 0 = off[0] <= ... <= off[n] = m).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void graph_row_start_kernel(
    const OffT* __restrict__ off, const int32_t* __restrict__ adj, long long n,
    unsigned long long* __restrict__ counters) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v0 = (long long)blockIdx.x * blockDim.x; v0 < n; v0 += stride) {
    const long long v = v0 + threadIdx.x;
    bool descent = false;
    if (v < n) {
      const OffT o0 = off[v], o1 = off[v + 1];
      if (o1 > o0 && o0 > 0) descent = adj[o0 - 1] > adj[o0];
    }
    const unsigned long long dm = __ballot(descent);
    if (lane_id() == 0 && dm) atomicAdd(&counters[3], (unsigned long long)__popcll(dm));
  }
}

__global__ __launch_bounds__(kBlock) void graph_edge_meta_kernel(
    const int32_t* __restrict__ adj, long long m, long long n,
    unsigned long long* __restrict__ counters) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long j0 = (long long)blockIdx.x * blockDim.x; j0 < m; j0 += stride) {
    const long long j = j0 + threadIdx.x;
    bool descent = false, out_of_range = false;
    if (j < m) {
      const int32_t x = adj[j];
      out_of_range = x < 0 || (long long)x >= n;
      if (j + 1 < m) descent = x > adj[j + 1];
    }
    const unsigned long long dm = __ballot(descent);
    const unsigned long long om = __ballot(out_of_range);
    if (lane_id() == 0) {
      if (dm) atomicAdd(&counters[4], (unsigned long long)__popcll(dm));
      if (om) atomicAdd(&counters[5], (unsigned long long)__popcll(om));
    }
  }
}

// Needs validated offsets, in-range ids and sorted rows.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void graph_symmetry_kernel(
    const OffT* __restrict__ off, const int32_t* __restrict__ adj, long long n, long long m,
    unsigned long long* __restrict__ counters) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long j0 = (long long)blockIdx.x * blockDim.x; j0 < m; j0 += stride) {
    const long long j = j0 + threadIdx.x;
    bool missing = false;
    if (j < m) {
      // source row u: largest u with off[u] <= j
      long long lo = 0, hi = n;  // off[lo] <= j < off[hi]
      while (hi - lo > 1) {
        const long long mid = (lo + hi) >> 1;
        if ((long long)off[mid] <= j) lo = mid; else hi = mid;
      }
      const int32_t u = (int32_t)lo;
      const int32_t v = adj[j];
      long long a0 = (long long)off[v], a1 = (long long)off[v + 1];
      while (a0 < a1) {
        const long long mid = (a0 + a1) >> 1;
        if (adj[mid] < u) a0 = mid + 1; else a1 = mid;
      }
      missing = !(a0 < (long long)off[v + 1] && adj[a0] == u);
    }
    const unsigned long long mm = __ballot(missing);
    if (lane_id() == 0 && mm) atomicAdd(&counters[6], (unsigned long long)__popcll(mm));
  }
}

// ------------------------------------------------------ bottom-up shares
// Per frontier word: sum over its vertices of degree 0 ? 0 : 1 + floor(log2(1 + deg)).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void word_weight_kernel(const OffT* __restrict__ off,
                                                             long long n, long long nwords,
                                                             unsigned long long* __restrict__ ww) {
  const unsigned lane = lane_id();
  for (long long b = global_wave(); b < nwords; b += num_waves()) {
    const long long v = b * kWave + lane;
    unsigned x = 0;
    if (v < n) {
      const unsigned long long deg = (unsigned long long)(off[v + 1] - off[v]);
      if (deg) x = 1u + (63u - (unsigned)__builtin_clzll(deg + 1));
    }
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) x += __shfl_xor(x, d);
    if (lane == 0) ww[b] = x;
  }
}

// bounds[i] = first word j with prefix[j] >= total * i / waves (prefix has
// nwords + 1 entries, exclusive).
__global__ __launch_bounds__(kBlock) void share_bounds_kernel(const unsigned long long* __restrict__ prefix,
                                                              long long nwords, long long waves,
                                                              uint32_t* __restrict__ bounds) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > waves) return;
  const unsigned long long total = prefix[nwords];
  const unsigned long long target = (unsigned long long)((__uint128_t)total * i / waves);
  long long lo = 0, hi = nwords;  // smallest j in [0, nwords] with prefix[j] >= target
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (prefix[mid] >= target) hi = mid; else lo = mid + 1;
  }
  bounds[i] = (uint32_t)(i == waves ? nwords : lo);
}

// ------------------------------------------------------------ host launchers
template <typename OffT>
hipError_t launch_bfs_init(const BfsArgs<OffT>& a, BfsArgs<OffT>* dst,
                           const unsigned long long* iso, int32_t s, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(bfs_init_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, a, dst, iso, s);
  return hipGetLastError();
}

template <typename OffT>
const void* level_kernel_symbol(bool sorted) {
  return sorted ? reinterpret_cast<const void*>(&bfs_level_kernel<OffT, true>)
                : reinterpret_cast<const void*>(&bfs_level_kernel<OffT, false>);
}

template <typename OffT>
hipError_t launch_bfs_level(const BfsArgs<OffT>* dargs, bool sorted, int grid, hipStream_t stream) {
  if (sorted)
    hipLaunchKernelGGL((bfs_level_kernel<OffT, true>), dim3(grid), dim3(kBlock), 0, stream, dargs);
  else
    hipLaunchKernelGGL((bfs_level_kernel<OffT, false>), dim3(grid), dim3(kBlock), 0, stream, dargs);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_graph_offsets(const OffT* off, long long n, long long m,
                                unsigned long long* iso, unsigned long long* counters, int grid,
                                hipStream_t stream) {
  const long long nwords = (n + kWave - 1) / kWave;
  hipLaunchKernelGGL(graph_offsets_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, off, n, m,
                     nwords, iso, counters);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_graph_rows(const OffT* off, const int32_t* adj, long long n, long long m,
                             unsigned long long* counters, int grid, hipStream_t stream) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(graph_row_start_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, off, adj,
                     n, counters);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(graph_edge_meta_kernel, dim3(grid), dim3(kBlock), 0, stream, adj, m, n,
                     counters);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_graph_symmetry(const OffT* off, const int32_t* adj, long long n, long long m,
                                 unsigned long long* counters, int grid, hipStream_t stream) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(graph_symmetry_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, off, adj,
                     n, m, counters);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_word_weights(const OffT* off, long long n, unsigned long long* ww, int grid,
                               hipStream_t stream) {
  const long long nwords = (n + kWave - 1) / kWave;
  hipLaunchKernelGGL(word_weight_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, off, n, nwords,
                     ww);
  return hipGetLastError();
}

// out[i] = sum of in[0..i) for i <= n (in has n + 1 entries, in[n] ignored).
hipError_t exclusive_scan_u64(const unsigned long long* in, unsigned long long* out, long long n,
                              hipStream_t stream) {
  size_t bytes = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)(n + 1), stream);
  if (e != hipSuccess) return e;
  void* tmp = nullptr;
  e = hipMalloc(&tmp, bytes);
  if (e != hipSuccess) return e;
  e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)(n + 1), stream);
  const hipError_t s = hipStreamSynchronize(stream);
  (void)hipFree(tmp);
  return e != hipSuccess ? e : s;
}

hipError_t launch_share_bounds(const unsigned long long* prefix, long long nwords, long long waves,
                               uint32_t* bounds, hipStream_t stream) {
  const long long threads = waves + 1;
  hipLaunchKernelGGL(share_bounds_kernel, dim3((unsigned)((threads + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, stream, prefix, nwords, waves, bounds);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_mbfs_setup(const MbfsArgs<OffT>& a, MbfsArgs<OffT>* dst, int grid,
                             hipStream_t stream) {
  hipLaunchKernelGGL(mbfs_clear_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mbfs_seed_kernel<OffT>, dim3(1), dim3(kBlock), 0, stream, a, dst);
  return hipGetLastError();
}

template <typename OffT>
const void* mbfs_level_symbol() {
  return reinterpret_cast<const void*>(&mbfs_level_kernel<OffT>);
}

template <typename OffT>
hipError_t launch_mbfs_level(const MbfsArgs<OffT>* dargs, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(mbfs_level_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, dargs);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_mbfs_stats(const MbfsArgs<OffT>& a, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(mbfs_stats_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

#ifdef PBBS_WAVE_TIMING
extern "C" int pbbs_debug_wave_times(unsigned long long* out, long long count) {
  const long long cap = (long long)kTimingLevels * kTimingWaves * 3;
  if (count > cap) count = cap;
  return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_wave_times), count * 8, 0, hipMemcpyDeviceToHost) ==
                 hipSuccess
             ? 0
             : -2;
}
#endif

#define PBBS_INSTANTIATE(OffT)                                                                   \
  template hipError_t launch_bfs_init<OffT>(const BfsArgs<OffT>&, BfsArgs<OffT>*,                \
                                            const unsigned long long*, int32_t, int,             \
                                            hipStream_t);                                        \
  template const void* level_kernel_symbol<OffT>(bool);                                         \
  template hipError_t launch_bfs_level<OffT>(const BfsArgs<OffT>*, bool, int, hipStream_t);      \
  template hipError_t launch_graph_offsets<OffT>(const OffT*, long long, long long,             \
                                                 unsigned long long*, unsigned long long*, int,  \
                                                 hipStream_t);                                   \
  template hipError_t launch_graph_rows<OffT>(const OffT*, const int32_t*, long long, long long, \
                                              unsigned long long*, int, hipStream_t);            \
  template hipError_t launch_mbfs_setup<OffT>(const MbfsArgs<OffT>&, MbfsArgs<OffT>*, int,      \
                                             hipStream_t);                                      \
  template const void* mbfs_level_symbol<OffT>();                                               \
  template hipError_t launch_mbfs_level<OffT>(const MbfsArgs<OffT>*, int, hipStream_t);          \
  template hipError_t launch_mbfs_stats<OffT>(const MbfsArgs<OffT>&, int, hipStream_t);          \
  template hipError_t launch_word_weights<OffT>(const OffT*, long long, unsigned long long*, int, \
                                               hipStream_t);                                    \
  template hipError_t launch_graph_symmetry<OffT>(const OffT*, const int32_t*, long long,        \
                                                  long long, unsigned long long*, int,           \
                                                  hipStream_t);

PBBS_INSTANTIATE(uint32_t)
PBBS_INSTANTIATE(int64_t)

}  // namespace pbbs
