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
#include <climits>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "bfs_device.hpp"

namespace pbbs {

// Device pointers typed in the global address space, so loads, stores and
// atomics through them compile to global_* instructions.  Generic (flat)
// accesses also count against the LDS counter, so every LDS or cross-lane
// wait would otherwise wait for the outstanding memory loads too.
#define PBBS_GLOBAL __attribute__((address_space(1)))
template <typename T>
using gptr = PBBS_GLOBAL T*;
template <typename T>
__device__ __forceinline__ gptr<T> glob(T* p) {
  return (gptr<T>)p;
}
// Generic pointer for the atomic builtins; the compiler still sees the
// global origin and emits global atomics.
template <typename T>
__device__ __forceinline__ T* gen(gptr<T> p) {
  return (T*)p;
}


// ---------------------------------------------------------------- wave utils
__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ unsigned lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Cross-lane moves through DPP (a VALU operand modifier, a few cycles)
// instead of ds_bpermute (an LDS round trip).  `old` is what lanes without a
// source lane (or masked rows) receive.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ unsigned dpp_u32(unsigned x, unsigned old) {
  return (unsigned)__builtin_amdgcn_update_dpp((int)old, (int)x, kCtrl, kRowMask, 0xf, false);
}
template <int kCtrl, int kRowMask>
__device__ __forceinline__ int dpp_i32(int x, int old) {
  return __builtin_amdgcn_update_dpp(old, x, kCtrl, kRowMask, 0xf, false);
}
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

// One step of a keyed (segmented) inclusive OR scan: lanes take the value
// of the source lane when it carries the same key.  Keys are non-decreasing
// over the lanes, so equal keys at both ends mean one segment in between.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ void seg_or_step(unsigned long long& x, int key) {
  const unsigned lo = dpp_u32<kCtrl, kRowMask>((unsigned)x, 0u);
  const unsigned hi = dpp_u32<kCtrl, kRowMask>((unsigned)(x >> 32), 0u);
  const int k2 = dpp_i32<kCtrl, kRowMask>(key, INT32_MIN);
  if (k2 == key) x |= ((unsigned long long)hi << 32) | lo;
}

// Inclusive OR of x over the lanes at or below this one with the same key.
__device__ __forceinline__ unsigned long long seg_or_scan(unsigned long long x, int key) {
  seg_or_step<kRowShr1, 0xf>(x, key);
  seg_or_step<kRowShr2, 0xf>(x, key);
  seg_or_step<kRowShr4, 0xf>(x, key);
  seg_or_step<kRowShr8, 0xf>(x, key);
  seg_or_step<kRowBcast15, 0xa>(x, key);
  seg_or_step<kRowBcast31, 0xc>(x, key);
  return x;
}

// The value of the lane below (0 for lane 0 or a different key).
__device__ __forceinline__ unsigned long long seg_prev(unsigned long long x, int key) {
  const unsigned lo = dpp_u32<kWaveShr1, 0xf>((unsigned)x, 0u);
  const unsigned hi = dpp_u32<kWaveShr1, 0xf>((unsigned)(x >> 32), 0u);
  const int k2 = dpp_i32<kWaveShr1, 0xf>(key, INT32_MIN);
  return k2 == key ? (((unsigned long long)hi << 32) | lo) : 0ull;
}

// Segmented inclusive OR where each lane knows `dist`, its distance to the
// first lane of its segment (segments are runs of lanes).  Row-local steps
// read 0 from lanes outside the 16-lane row (bound_ctrl), the two broadcast
// steps bring in the previous rows' last lanes.
template <int kCtrl, int kRowMask, bool kBound>
__device__ __forceinline__ unsigned dpp_take(unsigned x) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, kCtrl, kRowMask, 0xf, kBound);
}
__device__ __forceinline__ unsigned long long seg_or_scan_d(unsigned long long x, int dist,
                                                            int lane16, int lane32) {
  unsigned lo = (unsigned)x, hi = (unsigned)(x >> 32);
#define PBBS_SEG_STEP(ctrl, rm, bound, cond)                        \
  {                                                                 \
    const unsigned l2 = dpp_take<ctrl, rm, bound>(lo);             \
    const unsigned h2 = dpp_take<ctrl, rm, bound>(hi);             \
    if (cond) {                                                     \
      lo |= l2;                                                     \
      hi |= h2;                                                     \
    }                                                               \
  }
  PBBS_SEG_STEP(kRowShr1, 0xf, true, dist >= 1)
  PBBS_SEG_STEP(kRowShr2, 0xf, true, dist >= 2)
  PBBS_SEG_STEP(kRowShr4, 0xf, true, dist >= 4)
  PBBS_SEG_STEP(kRowShr8, 0xf, true, dist >= 8)
  PBBS_SEG_STEP(kRowBcast15, 0xa, false, dist > lane16)
  PBBS_SEG_STEP(kRowBcast31, 0xc, false, dist > lane32)
#undef PBBS_SEG_STEP
  return ((unsigned long long)hi << 32) | lo;
}

// Inclusive sum over the wavefront through DPP.
__device__ __forceinline__ unsigned wave_inclusive_sum_dpp(unsigned x) {
  x += dpp_u32<kRowShr1, 0xf>(x, 0u);
  x += dpp_u32<kRowShr2, 0xf>(x, 0u);
  x += dpp_u32<kRowShr4, 0xf>(x, 0u);
  x += dpp_u32<kRowShr8, 0xf>(x, 0u);
  x += dpp_u32<kRowBcast15, 0xa>(x, 0u);
  x += dpp_u32<kRowBcast31, 0xc>(x, 0u);
  return x;
}

__device__ __forceinline__ unsigned wave_inclusive_sum(unsigned x) {
  return wave_inclusive_sum_dpp(x);
}

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) x = max(x, __shfl_xor(x, d));
  return x;
}

// Orders this wavefront's LDS writes before its later LDS reads by other
// lanes.  The fences name LDS only: a fence on all memory would also wait for
// every global store in flight (loads and stores share one counter here).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__device__ __forceinline__ uint32_t shfl_off(uint32_t x, int src) { return __shfl(x, src); }
__device__ __forceinline__ int64_t shfl_off(int64_t x, int src) {
  return (int64_t)__shfl((long long)x, src);
}

__device__ __forceinline__ bool in_set(gptr<const unsigned> bits, int u) {
  return (bits[u >> 5] >> (u & 31)) & 1u;
}

// Bit u of a bitmap, false for u < 0.  The word is loaded whatever u is (the
// index is clamped to 0), so a run of these tests issues all its loads before
// one wait; `u >= 0 && in_set(...)` compiles to one guarded load and one
// wait per test.
__device__ __forceinline__ bool in_set_or_none(gptr<const unsigned> bits, int u) {
  const unsigned w = bits[(u < 0 ? 0 : u) >> 5];
  return ((w >> (u & 31)) & (unsigned)(u >= 0)) != 0u;   // no short circuit
}

// Workgroups are dispatched in index order, so the CU slot a workgroup
// occupies is about blockIdx / num_cus.  A SIMD issues for its oldest
// wavefront first: with equal static shares the youngest slot falls behind
// and finishes the level last (measured at rmat24's first bottom-up level:
// 6.4 us per probe pass in slot 0 against 15.5 us in slot 3).  Raising the
// issue priority of later slots evens their progress.
__device__ __forceinline__ void slot_priority(int num_cus) {
  const unsigned g = num_cus > 0 ? blockIdx.x / (unsigned)num_cus : 0u;
  if (g == 1) __builtin_amdgcn_s_setprio(1);
  else if (g == 2) __builtin_amdgcn_s_setprio(2);
  else if (g >= 3) __builtin_amdgcn_s_setprio(3);
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
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <typename OffT>
struct LevelView {
  gptr<const OffT> off;
  gptr<const int32_t> adj;
  gptr<int32_t> parent;
  gptr<int32_t> depth;
  gptr<unsigned long long> visited;
  gptr<const unsigned> fb_cur;          // F_k
  gptr<unsigned long long> fb_next;     // F_{k+1}
  gptr<const int32_t> q_cur;            // sub-queue regions of F_k
  gptr<const uint32_t> qe_cur;
  gptr<int32_t> q_next;                 // this workgroup's sub-queue region of F_{k+1}
  gptr<uint32_t> qe_next;
  gptr<unsigned long long> ctr_next;    // its counter
  gptr<int32_t> qover_next;             // set when an append finds the sub-queue full
  gptr<uint32_t> d0_next;
  gptr<const unsigned long long> ctr_cur;  // F_k's counters (all sub-queues)
  gptr<const uint32_t> bu_bounds;       // weighted wavefront shares, or null for equal shares
  gptr<const u32x4> tops;               // per vertex {six largest neighbours, row end, degree}, or null
  int small_rows;                       // every row has at most 64 entries
  long long n;
  long long m;
  long long nwords;
  long long qcap;
  long long gw;                         // this wavefront's index in the level's grid
  long long W;                          // wavefronts in the level's grid
  int k;
  int serial_depth;
};

// --------------------------------------------------------- frontier staging
// Per-wavefront LDS buffer of (vertex, degree offset within the buffer),
// kCap entries per wavefront.
template <int kCap>
struct StageBufT {
  static constexpr int kCapacity = kCap;
  int32_t v[kWavesPerBlock][kCap];
  uint32_t eo[kWavesPerBlock][kCap];
  uint32_t cnt[kWavesPerBlock];
  uint32_t deg[kWavesPerBlock];
  unsigned long long base;
};
using StageBuf = StageBufT<kWaveBuf>;
// The batched kernel keeps a smaller buffer: its LDS budget sets how many
// workgroups share a CU, and its appends are few per wavefront.
using BatchStageBuf = StageBufT<128>;

struct Stage {
  unsigned cnt = 0;   // wave-uniform
  unsigned deg = 0;   // wave-uniform
};

// Writes buffer entries [0, cnt) of wavefront w to the next frontier at
// (qbase, ebase) and empties it.  A sub-queue region holds qcap entries:
// entries past it are dropped and the frontier is marked incomplete, so the
// level that consumes it reads its bitmap instead.
template <typename View, typename SB>
__device__ __forceinline__ void stage_copy_out(const View& a, SB& sb, int w,
                                               unsigned cnt, unsigned qbase, unsigned ebase) {
  const unsigned long long room =
      (unsigned long long)a.qcap > qbase ? (unsigned long long)a.qcap - qbase : 0ull;
  const unsigned lim = room < cnt ? (unsigned)room : cnt;
  for (unsigned i = lane_id(); i < lim; i += kWave) {
    a.q_next[qbase + i] = sb.v[w][i];
    a.qe_next[qbase + i] = ebase + sb.eo[w][i];
  }
  if (lim < cnt && lane_id() == 0) *a.qover_next = 1;
}

template <typename View, typename SB>
__device__ __forceinline__ void stage_flush_wave(const View& a, SB& sb, Stage& s) {
  const int w = threadIdx.x / kWave;
  unsigned long long base = 0;
  if (lane_id() == 0)
    base = atomicAdd(gen(a.ctr_next), ((unsigned long long)s.deg << 32) | s.cnt);
  base = __shfl(base, 0);
  wave_lds_sync();
  stage_copy_out(a, sb, w, s.cnt, (unsigned)base, (unsigned)(base >> 32));
  wave_lds_sync();
  s.cnt = 0;
  s.deg = 0;
}

// Stages the lanes with `take` (vertex v of degree deg) for F_{k+1}.
template <typename View, typename SB>
__device__ __forceinline__ void stage_push(const View& a, SB& sb, Stage& s, bool take, int v,
                                           uint32_t deg) {
  const unsigned long long mask = __ballot(take);
  if (mask == 0) return;
  const unsigned c = (unsigned)__popcll(mask);
  const unsigned incl = wave_inclusive_sum(take ? deg : 0u);
  const unsigned total = (unsigned)__builtin_amdgcn_readlane((int)incl, kWave - 1);
  if (s.cnt + c > (unsigned)SB::kCapacity) stage_flush_wave(a, sb, s);
  if (take) {
    const int w = threadIdx.x / kWave;
    const unsigned pos = s.cnt + lanes_below(mask);
    sb.v[w][pos] = v;
    sb.eo[w][pos] = s.deg + incl - deg;
  }
  s.cnt += c;
  s.deg += total;
}

// Bottom-up levels only count their finds (count, degree sum): F_{k+1} is
// their bitmap, and the level after one reads that bitmap, so no queue
// entry is staged.
__device__ __forceinline__ void stage_count(Stage& s, bool take, uint32_t deg) {
  const unsigned long long mask = __ballot(take);
  if (!mask) return;
  const unsigned incl = wave_inclusive_sum(take ? deg : 0u);
  s.cnt += (unsigned)__popcll(mask);
  s.deg += (unsigned)__builtin_amdgcn_readlane((int)incl, kWave - 1);
}

// End of level: one reservation for the whole workgroup, then copy-out
// (none when the level only counted).
template <typename View, typename SB>
__device__ __forceinline__ void stage_flush_block(const View& a, SB& sb, const Stage& s,
                                                  bool staged = true) {
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
    sb.base = tc ? atomicAdd(gen(a.ctr_next), ((unsigned long long)td << 32) | tc) : 0ull;
  }
  __syncthreads();
  unsigned qb = (unsigned)sb.base, eb = (unsigned)(sb.base >> 32);
  for (int i = 0; i < w; ++i) {
    qb += sb.cnt[i];
    eb += sb.deg[i];
  }
  if (staged) stage_copy_out(a, sb, w, s.cnt, qb, eb);
}

// ------------------------------------------------------------------ top-down
// Edges [eb, ee) of one sub-queue (nf entries, E edges; E < 2^32 since m is):
// find the entry holding edge eb with a 64-ary search over qe, then walk
// windows of 64 entries; lane l of a batch takes edge base+l and finds its
// source by a register binary search over the window's qe values.
// Edges [e, lim) of a window of up to 64 frontier entries held one per lane:
// vertex qv, exclusive degree prefix qo (~0 past the window's entries), row
// start rs.  Lane l takes edge base + l and finds its entry by a register
// binary search over qo; a claim that returns -1 makes the neighbour new.
template <typename OffT>
__device__ __forceinline__ void td_window(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                                          int32_t qv, uint32_t qo, OffT rs, uint32_t e,
                                          uint32_t lim) {
  const unsigned lane = lane_id();
  const int k = a.k;
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
    uint32_t deg = 0;
    if (valid) {
      const OffT ei = rsj + (OffT)(my - qoj);
      v = a.adj[(unsigned long long)ei < (unsigned long long)a.m ? ei : (OffT)0];
      // the row bounds load with the depth (one round trip, not two after a claim)
      const int dv = a.depth[v];
      const OffT o0 = a.off[v], o1 = a.off[v + 1];
      if (dv > k) {
        const int old = atomicMax(gen(&a.parent[v]), u);
        if (old == -1) {
          fresh = true;
          a.depth[v] = k + 1;
          deg = (uint32_t)(o1 - o0);
        }
      }
    }
    if (__ballot(fresh)) {
      if (fresh && a.fb_next)   // null while a persistent launch keeps no bitmaps
        atomicOr(reinterpret_cast<unsigned*>(gen(a.fb_next)) + (v >> 5), 1u << (v & 31));
      const unsigned long long z = __ballot(fresh && deg == 0);
      if (z && lane == 0) atomicAdd(gen(a.d0_next), (unsigned)__popcll(z));
      stage_push(a, sb, stg, fresh && deg > 0, v, deg);
    }
  }
}

template <typename OffT>
__device__ void td_range(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                         gptr<const int32_t> q, gptr<const uint32_t> qe,
                         unsigned nf, uint32_t E, uint32_t eb, uint32_t ee, int start) {
  const unsigned lane = lane_id();
  unsigned i;
  if (start >= 0) {
    i = (unsigned)start;   // eb is entry start's first edge
  } else {
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
    const unsigned idx = lo + lane;
    const bool p = idx < hi && qe[idx] <= eb;
    i = lo + (unsigned)__popcll(__ballot(p)) - 1;
  }

  // The callers pass a non-empty range (every queued vertex has edges), so
  // the first window is loaded without waiting for eb and ee; were the range
  // empty, the window would find no edge in it and the loop would end.
  uint32_t e = eb;
  do {
    const unsigned idx = i + lane;
    int32_t qv = 0;
    uint32_t qo = 0xffffffffu;
    OffT rs = 0;
    if (idx < nf) {
      qv = q[idx];
      // entries are vertex ids written by the level before; the clamp keeps
      // every later index in bounds whatever the queue holds
      qv = (unsigned)qv < (unsigned long long)a.n ? qv : 0;
      qo = qe[idx];
      rs = a.off[qv];
    }
    const uint32_t wend = (i + kWave < nf) ? qe[i + kWave] : E;
    const uint32_t lim = min(wend, ee);
    td_window(a, sb, stg, qv, qo, rs, e, lim);
    e = max(e, lim);
    i += kWave;
  } while (e < ee && i < nf);
}

// The frontier's edge space is the concatenation of the sub-queues; each
// wavefront takes an equal slice of it.
template <typename OffT>
__device__ void top_down(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                         unsigned long long E, const unsigned long long (&qc)[kSubQueues]) {
  const unsigned long long W = (unsigned long long)a.W;
  const unsigned long long gw = (unsigned long long)a.gw;
  if (a.small_rows) {
    // Every row has at most 64 entries: slices of equal entry counts are
    // within one window of balanced, and a slice's first entry is known, so
    // the search over the degree prefix is skipped.
    unsigned long long N = 0;
    for (int s = 0; s < kSubQueues; ++s) N += (uint32_t)qc[s];
    const unsigned long long ib = N * gw / W, ie = N * (gw + 1) / W;
    if (ib >= ie) return;
    unsigned long long P = 0;
    for (int s = 0; s < kSubQueues; ++s) {
      const unsigned long long c = qc[s];
      const unsigned nf = (unsigned)c;
      const unsigned long long lo = ib > P ? ib : P;
      const unsigned long long hi = ie < P + nf ? ie : P + nf;
      if (lo < hi) {
        const gptr<const uint32_t> qe = a.qe_cur + s * a.qcap;
        const unsigned i0 = (unsigned)(lo - P), i1 = (unsigned)(hi - P);
        const uint32_t Es = (uint32_t)(c >> 32);
        const uint32_t eb = qe[i0], ee = i1 < nf ? qe[i1] : Es;
        td_range(a, sb, stg, a.q_cur + s * a.qcap, qe, nf, Es, eb, ee, (int)i0);
      }
      P += nf;
    }
    return;
  }
  const unsigned long long eb = E * gw / W, ee = E * (gw + 1) / W;
  if (eb >= ee) return;
  unsigned long long P = 0;
  for (int s = 0; s < kSubQueues; ++s) {
    const unsigned long long c = qc[s];
    const unsigned long long Es = c >> 32;
    const unsigned long long lo = eb > P ? eb : P;
    const unsigned long long hi = ee < P + Es ? ee : P + Es;
    if (lo < hi)
      td_range(a, sb, stg, a.q_cur + s * a.qcap, a.qe_cur + s * a.qcap, (unsigned)c,
               (uint32_t)Es, (uint32_t)(lo - P), (uint32_t)(hi - P), -1);
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
  const gptr<const unsigned> fbc = a.fb_cur;
  const OffT have = p - lo;
  // unconditional loads (entry 0 stands in past the row's start) so that all
  // N are in flight together
  int c[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const int x = a.adj[(OffT)j < have ? p - 1 - (OffT)j : (OffT)0];
    c[j] = (OffT)j < have ? x : -1;
  }
  bool h[N];
#pragma unroll
  for (int j = 0; j < N; ++j) h[j] = in_set_or_none(fbc, c[j]);
  int found = -1;
#pragma unroll
  for (int j = N - 1; j >= 0; --j)
    if (h[j]) found = c[j];
  p = have > (OffT)N ? p - (OffT)N : lo;
  return found;
}

// Any-parent levels: tests entries [p, p+N) of row [p, hi) from the bottom
// with N independent loads; returns the first hit (or -1) and moves p up
// past them.
template <int N, typename OffT>
__device__ __forceinline__ int scan_bot(const LevelView<OffT>& a, OffT hi, OffT& p) {
  const gptr<const unsigned> fbc = a.fb_cur;
  const OffT have = hi - p;
  int c[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const int x = a.adj[(OffT)j < have ? p + (OffT)j : (OffT)0];
    c[j] = (OffT)j < have ? x : -1;
  }
  bool h[N];
#pragma unroll
  for (int j = 0; j < N; ++j) h[j] = in_set_or_none(fbc, c[j]);
  int found = -1;
#pragma unroll
  for (int j = N - 1; j >= 0; --j)
    if (h[j]) found = c[j];
  p = have > (OffT)N ? p + (OffT)N : hi;
  return found;
}

// Whole-wavefront scan of the rows of the lanes in `left`, 64 entries per
// step: from the top (the first hit, lowest lane, is the highest entry), or
// with kAsc from the bottom (entries [p, bound) left; any hit is a parent).
template <typename OffT, bool kAsc = false>
__device__ __forceinline__ void scan_rows_together(const LevelView<OffT>& a,
                                                   unsigned long long left, OffT bound, OffT p,
                                                   int& found) {
  const unsigned lane = lane_id();
  const gptr<const unsigned> fbc = a.fb_cur;
  while (left) {
    const int L = __builtin_ctzll(left);
    left &= left - 1;
    const OffT bL = shfl_off(bound, L);
    OffT pL = shfl_off(p, L);
    int fu = -1;
    while (kAsc ? pL < bL : pL > bL) {
      const long long pos = kAsc ? (long long)pL + (long long)lane
                                 : (long long)pL - 1 - (long long)lane;
      bool hit = false;
      int u = -1;
      if (kAsc ? pos < (long long)bL : pos >= (long long)bL) {
        u = a.adj[pos];
        hit = in_set(fbc, u);
      }
      const unsigned long long hm = __ballot(hit);
      if (hm) {
        fu = __shfl(u, __builtin_ctzll(hm));
        break;
      }
      if (kAsc) pL = (bL - pL > (OffT)kWave) ? pL + (OffT)kWave : bL;
      else pL = (pL - bL > (OffT)kWave) ? pL - (OffT)kWave : bL;
    }
    if (lane == (unsigned)L) found = fu;
  }
}

// Full-row max scan for unsorted rows (no early exit).
template <typename OffT>
__device__ __forceinline__ int scan_unsorted(const LevelView<OffT>& a, bool mine, OffT lo, OffT hi) {
  const unsigned lane = lane_id();
  const gptr<const unsigned> fbc = a.fb_cur;
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

// Unreached vertices a lane probes per pass of a sorted bottom-up level.  The
// pass issues every load of all of them (row-top records, or offsets then row
// tops, then bitmap words) before it uses any, so a wavefront keeps kBuPer
// times as many independent fetches in flight: the first bottom-up level is
// bound by the latency of these round trips (3 measured best on RMAT 2^24,
// against 1 and 2).
constexpr int kBuPer = 3;
constexpr int kBuCap = (kBuPer + 1) * kWave;   // queue and park capacity per wavefront
constexpr int kTopEntries = 6;                 // neighbours held by a row-top record
static_assert(kBuCap >= 2 * kWave, "the bottom-up and bitmap top-down lists take 64 + 63 entries");

template <typename OffT>
struct BuLds {
  int32_t q[kWavesPerBlock][kBuCap];      // unreached vertices, ascending
  int32_t lv[kWavesPerBlock][kBuCap];     // parked vertices
  uint32_t llo[kWavesPerBlock][kBuCap];   // their row starts (any-parent levels: row ends)
  uint32_t lp[kWavesPerBlock][kBuCap];    // their scan positions (entries [lo, p) untested;
                                          // any-parent levels: [p, end))
};

// Records the found lanes: parent/depth, their F_{k+1} bits (one atomicOr per
// distinct word; lanes hold ascending ids so a word's lanes are adjacent),
// and their count and degree sum.
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
    // OR of the bits over each run of lanes with the same word (DPP scan);
    // the run's last lane holds the total and sets it.  Idle lanes (all at
    // the end) take a key of their own.
    const int key = valid ? (v >> 6) : INT32_MAX;
    unsigned long long bits = nw ? (1ull << (v & 63)) : 0ull;
    bits = seg_or_scan(bits, key);
    const int kp = dpp_i32<kWaveShr1, 0xf>(key, INT32_MIN);
    const unsigned long long heads = __ballot(kp != key);
    const bool tail = lane == (unsigned)(kWave - 1) || ((heads >> (lane + 1)) & 1ull);
    if (tail && bits) atomicOr(gen(&a.fb_next[key]), bits);
  }
  stage_count(stg, nw, deg);
}

template <typename OffT, bool kAsc = false>
__device__ void bu_finish_parked(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                                 BuLds<OffT>& L, unsigned& npark, int tested) {
  const unsigned lane = lane_id();
  const int w = threadIdx.x / kWave;
  wave_lds_sync();
  for (unsigned base = 0; base < npark; base += kWave) {
    const bool mine = base + lane < npark;
    int v = 0;
    OffT bound = 0, p = 0;   // untested entries [bound, p) (from the top) or [p, bound)
    uint32_t deg = 0;
    if (mine) {
      v = L.lv[w][base + lane];
      bound = (OffT)L.llo[w][base + lane];
      p = (OffT)L.lp[w][base + lane];
      deg = kAsc ? (uint32_t)(bound - (p - (OffT)tested)) : (uint32_t)(p + (OffT)tested - bound);
    }
    int found = -1;
    // parked rows are the deep ones: sixteen loads in flight per pass
    for (int t = tested; t < a.serial_depth; t += 16) {
      const bool active = mine && found < 0 && (kAsc ? p < bound : p > bound);
      if (!__ballot(active)) break;
      if (active) found = kAsc ? scan_bot<16>(a, bound, p) : scan_top<16>(a, bound, p);
    }
    scan_rows_together<OffT, kAsc>(
        a, __ballot(mine && found < 0 && (kAsc ? p < bound : p > bound)), bound, p, found);
    bu_commit(a, sb, stg, mine, mine && found >= 0, v, found, deg);
  }
  wave_lds_sync();
  npark = 0;
}

// Parks the lanes with `park` set: vertex v with row [lo, hi), its top
// `tested` entries already tested.  Finishes the parked rows first when they
// would overflow the list.  Offsets fit 32 bits (m < 2^32).
template <typename OffT, bool kAsc = false>
__device__ __forceinline__ void bu_park(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                                        BuLds<OffT>& L, unsigned& npark, bool park, int v,
                                        uint32_t lo, uint32_t hi, int tested) {
  const unsigned long long pm = __ballot(park);
  if (!pm) return;
  const int w = threadIdx.x / kWave;
  if (npark + (unsigned)__popcll(pm) > (unsigned)kBuCap)
    bu_finish_parked<OffT, kAsc>(a, sb, stg, L, npark, tested);
  if (park) {
    const unsigned slot = npark + lanes_below(pm);
    L.lv[w][slot] = v;
    L.llo[w][slot] = kAsc ? hi : lo;
    L.lp[w][slot] = kAsc ? lo + (uint32_t)tested : hi - (uint32_t)tested;
  }
  npark += (unsigned)__popcll(pm);
}

// One probe pass of a sorted level over queue entries [0, cnt), cnt <=
// kBuPer * 64: lane l takes entries l, l + 64, ...  Each row's top four
// entries are tested; a row with no hit there and entries left is parked.
template <typename OffT, bool kAsc = false>
__device__ void bu_probe(const LevelView<OffT>& a, StageBuf& sb, Stage& stg, BuLds<OffT>& L,
                         unsigned cnt, unsigned& npark) {
  constexpr int N = 4;
  const unsigned lane = lane_id();
  const int w = threadIdx.x / kWave;
  const gptr<const unsigned> fbc = a.fb_cur;
  wave_lds_sync();
  bool mine[kBuPer];
  int v[kBuPer];
#pragma unroll
  for (int b = 0; b < kBuPer; ++b) {
    mine[b] = lane + (unsigned)(b * kWave) < cnt;
    v[b] = mine[b] ? L.q[w][lane + b * kWave] : 0;
  }
  OffT lo[kBuPer], hi[kBuPer];
#pragma unroll
  for (int b = 0; b < kBuPer; ++b) {
    lo[b] = 0;
    hi[b] = 0;
    if (mine[b]) {
      lo[b] = a.off[v[b]];
      hi[b] = a.off[v[b] + 1];
    }
  }
  int c[kBuPer][N];
#pragma unroll
  for (int b = 0; b < kBuPer; ++b)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const bool in_row = (OffT)j < hi[b] - lo[b];
      const int x = a.adj[in_row ? (kAsc ? lo[b] + (OffT)j : hi[b] - 1 - (OffT)j) : (OffT)0];
      c[b][j] = in_row ? x : -1;
    }
  bool h[kBuPer][N];
#pragma unroll
  for (int b = 0; b < kBuPer; ++b)
#pragma unroll
    for (int j = 0; j < N; ++j) h[b][j] = in_set_or_none(fbc, c[b][j]);
#pragma unroll
  for (int b = 0; b < kBuPer; ++b) {
    int found = -1;
#pragma unroll
    for (int j = N - 1; j >= 0; --j)
      if (h[b][j]) found = c[b][j];
    const bool park = mine[b] && found < 0 && hi[b] - lo[b] > (OffT)N;
    bu_commit(a, sb, stg, mine[b], mine[b] && found >= 0, v[b], found, (uint32_t)(hi[b] - lo[b]));
    bu_park<OffT, kAsc>(a, sb, stg, L, npark, park, v[b], (uint32_t)lo[b], (uint32_t)hi[b], N);
  }
}

// Probe pass that reads each vertex's 32-byte row-top record {six largest
// neighbours (~0 where the row is shorter), row end, degree} in place of its
// offsets and row: two coalesced loads and six bitmap tests per vertex.  A
// row longer than six with no hit is parked with its position from the
// record, so no offsets are read at all.
template <typename OffT>
__device__ void bu_probe_tops(const LevelView<OffT>& a, StageBuf& sb, Stage& stg, BuLds<OffT>& L,
                              unsigned cnt, unsigned& npark) {
  constexpr int N = kTopEntries;
  const unsigned lane = lane_id();
  const int w = threadIdx.x / kWave;
  const gptr<const unsigned> fbc = a.fb_cur;
  wave_lds_sync();
  bool mine[kBuPer];
  int v[kBuPer];
#pragma unroll
  for (int b = 0; b < kBuPer; ++b) {
    mine[b] = lane + (unsigned)(b * kWave) < cnt;
    v[b] = mine[b] ? L.q[w][lane + b * kWave] : 0;
  }
  u32x4 r0[kBuPer], r1[kBuPer];
#pragma unroll
  for (int b = 0; b < kBuPer; ++b) {
    r0[b] = (u32x4){~0u, ~0u, ~0u, ~0u};
    r1[b] = (u32x4){~0u, ~0u, 0u, 0u};
    if (mine[b]) {
      r0[b] = a.tops[2 * (long long)v[b]];
      r1[b] = a.tops[2 * (long long)v[b] + 1];
    }
  }
  bool h[kBuPer][N];
#pragma unroll
  for (int b = 0; b < kBuPer; ++b) {
    const unsigned t[N] = {r0[b].x, r0[b].y, r0[b].z, r0[b].w, r1[b].x, r1[b].y};
#pragma unroll
    for (int j = 0; j < N; ++j) h[b][j] = in_set_or_none(fbc, (int)t[j]);
  }
#pragma unroll
  for (int b = 0; b < kBuPer; ++b) {
    const unsigned t[N] = {r0[b].x, r0[b].y, r0[b].z, r0[b].w, r1[b].x, r1[b].y};
    int found = -1;
#pragma unroll
    for (int j = N - 1; j >= 0; --j)
      if (h[b][j]) found = (int)t[j];
    const uint32_t hi = r1[b].z, deg = r1[b].w;
    const bool park = mine[b] && found < 0 && deg > (uint32_t)N;
    bu_commit(a, sb, stg, mine[b], mine[b] && found >= 0, v[b], found, deg);
    bu_park(a, sb, stg, L, npark, park, v[b], hi - deg, hi, N);
  }
}

// One pass of an unsorted level over queue entries [0, cnt), cnt <= 64.
template <typename OffT>
__device__ void bu_batch_unsorted(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                                  BuLds<OffT>& L, unsigned cnt) {
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
  const int found = scan_unsorted(a, mine, lo, hi);
  bu_commit(a, sb, stg, mine, mine && found >= 0, v, found, (uint32_t)(hi - lo));
}

template <typename OffT, bool kSorted, bool kTops, bool kAsc>
__device__ __forceinline__ void bu_pass(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                                        BuLds<OffT>& L, unsigned cnt, unsigned& npark) {
  if constexpr (kAsc)
    bu_probe<OffT, true>(a, sb, stg, L, cnt, npark);
  else if constexpr (!kSorted)
    bu_batch_unsorted(a, sb, stg, L, cnt);
  else if constexpr (kTops)
    bu_probe_tops(a, sb, stg, L, cnt, npark);
  else
    bu_probe(a, sb, stg, L, cnt, npark);
}

// kAsc (any-parent levels): every row is probed from its smallest entry up and
// the first frontier neighbour found is the parent, sorted rows or not.
template <typename OffT, bool kSorted, bool kTops, bool kAsc = false>
__device__ void bottom_up(const LevelView<OffT>& a, StageBuf& sb, Stage& stg, BuLds<OffT>& L) {
  // entries per pass: kBuPer per lane on sorted rows, one on unsorted rows
  constexpr unsigned kPass = (unsigned)((kSorted || kAsc ? kBuPer : 1) * kWave);
  const unsigned lane = lane_id();
  const int w = threadIdx.x / kWave;
  const gptr<const unsigned long long> fbc64 = (gptr<const unsigned long long>)a.fb_cur;
  const long long W = a.W;
  const long long gw = a.gw;
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
      // nq < kPass here, so the queue (kPass + 64 entries) has room for c <= 64
      if (take) L.q[w][nq + lanes_below(word)] = (int)((g0 + j) * kWave + lane);
      nq += c;
      if (nq >= kPass) {
        bu_pass<OffT, kSorted, kTops, kAsc>(a, sb, stg, L, kPass, npark);
        wave_lds_sync();
        const unsigned rest = nq - kPass;   // < 64
        int tmp = lane < rest ? L.q[w][kPass + lane] : 0;
        wave_lds_sync();
        if (lane < rest) L.q[w][lane] = tmp;
        nq = rest;
      }
    }
  }
  if (nq) bu_pass<OffT, kSorted, kTops, kAsc>(a, sb, stg, L, nq, npark);
  if ((kSorted || kAsc) && npark)
    bu_finish_parked<OffT, kAsc>(a, sb, stg, L, npark, kTops && !kAsc ? kTopEntries : 4);
}

// A top-down level after a bottom-up one: F_k exists only as its bitmap.
// Each wavefront takes an equal share of the bitmap words, gathers its
// frontier vertices 64 at a time (one per lane, with row start and degree)
// and expands their edges in windows.  Such frontiers are small and of low
// degree (the switch back needs |F| <= n / bu_to_td_div).
template <typename OffT>
__device__ void td_list(const LevelView<OffT>& a, StageBuf& sb, Stage& stg, BuLds<OffT>& L,
                        unsigned cnt) {
  const unsigned lane = lane_id();
  const int w = threadIdx.x / kWave;
  wave_lds_sync();
  const bool mine = lane < cnt;
  const int u = mine ? L.q[w][lane] : 0;
  OffT rs = 0;
  uint32_t deg = 0;
  if (mine) {
    rs = a.off[u];
    deg = (uint32_t)(a.off[u + 1] - rs);
  }
  const uint32_t incl = wave_inclusive_sum(deg);
  const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
  td_window(a, sb, stg, u, mine ? incl - deg : 0xffffffffu, rs, 0u, total);
}

template <typename OffT>
__device__ void td_from_bitmap(const LevelView<OffT>& a, StageBuf& sb, Stage& stg,
                               BuLds<OffT>& L) {
  const unsigned lane = lane_id();
  const int w = threadIdx.x / kWave;
  const gptr<const unsigned long long> fb = (gptr<const unsigned long long>)a.fb_cur;
  const long long W = a.W;
  const long long gw = a.gw;
  const long long w0 = a.nwords * gw / W, w1 = a.nwords * (gw + 1) / W;
  unsigned nq = 0;
  for (long long g0 = w0; g0 < w1; g0 += kWave) {
    const long long mb = g0 + lane;
    const unsigned long long word = mb < w1 ? fb[mb] : 0ull;
    unsigned long long todo = __ballot(word != 0ull);
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1;
      const unsigned long long wd = __shfl(word, j);
      // nq < 64 here and the list holds kBuCap >= 128 entries
      if ((wd >> lane) & 1ull) L.q[w][nq + lanes_below(wd)] = (int)((g0 + j) * kWave + lane);
      nq += (unsigned)__popcll(wd);
      if (nq >= (unsigned)kWave) {
        td_list(a, sb, stg, L, kWave);
        wave_lds_sync();
        const unsigned rest = nq - kWave;
        const int tmp = lane < rest ? L.q[w][kWave + lane] : 0;
        wave_lds_sync();
        if (lane < rest) L.q[w][lane] = tmp;
        nq = rest;
      }
    }
  }
  if (nq) td_list(a, sb, stg, L, nq);
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

// Level k of the BFS, as workgroup vb of a grid of vg workgroups (the
// launch's own grid, or this BFS's part of a launch shared with others).
// Returns false if F_k is empty (the BFS is over), true once this
// workgroup's share of level k is done.
// A persistent launch keeps no frontier or visited bitmap while its levels
// are top-down and their queues complete: nothing reads them then, and the
// whole-graph clear and fold of every level is skipped, with the per-claim
// bitmap atomicOr.  The first level that needs them (bottom-up, or a
// frontier whose queue overflowed) first rebuilds them from the depths,
// which every claim writes, behind one more grid barrier; the launch keeps
// them from then on.
struct BarView;
__device__ bool grid_barrier(BfsState* st, int32_t* host_status, long long wait_ticks,
                             BarView& b);
struct LazyBitmaps {
  BarView* bv;   // the launch's barrier
  bool on;       // bitmaps not kept so far
};

// visited = {v : depth[v] < k} | degree-0 | pad bits, F_k = {v : depth[v] == k},
// F_{k+1} = {}: the bitmaps as a level k that keeps them would find them.
// Wavefront w of the grid builds words w*8 .. w*8+7, then every W*8-th after:
// lane l tests vertex 64 word + l.
template <typename OffT>
__device__ void rebuild_bitmaps(const BfsArgs<OffT>* __restrict__ A, const int k,
                                const unsigned vb, const unsigned vg) {
  constexpr int kU = 8;
  const unsigned lane = lane_id();
  const long long W = (long long)vg * kWavesPerBlock;
  const long long gw = (long long)vb * kWavesPerBlock + threadIdx.x / kWave;
  const long long n = A->n, nwords = A->nwords;
  const gptr<const int32_t> depth = glob((const int32_t*)A->depth);
  const gptr<const unsigned long long> iso = glob(A->iso);
  const gptr<unsigned long long> vis = glob(A->visited);
  const gptr<unsigned long long> f0 = glob(A->fb[k % 3]);
  const gptr<unsigned long long> f1 = glob(A->fb[(k + 1) % 3]);
  const unsigned long long pad = (n & 63) ? (~0ull << (n & 63)) : 0ull;
  for (long long w0 = gw * kU; w0 < nwords; w0 += W * kU) {
    int d[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const long long v = (w0 + u) * 64 + lane;
      d[u] = v < n ? depth[v] : INT_MAX;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const long long w = w0 + u;
      if (w >= nwords) break;   // wave-uniform
      const unsigned long long lt = __ballot(d[u] < k), eq = __ballot(d[u] == k);
      if (lane == 0) {
        vis[w] = lt | iso[w] | (w == nwords - 1 ? pad : 0ull);
        f0[w] = eq;
        f1[w] = 0ull;
      }
    }
  }
}

template <typename OffT, bool kSorted, bool kTops, bool kLazy = false, bool kAsc = false>
__device__ __forceinline__ bool level_step(const BfsArgs<OffT>* __restrict__ A, const int k,
                                           StageBuf& sb, BuLds<OffT>& bl, const unsigned vb,
                                           const unsigned vg, LazyBitmaps* lz = nullptr) {
  BfsState* st = A->st;
  const int c0 = k % 3, c1 = (k + 1) % 3, c2 = (k + 2) % 3;
  unsigned long long cnt = 0, E = 0;
  unsigned long long qc[kSubQueues];   // F_k's sub-queue counters, kept for the expansion
#pragma unroll
  for (int s = 0; s < kSubQueues; ++s) {
    qc[s] = st->ctr[c0][s];
    cnt += (uint32_t)qc[s];
    E += qc[s] >> 32;
  }
  if (cnt == 0) {
    // F_k has no vertex to expand: the BFS is over.  The first kernel to see
    // it closes the books; later (overshoot) launches return at once.
    if (vb == 0 && threadIdx.x == 0 && !st->finished) {
      const uint32_t z = st->d0[c0];
      st->reached += z;
      if (z) st->max_depth = k;
      st->finished = 1;
      BfsState* hst = A->host_state;   // the host reads these once host_status[0] is set
      hst->reached = st->reached;
      hst->edges_reached = st->edges_reached;
      hst->max_depth = st->max_depth;
      hst->td_levels = st->td_levels;
      hst->bu_levels = st->bu_levels;
      __threadfence_system();
      __hip_atomic_store(&A->host_status[1], k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&A->host_status[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return false;
  }
  const int prev = k == 0 ? kTopDown : st->modes[(k - 1) & 3];
  const int mode = level_mode(*A, prev, cnt, E);
  const bool qover = st->qover[c0] != 0;
  bool keep = true;   // this level keeps the bitmaps
  if constexpr (kLazy) {
    if (lz->on) {
      if (mode == kBottomUp || qover) {
        rebuild_bitmaps(A, k, vb, vg);
        if (!grid_barrier(st, A->host_status, A->wait_ticks, *lz->bv)) return false;
        lz->on = false;
      } else {
        keep = false;
      }
    }
  }
  if (vb == 0 && threadIdx.x == 0) {
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
    st->qover[c2] = 0;
    __hip_atomic_store(&A->host_status[1], k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  if (keep) {
    // F_{k+2}'s bitmap held F_{k-1}.  A top-down level also folds F_k into
    // the visited bitmap, which only bottom-up levels read.
    const gptr<unsigned long long> z = glob(A->fb[c2]);
    const gptr<const unsigned long long> f = glob((const unsigned long long*)A->fb[c0]);
    const gptr<unsigned long long> vis = glob(A->visited);
    const bool fold = mode == kTopDown && A->symmetric && A->td_to_bu_div > 0;
    const long long nwords = A->nwords;
    const long long stride = (long long)vg * blockDim.x;
    for (long long w = (long long)vb * blockDim.x + threadIdx.x; w < nwords; w += stride) {
      z[w] = 0;
      if (fold) {
        const unsigned long long fw = f[w];
        if (fw) vis[w] |= fw;
      }
    }
  }
  const int sq = (int)(vb % kSubQueues);
  LevelView<OffT> a;
  a.off = glob(A->off);
  a.adj = glob(A->adj);
  a.parent = glob(A->parent);
  a.depth = glob(A->depth);
  a.visited = glob(A->visited);
  a.fb_cur = glob(reinterpret_cast<const unsigned*>(A->fb[c0]));
  a.fb_next = keep ? glob(A->fb[c1]) : nullptr;
  a.qcap = A->qcap;
  a.q_cur = glob(A->q[k & 1]);
  a.qe_cur = glob(A->qe[k & 1]);
  a.q_next = glob(A->q[(k + 1) & 1] + sq * a.qcap);
  a.qe_next = glob(A->qe[(k + 1) & 1] + sq * a.qcap);
  a.ctr_next = glob(&st->ctr[c1][sq]);
  a.qover_next = glob(&st->qover[c1]);
  a.d0_next = glob(&st->d0[c1]);
  a.ctr_cur = glob(st->ctr[c0]);
  // The first bottom-up level after top-down meets almost every vertex
  // unreached; its per-vertex cost grows with the degree, so the shares are
  // cut at equal sums of 1 + 2 log2(1 + degree).  Later bottom-up levels meet
  // few, low-degree vertices and split the words evenly.
  a.bu_bounds = glob((prev == kTopDown) ? A->bu_bounds : nullptr);
  a.tops = glob(kTops ? reinterpret_cast<const u32x4*>(A->bu_tops) : nullptr);
  a.small_rows = A->small_rows;
  a.n = A->n;
  a.m = A->m;
  a.nwords = A->nwords;
  a.gw = (long long)vb * kWavesPerBlock + threadIdx.x / kWave;
  a.W = (long long)vg * kWavesPerBlock;
  a.k = k;
  a.serial_depth = A->bu_serial_depth;
  Stage stg;
  if (mode == kTopDown) {
    if (prev == kBottomUp || qover)
      td_from_bitmap(a, sb, stg, bl);
    else
      top_down(a, sb, stg, E, qc);
  } else
  {
    slot_priority(A->num_cus);   // bottom-up only: top-down levels measured the same with it
    bottom_up<OffT, kSorted, kTops, kAsc>(a, sb, stg, bl);
    __builtin_amdgcn_s_setprio(0);
  }
  stage_flush_block(a, sb, stg, mode == kTopDown);
  return true;
}

// One level per launch: the level number comes from the workgroup's own
// counter, so a captured sequence of launches needs no host input.
template <typename OffT, bool kSorted, bool kTops, bool kAsc = false>
__global__ __launch_bounds__(kBlock) void bfs_level_kernel(const BfsArgs<OffT>* __restrict__ A) {
  __shared__ StageBuf sb;
  __shared__ BuLds<OffT> bl;
  const int k = (int)A->wg_level[blockIdx.x];
  if (level_step<OffT, kSorted, kTops, false, kAsc>(A, k, sb, bl, blockIdx.x, gridDim.x) &&
      threadIdx.x == 0)
    A->wg_level[blockIdx.x] = (uint32_t)(k + 1);
}

// Grid barrier between the levels of a persistent launch.  It does what a
// kernel boundary does for the data the levels share, by the agent-scope
// release/acquire rules of the memory model: each workgroup, once its
// stores are in, makes them visible device-wide (release: L2 write-back),
// arrives, waits for the barrier to complete, then drops stale cached
// lines (acquire: L1 and L2 invalidate) before it reads again.  The wait
// itself polls the generation word with device-scope loads and no cache
// maintenance (an acquire inside the loop would invalidate the L2 on every
// poll).  Arrivals go to kBarGroups first-tier counters, whose last
// arrivals arrive on bar_top; the last of those bumps bar_gen.
//
// Every wait is bounded by A->wait_ticks of the device wall clock: a
// workgroup that runs out sets st->abort (and host_status[2] for the host),
// and every workgroup leaves at its next wait or level start, so the launch
// always ends; the host then runs the BFS again one level per launch.
__device__ __forceinline__ unsigned ld_dev(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned add_dev(unsigned* p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until *w >= target (wrap-safe); false if the launch aborts or the
// wait runs out (which aborts it).
__device__ bool bar_wait(const unsigned* w, unsigned target, BfsState* st, int32_t* host_status,
                         long long wait_ticks) {
  const long long t0 = (long long)wall_clock64();
  while ((int)(ld_dev(w) - target) < 0) {
    if (__hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    if ((long long)wall_clock64() - t0 > wait_ticks) {
      __hip_atomic_store(&st->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&host_status[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// A workgroup's view of the barrier: its index in the grid that meets at
// it, that grid's size, the barriers passed so far and the last result.
struct BarView {
  unsigned vb, vg, nb;
  int ok;
};

__device__ __forceinline__ void barrier_join(unsigned vb, unsigned vg, BarView& b) {
  if (threadIdx.x == 0) {
    b.vb = vb;
    b.vg = vg;
    b.nb = 0;
    b.ok = 1;
  }
  __syncthreads();
}

// Returns false when the launch is to end.
__device__ bool grid_barrier(BfsState* st, int32_t* host_status, long long wait_ticks,
                             BarView& b) {
  __builtin_amdgcn_s_waitcnt(0);   // this wavefront's stores are in
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned nb = ++b.nb;
    const unsigned G = b.vg < (unsigned)kBarGroups ? b.vg : (unsigned)kBarGroups;
    const unsigned grp = b.vb % G;
    const unsigned gsize = (b.vg - grp + G - 1) / G;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");   // the block's stores, device-wide
    if (add_dev(&st->bar_grp[grp].v, 1u) == gsize - 1) {
      // last of its group: reset the group counter before the group can
      // arrive again (that needs bar_gen to move, which needs this arrival)
      st_dev(&st->bar_grp[grp].v, 0u);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");   // the reset lands first
      if (add_dev(&st->bar_top.v, 1u) == G - 1) {
        st_dev(&st->bar_top.v, 0u);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        add_dev(&st->bar_gen.v, 1u);
      }
    }
    const bool ok = bar_wait(&st->bar_gen.v, nb, st, host_status, wait_ticks);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // drop lines others have written
    b.ok = ok ? 1 : 0;
  }
  __syncthreads();
  return b.ok != 0;
}

// Every level of the BFS in one launch, separated by grid barriers.  The
// grid must fit on the device at once (the host launches at most the
// kernel's occupancy times the CU count); a launch that cannot gather at a
// barrier within wait_ticks gives up as described above.  On leaving, each
// workgroup records the level it reached, as the one-level kernel does.
template <typename OffT, bool kSorted, bool kTops, bool kAsc = false>
__global__ __launch_bounds__(kBlock) void bfs_persist_kernel(const BfsArgs<OffT>* __restrict__ A) {
  __shared__ StageBuf sb;
  __shared__ BuLds<OffT> bl;
  __shared__ int stop;
  __shared__ BarView bv;
  BfsState* st = A->st;
  int k = (int)A->wg_level[blockIdx.x];
  barrier_join(blockIdx.x, gridDim.x, bv);
  LazyBitmaps lz{&bv, true};
  {
    for (;;) {
      if (threadIdx.x == 0)
        stop = __hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      if (stop) break;
      if (!level_step<OffT, kSorted, kTops, true, kAsc>(A, k, sb, bl, blockIdx.x, gridDim.x, &lz))
        break;
      ++k;
      if (!grid_barrier(st, A->host_status, A->wait_ticks, bv)) break;
    }
  }
  if (threadIdx.x == 0) A->wg_level[blockIdx.x] = (uint32_t)k;
}

// Several single-source BFS runs in one launch, each on its own graph: job
// j's workgroups [wg_begin, wg_begin + wg_count) run its levels as the
// persistent kernel does, meeting at a barrier of their own, so jobs of
// different depths and sizes proceed independently.  All jobs of a launch
// have sorted rows (kSorted: read from the CSR, no row-top records) or all
// have unsorted rows.  The whole grid fits on
// the device at once (the host sizes it by occupancy); a job whose barrier
// wait runs out ends as in the persistent kernel and the host runs it again.
template <typename OffT, bool kSorted>
__global__ __launch_bounds__(kBlock) void bfs_multi_kernel(
    const BfsArgs<OffT>* __restrict__ args, const uint32_t* __restrict__ wg_begin, int njobs) {
  __shared__ StageBuf sb;
  __shared__ BuLds<OffT> bl;
  __shared__ int stop;
  __shared__ BarView bv;
  int lo = 0, hi = njobs - 1;   // the last job with wg_begin <= blockIdx.x
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (wg_begin[mid] <= blockIdx.x) lo = mid; else hi = mid - 1;
  }
  const unsigned vb = blockIdx.x - wg_begin[lo], vg = wg_begin[lo + 1] - wg_begin[lo];
  if (vb >= vg) return;   // not part of any job (the host sizes the grid to the jobs)
  const BfsArgs<OffT>* __restrict__ A = args + lo;
  BfsState* st = A->st;
  int k = (int)A->wg_level[vb];
  barrier_join(vb, vg, bv);
  {
    for (;;) {
      if (threadIdx.x == 0)
        stop = __hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      if (stop) break;
      if (!level_step<OffT, kSorted, false>(A, k, sb, bl, vb, vg)) break;
      ++k;
      if (!grid_barrier(st, A->host_status, A->wait_ticks, bv)) break;
    }
  }
  if (threadIdx.x == 0) A->wg_level[vb] = (uint32_t)k;
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
    BfsArgs<OffT> b = a;
    b.iso = iso;
    *dst = b;
    const unsigned long long deg = (unsigned long long)(a.off[s + 1] - a.off[s]);
    BfsState* st = a.st;
    a.q[0][0] = s;
    a.qe[0][0] = 0;
    for (int j = 0; j < 3; ++j) {
      for (int q = 0; q < kSubQueues; ++q) st->ctr[j][q] = 0;
      st->d0[j] = 0;
      st->qover[j] = 0;
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
    for (int j = 0; j < kBarGroups; ++j) st->bar_grp[j].v = 0;
    st->bar_top.v = 0;
    st->bar_gen.v = 0;
    st->abort = 0;
    __threadfence_system();
    __hip_atomic_store(&a.host_status[2], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
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
// "Reached by level k" is vis[v] | F_k[v]: level k's kernel ORs F_k into vis
// for the queue entries of F_k while other waves read it, and the union is
// the same before and after that store.
//
//   * push level k (small union frontier): every edge (u,v) of the frontier
//     ORs F_k[u] & ~reached(v) into F_{k+1}[v] (the lane whose atomicOr saw 0
//     appends v) and does atomicMax(&parent[i][v], u) for each of those
//     bits i, as the single-source top-down claim does.
//   * pull level k: the work axis (edges plus a fixed cost per row) is cut
//     into equal slices, one per wavefront; every vertex with sources not yet
//     reached scans its row (its part of the row, when a slice boundary cuts
//     it) against F_k.
//
// The pull walks rows in windows of kWinRows rows, 64 x kUnroll edges per
// pass, each row from its largest neighbour down (rows are sorted), so the
// first hit of a source bit in a row is its parent.  The first hit is found
// with a segmented prefix-OR over the lanes holding that row's edges; a row
// leaves the scan once every bit it needs is found.  Parents of a window go
// to an LDS tile [row][source] and leave as one coalesced store per source
// bit.  A row cut by a slice boundary is scanned by several wavefronts, each
// combining its local first hits with atomicMax (and atomicOr for F_{k+1}).
// Unsorted rows are scanned completely, every hit taking a max.
//
// Mask layout: the F arrays are stored by degree rank (rank[v] = position of
// v in decreasing-degree order), and adj_rank[e] = rank[adj[e]] is a second
// copy of the adjacency in rank ids.  The scan gathers F[adj_rank[e]] once per
// edge, and on skewed graphs most of those gathers hit the few high-degree
// vertices, whose masks are then packed into a few MB of L2 instead of one
// cache line each.  vis, the queues and the outputs keep the original ids;
// a hit reads inv_rank[rank] for the parent's original id (parents are
// mostly high-degree vertices, so those reads stay in L2; reading adj[e]
// instead cost one HBM line per hit).
// 32 rows per window: 16-row windows (LDS 37.4 -> 20.8 KB, 5 or 6 workgroups
// per CU with 96 or 80 VGPRs) measured 4.9-5.8 ms per batch against 4.47.
constexpr int kWinRows = 32;      // rows per window (lanes 0..31 own one each)
constexpr int kTileStride = 65;   // LDS words per tile row: column reads hit distinct banks
// 64-edge groups in flight per pass (3, 4, 6, 8 measured 4.49, 4.50, 4.92,
// 5.02 ms per batch on RMAT 2^22 x 64; from 6 up occupancy falls to 3
// wavefronts per SIMD)
constexpr int kUnroll = 4;
constexpr int kPass = kWave * kUnroll;
// Work units are cut at equal sums of (degree + kRowCost) per row: a row
// costs about as much as kRowCost edges (its window slot, the mask and
// offset loads, the output write-out) whatever its degree, and splitting by
// edges alone gave the wavefronts over low-degree ranges ~30x the rows.
constexpr int kRowCost = 32;   // swept over 8..128 on RMAT 2^22 x 64: 32 best
// A pull level tests the frontier bitmap before each mask gather when
// |F_k| < n / kFilterDiv.
// (1, 2, 4, 16 measured 4.82, 4.68, 4.48, 4.48 ms per batch)
constexpr int kFilterDiv = 4;
// When the previous pull left fewer than nrows / kFewRowsDiv rows with
// sources to find, a pull level splits whole rows evenly over the wavefronts:
// most rows are then skipped, at a cost per row, and an edge-weighted split
// piles the many rows of low-degree id ranges on a few wavefronts.
// (2, 4, 8 measured 4.49, 4.48, 4.49 ms per batch)
constexpr int kFewRowsDiv = 4;

template <typename OffT>
struct MbfsView {
  gptr<const OffT> off;
  gptr<const int32_t> adj;
  gptr<const int32_t> adj_rank;
  gptr<const int32_t> rank;
  gptr<const int32_t> inv_rank;
  gptr<int32_t> parent;
  gptr<int32_t> depth;
  gptr<uint8_t> dlev;                       // [n][64] level found, per vertex and source
  gptr<unsigned long long> vis;
  gptr<const unsigned long long> fr_cur;    // F_k, by rank
  gptr<unsigned long long> fr_next;         // F_{k+1}
  gptr<const unsigned> bm_cur;              // bitmaps (by rank) of the non-zero masks
  gptr<unsigned> bm_next;
  gptr<const int32_t> q_cur;
  gptr<const uint32_t> qe_cur;
  gptr<const unsigned long long> ctr_cur;
  gptr<int32_t> q_next;                     // this workgroup's sub-queue region
  gptr<uint32_t> qe_next;
  gptr<unsigned long long> ctr_next;
  gptr<int32_t> qover_next;                 // never set: sub-queues hold n entries each
  gptr<const int32_t> pull_rows;            // first list entry of each wavefront's slice
  gptr<const int32_t> rows;                 // vertices of degree > 0, ascending
  long long nrows;
  long long n;
  long long m;
  long long qcap;
  unsigned long long srcmask;
  int k;
};

struct MbfsLds {
  BatchStageBuf sb;
  unsigned unf[kWavesPerBlock];   // unfinished rows per wavefront
  int32_t tile[kWavesPerBlock][kWinRows * kTileStride];
};

// One row of a window, held by lane p < kWinRows.
template <typename OffT>
struct WinRow {
  int v;                      // vertex
  OffT lo, hi;                // adj[lo, hi) still to scan, from hi - 1 down
  unsigned long long rem;     // sources this row looks for
  unsigned long long found;   // sources found so far (this level)
  bool split;                 // a slice boundary cuts the row
};

__device__ __forceinline__ unsigned long long shfl64(unsigned long long x, int src) {
  return __shfl(x, src);
}

// Depths of the batched BFS go to a per-vertex byte table dlev[v][source]
// during the levels (a row's 64 bytes are contiguous, so a level touches one
// 64-byte piece per vertex it reaches), and reach depth[source][v] in one
// coalesced pass at the end (mbfs_depth_kernel).  Levels from kDepthEscape on
// are written straight to depth, with kDepthEscape in the table.
constexpr int kDepthEscape = 255;

template <typename OffT>
__device__ __forceinline__ void put_depth(const MbfsView<OffT>& a, int i, long long v, int d) {
  if (d < kDepthEscape) {
    a.dlev[v * kMaxSources + i] = (uint8_t)d;
  } else {
    a.dlev[v * kMaxSources + i] = (uint8_t)kDepthEscape;
    a.depth[(long long)i * a.n + v] = d;
  }
}


// 0xFF in each byte of a 32-bit word whose bit is set in the low 4 bits of b.
__device__ __forceinline__ unsigned byte_mask4(unsigned b) {
  return ((b & 1u) ? 0x000000FFu : 0u) | ((b & 2u) ? 0x0000FF00u : 0u) |
         ((b & 4u) ? 0x00FF0000u : 0u) | ((b & 8u) ? 0xFF000000u : 0u);
}

__device__ __forceinline__ unsigned long long wave_or64(unsigned long long x) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) x |= __shfl_xor(x, d);
  return x;
}

// One pass of a window scan: which row and adjacency position each lane
// takes (64 x kUnroll positions, rows in lane order), and the rank ids read
// there.
template <typename OffT>
struct ScanPass {
  uint32_t excl;   // lane p < kWinRows: first position of row p
  uint32_t cnt;    // lane p: positions row p takes
  uint32_t total;  // positions in use (wave-uniform)
  int rowp[kUnroll];
  OffT te[kUnroll];
  int ur[kUnroll];
};

// Assigns the next positions of the rows that still look for sources and
// issues their adjacency loads.  Rows are walked down from r.hi.
template <typename OffT>
__device__ __forceinline__ void scan_assign(const MbfsView<OffT>& a, const WinRow<OffT>& r,
                                            unsigned long long need, ScanPass<OffT>& P) {
  const int lane = (int)lane_id();
  P.cnt = (need != 0 && r.hi > r.lo) ? (uint32_t)min((OffT)(r.hi - r.lo), (OffT)kPass) : 0u;
  const uint32_t incl = wave_inclusive_sum_dpp(P.cnt);
  P.total = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
  P.excl = incl - P.cnt;
#pragma unroll
  for (int s = 0; s < kUnroll; ++s) {
    const uint32_t pos = (uint32_t)(s * kWave + lane);
    int p = 0;
#pragma unroll
    for (int st = kWinRows / 2; st >= 1; st >>= 1) {
      const uint32_t e = __shfl(P.excl, p + st);
      if (e <= pos) p += st;
    }
    const OffT hp = shfl_off(r.hi, p);
    const uint32_t ep = __shfl(P.excl, p);
    const bool valid = pos < P.total;
    P.rowp[s] = valid ? p : -1;
    P.te[s] = hp - 1 - (OffT)(pos - ep);
    // streamed once per level: non-temporal, so it does not evict the hot masks
    P.ur[s] = valid ? __builtin_nontemporal_load(&a.adj_rank[P.te[s]]) : -1;
  }
}

// The positions a pass took come off the top of each row.
template <typename OffT>
__device__ __forceinline__ void scan_consume(WinRow<OffT>& r, const ScanPass<OffT>& P) {
  const long long used = min((long long)P.excl + P.cnt, (long long)kPass) - (long long)P.excl;
  if (used > 0) r.hi -= (OffT)used;
}

// Scans the window's rows against the masks M; parents go to the tile (or,
// for split rows, straight to the outputs with atomicMax) with depth d.
// bm, when given, is the bitmap (by rank) of the vertices with a non-zero
// mask in M: a neighbour outside it is skipped without gathering its mask.
//
// Software-pipelined: pass p+1's positions are assigned, and its adjacency
// loads issued, before pass p's masks are gathered and its hits processed.
// Pass p+1 is assigned as if no row finished in pass p; a row that did finish
// finds nothing more in it (its need is 0) and drops out of pass p+2.
template <typename OffT, bool kSorted>
__device__ void scan_window(const MbfsView<OffT>& a, gptr<const unsigned long long> M,
                            gptr<const unsigned> bm, WinRow<OffT>& r, int32_t* tile, int d) {
  const int lane = (int)lane_id();
  if (!kSorted) {
    // every hit takes a max, so the tile cells this window uses start at -1
    for (unsigned long long b = r.rem; b; b &= b - 1)
      tile[lane * kTileStride + __builtin_ctzll(b)] = -1;
    wave_lds_sync();
  }
  ScanPass<OffT> cur, nxt;
  scan_assign(a, r, r.rem, cur);
  if (cur.total == 0) return;
  scan_consume(r, cur);
  for (;;) {
    const unsigned long long need = kSorted ? (r.rem & ~r.found) : r.rem;
    scan_assign(a, r, need, nxt);
    unsigned long long h[kUnroll];
    if (bm) {
      // the filter words are gathered unconditionally (idle positions read
      // word 0): a guarded load compiles to a branch with its own wait
      unsigned fw[kUnroll];
#pragma unroll
      for (int s = 0; s < kUnroll; ++s) fw[s] = bm[(cur.ur[s] < 0 ? 0 : cur.ur[s]) >> 5];
#pragma unroll
      for (int s = 0; s < kUnroll; ++s)
        cur.ur[s] = ((fw[s] >> (cur.ur[s] & 31)) & 1u) ? cur.ur[s] : -1;
    }
#pragma unroll
    for (int s = 0; s < kUnroll; ++s) {
      const unsigned long long np = shfl64(need, cur.rowp[s] < 0 ? 0 : cur.rowp[s]);
      h[s] = cur.ur[s] >= 0 ? (M[cur.ur[s]] & np) : 0ull;
    }
#pragma unroll
    for (int s = 0; s < kUnroll; ++s) {
      if (!__ballot(h[s] != 0ull)) continue;   // no hit in this group: nothing changes
      const int p = cur.rowp[s];
      const int ps = p < 0 ? 0 : p;
      // row p's lanes in this group: [first, last]; dist = lane - first
      const int g0 = s * kWave;
      const int ep = (int)__shfl(cur.excl, ps);
      const int first = max(ep - g0, 0);
      const int dist = p < 0 ? 0 : lane - first;
      // inclusive OR of h over the lanes of this group holding row p
      const unsigned long long x = seg_or_scan_d(h[s], dist, lane & 15, lane & 31);
      unsigned long long nb;
      if (kSorted) {
        // exclusive part: the lane below, if it holds the same row
        const unsigned plo = dpp_take<kWaveShr1, 0xf, true>((unsigned)x);
        const unsigned phi = dpp_take<kWaveShr1, 0xf, true>((unsigned)(x >> 32));
        const unsigned long long pre = dist >= 1 ? (((unsigned long long)phi << 32) | plo) : 0ull;
        nb = h[s] & ~pre & ~shfl64(r.found, ps);
      } else {
        nb = h[s];
      }
      const int sp = __shfl((int)r.split, ps);
      const int vp = __shfl(r.v, ps);   // shuffles stay outside divergent code
      if (nb) {
        const int u = a.inv_rank[cur.ur[s]];
        if (sp) {
          for (unsigned long long b = nb; b; b &= b - 1) {
            const int i = __builtin_ctzll(b);
            atomicMax(gen(&a.parent[(long long)i * a.n + vp]), u);
            put_depth(a, i, vp, d);
          }
        } else {
          int32_t* trow = tile + p * kTileStride;
          for (unsigned b = (unsigned)nb; b; b &= b - 1) {
            if (kSorted) trow[__builtin_ctz(b)] = u;
            else atomicMax(&trow[__builtin_ctz(b)], u);
          }
          for (unsigned b = (unsigned)(nb >> 32); b; b &= b - 1) {
            if (kSorted) trow[32 + __builtin_ctz(b)] = u;
            else atomicMax(&trow[32 + __builtin_ctz(b)], u);
          }
        }
      }
      // the last lane of row q's segment in this group holds the row's OR
      const int qf = max((int)cur.excl - g0, 0);
      const int qe = min((int)(cur.excl + cur.cnt) - g0, kWave);
      const int src = qe > qf ? qe - 1 : lane;
      const unsigned long long t = shfl64(x, src);
      if (qe > qf) r.found |= t;
    }
    if (nxt.total == 0) break;
    scan_consume(r, nxt);
    cur = nxt;
  }
}

// Stores the window's unsplit rows' parents (from the tile) and depths.
template <typename OffT>
__device__ __forceinline__ void write_window(const MbfsView<OffT>& a, const WinRow<OffT>& r,
                                             const int32_t* tile, int d) {
  const int lane = (int)lane_id();
  const unsigned long long my = (lane < kWinRows && !r.split) ? r.found : 0ull;
  unsigned long long all = wave_or64(my);
  wave_lds_sync();
  while (all) {
    const int i = __builtin_ctzll(all);
    all &= all - 1;
    if ((my >> i) & 1ull)
      __builtin_nontemporal_store(tile[lane * kTileStride + i], &a.parent[(long long)i * a.n + r.v]);
  }
  wave_lds_sync();
  if (my) {
    // this row's depth bytes: read, set the bytes of the sources found, write
    const gptr<u32x4> row = (gptr<u32x4>)(a.dlev + (long long)r.v * kMaxSources);
    const unsigned dv = d < kDepthEscape ? (unsigned)d : (unsigned)kDepthEscape;
    const unsigned d4 = dv * 0x01010101u;
    u32x4 w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) w[j] = row[j];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const unsigned m0 = byte_mask4((unsigned)(my >> (16 * j)) & 15u);
      const unsigned m1 = byte_mask4((unsigned)(my >> (16 * j + 4)) & 15u);
      const unsigned m2 = byte_mask4((unsigned)(my >> (16 * j + 8)) & 15u);
      const unsigned m3 = byte_mask4((unsigned)(my >> (16 * j + 12)) & 15u);
      w[j].x = (w[j].x & ~m0) | (d4 & m0);
      w[j].y = (w[j].y & ~m1) | (d4 & m1);
      w[j].z = (w[j].z & ~m2) | (d4 & m2);
      w[j].w = (w[j].w & ~m3) | (d4 & m3);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) row[j] = w[j];
    if (d >= kDepthEscape)
      for (unsigned long long b = my; b; b &= b - 1)
        a.depth[(long long)__builtin_ctzll(b) * a.n + r.v] = d;
  }
}

// vis |= F_k over F_k's queue entries.
template <typename OffT>
__device__ void mbfs_mark(const MbfsView<OffT>& a, unsigned long long total) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < (long long)total;
       t += stride) {
    long long idx = t;
    int sq = 0;
    for (; sq < kSubQueues - 1; ++sq) {
      const long long c = (uint32_t)a.ctr_cur[sq];
      if (idx < c) break;
      idx -= c;
    }
    const int v = a.q_cur[sq * a.qcap + idx];
    a.vis[v] = a.vis[v] | a.fr_cur[a.rank[v]];
  }
}

// Largest entry i < nf of a sub-queue with qe[i] + cost*i <= eb (64-ary
// search; the key is strictly increasing when cost > 0, non-decreasing else).
__device__ __forceinline__ unsigned find_entry(gptr<const uint32_t> qe, unsigned nf,
                                               unsigned long long eb, unsigned cost) {
  const unsigned lane = lane_id();
  unsigned lo = 0, hi = nf;
  while (hi - lo > (unsigned)kWave) {
    const unsigned step = (hi - lo + kWave - 1) / kWave;
    const unsigned idx = lo + lane * step;
    const bool pr = idx < hi && qe[idx] + (unsigned long long)cost * idx <= eb;
    const unsigned cnt = (unsigned)__popcll(__ballot(pr));
    lo += (cnt - 1) * step;
    hi = min(hi, lo + step);
  }
  const unsigned idx = lo + lane;
  const bool pr = idx < hi && qe[idx] + (unsigned long long)cost * idx <= eb;
  return lo + (unsigned)__popcll(__ballot(pr)) - 1;
}

// edges per lane per push step (1, 2, 4 measured 4.49, 4.50, 4.49 ms per
// batch: the first push is bound by its scattered atomics, not its loads)
constexpr int kPushPer = 2;

template <typename OffT>
__device__ void mbfs_push(const MbfsView<OffT>& a, BatchStageBuf& sb, Stage& stg,
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
    const gptr<const int32_t> q = a.q_cur + sq * a.qcap;
    const gptr<const uint32_t> qe = a.qe_cur + sq * a.qcap;
    const uint32_t eb = (uint32_t)(lo_e - (P - Es)), ee = (uint32_t)(hi_e - (P - Es));
    unsigned i = find_entry(qe, nf, eb, 0u);
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
        qm = a.fr_cur[a.rank[qv]];
      }
      const uint32_t wend = (i + kWave < nf) ? qe[i + kWave] : (uint32_t)Es;
      const uint32_t lim = min(wend, ee);
      // kPushPer edges per lane per step: every step issues the loads of all
      // of them (adjacency, then visited and frontier masks) before it uses
      // any, since a step is a chain of dependent round trips.
      for (uint32_t base = e; base < lim; base += kPushPer * kWave) {
        bool valid[kPushPer];
        int v[kPushPer], rv[kPushPer], uj[kPushPer];
        unsigned long long um[kPushPer];
#pragma unroll
        for (int t = 0; t < kPushPer; ++t) {
          const uint32_t my = base + (uint32_t)(t * kWave) + lane;
          valid[t] = (unsigned long long)base + t * kWave + lane < (unsigned long long)lim;
          int j = 0;
#pragma unroll
          for (int st = kWave / 2; st >= 1; st >>= 1) {
            const uint32_t qt = __shfl(qo, j + st);
            if (qt <= my) j += st;
          }
          const OffT rsj = shfl_off(rs, j);
          const uint32_t qoj = __shfl(qo, j);
          um[t] = __shfl(qm, j);
          uj[t] = __shfl(qv, j);
          // idle lanes read entry 0 (m >= 1 whenever a push level has edges;
          // upload checked every entry is a vertex id) and are masked below
          const OffT ed = valid[t] ? rsj + (OffT)(my - qoj) : (OffT)0;
          v[t] = a.adj[ed];
          rv[t] = a.adj_rank[ed];
        }
        unsigned long long nbv[kPushPer];
#pragma unroll
        for (int t = 0; t < kPushPer; ++t) {
          const unsigned long long seen = a.vis[v[t]] | a.fr_cur[rv[t]];
          nbv[t] = um[t] & ~seen & a.srcmask & (0ull - (unsigned long long)valid[t]);
        }
#pragma unroll
        for (int t = 0; t < kPushPer; ++t) {
          bool fresh = false;
          const unsigned long long nb = nbv[t];
          if (nb) {
            const unsigned long long old = atomicOr(gen(&a.fr_next[rv[t]]), nb);
            fresh = old == 0ull;
            if (fresh) atomicOr(gen(&a.bm_next[rv[t] >> 5]), 1u << (rv[t] & 31));
            // every frontier edge into a new bit competes; the max stays
            for (unsigned long long b = nb; b; b &= b - 1) {
              const int i = __builtin_ctzll(b);
              const long long o = (long long)i * a.n + v[t];
              atomicMax(gen(&a.parent[o]), uj[t]);
              if (!((old >> i) & 1ull)) put_depth(a, i, v[t], a.k + 1);   // the bit's first setter
            }
          }
          if (__ballot(fresh)) {
            uint32_t deg = 0;
            if (fresh) deg = (uint32_t)(a.off[v[t] + 1] - a.off[v[t]]);
            stage_push(a, sb, stg, fresh, v[t], deg);
          }
        }
      }
      e = lim;
      i += kWave;
    }
  }
}

// Pull over this wavefront's slice of the work axis.  The rows are the
// vertices of degree > 0 in ascending order (a degree-0 vertex is never
// reached from another), and entry i (vertex v) spans
// [off[v] + kRowCost*i, off[v'] + kRowCost*(i+1)) with v' the next entry:
// its edges, then kRowCost units for the row itself.
//
// by_rows: wavefront w takes whole rows [nrows*w/W, nrows*(w+1)/W) instead.
// Rows left with sources to find are counted into *unfinished.
template <typename OffT, bool kSorted>
__device__ void mbfs_pull(const MbfsView<OffT>& a, BatchStageBuf& sb, Stage& stg, int32_t* tile,
                          gptr<const unsigned> filter, bool by_rows,
                          unsigned long long* unfinished, unsigned* unf_lds) {
  const int lane = (int)lane_id();
  const int w = threadIdx.x / kWave;
  const unsigned long long W = (unsigned long long)num_waves();
  const unsigned long long gw = (unsigned long long)global_wave();
  const unsigned long long X = (unsigned long long)a.m + (unsigned long long)kRowCost * a.nrows;
  long long x0, x1, iend;
  if (by_rows) {
    x0 = -1;   // rows are whole: no edge slice cuts them
    x1 = LLONG_MAX;
    iend = (long long)((unsigned long long)a.nrows * (gw + 1) / W);
  } else {
    x0 = (long long)(X * gw / W);
    x1 = (long long)(X * (gw + 1) / W);
    iend = a.nrows;
  }
  unsigned left = 0;   // rows still looking for sources (lane 0's count)
  if (lane == 0) unf_lds[w] = 0;
  // Row data of the next window is loaded while the current one is scanned:
  // vertex, offsets, reached mask and rank first, the frontier mask (which
  // needs the rank) after the scan.
  long long p0 = 0, p1 = 0;
  unsigned long long pvis = 0, pfk = 0;
  int prk = 0, pv = 0;
  auto load_rows = [&](long long ibase) {
    const long long i = ibase + lane;
    if (lane < kWinRows && i < a.nrows) {
      pv = a.rows[i];
      p0 = (long long)a.off[pv];
      p1 = (long long)a.off[pv + 1];
      pvis = a.vis[pv];
      prk = a.rank[pv];
    } else {
      pv = 0;
      p0 = p1 = 0;
      pvis = ~0ull;
      prk = 0;
    }
  };
  auto load_mask = [&]() {
    pfk = 0;
    if (p1 > p0 && (a.srcmask & ~pvis)) pfk = a.fr_cur[prk];
  };
  long long i0 = by_rows ? (long long)((unsigned long long)a.nrows * gw / W)
                         : (x0 < x1 ? (long long)a.pull_rows[gw] : iend);
  load_rows(i0);
  load_mask();
  for (;; i0 += kWinRows) {
    const long long i = i0 + lane;
    const int v = pv;
    const long long o0 = p0, o1 = p1;
    const unsigned long long vis = pvis, fk = pfk;
    const int rv = prk;
    const bool mine = lane < kWinRows && i < iend && o0 + kRowCost * i < x1;
    if (!__ballot(mine)) break;
    load_rows(i0 + kWinRows);
    WinRow<OffT> r;
    r.v = v;
    r.lo = r.hi = 0;
    r.rem = 0;
    r.found = 0;
    r.split = false;
    if (mine) {
      const long long lo = max(o0, x0 - kRowCost * i);
      const long long hi = min(o1, x1 - kRowCost * i);
      if (hi > lo) {
        r.lo = (OffT)lo;
        r.hi = (OffT)hi;
        r.split = lo > o0 || hi < o1;
        r.rem = a.srcmask & ~(vis | fk);
      }
    }
    if (!__ballot(r.rem != 0ull)) {   // every row of the window is done
      load_mask();
      continue;
    }
    scan_window<OffT, kSorted>(a, a.fr_cur, filter, r, tile, a.k + 1);
    left += (unsigned)__popcll(__ballot((r.rem & ~r.found) != 0ull && r.lo == (OffT)o0));
    load_mask();
    bool take = false;
    if (r.found) {
      if (!r.split) {
        a.fr_next[rv] = r.found;
        take = true;
      } else {
        take = atomicOr(gen(&a.fr_next[rv]), r.found) == 0ull;
      }
      if (take) atomicOr(gen(&a.bm_next[rv >> 5]), 1u << (rv & 31));
    }
    stage_push(a, sb, stg, take, v, take ? (uint32_t)(o1 - o0) : 0u);
    write_window(a, r, tile, a.k + 1);
  }
  if (lane == 0) unf_lds[w] = left;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int j = 0; j < kWavesPerBlock; ++j) t += unf_lds[j];
    if (t) atomicAdd(unfinished, t);
  }
}

template <typename OffT>
__device__ __forceinline__ int mbfs_mode(const MbfsArgs<OffT>& A, unsigned long long cnt,
                                         unsigned long long E) {
  if (A.td_to_bu_div <= 0 || A.pull_div <= 0) return kTopDown;
  return (cnt + E) * (unsigned long long)A.pull_div > (unsigned long long)A.m ? kBottomUp
                                                                               : kTopDown;
}

template <typename OffT, bool kSorted>
__global__ __launch_bounds__(kBlock) void mbfs_level_kernel(const MbfsArgs<OffT>* __restrict__ A) {
  __shared__ MbfsLds lds;
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
  const int mode = mbfs_mode(*A, cnt, E);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->modes[k & 3] = mode;
    st->max_depth = k;
    if (mode == kTopDown) st->push_levels += 1; else st->pull_levels += 1;
#pragma unroll
    for (int s = 0; s < kSubQueues; ++s) st->ctr[c2][s] = 0;
    st->unfinished[c2] = 0;
    __hip_atomic_store(&A->host_status[1], k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  {
    // F_{k+2}'s masks held F_{k-2}: level k reads F_k's masks and writes F_{k+1}'s only.
    const gptr<unsigned long long> z = glob(A->fr[(k + 2) & 3]);
    const gptr<unsigned long long> zb = glob(A->fbm[(k + 2) & 3]);
    const long long n = A->n, nw = A->nwords;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
      z[v] = 0;
      if (v < nw) zb[v] = 0;
    }
  }
  const int sq = blockIdx.x % kSubQueues;
  MbfsView<OffT> a;
  a.off = glob(A->off);
  a.adj = glob(A->adj);
  a.adj_rank = glob(A->adj_rank);
  a.rank = glob(A->rank);
  a.inv_rank = glob(A->inv_rank);
  a.parent = glob(A->parent);
  a.depth = glob(A->depth);
  a.dlev = glob(A->dlev);
  a.vis = glob(A->vis);
  a.fr_cur = glob(A->fr[k & 3]);
  a.fr_next = glob(A->fr[(k + 1) & 3]);
  a.bm_cur = glob((const unsigned*)A->fbm[k & 3]);
  a.bm_next = glob((unsigned*)A->fbm[(k + 1) & 3]);
  a.qcap = A->qcap;
  a.q_cur = glob(A->q[k & 1]);
  a.qe_cur = glob(A->qe[k & 1]);
  a.ctr_cur = glob(st->ctr[c0]);
  a.q_next = glob(A->q[(k + 1) & 1] + sq * a.qcap);
  a.qe_next = glob(A->qe[(k + 1) & 1] + sq * a.qcap);
  a.ctr_next = glob(&st->ctr[c1][sq]);
  a.qover_next = glob(&st->qover[c1]);
  a.pull_rows = glob(A->pull_rows);
  a.rows = glob(A->rows);
  a.nrows = A->nrows;
  a.n = A->n;
  a.m = A->m;
  a.srcmask = A->srcmask;
  a.k = k;
  int32_t* tile = lds.tile[threadIdx.x / kWave];
  slot_priority(A->num_cus);
  mbfs_mark(a, cnt);
  Stage stg;
  if (mode == kTopDown)
    mbfs_push(a, lds.sb, stg, E);
  else
  {
    // a sparse union frontier is tested by its bitmap before gathering masks;
    // after a pull that left few rows unfinished, rows are split evenly
    const int prev = k == 0 ? kTopDown : st->modes[(k - 1) & 3];
    const bool by_rows =
        prev == kBottomUp && st->unfinished[c0] * kFewRowsDiv < (unsigned long long)a.nrows;
    mbfs_pull<OffT, kSorted>(a, lds.sb, stg, tile,
                             cnt * kFilterDiv < (unsigned long long)a.n ? a.bm_cur : nullptr,
                             by_rows, &st->unfinished[c1], lds.unf);
  }
  stage_flush_block(a, lds.sb, stg);
  if (threadIdx.x == 0) A->wg_level[blockIdx.x] = (uint32_t)(k + 1);
}

// Degree ranks: keys[v] = ~deg(v) so an ascending sort is decreasing degree.
template <typename OffT>
__global__ void degree_key_kernel(const OffT* __restrict__ off, long long n,
                                  uint32_t* __restrict__ keys, int32_t* __restrict__ ids) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    keys[v] = ~(uint32_t)(off[v + 1] - off[v]);
    ids[v] = (int32_t)v;
  }
}

__global__ void rank_scatter_kernel(const int32_t* __restrict__ order, long long n,
                                    int32_t* __restrict__ rank) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride)
    rank[order[r]] = (int32_t)r;
}

__global__ void adj_rank_kernel(const int32_t* __restrict__ adj, long long m,
                                const int32_t* __restrict__ rank, int32_t* __restrict__ out) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride)
    out[e] = rank[adj[e]];
}

// First list entry of each wavefront's slice of the pull work axis (entry i,
// vertex rows[i], starts at off[rows[i]] + kRowCost*i): out[w] = the entry
// whose span holds floor(X*w/W), X = m + kRowCost*nrows; out[W] = nrows.
template <typename OffT>
__global__ void pull_rows_kernel(const OffT* __restrict__ off, const int32_t* __restrict__ rows,
                                 long long nrows, long long m, long long waves,
                                 int32_t* __restrict__ out) {
  const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (w > waves) return;
  const unsigned long long X = (unsigned long long)m + (unsigned long long)kRowCost * nrows;
  const unsigned long long x = X * (unsigned long long)w / (unsigned long long)waves;
  if (w == waves || x >= X || nrows == 0) {
    out[w] = (int32_t)nrows;
    return;
  }
  long long lo = 0, hi = nrows - 1;   // largest i with off[rows[i]] + kRowCost*i <= x
  while (lo < hi) {
    const long long mid = (lo + hi + 1) / 2;
    if ((unsigned long long)off[rows[mid]] + (unsigned long long)kRowCost * mid <= x) lo = mid;
    else hi = mid - 1;
  }
  out[w] = (int32_t)lo;
}

__global__ void nonzero_flag_kernel(const unsigned* __restrict__ keys, long long n,
                                    uint8_t* __restrict__ flags) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
    flags[v] = keys[v] != ~0u;   // key = ~degree
}

// Clears the masks and presets every output to "unreached".
template <typename OffT>
__global__ __launch_bounds__(kBlock) void mbfs_clear_kernel(MbfsArgs<OffT> a) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  // F_0 (seeded next) and F_1 (written by level 0) start empty.  F_2 and F_3
  // are cleared by levels 0 and 1, as every level k clears F_{k+2} before
  // level k+1 writes it, and no level reads F_{k-1}.
  for (long long v = tid; v < a.n; v += nth) {
    a.vis[v] = 0;
    a.fr[0][v] = 0;
    a.fr[1][v] = 0;
  }
  for (long long w = tid; w < a.nwords; w += nth) {
    a.fbm[0][w] = 0;
    a.fbm[1][w] = 0;
  }
  const long long total = a.preset_parents ? 0 : (long long)a.nsrc * a.n;   // depth: mbfs_depth_kernel
  if ((reinterpret_cast<uintptr_t>(a.parent) & 15) == 0 && (total & 3) == 0) {
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    const i32x4 m1 = {-1, -1, -1, -1};
    i32x4* p4 = reinterpret_cast<i32x4*>(a.parent);
    for (long long t = tid; t < total / 4; t += nth) __builtin_nontemporal_store(m1, &p4[t]);
  } else {
    for (long long t = tid; t < total; t += nth) __builtin_nontemporal_store(-1, &a.parent[t]);
  }
  for (long long i = tid; i < kMaxGrid; i += nth) a.wg_level[i] = 0;
}

// Fills p[0, count) with -1 on a small grid (16-byte non-temporal stores):
// it runs on a side stream beside a batch's levels, so it takes few
// workgroup slots and trickles over the batch's duration.
__global__ __launch_bounds__(kBlock) void fill_minus1_kernel(int32_t* __restrict__ p,
                                                             long long count) {
  typedef int i32x4 __attribute__((ext_vector_type(4)));
  const i32x4 m1 = {-1, -1, -1, -1};
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nth = (long long)gridDim.x * blockDim.x;
  const long long head = (long long)((16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15) / 4;
  const long long h = head < count ? head : count;
  for (long long i = tid; i < h; i += nth) p[i] = -1;
  i32x4* p4 = reinterpret_cast<i32x4*>(p + h);
  const long long n4 = (count - h) / 4;
  for (long long i = tid; i < n4; i += nth) __builtin_nontemporal_store(m1, &p4[i]);
  for (long long i = h + 4 * n4 + tid; i < count; i += nth) p[i] = -1;
}

hipError_t launch_fill_minus1(int32_t* p, long long count, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(fill_minus1_kernel, dim3(grid), dim3(kBlock), 0, stream, p, count);
  return hipGetLastError();
}

// One workgroup: marks the sources, queues the distinct ones as F_0 and
// resets the control block.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void mbfs_seed_kernel(MbfsArgs<OffT> a, MbfsArgs<OffT>* dst) {
  __shared__ int first[kMaxSources];
  __shared__ int src[kMaxSources];
  __shared__ unsigned long long deg[kMaxSources];
  const int t = threadIdx.x;
  if (t < a.nsrc) {
    const int s = a.sources[t];
    src[t] = s;
    deg[t] = (unsigned long long)(a.off[s + 1] - a.off[s]);
    const unsigned long long bit = 1ull << t;
    a.parent[(long long)t * a.n + s] = s;
    a.dlev[(long long)s * kMaxSources + t] = 0;
    const int rs = a.rank[s];
    atomicOr(&a.vis[s], bit);
    first[t] = atomicOr(&a.fr[0][rs], bit) == 0ull;
    atomicOr(reinterpret_cast<unsigned*>(a.fbm[0]) + (rs >> 5), 1u << (rs & 31));
  }
  __syncthreads();
  if (t == 0) {
    *dst = a;
    MbfsState* st = a.st;
    unsigned cnt = 0;
    unsigned long long E = 0;
    // the sources' degrees were loaded by their own threads above: this
    // loop reads LDS only
    for (int i = 0; i < a.nsrc; ++i) {
      if (!first[i]) continue;
      a.q[0][cnt] = src[i];
      a.qe[0][cnt] = (uint32_t)E;
      E += deg[i];
      ++cnt;
    }
    for (int j = 0; j < 3; ++j) {
      for (int q = 0; q < kSubQueues; ++q) st->ctr[j][q] = 0;
      st->unfinished[j] = 0;
      st->qover[j] = 0;
    }
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

// End of a batched BFS: depth[i][v] for every source i and vertex v, from
// the level table (n where source i never reached v; entries from
// kDepthEscape on were written during the levels).  A lane takes 4
// consecutive vertices, so each source's store is 16 bytes per lane and 1 KB
// per wavefront.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void mbfs_depth_kernel(MbfsArgs<OffT> a) {
  // A workgroup takes tiles of 256 vertices: their 64-byte depth rows are
  // staged in LDS with coalesced 16-byte loads (row stride 17 dwords), then
  // wavefront w writes sources [16w, 16w + 16) for the tile, lane l taking
  // vertices 4l .. 4l + 3 (16-byte stores, 1 KB contiguous per source).
  constexpr int kRowDw = 17;
  static_assert(kWavesPerBlock * 16 == kMaxSources, "four wavefronts of 16 sources each");
  __shared__ uint32_t rows[kBlock * kRowDw];
  __shared__ unsigned long long rvm_lds[kBlock];
  const gptr<const unsigned long long> vis = glob((const unsigned long long*)a.vis);
  const gptr<const u32x4> dl = glob((const u32x4*)a.dlev);
  const gptr<int32_t> depth = glob(a.depth);
  const int ns = a.nsrc;
  const int32_t nval = (int32_t)a.n;
  const unsigned lane = lane_id();
  const int wv = threadIdx.x / kWave;
  const long long tiles = (a.n + kBlock - 1) / kBlock;
  unsigned long long pairs = 0, edges = 0;   // reached (source, vertex) pairs, their degrees
  for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const long long vb = tile * kBlock;
    __syncthreads();   // the previous tile's readers are done with the LDS rows
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int c = r * kBlock + (int)threadIdx.x;   // 16-byte chunk c of the tile
      const long long v = vb + c / 4;
      const u32x4 x = dl[(v < a.n ? v : 0) * 4 + (c & 3)];
      uint32_t* d = &rows[(c / 4) * kRowDw + (c & 3) * 4];
      d[0] = x.x;
      d[1] = x.y;
      d[2] = x.z;
      d[3] = x.w;
    }
    {
      const long long v = vb + threadIdx.x;
      const unsigned long long m =
          vis[v < a.n ? v : 0] & a.srcmask & (0ull - (unsigned long long)(v < a.n));
      rvm_lds[threadIdx.x] = m;
      if (m) {
        const unsigned long long c = (unsigned long long)__popcll(m);
        pairs += c;
        edges += c * (unsigned long long)(a.off[v + 1] - a.off[v]);
      }
    }
    __syncthreads();
    const int vl = (int)lane * 4;
    const long long v0 = vb + vl;
    const bool whole = v0 + 4 <= a.n && (a.n & 3) == 0;   // 16-byte aligned rows of depth
    unsigned long long rvm[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) rvm[t] = rvm_lds[vl + t];
#pragma unroll
    for (int jd = 0; jd < 4; ++jd) {   // the wavefront's 16 sources, one dword of 4 at a time
      const int i0 = 16 * wv + 4 * jd;
      if (i0 >= ns) break;
      uint32_t word[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) word[t] = rows[(vl + t) * kRowDw + i0 / 4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int i = i0 + b;
        if (i >= ns) break;
        int32_t d[4];
        bool keep[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const unsigned lv = (word[t] >> (8 * b)) & 0xFFu;
          const bool hit = (rvm[t] >> i) & 1ull;
          d[t] = hit ? (int32_t)lv : nval;
          keep[t] = hit && lv == (unsigned)kDepthEscape;   // written during the levels
        }
        const long long o = (long long)i * a.n + v0;
        if (whole && !(keep[0] | keep[1] | keep[2] | keep[3])) {
          typedef int i32x4 __attribute__((ext_vector_type(4)));
          const i32x4 dv = {d[0], d[1], d[2], d[3]};
          __builtin_nontemporal_store(dv, (gptr<i32x4>)(depth + o));
        } else {
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (v0 + t < a.n && !keep[t]) depth[o + t] = d[t];
        }
      }
    }
  }
  // totals over the reached (source, vertex) pairs: count and degree sum
  __shared__ unsigned long long sp[kWavesPerBlock], se[kWavesPerBlock];
#pragma unroll
  for (int dd = 1; dd < kWave; dd <<= 1) {
    pairs += __shfl_xor(pairs, dd);
    edges += __shfl_xor(edges, dd);
  }
  if (lane == 0) {
    sp[wv] = pairs;
    se[wv] = edges;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long P = 0, Ed = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
      P += sp[i];
      Ed += se[i];
    }
    if (P) {
      atomicAdd(&a.st->reached_pairs, P);
      atomicAdd(&a.st->edges_pairs, Ed);
    }
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
// Per frontier word: sum over its vertices of degree 0 ? 0 : 1 + 2 floor(log2(1 + deg)).
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
      if (deg) x = 1u + 2u * (63u - (unsigned)__builtin_clzll(deg + 1));
    }
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) x += __shfl_xor(x, d);
    if (lane == 0) ww[b] = x;
  }
}

// Row-top records for sorted rows, two 16-byte halves per vertex:
// out[2v] = {adj[hi-1], adj[hi-2], adj[hi-3], adj[hi-4]}, out[2v+1] =
// {adj[hi-5], adj[hi-6], hi, degree}, hi = off[v+1], ~0 where the row is
// shorter.  hi fits 32 bits (m < 2^32).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void row_tops_kernel(const OffT* __restrict__ off,
                                                          const int32_t* __restrict__ adj,
                                                          long long n, u32x4* __restrict__ out) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    const OffT lo = off[v], hi = off[v + 1];
    const OffT d = hi - lo;
    unsigned t[kTopEntries];
#pragma unroll
    for (int j = 0; j < kTopEntries; ++j) t[j] = d > (OffT)j ? (unsigned)adj[hi - 1 - j] : ~0u;
    out[2 * v] = (u32x4){t[0], t[1], t[2], t[3]};
    out[2 * v + 1] = (u32x4){t[4], t[5], (unsigned)hi, (unsigned)d};
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
const void* level_kernel_symbol(int variant) {
  switch (variant) {
    case kLevelAnyParent:
      return reinterpret_cast<const void*>(&bfs_level_kernel<OffT, true, false, true>);
    case kLevelSortedTops: return reinterpret_cast<const void*>(&bfs_level_kernel<OffT, true, true>);
    case kLevelSorted: return reinterpret_cast<const void*>(&bfs_level_kernel<OffT, true, false>);
    default: return reinterpret_cast<const void*>(&bfs_level_kernel<OffT, false, false>);
  }
}

template <typename OffT>
const void* persist_kernel_symbol(int variant) {
  switch (variant) {
    case kLevelAnyParent:
      return reinterpret_cast<const void*>(&bfs_persist_kernel<OffT, true, false, true>);
    case kLevelSortedTops: return reinterpret_cast<const void*>(&bfs_persist_kernel<OffT, true, true>);
    case kLevelSorted: return reinterpret_cast<const void*>(&bfs_persist_kernel<OffT, true, false>);
    default: return reinterpret_cast<const void*>(&bfs_persist_kernel<OffT, false, false>);
  }
}

template <typename OffT>
hipError_t launch_bfs_persist(const BfsArgs<OffT>* dargs, int variant, int grid,
                              hipStream_t stream) {
  if (variant == kLevelAnyParent)
    hipLaunchKernelGGL((bfs_persist_kernel<OffT, true, false, true>), dim3(grid), dim3(kBlock), 0,
                       stream, dargs);
  else if (variant == kLevelSortedTops)
    hipLaunchKernelGGL((bfs_persist_kernel<OffT, true, true>), dim3(grid), dim3(kBlock), 0, stream,
                       dargs);
  else if (variant == kLevelSorted)
    hipLaunchKernelGGL((bfs_persist_kernel<OffT, true, false>), dim3(grid), dim3(kBlock), 0, stream,
                       dargs);
  else
    hipLaunchKernelGGL((bfs_persist_kernel<OffT, false, false>), dim3(grid), dim3(kBlock), 0,
                       stream, dargs);
  return hipGetLastError();
}

template <typename OffT>
const void* multi_kernel_symbol(bool sorted) {
  return sorted ? reinterpret_cast<const void*>(&bfs_multi_kernel<OffT, true>)
                : reinterpret_cast<const void*>(&bfs_multi_kernel<OffT, false>);
}

template <typename OffT>
hipError_t launch_bfs_multi(const BfsArgs<OffT>* args, const uint32_t* wg_begin, int njobs,
                            bool sorted, int grid, hipStream_t stream) {
  if (sorted)
    hipLaunchKernelGGL((bfs_multi_kernel<OffT, true>), dim3(grid), dim3(kBlock), 0, stream, args,
                       wg_begin, njobs);
  else
    hipLaunchKernelGGL((bfs_multi_kernel<OffT, false>), dim3(grid), dim3(kBlock), 0, stream, args,
                       wg_begin, njobs);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_bfs_level(const BfsArgs<OffT>* dargs, int variant, int grid, hipStream_t stream) {
  if (variant == kLevelAnyParent)
    hipLaunchKernelGGL((bfs_level_kernel<OffT, true, false, true>), dim3(grid), dim3(kBlock), 0,
                       stream, dargs);
  else if (variant == kLevelSortedTops)
    hipLaunchKernelGGL((bfs_level_kernel<OffT, true, true>), dim3(grid), dim3(kBlock), 0, stream, dargs);
  else if (variant == kLevelSorted)
    hipLaunchKernelGGL((bfs_level_kernel<OffT, true, false>), dim3(grid), dim3(kBlock), 0, stream, dargs);
  else
    hipLaunchKernelGGL((bfs_level_kernel<OffT, false, false>), dim3(grid), dim3(kBlock), 0, stream, dargs);
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
  hipError_t e = rocprim::exclusive_scan(nullptr, bytes, in, out, 0ull, (size_t)(n + 1),
                                         rocprim::plus<unsigned long long>(), stream);
  if (e != hipSuccess) return e;
  void* tmp = nullptr;
  e = hipMalloc(&tmp, bytes);
  if (e != hipSuccess) return e;
  e = rocprim::exclusive_scan(tmp, bytes, in, out, 0ull, (size_t)(n + 1),
                              rocprim::plus<unsigned long long>(), stream);
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
hipError_t launch_row_tops(const OffT* off, const int32_t* adj, long long n, uint32_t* out,
                           int grid, hipStream_t stream) {
  hipLaunchKernelGGL(row_tops_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, off, adj, n,
                     reinterpret_cast<u32x4*>(out));
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
const void* mbfs_level_symbol(bool sorted) {
  return sorted ? reinterpret_cast<const void*>(&mbfs_level_kernel<OffT, true>)
                : reinterpret_cast<const void*>(&mbfs_level_kernel<OffT, false>);
}

template <typename OffT>
hipError_t launch_mbfs_level(const MbfsArgs<OffT>* dargs, bool sorted, int grid,
                             hipStream_t stream) {
  if (sorted)
    hipLaunchKernelGGL((mbfs_level_kernel<OffT, true>), dim3(grid), dim3(kBlock), 0, stream, dargs);
  else
    hipLaunchKernelGGL((mbfs_level_kernel<OffT, false>), dim3(grid), dim3(kBlock), 0, stream,
                       dargs);
  return hipGetLastError();
}

// rank[v] = position of v in decreasing-degree order (ties by id),
// inv_rank[r] = the vertex of rank r, adj_rank[e] = rank[adj[e]], and
// rows[0..*nrows) = the vertices of degree > 0 in ascending order.
// Synchronises the stream.
template <typename OffT>
hipError_t build_degree_ranks(const OffT* off, const int32_t* adj, long long n, long long m,
                              int32_t* rank, int32_t* inv_rank, int32_t* adj_rank, int32_t* rows,
                              long long* nrows, int grid, hipStream_t stream) {
  uint32_t *keys = nullptr, *keys2 = nullptr;
  int32_t *ids = nullptr;
  int32_t* order = inv_rank;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  hipError_t e = hipMalloc(&keys, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&keys2, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&ids, n * sizeof(int32_t));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(degree_key_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, off, n, keys,
                       ids);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, ids, order, (size_t)n, 0, 32,
                                  stream);
  if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes);
  if (e == hipSuccess)
    e = rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys2, ids, order, (size_t)n, 0, 32,
                                  stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(rank_scatter_kernel, dim3(grid), dim3(kBlock), 0, stream, order, n, rank);
    e = hipGetLastError();
  }
  if (e == hipSuccess && m > 0) {
    hipLaunchKernelGGL(adj_rank_kernel, dim3(grid), dim3(kBlock), 0, stream, adj, m, rank,
                       adj_rank);
    e = hipGetLastError();
  }
  // the degree-0 test reuses keys (~degree); keys2 is free again for flags
  uint8_t* flags = reinterpret_cast<uint8_t*>(keys2);
  int32_t* count = reinterpret_cast<int32_t*>(ids);   // ids is free after the sort
  void* tmp2 = nullptr;
  size_t tmp2_bytes = 0;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(nonzero_flag_kernel, dim3(grid), dim3(kBlock), 0, stream, keys, n, flags);
    e = hipGetLastError();
  }
  const rocprim::counting_iterator<int32_t> it(0);
  if (e == hipSuccess)
    e = rocprim::select(nullptr, tmp2_bytes, it, flags, rows, count, (size_t)n, stream);
  if (e == hipSuccess) e = hipMalloc(&tmp2, tmp2_bytes);
  if (e == hipSuccess)
    e = rocprim::select(tmp2, tmp2_bytes, it, flags, rows, count, (size_t)n, stream);
  int32_t hcount = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&hcount, count, sizeof hcount, hipMemcpyDeviceToHost, stream);
  const hipError_t se = hipStreamSynchronize(stream);
  *nrows = hcount;
  (void)hipFree(tmp2);
  (void)hipFree(keys);
  (void)hipFree(keys2);
  (void)hipFree(ids);
  (void)hipFree(tmp);
  return e != hipSuccess ? e : se;
}

template <typename OffT>
hipError_t launch_pull_rows(const OffT* off, const int32_t* rows, long long nrows, long long m,
                            long long waves, int32_t* out, hipStream_t stream) {
  const long long threads = waves + 1;
  hipLaunchKernelGGL(pull_rows_kernel<OffT>, dim3((unsigned)((threads + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, stream, off, rows, nrows, m, waves, out);
  return hipGetLastError();
}

template <typename OffT>
hipError_t launch_mbfs_depth(const MbfsArgs<OffT>& a, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(mbfs_depth_kernel<OffT>, dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}



#define PBBS_INSTANTIATE(OffT)                                                                   \
  template hipError_t launch_bfs_init<OffT>(const BfsArgs<OffT>&, BfsArgs<OffT>*,                \
                                            const unsigned long long*, int32_t, int,             \
                                            hipStream_t);                                        \
  template const void* level_kernel_symbol<OffT>(int);                                          \
  template hipError_t launch_bfs_level<OffT>(const BfsArgs<OffT>*, int, int, hipStream_t);       \
  template const void* persist_kernel_symbol<OffT>(int);                                        \
  template const void* multi_kernel_symbol<OffT>(bool);                                         \
  template hipError_t launch_bfs_multi<OffT>(const BfsArgs<OffT>*, const uint32_t*, int, bool,   \
                                            int, hipStream_t);                                   \
  template hipError_t launch_bfs_persist<OffT>(const BfsArgs<OffT>*, int, int, hipStream_t);     \
  template hipError_t launch_graph_offsets<OffT>(const OffT*, long long, long long,             \
                                                 unsigned long long*, unsigned long long*, int,  \
                                                 hipStream_t);                                   \
  template hipError_t launch_graph_rows<OffT>(const OffT*, const int32_t*, long long, long long, \
                                              unsigned long long*, int, hipStream_t);            \
  template hipError_t launch_mbfs_setup<OffT>(const MbfsArgs<OffT>&, MbfsArgs<OffT>*, int,      \
                                             hipStream_t);                                      \
  template const void* mbfs_level_symbol<OffT>(bool);                                           \
  template hipError_t launch_mbfs_level<OffT>(const MbfsArgs<OffT>*, bool, int, hipStream_t);    \
  template hipError_t launch_pull_rows<OffT>(const OffT*, const int32_t*, long long, long long,  \
                                             long long, int32_t*, hipStream_t);                  \
  template hipError_t build_degree_ranks<OffT>(const OffT*, const int32_t*, long long, long long, \
                                               int32_t*, int32_t*, int32_t*, int32_t*,           \
                                               long long*, int, hipStream_t);                    \
  template hipError_t launch_mbfs_depth<OffT>(const MbfsArgs<OffT>&, int, hipStream_t);          \
  template hipError_t launch_word_weights<OffT>(const OffT*, long long, unsigned long long*, int, \
                                               hipStream_t);                                    \
  template hipError_t launch_graph_symmetry<OffT>(const OffT*, const int32_t*, long long,        \
                                                  long long, unsigned long long*, int,           \
                                                  hipStream_t);                                  \
  template hipError_t launch_row_tops<OffT>(const OffT*, const int32_t*, long long, uint32_t*,   \
                                            int, hipStream_t);

PBBS_INSTANTIATE(uint32_t)
PBBS_INSTANTIATE(int64_t)

}  // namespace pbbs
