// Shared between the HIP kernels (bfs_kernels.hip) and the host engine
// (engine.cpp): the device-resident BFS state and the argument block the
// level kernels read.
#pragma once
#include <stdint.h>

namespace pbbs {

constexpr int kWave = 64;             // CDNA wavefront width
constexpr int kBlock = 256;           // threads per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kSubQueues = 8;         // frontier sub-queues (one append counter each)
constexpr int kWaveBuf = 512;         // LDS staging entries per wavefront
constexpr int kMaxGrid = 1 << 16;     // level-kernel workgroups at most
constexpr int kLevelLogMax = 4096;    // per-level records kept for stats/tuning
constexpr int kBarGroups = 32;        // first-tier counters of the persistent launch's grid barrier

enum : int32_t { kTopDown = 0, kBottomUp = 1 };

// One record per executed level: direction and the frontier it consumed.
struct LevelRecord {
  int32_t mode;
  uint32_t count;   // |F_k| (vertices of degree > 0)
  uint64_t edges;   // sum of degrees over F_k
};

// A counter alone on its 128-byte line (same-line atomics serialise).
struct alignas(128) LineCounter {
  unsigned v;
  unsigned pad[31];
};

// Device-resident control block of one BFS run.
//
// A frontier F_k is the concatenation of kSubQueues sub-queues; ctr[k % 3][s]
// packs sub-queue s's vertex count (low 32 bits) and degree sum (high 32
// bits), so one 64-bit atomicAdd reserves queue slots and edge offsets
// together.  A sub-queue region holds qcap entries (a quarter of n over all
// sub-queues by default); appends past it are counted but not stored, and
// qover[k % 3] tells level k to expand F_k from its bitmap instead.  Degree-0 vertices are reached but not queued; d0[k % 3] counts
// those at depth k.  Workgroup 0 of level k's kernel does the bookkeeping
// (stats, modes[k & 3], clearing the counters level k+1 will fill); every
// workgroup derives the level's direction from the same inputs itself.
struct BfsState {
  unsigned long long ctr[3][kSubQueues];
  uint32_t d0[3];
  int32_t qover[3];                  // F_k's queue overflowed (entries dropped): expand its bitmap
  int32_t modes[4];
  int32_t finished;                  // set by the first kernel that sees F_k empty
  int32_t td_levels;
  int32_t bu_levels;
  int32_t max_depth;
  int32_t pad;
  unsigned long long reached;        // n_r
  unsigned long long edges_reached;  // m_r
  // grid barrier of a persistent launch (grid_barrier in bfs_kernels.hip):
  // workgroup b arrives on bar_grp[b % kBarGroups]; a group's last arrival
  // arrives on bar_top, whose last arrival bumps bar_gen (the barriers
  // completed).  abort is set by a workgroup whose wait ran out; every
  // workgroup then leaves the launch.
  LineCounter bar_grp[kBarGroups];
  LineCounter bar_top;
  LineCounter bar_gen;
  int32_t abort;
  int32_t pad2;
  LevelRecord log[kLevelLogMax];
};

// Single-source level kernel variants: unsorted rows (full-row bottom-up
// scans), sorted rows read from the CSR, sorted rows with row-top records;
// and any-parent levels (bottom-up rows probed from their smallest entry up,
// the first frontier neighbour found taken as the parent, sorted or not).
enum LevelVariant { kLevelUnsorted = 0, kLevelSorted = 1, kLevelSortedTops = 2, kLevelAnyParent = 3 };

// Everything a level kernel needs.  The init kernel copies it into device
// memory so one captured HIP graph of level launches serves every run that
// uses the same workspace.

template <typename OffT>
struct BfsArgs {
  const OffT* off;             // offsets[n+1]
  const int32_t* adj;          // adjacency[m]
  int32_t* parent;             // out: parent[n]
  int32_t* depth;              // out: depth[n]
  unsigned long long* visited; // bitmap, ceil(n/64) words; pad bits set
  unsigned long long* fb[3];   // frontier bitmaps, fb[k % 3] = F_k
  int32_t* q[2];               // queues, q[k & 1] = F_k: kSubQueues regions of qcap
  uint32_t* qe[2];             // per entry: exclusive degree prefix within its sub-queue
  BfsState* st;
  uint32_t* wg_level;          // per-workgroup level counter (kMaxGrid entries)
  int32_t* host_status;        // pinned host memory: [0] done, [1] levels started,
                               // [2] a persistent launch gave up waiting at its barrier
  BfsState* host_state;        // pinned host memory: the final counters, written by the
                               // kernel that ends the BFS before it sets host_status[0]
  const uint32_t* bu_bounds;   // first-word index of each wavefront's share in the first
                               // bottom-up level after top-down (num_waves + 1 entries)
  const uint32_t* bu_tops;     // [4n] row-top records {three largest neighbours or ~0,
                               // degree}, or null: bottom-up reads offsets and row tops
  const unsigned long long* iso;  // degree-0 bitmap (nwords), set by the init kernel: a
                               // persistent launch rebuilds visited from it and the depths
  int64_t n;
  int64_t m;
  int64_t nwords;              // ceil(n/64)
  int64_t qcap;                // entries per sub-queue region
  int32_t td_to_bu_div;        // 0 = never bottom-up
  int32_t bu_to_td_div;
  int32_t bu_serial_depth;
  int32_t symmetric;
  int32_t num_cus;             // CUs of the device (workgroup slot = blockIdx / num_cus)
  int32_t small_rows;          // 1 if no row has more than 64 entries (top-down slices by entry)
  long long wait_ticks;        // persistent launch: longest barrier wait (device wall-clock
                               // ticks) before the launch gives up
};

// Jobs of a multi-graph launch (several single-source BFS runs, each on its
// own graph, in one launch): an array of their BfsArgs and the first
// workgroup of each (njobs + 1 entries, the last one the grid size).
constexpr int kMaxMultiJobs = 64;

// ----------------------------------------------------------- batched BFS
// Up to 64 sources in one traversal: bit i of a per-vertex 64-bit mask
// stands for source i.  F_k of every source is fr[k % 4]; the union frontier
// (vertices with a non-zero mask) is queued like the single-source one.
constexpr int kMaxSources = 64;

struct MbfsState {
  unsigned long long ctr[3][kSubQueues];  // union-frontier sub-queue counters, [k % 3]
  int32_t qover[3];                       // never set: batched sub-queues hold n entries each
  int32_t pad;
  unsigned long long unfinished[3];       // rows a pull left with sources to find, [(k+1) % 3]
  int32_t modes[4];
  int32_t finished;
  int32_t push_levels;
  int32_t pull_levels;
  int32_t max_depth;
  unsigned long long reached_pairs;  // (source, vertex) pairs reached (summed by the depth pass)
  unsigned long long edges_pairs;    // sum over those pairs of deg(vertex)
};

template <typename OffT>
struct MbfsArgs {
  const OffT* off;
  const int32_t* adj;
  const int32_t* adj_rank;     // [m] rank of each adjacency entry (masks are stored by rank)
  const int32_t* rank;         // [n] rank of v in decreasing-degree order
  const int32_t* inv_rank;     // [n] vertex of each rank
  int32_t* parent;             // out: [nsrc][n], source-major
  int32_t* depth;              // out: [nsrc][n]
  uint8_t* dlev;               // [n][64] level at which each source reached v
  unsigned long long* vis;     // [n] sources that reached v (by vertex id)
  unsigned long long* fr[4];   // [n] each, indexed by rank
  unsigned long long* fbm[4];  // [nwords] each: bit r set iff fr[.][r] != 0
  int32_t* q[2];               // union queue, kSubQueues regions of qcap
  uint32_t* qe[2];
  MbfsState* st;
  uint32_t* wg_level;
  int32_t* host_status;
  const int32_t* sources;      // [nsrc]
  const int32_t* pull_rows;    // first list entry of each wavefront's pull slice (waves + 1)
  const int32_t* rows;         // vertices of degree > 0, ascending
  int64_t nrows;
  int64_t n;
  int64_t m;
  int64_t nwords;
  int64_t qcap;
  unsigned long long srcmask;  // bits of the sources in use
  int32_t nsrc;
  int32_t td_to_bu_div;        // 0 = push levels only
  int32_t pull_div;            // pull when (|F| + edges(F)) * pull_div > m
  int32_t num_cus;             // CUs of the device (workgroup slot = blockIdx / num_cus)
  int32_t preset_parents;      // 1: parent already holds -1 everywhere (filled ahead of
                               // the batch), so the clear kernel leaves it alone
  int32_t pad;
};

// ------------------------------------------------- maximal independent set
// One run of pbbs_mis (mis.hip): flags[v] 0 undecided, 1 in the set, 2 out;
// pos[v] the row position the next look at v starts from (written when v is
// deferred, read in later rounds only).  A round's deferred vertices go to
// kMisStripes regions of the next list, `region` entries each, with a count
// per region: three sets of counts rotate (ctr + set * kMisStripes * 16, one
// count per 128-byte line), so no one address takes every append.
constexpr int kMisStripes = 64;
constexpr int kMisMaxChunk = 1024;   // list entries one wavefront takes at a time

template <typename OffT>
struct MisArgs {
  const OffT* off;
  const int32_t* adj;
  uint8_t* flags;              // [n]
  uint32_t* pos;               // [n] relative to the row start
  int32_t* list[2];            // [kMisStripes * region] each
  unsigned long long* ctr;     // [3][kMisStripes][16]
  unsigned long long* part;    // [kMaxGrid] set members per output workgroup
  uint32_t* wg_round;          // kMaxGrid per-workgroup round counters
  int32_t* finished;
  int32_t* host_status;        // pinned: [0] done, [1] rounds started
  int64_t n;
  int64_t region;              // entries per list region: ceil(n / kMisStripes) + kMisMaxChunk
  int32_t sorted;              // rows ascending: the lower neighbours are a prefix
  int32_t pad;
};

// ------------------------------------------------------- spanning forest
// One run of pbbs_spanning_forest (st.hip).  The control words of the round
// schedule (common/speculative_for.h:30-100) live in device memory.
struct StCtl {
  long long done;          // iterations finished (numberDone)
  long long nkeep;         // failed iterates carried to the front of this round
  long long cur;           // currentRoundSize
  long long size;          // this round's iterates
  long long scatter_size;  // the round just scheduled, for its pack kernel (0: none)
  long long rounds;        // rounds scheduled so far
  long long count;         // forest edges, after the output scan
  int error;               // bits: 1 an edge end outside [0, n), 2 the schedule gave up,
                           // 4 an any-forest edge ran out of tries
  int finished;
};

constexpr int kStMaxGroups = 4096;   // workgroups of a round launch at most
constexpr int kStScanPer = 4;        // counts each thread of the one-workgroup scan sums

struct StArgs {
  const int32_t* eu;       // [m] edge ends
  const int32_t* ev;
  int32_t* parents;        // [n] union-find, -1 at roots
  uint32_t* R;             // [n] reservations, UINT_MAX when free
  uint32_t* I;             // [max_round] this round's iterates
  int32_t* su;             // [max_round] their roots, su < sv
  int32_t* sv;
  uint8_t* keep;           // [max_round] reserved and not yet committed
  uint32_t* hold[2];       // [max_round] failed iterates, by round parity
  unsigned* blk;           // [kStMaxGroups] per-workgroup counts, then offsets
  StCtl* ctl;
  int32_t* host_status;    // pinned: [0] done, [1] rounds
  long long n, m;
  long long max_round;     // m / granularity + 1
  long long max_tries;     // 100 + 200 * granularity
};

}  // namespace pbbs
