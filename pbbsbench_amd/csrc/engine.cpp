// Host side of the C ABI: contexts, graphs resident in HBM, the per-context
// BFS workspace, and the level loop (init launch, then captured HIP graphs of
// level launches until the device reports an empty frontier).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pbbs_bfs.h"
#include "bfs_device.hpp"
#include "engine.hpp"

namespace pbbs {

template <typename OffT>
hipError_t launch_bfs_init(const BfsArgs<OffT>& a, BfsArgs<OffT>* dst,
                           const unsigned long long* iso, int32_t s, int grid, hipStream_t stream);
template <typename OffT>
const void* level_kernel_symbol(int variant);
template <typename OffT>
hipError_t launch_bfs_level(const BfsArgs<OffT>* dargs, int variant, int grid, hipStream_t stream);
template <typename OffT>
const void* persist_kernel_symbol(int variant);
template <typename OffT>
const void* multi_kernel_symbol(bool sorted);
template <typename OffT>
hipError_t launch_bfs_multi(const BfsArgs<OffT>* args, const uint32_t* wg_begin, int njobs,
                            bool sorted, int grid, hipStream_t stream);
template <typename OffT>
hipError_t launch_bfs_persist(const BfsArgs<OffT>* dargs, int variant, int grid,
                              hipStream_t stream);
template <typename OffT>
hipError_t launch_graph_offsets(const OffT* off, long long n, long long m,
                                unsigned long long* iso, unsigned long long* counters, int grid,
                                hipStream_t stream);
template <typename OffT>
hipError_t launch_graph_rows(const OffT* off, const int32_t* adj, long long n, long long m,
                             unsigned long long* counters, int grid, hipStream_t stream);
template <typename OffT>
hipError_t launch_mbfs_setup(const MbfsArgs<OffT>& a, MbfsArgs<OffT>* dst, int grid,
                             hipStream_t stream);
template <typename OffT>
const void* mbfs_level_symbol(bool sorted);
template <typename OffT>
hipError_t launch_mbfs_level(const MbfsArgs<OffT>* dargs, bool sorted, int grid,
                             hipStream_t stream);
template <typename OffT>
hipError_t launch_mis_init(const MisArgs<OffT>& a, int grid, hipStream_t stream);
template <typename OffT>
hipError_t launch_mis_round(const MisArgs<OffT>& a, int step, int grid, hipStream_t stream);
template <typename OffT>
hipError_t launch_mis_output(const MisArgs<OffT>& a, int32_t* flags_out,
                             unsigned long long* members, int grid, hipStream_t stream);
hipError_t launch_st_init(const StArgs& a, int grid, hipStream_t stream);
hipError_t launch_st_round(const StArgs& a, int grid, hipStream_t stream);
hipError_t launch_st_any(const StArgs& a, int grid, hipStream_t stream);
hipError_t launch_st_output(const StArgs& a, int32_t* forest, int grid, hipStream_t stream);
template <typename OffT>
hipError_t launch_pull_rows(const OffT* off, const int32_t* rows, long long nrows, long long m,
                            long long waves, int32_t* out, hipStream_t stream);
template <typename OffT>
hipError_t build_degree_ranks(const OffT* off, const int32_t* adj, long long n, long long m,
                              int32_t* rank, int32_t* inv_rank, int32_t* adj_rank, int32_t* rows,
                              long long* nrows, int grid, hipStream_t stream);
template <typename OffT>
hipError_t launch_mbfs_depth(const MbfsArgs<OffT>& a, int grid, hipStream_t stream);
template <typename OffT>
hipError_t launch_word_weights(const OffT* off, long long n, unsigned long long* ww, int grid,
                               hipStream_t stream);
hipError_t exclusive_scan_u64(const unsigned long long* in, unsigned long long* out, long long n,
                              hipStream_t stream);
hipError_t launch_share_bounds(const unsigned long long* prefix, long long nwords, long long waves,
                               uint32_t* bounds, hipStream_t stream);
template <typename OffT>
hipError_t launch_row_tops(const OffT* off, const int32_t* adj, long long n, uint32_t* out,
                           int grid, hipStream_t stream);
hipError_t launch_fill_minus1(int32_t* p, long long count, int grid, hipStream_t stream);
template <typename OffT>
hipError_t launch_graph_symmetry(const OffT* off, const int32_t* adj, long long n, long long m,
                                 unsigned long long* counters, int grid, hipStream_t stream);

// ------------------------------------------------------------------- errors
static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }
const char* last_error() { return g_last_error.c_str(); }

}  // namespace pbbs

using namespace pbbs;

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(PBBS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                   \
  } while (0)

// ------------------------------------------------------------------ context
// Working memory of one single-source BFS (outputs aside): the four bitmaps
// (visited, F_k mod 3), the two frontier queues and their edge prefixes
// (kSubQueues regions of qcap entries each), the control block, the
// per-workgroup level counters and the device copy of the argument block.
// A context's own buffers, or a block the caller hands in.
struct SingleWs {
  unsigned long long* bits = nullptr;
  int32_t* q = nullptr;
  uint32_t* qe = nullptr;
  int64_t qcap = 0;
  BfsState* st = nullptr;
  uint32_t* wg = nullptr;
  void* args = nullptr;
};

struct pbbs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  int wall_khz = 0;          // device wall-clock rate (barrier wait bound)
  pbbs_tuning tuning{};
  // graphs uploaded through this context: destroying the context detaches
  // them (they stay usable from other contexts on the device, and freeable)
  std::mutex graphs_mu;
  std::set<pbbs_graph*> graphs;
  // workspace, sized for the largest graph seen so far
  int64_t cap_n = 0;
  int64_t cap_q = 0;                     // entries per sub-queue region w_q/w_qe hold
  int32_t* w_parent = nullptr;
  int32_t* w_depth = nullptr;
  unsigned long long* w_bits = nullptr;  // visited, fb0, fb1, fb2
  int32_t* w_q = nullptr;                // q0 | q1, each kSubQueues regions of qcap entries
  uint32_t* w_qe = nullptr;              // qe0 | qe1, same layout
  uint32_t* w_wg_level = nullptr;        // kMaxGrid per-workgroup level counters
  BfsState* w_state = nullptr;
  void* w_args = nullptr;                // device copy of BfsArgs<OffT>
  int32_t* host_status = nullptr;        // pinned: [0] done, [1] levels finished
  BfsState* host_state = nullptr;        // pinned: stats read-back
  hipEvent_t ev_start = nullptr, ev_end = nullptr, ev_c0 = nullptr, ev_c1 = nullptr;
  std::map<std::tuple<int, int, int, int>, hipGraphExec_t> level_graphs;
  // the BFS in flight on this context's stream (one at a time)
  struct Run {
    int kind = 0;  // 0 single-source, 1 batched
    bool off64 = false;
    int sorted = 0;   // single source: LevelVariant; batched: rows sorted
    int grid = 0;
    int K = 0;
    int64_t launched = 0;
    int64_t limit = 0;
    bool finalized = false;  // batched: the depth pass after the last level is queued
    bool persist = false;    // single source: every level in one persistent launch
    bool token = false;      // holds the device's persistent-launch token
    bool any_parent = false; // single source: PBBS_BFS_ANY_PARENT levels
    int32_t source = 0;
    int32_t* parent = nullptr;
    int32_t* depth = nullptr;
    pbbs_graph* graph = nullptr;  // learns the launch count when the run is done
    SingleWs ws;                  // single source: the run's working memory
    bool graphs = true;           // single source: captured launch graphs allowed (the
                                  // context's own workspace only; capture allocates)
  } run;
  // batched BFS
  int64_t cap_bn = 0;                    // vertices the mask buffers hold
  unsigned long long* b_masks = nullptr; // vis | fr0 | fr1 | fr2 | fr3
  MbfsState* b_state = nullptr;
  MbfsState* b_host_state = nullptr;     // pinned: the control block after the depth pass
  void* b_args = nullptr;                // device copy of MbfsArgs<OffT>
  int32_t* b_sources = nullptr;          // kMaxSources
  int64_t cap_bout = 0;                  // int32 entries of each output buffer
  int32_t* b_parent = nullptr;          // parents of the last batched run
  int32_t* b_depth = nullptr;
  // A second parent buffer, filled with -1 on the side stream while a batch
  // runs so that the next batch needs no clear of its parents (the two swap
  // each batch); alt_cleared = entries of it known to hold -1.
  int32_t* b_parent_alt = nullptr;
  int64_t alt_cleared = 0;
  bool prefill = true;                   // false in server slots (their runs overlap)
  hipStream_t side = nullptr;
  hipEvent_t ev_alt = nullptr;
  uint8_t* b_dlev = nullptr;             // [cap_bn][64] per-vertex level table
  int32_t* b_q = nullptr;                // union queues: q0 | q1, kSubQueues regions of cap_bn
  uint32_t* b_qe = nullptr;
  bool b_last_args_off64 = false;
  alignas(16) unsigned char b_last_args[sizeof(MbfsArgs<int64_t>)] = {};  // host copy for stats
  std::map<std::tuple<int, int, int>, hipGraphExec_t> batch_graphs;
  // multi-graph launches: the jobs' workspaces, their argument blocks and
  // first workgroups, and pinned status words and control-block copies
  unsigned char* m_ws = nullptr;
  int64_t m_ws_bytes = 0;
  void* m_args = nullptr;                // BfsArgs<int64_t>[kMaxMultiJobs]
  uint32_t* m_wg = nullptr;              // [kMaxMultiJobs + 1]
  int32_t* m_host_status = nullptr;      // pinned: 16 words per job
  BfsState* m_host_state = nullptr;      // pinned: one per job
  // maximal independent set: per-vertex flags and row positions, the two
  // worklists (kMisStripes regions each), the int32 output staged for a host
  // copy, the control words (three sets of region counts, the finished flag,
  // the set's size), per-workgroup member counts and round counters
  int64_t cap_sn = 0;
  uint8_t* s_flags = nullptr;
  uint32_t* s_pos = nullptr;
  int32_t* s_list = nullptr;             // list0 | list1, kMisStripes * mis_region(cap_sn) each
  int32_t* s_out = nullptr;
  unsigned long long* s_ctl = nullptr;   // kMisCtlWords
  unsigned long long* s_part = nullptr;  // kMaxGrid
  uint32_t* s_wg = nullptr;              // kMaxGrid
  // spanning forest: union-find and reservations per vertex, the round's
  // iterates and their roots, the two carry lists, per-workgroup counts, the
  // schedule's control words, and the edges when the caller's are in host
  // memory
  int64_t cap_tn = 0, cap_tr = 0, cap_te = 0;
  int32_t* t_parents = nullptr;
  uint32_t* t_R = nullptr;
  unsigned char* t_round = nullptr;      // I | su | sv | hold0 | hold1 | keep
  unsigned* t_blk = nullptr;
  StCtl* t_ctl = nullptr;
  int32_t* t_edges = nullptr;            // eu | ev, cap_te each
  int32_t* t_out = nullptr;              // cap_tn forest entries for a host copy
};

struct pbbs_graph {
  pbbs_ctx* ctx = nullptr;   // uploading context, null once it is destroyed
  int device = 0;
  int64_t n = 0, m = 0;
  bool off64 = false;
  bool sorted = false;
  bool symmetric = false;
  void* d_off = nullptr;
  int32_t* d_adj = nullptr;
  unsigned long long* d_iso = nullptr;  // degree-0 bitmap
  unsigned long long* d_wprefix = nullptr;  // exclusive prefix of per-word degree weights
  // Per launch shape (wavefront count), built on first use and kept until the
  // graph is freed: contexts with different shapes may share the graph.
  std::mutex meta_mu;
  std::map<int64_t, uint32_t*> bounds;   // weighted bottom-up shares (single source)
  std::map<int64_t, int32_t*> pull_rows; // first row of each edge slice (batched pull)
  uint32_t* d_tops = nullptr;            // single source: row-top records, 32 B per vertex
  int32_t* d_rank = nullptr;             // batched BFS: decreasing-degree rank of each vertex
  int32_t* d_inv_rank = nullptr;         // batched BFS: vertex of each rank
  int32_t* d_adj_rank = nullptr;         // batched BFS: adj in rank ids
  int32_t* d_rows = nullptr;             // batched BFS: vertices of degree > 0, ascending
  int64_t nrows = 0;
  int64_t isolated = 0;
  int64_t max_degree = 0;
  int64_t device_bytes = 0;
  // Level launches the last single-source [0] / batched [1] run needed (its
  // levels plus the launch that found the frontier empty); 0 = none yet.
  std::atomic<int> last_launches[2] = {0, 0};
};

// One persistent launch per device at a time: its grid fills the device and
// waits at barriers for all of its workgroups, so two such launches sharing
// the device could each hold part of it and wait for the rest.  A run that
// finds the token taken runs one level per launch instead.
constexpr int kMaxDevices = 64;
static std::atomic<int> g_persist_busy[kMaxDevices];

static bool take_persist(pbbs_ctx* c) {
  if (c->device < 0 || c->device >= kMaxDevices) return false;
  int expect = 0;
  if (!g_persist_busy[c->device].compare_exchange_strong(expect, 1)) return false;
  c->run.token = true;
  return true;
}

static void release_persist(pbbs_ctx* c) {
  if (c->run.token) {
    g_persist_busy[c->device].store(0);
    c->run.token = false;
  }
}

static void free_workspace(pbbs_ctx* c) {
  (void)hipFree(c->w_parent);
  (void)hipFree(c->w_depth);
  (void)hipFree(c->w_bits);
  (void)hipFree(c->w_q);
  (void)hipFree(c->w_qe);
  c->w_parent = c->w_depth = nullptr;
  c->w_bits = nullptr;
  c->w_q = nullptr;
  c->w_qe = nullptr;
  c->cap_n = 0;
  c->cap_q = 0;
}

// Control words of a set run: three sets of kMisStripes region counts, one
// per 128-byte line, then the finished flag and the set's size.
constexpr int kMisCtlFinished = 3 * kMisStripes * 16;
constexpr int kMisCtlMembers = kMisCtlFinished + 8;
constexpr int kMisCtlWords = kMisCtlMembers + 8;

static int64_t mis_region(int64_t n) { return (n + kMisStripes - 1) / kMisStripes + kMisMaxChunk; }

static void free_st_workspace(pbbs_ctx* c) {
  (void)hipFree(c->t_parents);
  (void)hipFree(c->t_R);
  (void)hipFree(c->t_round);
  (void)hipFree(c->t_blk);
  (void)hipFree(c->t_ctl);
  (void)hipFree(c->t_edges);
  (void)hipFree(c->t_out);
  c->t_parents = c->t_out = c->t_edges = nullptr;
  c->t_R = nullptr;
  c->t_round = nullptr;
  c->t_blk = nullptr;
  c->t_ctl = nullptr;
  c->cap_tn = c->cap_tr = c->cap_te = 0;
}

// Grows the spanning-forest workspace to n vertices, r iterates a round and
// (host edges) e edges; buffers only grow.
static int ensure_st_workspace(pbbs_ctx* c, int64_t n, int64_t r, int64_t e) {
  if (n <= c->cap_tn && r <= c->cap_tr && e <= c->cap_te && c->t_ctl) return PBBS_OK;
  const int64_t cn = std::max(n, c->cap_tn), cr = std::max(r, c->cap_tr),
                ce = std::max(e, c->cap_te);
  free_st_workspace(c);
  if (hipMalloc(&c->t_parents, std::max<int64_t>(cn, 1) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->t_R, std::max<int64_t>(cn, 1) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&c->t_out, std::max<int64_t>(cn, 1) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->t_round, std::max<int64_t>(cr, 1) * 21) != hipSuccess ||
      hipMalloc(&c->t_blk, kStMaxGroups * sizeof(unsigned)) != hipSuccess ||
      hipMalloc(&c->t_ctl, sizeof(StCtl)) != hipSuccess ||
      (ce > 0 && hipMalloc(&c->t_edges, 2 * ce * sizeof(int32_t)) != hipSuccess)) {
    (void)hipGetLastError();
    free_st_workspace(c);
    return fail(PBBS_ERR_NOMEM, "forest workspace allocation (n=%lld, round=%lld) failed",
                (long long)n, (long long)r);
  }
  c->cap_tn = cn;
  c->cap_tr = cr;
  c->cap_te = ce;
  return PBBS_OK;
}

static void free_mis_workspace(pbbs_ctx* c) {
  (void)hipFree(c->s_flags);
  (void)hipFree(c->s_pos);
  (void)hipFree(c->s_list);
  (void)hipFree(c->s_out);
  (void)hipFree(c->s_ctl);
  (void)hipFree(c->s_part);
  (void)hipFree(c->s_wg);
  c->s_part = nullptr;
  c->s_flags = nullptr;
  c->s_pos = nullptr;
  c->s_list = c->s_out = nullptr;
  c->s_ctl = nullptr;
  c->s_wg = nullptr;
  c->cap_sn = 0;
}

static int ensure_mis_workspace(pbbs_ctx* c, int64_t n) {
  if (n <= c->cap_sn) return PBBS_OK;
  free_mis_workspace(c);
  if (hipMalloc(&c->s_flags, n) != hipSuccess ||
      hipMalloc(&c->s_pos, n * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&c->s_list, 2 * kMisStripes * mis_region(n) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->s_out, n * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->s_ctl, kMisCtlWords * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->s_part, kMaxGrid * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->s_wg, kMaxGrid * sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    free_mis_workspace(c);
    return fail(PBBS_ERR_NOMEM, "set workspace allocation for n=%lld failed", (long long)n);
  }
  c->cap_sn = n;
  return PBBS_OK;
}

namespace pbbs {
// Entries per sub-queue region of the single-source frontier queues: n / div
// over all kSubQueues regions, at least 64 each (n at most).  A frontier
// that does not fit is expanded from its bitmap by the level that reads it.
int64_t queue_cap(int64_t n, int32_t div) {
  const int64_t per = (n + (int64_t)div * kSubQueues - 1) / ((int64_t)div * kSubQueues);
  return std::min<int64_t>(n, std::max<int64_t>(kWave, per));
}
}  // namespace pbbs

static int ensure_workspace(pbbs_ctx* c, int64_t n) {
  const int64_t qcap = queue_cap(n, c->tuning.queue_div);
  if (n <= c->cap_n && qcap <= c->cap_q) return PBBS_OK;
  const int64_t cn = std::max(n, c->cap_n), cq = std::max(qcap, c->cap_q);
  free_workspace(c);
  const int64_t nwords = (cn + kWave - 1) / kWave;
  if (hipMalloc(&c->w_parent, cn * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->w_depth, cn * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->w_bits, 4 * nwords * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->w_q, 2 * kSubQueues * cq * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->w_qe, 2 * kSubQueues * cq * sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    free_workspace(c);
    return fail(PBBS_ERR_NOMEM, "workspace allocation for n=%lld failed", (long long)n);
  }
  c->cap_n = cn;
  c->cap_q = cq;
  return PBBS_OK;
}

static int default_tuning(pbbs_tuning* t) {
  std::memset(t, 0, sizeof *t);
  t->td_to_bu_div = 20;
  t->bu_to_td_div = 20;
  t->grid_blocks = 0;
  t->bu_serial_depth = 128;
  t->chunk_levels = 16;
  t->use_graphs = 1;
  t->batch_pull_div = 10;
  t->bu_row_tops = 1;
  t->queue_div = 4;
  t->persist = 2;
  t->persist_wait_us = 200000;
  return PBBS_OK;
}

extern "C" {

const char* pbbs_last_error(void) { return last_error(); }
const char* pbbs_version(void) { return "pbbs-bfs-gfx950 0.1"; }

int pbbs_ctx_create(int device, pbbs_ctx** out) {
  clear_error();
  if (!out) return fail(PBBS_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int count = 0;
  HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count)
    return fail(PBBS_ERR_INVALID, "device %d out of range (%d devices)", device, count);
  HIP_TRY(hipSetDevice(device));
  pbbs_ctx* c = new pbbs_ctx();
  c->device = device;
  default_tuning(&c->tuning);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->w_state, sizeof(BfsState)) != hipSuccess ||
      hipMalloc(&c->w_wg_level, kMaxGrid * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&c->w_args, sizeof(BfsArgs<int64_t>)) != hipSuccess ||
      hipHostMalloc((void**)&c->host_status, 64, hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void**)&c->host_state, sizeof(BfsState), hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&c->ev_start) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess ||
      hipEventCreate(&c->ev_c0) != hipSuccess || hipEventCreate(&c->ev_c1) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_alt, hipEventDisableTiming) != hipSuccess) {
    pbbs_ctx_destroy(c);
    return fail(PBBS_ERR_HIP, "context setup on device %d failed", device);
  }
  if (scratch_init(device) != PBBS_OK) {
    pbbs_ctx_destroy(c);
    return PBBS_ERR_HIP;   // scratch_init's message stands
  }
  c->num_cus = prop.multiProcessorCount;
  if (hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
      c->wall_khz <= 0)
    c->wall_khz = 100000;   // gfx9 wall clock: 100 MHz
  std::memset(c->host_status, 0, 64);
  *out = c;
  return PBBS_OK;
}

int pbbs_ctx_destroy(pbbs_ctx* c) {
  if (!c) return PBBS_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->side) (void)hipStreamSynchronize(c->side);
  release_persist(c);
  {
    std::lock_guard<std::mutex> lk(c->graphs_mu);
    for (pbbs_graph* g : c->graphs) g->ctx = nullptr;
    c->graphs.clear();
  }
  for (auto& kv : c->level_graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& kv : c->batch_graphs) (void)hipGraphExecDestroy(kv.second);
  free_workspace(c);
  free_mis_workspace(c);
  free_st_workspace(c);
  (void)hipFree(c->b_masks);
  (void)hipFree(c->b_dlev);
  (void)hipFree(c->b_q);
  (void)hipFree(c->b_qe);
  (void)hipFree(c->b_state);
  if (c->b_host_state) (void)hipHostFree(c->b_host_state);
  (void)hipFree(c->b_args);
  (void)hipFree(c->b_sources);
  (void)hipFree(c->b_parent);
  (void)hipFree(c->b_parent_alt);
  (void)hipFree(c->b_depth);
  (void)hipFree(c->w_state);
  (void)hipFree(c->w_wg_level);
  (void)hipFree(c->w_args);
  (void)hipFree(c->m_ws);
  (void)hipFree(c->m_args);
  (void)hipFree(c->m_wg);
  if (c->m_host_status) (void)hipHostFree(c->m_host_status);
  if (c->m_host_state) (void)hipHostFree(c->m_host_state);
  if (c->host_status) (void)hipHostFree(c->host_status);
  if (c->host_state) (void)hipHostFree(c->host_state);
  if (c->ev_start) (void)hipEventDestroy(c->ev_start);
  if (c->ev_end) (void)hipEventDestroy(c->ev_end);
  if (c->ev_c0) (void)hipEventDestroy(c->ev_c0);
  if (c->ev_c1) (void)hipEventDestroy(c->ev_c1);
  if (c->ev_alt) (void)hipEventDestroy(c->ev_alt);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->side) (void)hipStreamDestroy(c->side);
  delete c;
  return PBBS_OK;
}

int pbbs_tuning_get(const pbbs_ctx* c, pbbs_tuning* out) {
  clear_error();
  if (!c || !out) return fail(PBBS_ERR_INVALID, "NULL argument");
  *out = c->tuning;
  return PBBS_OK;
}

int pbbs_tuning_set(pbbs_ctx* c, const pbbs_tuning* t) {
  clear_error();
  if (!c || !t) return fail(PBBS_ERR_INVALID, "NULL argument");
  if (t->td_to_bu_div < 0 || t->bu_to_td_div < 1 || t->grid_blocks < 0 ||
      t->bu_serial_depth < 0 || t->bu_serial_depth > 4096 || t->chunk_levels < 1 ||
      t->chunk_levels > 4096 || t->batch_pull_div < 0 || t->bu_row_tops < 0 ||
      t->bu_row_tops > 1 || t->queue_div < 1 || t->persist < 0 || t->persist > 2 ||
      t->persist_wait_us < 0)
    return fail(PBBS_ERR_INVALID,
                "tuning out of range (td_to_bu_div>=0, bu_to_td_div>=1, grid_blocks>=0, "
                "0<=bu_serial_depth<=4096, 1<=chunk_levels<=4096, batch_pull_div>=0, "
                "bu_row_tops 0 or 1, queue_div>=1, persist 0..2, persist_wait_us>=0)");
  c->tuning = *t;
  return PBBS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------- graphs
template <typename OffT>
static int upload_checks(pbbs_ctx* c, pbbs_graph* g, uint32_t flags) {
  hipStream_t s = c->stream;
  const int64_t nwords = (g->n + kWave - 1) / kWave;
  unsigned long long* d_cnt = nullptr;
  HIP_TRY(hipMalloc(&d_cnt, 8 * sizeof(unsigned long long)));
  unsigned long long h[8];
  int rc = PBBS_OK;
  const int grid = std::max(1, c->num_cus * 8);
  const OffT* off = static_cast<const OffT*>(g->d_off);
  OffT ends[2];
  do {
    if (hipMemsetAsync(d_cnt, 0, 8 * sizeof(unsigned long long), s) != hipSuccess ||
        launch_graph_offsets<OffT>(off, g->n, g->m, g->d_iso, d_cnt, grid, s) != hipSuccess ||
        hipMemcpyAsync(h, d_cnt, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&ends[0], off, sizeof(OffT), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&ends[1], off + g->n, sizeof(OffT), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = fail(PBBS_ERR_HIP, "graph offset check failed: %s", hipGetErrorString(hipGetLastError()));
      break;
    }
    if ((int64_t)ends[0] != 0 || (int64_t)ends[1] != g->m || h[2] != 0) {
      rc = fail(PBBS_ERR_FORMAT,
                "bad offsets: offsets[0]=%lld offsets[n]=%lld (m=%lld), %llu decreasing or >m",
                (long long)ends[0], (long long)ends[1], (long long)g->m, h[2]);
      break;
    }
    if (launch_graph_rows<OffT>(off, g->d_adj, g->n, g->m, d_cnt, grid, s) != hipSuccess ||
        hipMemcpyAsync(h, d_cnt, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = fail(PBBS_ERR_HIP, "graph row check failed");
      break;
    }
    if (h[5] != 0) {
      rc = fail(PBBS_ERR_FORMAT, "%llu adjacency entries outside [0, n)", h[5]);
      break;
    }
    g->isolated = (int64_t)h[0];
    g->max_degree = (int64_t)h[1];
    g->sorted = h[4] == h[3];
    if (flags & PBBS_GRAPH_ASSUME_SYMMETRIC) {
      g->symmetric = true;
    } else if (g->sorted) {
      if (launch_graph_symmetry<OffT>(off, g->d_adj, g->n, g->m, d_cnt, grid, s) != hipSuccess ||
          hipMemcpyAsync(h, d_cnt, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        rc = fail(PBBS_ERR_HIP, "graph symmetry check failed");
        break;
      }
      g->symmetric = h[6] == 0;
    } else {
      g->symmetric = false;
    }
    // per-word degree weights and their prefix, for the weighted bottom-up split
    unsigned long long* ww = nullptr;
    if (hipMalloc(&ww, (nwords + 1) * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&g->d_wprefix, (nwords + 1) * sizeof(unsigned long long)) != hipSuccess) {
      (void)hipFree(ww);
      rc = fail(PBBS_ERR_NOMEM, "weight prefix allocation failed");
      break;
    }
    const hipError_t we = launch_word_weights<OffT>(off, g->n, ww, grid, s);
    const hipError_t se = we == hipSuccess ? exclusive_scan_u64(ww, g->d_wprefix, nwords, s) : we;
    (void)hipFree(ww);
    if (se != hipSuccess) {
      rc = fail(PBBS_ERR_HIP, "weight prefix failed: %s", hipGetErrorString(se));
      break;
    }
    g->device_bytes += (nwords + 1) * sizeof(unsigned long long);
  } while (false);
  (void)hipFree(d_cnt);
  return rc;
}

extern "C" {

int pbbs_graph_free(pbbs_graph* g) {
  if (!g) return PBBS_OK;
  (void)hipSetDevice(g->device);
  if (g->ctx) {
    (void)hipStreamSynchronize(g->ctx->stream);
    std::lock_guard<std::mutex> lk(g->ctx->graphs_mu);
    g->ctx->graphs.erase(g);
  }
  (void)hipFree(g->d_off);
  (void)hipFree(g->d_adj);
  (void)hipFree(g->d_iso);
  (void)hipFree(g->d_wprefix);
  for (auto& kv : g->bounds) (void)hipFree(kv.second);
  for (auto& kv : g->pull_rows) (void)hipFree(kv.second);
  (void)hipFree(g->d_tops);
  (void)hipFree(g->d_rank);
  (void)hipFree(g->d_inv_rank);
  (void)hipFree(g->d_adj_rank);
  (void)hipFree(g->d_rows);
  delete g;
  return PBBS_OK;
}

int pbbs_graph_upload(pbbs_ctx* c, int64_t n, int64_t m, const void* offsets, const int32_t* adj,
                      uint32_t flags, pbbs_graph** out) {
  clear_error();
  if (!c || !out || !offsets) return fail(PBBS_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (n <= 0 || n > INT32_MAX) return fail(PBBS_ERR_LIMIT, "n=%lld outside [1, 2^31)", (long long)n);
  if (m < 0 || m > (int64_t)UINT32_MAX)
    return fail(PBBS_ERR_LIMIT, "m=%lld outside [0, 2^32)", (long long)m);
  if (m > 0 && !adj) return fail(PBBS_ERR_INVALID, "adj is NULL with m > 0");
  HIP_TRY(hipSetDevice(c->device));
  pbbs_graph* g = new pbbs_graph();
  g->ctx = c;
  g->device = c->device;
  {
    std::lock_guard<std::mutex> lk(c->graphs_mu);
    c->graphs.insert(g);
  }
  g->n = n;
  g->m = m;
  g->off64 = (flags & PBBS_GRAPH_OFFSETS_I64) != 0;
  const size_t osz = g->off64 ? sizeof(int64_t) : sizeof(uint32_t);
  const int64_t nwords = (n + kWave - 1) / kWave;
  const hipMemcpyKind kind =
      (flags & PBBS_GRAPH_DEVICE_INPUT) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  if (hipMalloc(&g->d_off, (n + 1) * osz) != hipSuccess ||
      hipMalloc(&g->d_adj, std::max<int64_t>(m, 1) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&g->d_iso, nwords * sizeof(unsigned long long)) != hipSuccess) {
    pbbs_graph_free(g);
    return fail(PBBS_ERR_NOMEM, "graph allocation (n=%lld, m=%lld) failed", (long long)n,
                (long long)m);
  }
  g->device_bytes = (n + 1) * osz + std::max<int64_t>(m, 1) * 4 + nwords * 8;
  if (hipMemcpyAsync(g->d_off, offsets, (n + 1) * osz, kind, c->stream) != hipSuccess ||
      (m > 0 && hipMemcpyAsync(g->d_adj, adj, m * sizeof(int32_t), kind, c->stream) != hipSuccess) ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    pbbs_graph_free(g);
    return fail(PBBS_ERR_HIP, "graph copy into HBM failed");
  }
  const int rc = g->off64 ? upload_checks<int64_t>(c, g, flags) : upload_checks<uint32_t>(c, g, flags);
  if (rc != PBBS_OK) {
    pbbs_graph_free(g);
    return rc;
  }
  *out = g;
  return PBBS_OK;
}

int pbbs_graph_info_get(const pbbs_graph* g, pbbs_graph_info* out) {
  clear_error();
  if (!g || !out) return fail(PBBS_ERR_INVALID, "NULL argument");
  std::memset(out, 0, sizeof *out);
  out->n = g->n;
  out->m = g->m;
  out->offsets_i64 = g->off64;
  out->sorted = g->sorted;
  out->symmetric = g->symmetric;
  out->isolated = g->isolated;
  out->max_degree = g->max_degree;
  out->device_bytes = g->device_bytes;
  return PBBS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------- BFS
static int level_grid(pbbs_ctx* c, bool off64, int variant) {
  if (c->tuning.grid_blocks > 0) return std::min(c->tuning.grid_blocks, kMaxGrid);
  int per_cu = 0;
  const void* fn =
      off64 ? level_kernel_symbol<int64_t>(variant) : level_kernel_symbol<uint32_t>(variant);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return std::min(per_cu * c->num_cus, kMaxGrid);
}

template <typename OffT>
static int get_level_graph(pbbs_ctx* c, int variant, int grid, int K, hipGraphExec_t* out) {
  const auto key = std::make_tuple((int)sizeof(OffT), variant, grid, K);
  auto it = c->level_graphs.find(key);
  if (it != c->level_graphs.end()) {
    *out = it->second;
    return PBBS_OK;
  }
  hipGraph_t graph = nullptr;
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int i = 0; i < K && le == hipSuccess; ++i)
    le = launch_bfs_level<OffT>(static_cast<const BfsArgs<OffT>*>(c->w_args), variant, grid,
                                c->stream);
  const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
  if (le != hipSuccess || ce != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(PBBS_ERR_HIP, "capturing %d level launches failed: %s", K,
                hipGetErrorString(le != hipSuccess ? le : ce));
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) return fail(PBBS_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ie));
  c->level_graphs[key] = exec;
  *out = exec;
  return PBBS_OK;
}

static int next_chunk(int base, int launched) {
  int k = base;
  while (k < launched && k < 1024) k *= 2;
  return k;
}

// Cuts the weighted bottom-up shares for `waves` wavefronts (cached per graph).
// With build false, bounds not built yet are an error (no allocation).
static int ensure_bounds(pbbs_ctx* c, pbbs_graph* g, int64_t waves, const uint32_t** out,
                         bool build = true) {
  std::lock_guard<std::mutex> lk(g->meta_mu);
  auto it = g->bounds.find(waves);
  if (it != g->bounds.end()) {
    *out = it->second;
    return PBBS_OK;
  }
  if (!build)
    return fail(PBBS_ERR_STATE, "graph not prepared for %lld wavefronts (pbbs_graph_prepare "
                "with PBBS_PREPARE_SINGLE on this context first)", (long long)waves);
  const int64_t nwords = (g->n + kWave - 1) / kWave;
  uint32_t* d = nullptr;
  HIP_TRY(hipMalloc(&d, (waves + 1) * sizeof(uint32_t)));
  hipError_t e = launch_share_bounds(g->d_wprefix, nwords, waves, d, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return fail(PBBS_ERR_HIP, "wavefront shares: %s", hipGetErrorString(e));
  }
  g->bounds[waves] = d;
  g->device_bytes += (waves + 1) * (int64_t)sizeof(uint32_t);
  *out = d;
  return PBBS_OK;
}

// Builds the row-top records of a sorted, symmetric graph on first use
// (bottom-up levels read them in place of offsets and row tops); null when
// the graph never runs bottom-up or the tuning turns them off.
// With build false, records not built yet are an error (no allocation).
template <typename OffT>
static int ensure_tops(pbbs_ctx* c, pbbs_graph* g, const uint32_t** out, bool build = true) {
  *out = nullptr;
  if (!g->sorted || !g->symmetric || c->tuning.bu_row_tops == 0 || c->tuning.td_to_bu_div <= 0)
    return PBBS_OK;
  std::lock_guard<std::mutex> lk(g->meta_mu);
  if (!g->d_tops && !build)
    return fail(PBBS_ERR_STATE, "graph has no row-top records (pbbs_graph_prepare with "
                "PBBS_PREPARE_SINGLE first)");
  if (!g->d_tops) {
    uint32_t* d = nullptr;
    const size_t bytes = (size_t)g->n * 32;
    if (hipMalloc(&d, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PBBS_ERR_NOMEM, "row-top records (%zu bytes)", bytes);
    }
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((g->n + kBlock - 1) / kBlock,
                                                                  c->num_cus * 8));
    hipError_t e = launch_row_tops<OffT>(static_cast<const OffT*>(g->d_off), g->d_adj, g->n, d,
                                         grid, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return fail(PBBS_ERR_HIP, "row-top records: %s", hipGetErrorString(e));
    }
    g->d_tops = d;
    g->device_bytes += (int64_t)bytes;
  }
  *out = g->d_tops;
  return PBBS_OK;
}

template <typename OffT>
static int ensure_pull_rows(pbbs_ctx* c, pbbs_graph* g, int64_t waves, const int32_t** out) {
  std::lock_guard<std::mutex> lk(g->meta_mu);
  auto it = g->pull_rows.find(waves);
  if (it != g->pull_rows.end()) {
    *out = it->second;
    return PBBS_OK;
  }
  int32_t* d = nullptr;
  HIP_TRY(hipMalloc(&d, (waves + 1) * sizeof(int32_t)));
  hipError_t e = launch_pull_rows<OffT>(static_cast<const OffT*>(g->d_off), g->d_rows, g->nrows,
                                        g->m, waves, d, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return fail(PBBS_ERR_HIP, "edge-slice rows: %s", hipGetErrorString(e));
  }
  g->pull_rows[waves] = d;
  g->device_bytes += (waves + 1) * (int64_t)sizeof(int32_t);
  *out = d;
  return PBBS_OK;
}

// Degree ranks and the rank-id adjacency of a graph, built on its first
// batched BFS and kept until the graph is freed.
template <typename OffT>
static int ensure_ranks(pbbs_ctx* c, pbbs_graph* g) {
  std::lock_guard<std::mutex> lk(g->meta_mu);
  if (g->d_rank) return PBBS_OK;
  int32_t *rank = nullptr, *inv = nullptr, *adj_rank = nullptr, *rows = nullptr;
  auto release = [&]() {
    (void)hipFree(rank);
    (void)hipFree(inv);
    (void)hipFree(adj_rank);
    (void)hipFree(rows);
  };
  if (hipMalloc(&rank, g->n * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&inv, g->n * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&rows, g->n * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&adj_rank, std::max<int64_t>(g->m, 1) * sizeof(int32_t)) != hipSuccess) {
    release();
    return fail(PBBS_ERR_NOMEM, "degree-rank arrays for n=%lld m=%lld failed", (long long)g->n,
                (long long)g->m);
  }
  long long nrows = 0;
  const hipError_t e =
      build_degree_ranks<OffT>(static_cast<const OffT*>(g->d_off), g->d_adj, g->n, g->m, rank,
                               inv, adj_rank, rows, &nrows, c->num_cus * 8, c->stream);
  if (e != hipSuccess) {
    release();
    return fail(PBBS_ERR_HIP, "degree ranks: %s", hipGetErrorString(e));
  }
  g->d_rank = rank;
  g->d_inv_rank = inv;
  g->d_adj_rank = adj_rank;
  g->d_rows = rows;
  g->nrows = nrows;
  g->device_bytes += (3 * g->n + g->m) * (int64_t)sizeof(int32_t);
  return PBBS_OK;
}

// The context's own single-source workspace, laid out for an n-vertex graph
// (ensure_workspace(c, n) has run).
static SingleWs own_ws(pbbs_ctx* c, int64_t n) {
  SingleWs w;
  w.bits = c->w_bits;
  w.q = c->w_q;
  w.qe = c->w_qe;
  w.qcap = queue_cap(n, c->tuning.queue_div);   // <= c->cap_q
  w.st = c->w_state;
  w.wg = c->w_wg_level;
  w.args = c->w_args;
  return w;
}

template <typename OffT>
static BfsArgs<OffT> make_args(pbbs_ctx* c, pbbs_graph* g, const SingleWs& ws, int32_t* d_parent,
                               int32_t* d_depth, const uint32_t* bounds, const uint32_t* tops) {
  const int64_t nwords = (g->n + kWave - 1) / kWave;
  BfsArgs<OffT> a;
  std::memset(&a, 0, sizeof a);
  a.off = static_cast<const OffT*>(g->d_off);
  a.adj = g->d_adj;
  a.parent = d_parent;
  a.depth = d_depth;
  a.visited = ws.bits;
  a.fb[0] = ws.bits + nwords;
  a.fb[1] = ws.bits + 2 * nwords;
  a.fb[2] = ws.bits + 3 * nwords;
  a.q[0] = ws.q;
  a.q[1] = ws.q + kSubQueues * ws.qcap;
  a.qe[0] = ws.qe;
  a.qe[1] = ws.qe + kSubQueues * ws.qcap;
  a.qcap = ws.qcap;
  a.st = ws.st;
  a.wg_level = ws.wg;
  a.host_status = c->host_status;
  a.host_state = c->host_state;
  a.bu_bounds = bounds;
  a.bu_tops = tops;
  a.n = g->n;
  a.m = g->m;
  a.nwords = nwords;
  a.td_to_bu_div = c->tuning.td_to_bu_div;
  a.bu_to_td_div = c->tuning.bu_to_td_div;
  a.bu_serial_depth = c->tuning.bu_serial_depth;
  a.symmetric = g->symmetric ? 1 : 0;
  a.num_cus = c->num_cus;
  a.small_rows = g->max_degree <= kWave ? 1 : 0;
  a.wait_ticks = (long long)c->tuning.persist_wait_us * c->wall_khz / 1000;
  return a;
}

// --- the level loop as a state machine: begin_* launches the init work and
// the first chunk of level launches; poll_run checks the chunk's completion
// event and either reports the BFS done (the device set host_status[0]) or
// launches the next chunk.  pbbs_bfs drives it to completion; the server
// drives many contexts at once without blocking.

template <typename OffT>
static int get_batch_graph(pbbs_ctx* c, bool sorted, int grid, int K, hipGraphExec_t* out);

static int launch_chunk(pbbs_ctx* c) {
  pbbs_ctx::Run& r = c->run;
  if (c->tuning.use_graphs && (r.kind == 1 || r.graphs)) {
    hipGraphExec_t exec = nullptr;
    int rc;
    if (r.kind == 0)
      rc = r.off64 ? get_level_graph<int64_t>(c, r.sorted, r.grid, r.K, &exec)
                   : get_level_graph<uint32_t>(c, r.sorted, r.grid, r.K, &exec);
    else
      rc = r.off64 ? get_batch_graph<int64_t>(c, r.sorted, r.grid, r.K, &exec)
                   : get_batch_graph<uint32_t>(c, r.sorted, r.grid, r.K, &exec);
    if (rc != PBBS_OK) return rc;
    HIP_TRY(hipGraphLaunch(exec, c->stream));
  } else {
    for (int i = 0; i < r.K; ++i) {
      if (r.kind == 0) {
        if (r.off64)
          HIP_TRY(launch_bfs_level<int64_t>(static_cast<const BfsArgs<int64_t>*>(r.ws.args),
                                            r.sorted, r.grid, c->stream));
        else
          HIP_TRY(launch_bfs_level<uint32_t>(static_cast<const BfsArgs<uint32_t>*>(r.ws.args),
                                             r.sorted, r.grid, c->stream));
      } else {
        if (r.off64)
          HIP_TRY(launch_mbfs_level<int64_t>(static_cast<const MbfsArgs<int64_t>*>(c->b_args),
                                             r.sorted, r.grid, c->stream));
        else
          HIP_TRY(launch_mbfs_level<uint32_t>(static_cast<const MbfsArgs<uint32_t>*>(c->b_args),
                                              r.sorted, r.grid, c->stream));
      }
    }
  }
  r.launched += r.K;
  HIP_TRY(hipEventRecord(c->ev_end, c->stream));
  return PBBS_OK;
}

static int restart_levels(pbbs_ctx* c);

// Advances the BFS in flight.  With wait, blocks for the current chunk.
static int poll_run_impl(pbbs_ctx* c, bool wait, bool* done) {
  *done = false;
  if (wait) {
    HIP_TRY(hipEventSynchronize(c->ev_end));
  } else {
    const hipError_t e = hipEventQuery(c->ev_end);
    if (e == hipErrorNotReady) return PBBS_OK;
    if (e != hipSuccess) return fail(PBBS_ERR_HIP, "event query: %s", hipGetErrorString(e));
  }
  if (c->host_status[0]) {
    pbbs_ctx::Run& r = c->run;
    if (r.kind == 1 && !r.finalized) {
      // batched: write the depth outputs, then report done once that is through
      const int grid = std::max(1, c->num_cus * 8);
      if (r.off64)
        HIP_TRY(launch_mbfs_depth<int64_t>(
            *reinterpret_cast<MbfsArgs<int64_t>*>(c->b_last_args), grid, c->stream));
      else
        HIP_TRY(launch_mbfs_depth<uint32_t>(
            *reinterpret_cast<MbfsArgs<uint32_t>*>(c->b_last_args), grid, c->stream));
      HIP_TRY(hipMemcpyAsync(c->b_host_state, c->b_state, sizeof(MbfsState),
                             hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipEventRecord(c->ev_end, c->stream));
      r.finalized = true;
      if (!wait) return PBBS_OK;
      HIP_TRY(hipEventSynchronize(c->ev_end));
    }
    if (r.graph) r.graph->last_launches[r.kind] = c->host_status[1] + 1;
    release_persist(c);
    r.persist = false;
    *done = true;
    return PBBS_OK;
  }
  pbbs_ctx::Run& r = c->run;
  if (r.persist) {
    release_persist(c);
    r.persist = false;
    if (!c->host_status[2])
      return fail(PBBS_ERR_STATE, "persistent BFS launch ended before the BFS did");
    // the launch gave up waiting at a barrier (the device was shared): run the
    // BFS again from its start, one level per launch
    return restart_levels(c);
  }
  if (r.launched > r.limit)
    return fail(PBBS_ERR_STATE, "BFS did not finish after %lld level launches",
                (long long)r.launched);
  r.K = next_chunk(c->tuning.chunk_levels, (int)std::min<int64_t>(r.launched, 1 << 20));
  return launch_chunk(c);
}

static int poll_run(pbbs_ctx* c, bool wait, bool* done) {
  const int rc = poll_run_impl(c, wait, done);
  if (rc != PBBS_OK) release_persist(c);
  return rc;
}

// Launches in a run's first chunk: BFS depths vary little between sources of
// one graph, so after the first run the chunk covers the last run's launches
// plus one (a launch past the end costs a few microseconds; one short costs
// a host round trip), at most kMaxFirstChunk.  A first run starts with the
// tuning's chunk_levels.
constexpr int kMaxFirstChunk = 1024;
static int first_chunk(const pbbs_ctx* c, const pbbs_graph* g, int kind) {
  const int last = g->last_launches[kind];
  return last > 0 ? std::min(last + 1, kMaxFirstChunk) : c->tuning.chunk_levels;
}

static int level_variant(const pbbs_graph* g, const uint32_t* tops) {
  return !g->sorted ? kLevelUnsorted : tops ? kLevelSortedTops : kLevelSorted;
}

// Workgroups of the persistent kernel that fit on the device at once.
static int persist_capacity(pbbs_ctx* c, bool off64, int variant) {
  int per_cu = 0;
  const void* fn =
      off64 ? persist_kernel_symbol<int64_t>(variant) : persist_kernel_symbol<uint32_t>(variant);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0) != hipSuccess)
    return 0;
  return std::min(per_cu * c->num_cus, kMaxGrid);
}

// Workgroups of a persistent launch: the tuning's grid_blocks if set, else
// one per CU (on the 216^3 torus, 325 levels: 6.97 ms on 256 workgroups,
// 7.87 on 128, 8.77 on 512, 13.9 on 1024; the barrier's cost grows with
// the workgroups that meet at it); never more than fit at once.
static int persist_grid(pbbs_ctx* c, int cap) {
  const int want = c->tuning.grid_blocks > 0 ? std::min(c->tuning.grid_blocks, kMaxGrid)
                                             : c->num_cus;
  return std::min(want, cap);
}

// persist = 2 (auto) runs every level in one launch once the graph's
// previous single-source run needed at least this many levels: there the
// per-level launch and drain cost dominates (C3, 325 levels: 6.97 ms
// persistent against 9.36 ms one launch per level), while shallow graphs
// run faster one launch per level (C4: 0.53 against 0.79 ms).
constexpr int kPersistMinLevels = 64;

static bool want_persist(const pbbs_ctx* c, const pbbs_graph* g) {
  if (c->tuning.persist == 1) return true;
  if (c->tuning.persist != 2) return false;
  return g->last_launches[0] - 1 >= kPersistMinLevels;   // levels + 1 launches
}

// Starts a single-source BFS.  With allow_persist, every level runs in one
// persistent launch when the tuning asks for it and this context can take
// the device's token; otherwise the levels run one per launch, in captured
// chunks.  restart: the run's first attempt gave up; keep its start event.
// ws: the run's working memory; own: it is the context's own workspace (so
// captured launch graphs may be used, and per-graph search data built on
// first use); otherwise nothing is allocated: the graph must be prepared and
// the levels run in one persistent launch or as plain launches.
template <typename OffT>
static int begin_bfs(pbbs_ctx* c, pbbs_graph* g, int32_t s, int32_t* d_parent, int32_t* d_depth,
                     const SingleWs& ws, bool own, bool allow_persist = true,
                     bool restart = false, bool any_parent = false) {
  const uint32_t* tops = nullptr;
  int rc = any_parent ? PBBS_OK : ensure_tops<OffT>(c, g, &tops, own);
  if (rc != PBBS_OK) return rc;
  const int variant = any_parent ? kLevelAnyParent : level_variant(g, tops);
  int grid = level_grid(c, g->off64, variant);
  bool persist = false;
  if (allow_persist && want_persist(c, g)) {
    const int cap = persist_capacity(c, g->off64, variant);
    if (cap >= 1 && take_persist(c)) {
      persist = true;
      grid = persist_grid(c, cap);
    }
  }
  const uint32_t* bounds = nullptr;
  rc = ensure_bounds(c, g, (int64_t)grid * kWavesPerBlock, &bounds, own);
  if (rc != PBBS_OK) {
    release_persist(c);
    return rc;
  }
  BfsArgs<OffT> a = make_args<OffT>(c, g, ws, d_parent, d_depth, bounds, tops);
  const int init_grid =
      (int)std::max<int64_t>(1, std::min<int64_t>((g->n + kBlock - 1) / kBlock, c->num_cus * 16));
  pbbs_ctx::Run& r = c->run;
  r.kind = 0;
  r.off64 = g->off64;
  r.sorted = variant;
  r.grid = grid;
  r.K = first_chunk(c, g, r.kind);
  r.graph = g;
  r.launched = 0;
  r.finalized = false;
  r.limit = g->n + 2;
  r.persist = persist;
  r.any_parent = any_parent;
  r.source = s;
  r.parent = d_parent;
  r.depth = d_depth;
  r.ws = ws;
  r.graphs = own;
  if (c->tuning.use_graphs && own && !persist) {  // capture (first use only) while the stream is idle
    hipGraphExec_t exec = nullptr;
    rc = get_level_graph<OffT>(c, variant, grid, r.K, &exec);
    if (rc != PBBS_OK) return rc;
  }
  volatile int32_t* hs = c->host_status;
  hs[0] = 0;
  hs[1] = 0;
  hs[2] = 0;
  if (!restart) HIP_TRY(hipEventRecord(c->ev_start, c->stream));
  HIP_TRY(launch_bfs_init<OffT>(a, static_cast<BfsArgs<OffT>*>(ws.args), g->d_iso, s, init_grid,
                                c->stream));
  if (persist) {
    HIP_TRY(launch_bfs_persist<OffT>(static_cast<const BfsArgs<OffT>*>(ws.args), variant, grid,
                                     c->stream));
    r.launched = 1;
    HIP_TRY(hipEventRecord(c->ev_end, c->stream));
    return PBBS_OK;
  }
  return launch_chunk(c);
}

static int restart_levels(pbbs_ctx* c) {
  pbbs_ctx::Run& r = c->run;
  const SingleWs ws = r.ws;
  const bool own = r.graphs;
  return r.off64 ? begin_bfs<int64_t>(c, r.graph, r.source, r.parent, r.depth, ws, own, false,
                                      true, r.any_parent)
                 : begin_bfs<uint32_t>(c, r.graph, r.source, r.parent, r.depth, ws, own, false,
                                       true, r.any_parent);
}

static int single_stats(pbbs_ctx* c, pbbs_bfs_stats* stats) {
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev_start, c->ev_end));
  // the kernel that ended the BFS wrote the counters to pinned host_state
  // (ordered before host_status[0], which the completed event covers)
  std::memset(stats, 0, sizeof *stats);
  stats->levels = c->host_state->max_depth + 1;
  stats->td_levels = c->host_state->td_levels;
  stats->bu_levels = c->host_state->bu_levels;
  stats->launches = (int32_t)c->run.launched;
  stats->reached = (int64_t)c->host_state->reached;
  stats->edges_reached = (int64_t)c->host_state->edges_reached;
  stats->device_ms = ms;
  return PBBS_OK;
}

template <typename OffT>
static int run_bfs(pbbs_ctx* c, pbbs_graph* g, int32_t s, int32_t* d_parent, int32_t* d_depth,
                   pbbs_bfs_stats* stats, bool any_parent) {
  int rc = begin_bfs<OffT>(c, g, s, d_parent, d_depth, own_ws(c, g->n), true, true, false,
                           any_parent);
  bool done = false;
  while (rc == PBBS_OK && !done) rc = poll_run(c, true, &done);
  if (rc != PBBS_OK) {
    release_persist(c);
    return rc;
  }
  return stats ? single_stats(c, stats) : PBBS_OK;
}

extern "C" int pbbs_bfs(pbbs_ctx* c, pbbs_graph* g, int32_t source, int32_t* parent,
                        int32_t* depth, uint32_t out_flags, pbbs_bfs_stats* stats) {
  clear_error();
  if (!c || !g) return fail(PBBS_ERR_INVALID, "NULL context or graph");
  if (g->device != c->device) return fail(PBBS_ERR_INVALID, "graph lives on another device");
  if (source < 0 || (int64_t)source >= g->n)
    return fail(PBBS_ERR_INVALID, "source %d outside [0, %lld)", source, (long long)g->n);
  HIP_TRY(hipSetDevice(c->device));
  int rc = ensure_workspace(c, g->n);
  if (rc != PBBS_OK) return rc;
  if (out_flags & ~(PBBS_OUT_DEVICE | PBBS_BFS_ANY_PARENT))
    return fail(PBBS_ERR_INVALID, "unknown flags 0x%x", out_flags);
  const bool dev_out = (out_flags & PBBS_OUT_DEVICE) != 0;
  const bool any = (out_flags & PBBS_BFS_ANY_PARENT) != 0;
  int32_t* dp = (dev_out && parent) ? parent : c->w_parent;
  int32_t* dd = (dev_out && depth) ? depth : c->w_depth;
  rc = g->off64 ? run_bfs<int64_t>(c, g, source, dp, dd, stats, any)
                : run_bfs<uint32_t>(c, g, source, dp, dd, stats, any);
  if (rc != PBBS_OK) return rc;
  if (!dev_out && (parent || depth)) {
    HIP_TRY(hipEventRecord(c->ev_c0, c->stream));
    if (parent)
      HIP_TRY(hipMemcpyAsync(parent, dp, g->n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    if (depth)
      HIP_TRY(hipMemcpyAsync(depth, dd, g->n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(c->ev_c1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev_c1));
    if (stats) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev_c0, c->ev_c1));
      stats->copy_ms = ms;
    }
  }
  return PBBS_OK;
}

// ------------------------------------------------- maximal independent set
// Rounds run in chunks of 8 launches on the context's stream (the
// relabelled benchmark graphs finish in 6-9 rounds, and a launch after the
// last round still costs its dispatch); after each chunk the host reads whether a round found the worklist
// empty.  A round decides at least the lowest undecided vertex (its lower
// neighbours were all decided before the round began), so n + 1 rounds end
// any run; more means a fault, reported as an error.
template <typename OffT>
static int run_mis(pbbs_ctx* c, pbbs_graph* g, int32_t* d_out, pbbs_mis_stats* stats) {
  MisArgs<OffT> a{};
  a.off = static_cast<const OffT*>(g->d_off);
  a.adj = g->d_adj;
  a.flags = c->s_flags;
  a.pos = c->s_pos;
  a.region = mis_region(c->cap_sn);
  a.list[0] = c->s_list;
  a.list[1] = c->s_list + kMisStripes * a.region;
  a.ctr = c->s_ctl;
  a.part = c->s_part;
  a.wg_round = c->s_wg;
  a.finished = reinterpret_cast<int32_t*>(c->s_ctl + kMisCtlFinished);
  a.host_status = c->host_status;
  a.n = g->n;
  a.sorted = g->sorted ? 1 : 0;
  const int64_t need = (g->n + kBlock - 1) / kBlock;
  const int grid = (int)std::min<int64_t>(std::max<int64_t>(need, 1), 8 * c->num_cus);
  // row entries per look: 4 where rows average 16 entries or more, else 2
  // (measured on the benchmark's rMat, randLocal and 3D-grid graphs)
  const int step = g->m >= 16 * g->n ? 4 : 2;
  volatile int32_t* hs = c->host_status;
  HIP_TRY(hipEventRecord(c->ev_start, c->stream));
  HIP_TRY(launch_mis_init<OffT>(a, grid, c->stream));
  const int64_t limit = g->n + 2;
  int64_t launched = 0;
  const int K = 8;
  for (;;) {
    const int k = (int)std::min<int64_t>(K, limit - launched);
    for (int i = 0; i < k; ++i) HIP_TRY(launch_mis_round<OffT>(a, step, grid, c->stream));
    launched += k;
    HIP_TRY(hipEventRecord(c->ev_end, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev_end));
    if (hs[2]) return fail(PBBS_ERR_STATE, "set worklist region overflow");
    if (hs[0]) break;
    if (launched >= limit)
      return fail(PBBS_ERR_STATE, "set rounds did not finish in %lld launches",
                  (long long)launched);
  }
  const int32_t rounds = hs[1];
  HIP_TRY(launch_mis_output<OffT>(a, d_out, c->s_ctl + kMisCtlMembers, grid, c->stream));
  HIP_TRY(hipMemcpyAsync(c->host_status + 4, c->s_ctl + kMisCtlMembers, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipEventRecord(c->ev_end, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev_end));
  if (stats) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev_start, c->ev_end));
    unsigned long long members = 0;
    std::memcpy(&members, c->host_status + 4, sizeof(members));
    stats->rounds = rounds;
    stats->pad = 0;
    stats->members = (int64_t)members;
    stats->device_ms = ms;
    stats->copy_ms = 0.0;
  }
  return PBBS_OK;
}

extern "C" int pbbs_mis(pbbs_ctx* c, pbbs_graph* g, int32_t* flags, uint32_t out_flags,
                        pbbs_mis_stats* stats) {
  clear_error();
  if (!c || !g) return fail(PBBS_ERR_INVALID, "NULL context or graph");
  if (g->device != c->device) return fail(PBBS_ERR_INVALID, "graph lives on another device");
  if (out_flags & ~PBBS_OUT_DEVICE) return fail(PBBS_ERR_INVALID, "unknown flags 0x%x", out_flags);
  HIP_TRY(hipSetDevice(c->device));
  int rc = ensure_mis_workspace(c, g->n);
  if (rc != PBBS_OK) return rc;
  const bool dev_out = (out_flags & PBBS_OUT_DEVICE) != 0;
  int32_t* d_out = (dev_out && flags) ? flags : c->s_out;
  rc = g->off64 ? run_mis<int64_t>(c, g, d_out, stats) : run_mis<uint32_t>(c, g, d_out, stats);
  if (rc != PBBS_OK) return rc;
  if (!dev_out && flags) {
    HIP_TRY(hipEventRecord(c->ev_c0, c->stream));
    HIP_TRY(hipMemcpyAsync(flags, d_out, g->n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(c->ev_c1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev_c1));
    if (stats) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev_c0, c->ev_c1));
      stats->copy_ms = ms;
    }
  }
  return PBBS_OK;
}

// ----------------------------------------------------------- spanning forest
// The round schedule's constants for granularity 100 (spanningForest/
// incrementalST/ST.C:69; common/speculative_for.h:33-38).
constexpr int64_t kStGranularity = 100;

extern "C" int pbbs_spanning_forest(pbbs_ctx* c, int64_t n, int64_t m, const int32_t* u,
                                    const int32_t* v, uint32_t flags, int32_t* forest,
                                    int64_t* count, pbbs_st_stats* stats) {
  clear_error();
  if (!c || !count) return fail(PBBS_ERR_INVALID, "NULL context or count");
  if (flags & ~(PBBS_OUT_DEVICE | PBBS_EDGES_DEVICE | PBBS_ST_ANY_FOREST))
    return fail(PBBS_ERR_INVALID, "unknown flags 0x%x", flags);
  const bool any = (flags & PBBS_ST_ANY_FOREST) != 0;
  if (n <= 0 || n > INT32_MAX) return fail(PBBS_ERR_LIMIT, "n=%lld outside [1, 2^31)", (long long)n);
  if (m < 0 || m >= (int64_t)UINT32_MAX)
    return fail(PBBS_ERR_LIMIT, "m=%lld outside [0, 2^32 - 1)", (long long)m);
  if (m > 0 && (!u || !v)) return fail(PBBS_ERR_INVALID, "NULL edge array");
  const int64_t max_round = m / kStGranularity + 1;
  if (!any && m > 0 && max_round / 4 == 0)
    return fail(PBBS_ERR_LIMIT,
                "m=%lld < %lld: the reference's round schedule starts at size 0 and gives up "
                "(common/speculative_for.h:36-40)",
                (long long)m, (long long)(3 * kStGranularity));
  HIP_TRY(hipSetDevice(c->device));
  const bool dev_edges = (flags & PBBS_EDGES_DEVICE) != 0;
  int rc = ensure_st_workspace(c, n, max_round, dev_edges ? 0 : m);
  if (rc != PBBS_OK) return rc;
  StArgs a{};
  a.n = n;
  a.m = m;
  a.max_round = max_round;
  a.max_tries = 100 + 200 * kStGranularity;
  a.parents = c->t_parents;
  a.R = c->t_R;
  a.I = reinterpret_cast<uint32_t*>(c->t_round);
  a.su = reinterpret_cast<int32_t*>(c->t_round + 4 * c->cap_tr);
  a.sv = reinterpret_cast<int32_t*>(c->t_round + 8 * c->cap_tr);
  a.hold[0] = reinterpret_cast<uint32_t*>(c->t_round + 12 * c->cap_tr);
  a.hold[1] = reinterpret_cast<uint32_t*>(c->t_round + 16 * c->cap_tr);
  a.keep = c->t_round + 20 * c->cap_tr;
  a.blk = c->t_blk;
  a.ctl = c->t_ctl;
  a.host_status = c->host_status;
  double copy_ms = 0.0;
  if (dev_edges) {
    a.eu = u;
    a.ev = v;
  } else {
    a.eu = c->t_edges;
    a.ev = c->t_edges + c->cap_te;
    HIP_TRY(hipEventRecord(c->ev_c0, c->stream));
    if (m > 0) {
      HIP_TRY(hipMemcpyAsync(c->t_edges, u, m * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(c->t_edges + c->cap_te, v, m * sizeof(int32_t),
                             hipMemcpyHostToDevice, c->stream));
    }
    HIP_TRY(hipEventRecord(c->ev_c1, c->stream));
  }
  // round launches: one thread per iterate of the largest round, up to
  // kStMaxGroups workgroups (the finds are chains of dependent loads, so
  // more of them in flight is what hides their latency)
  const int64_t tiles = (max_round + kBlock - 1) / kBlock;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, kStMaxGroups));
  const int init_grid = (int)std::min<int64_t>(std::max<int64_t>((n + kBlock - 1) / kBlock, 1),
                                               8 * c->num_cus);
  volatile int32_t* hs = c->host_status;
  HIP_TRY(hipEventRecord(c->ev_start, c->stream));
  HIP_TRY(launch_st_init(a, init_grid, c->stream));
  // at least ceil(m / max_round) rounds run (granularity of them at full
  // size), so the first chunk launches that many; then chunks of 8
  const int64_t limit = a.max_tries + 2;
  int64_t launched = 0;
  int K = (int)std::min<int64_t>(std::max<int64_t>((m + max_round - 1) / max_round, 8), limit);
  if (any) {   // one pass over the edges, no schedule
    const int any_grid = (int)std::min<int64_t>(std::max<int64_t>((m + kBlock - 1) / kBlock, 1),
                                                16 * c->num_cus);
    if (m > 0) HIP_TRY(launch_st_any(a, any_grid, c->stream));
    K = 0;
  }
  for (; K > 0;) {
    const int k = (int)std::min<int64_t>(K, limit - launched);
    for (int i = 0; i < k; ++i) HIP_TRY(launch_st_round(a, grid, c->stream));
    launched += k;
    HIP_TRY(hipEventRecord(c->ev_end, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev_end));
    if (hs[0]) break;
    if (launched >= limit)
      return fail(PBBS_ERR_STATE, "forest rounds did not finish in %lld launches",
                  (long long)launched);
    K = 8;
  }
  int32_t* d_out = (flags & PBBS_OUT_DEVICE) && forest ? forest : c->t_out;
  HIP_TRY(launch_st_output(a, d_out, grid, c->stream));
  StCtl hc{};
  HIP_TRY(hipMemcpyAsync(&hc, c->t_ctl, sizeof hc, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipEventRecord(c->ev_end, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev_end));
  if (hc.error & 1) return fail(PBBS_ERR_INVALID, "an edge names a vertex outside [0, %lld)",
                                (long long)n);
  if (hc.error & 2) return fail(PBBS_ERR_STATE, "the round schedule gave up after %lld rounds",
                                (long long)hc.rounds);
  if (hc.error & 4) return fail(PBBS_ERR_STATE, "an edge's hook did not settle");
  float dev_ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&dev_ms, c->ev_start, c->ev_end));
  if (!dev_edges) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev_c0, c->ev_c1));
    copy_ms = ms;
  }
  *count = hc.count;
  if (!(flags & PBBS_OUT_DEVICE) && forest && hc.count > 0) {
    HIP_TRY(hipEventRecord(c->ev_c0, c->stream));
    HIP_TRY(hipMemcpyAsync(forest, d_out, hc.count * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipEventRecord(c->ev_c1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev_c1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev_c0, c->ev_c1));
    copy_ms += ms;
  }
  if (stats) {
    stats->rounds = (int32_t)hc.rounds;
    stats->pad = 0;
    stats->edges = hc.count;
    stats->device_ms = dev_ms;
    stats->copy_ms = copy_ms;
  }
  return PBBS_OK;
}

// ------------------------------------------------------------ level profile
template <typename OffT>
static int run_profile(pbbs_ctx* c, pbbs_graph* g, int32_t s, pbbs_level_profile* out,
                       int32_t max_levels, int32_t* count, double* init_ms) {
  const uint32_t* tops = nullptr;
  int brc = ensure_tops<OffT>(c, g, &tops);
  if (brc != PBBS_OK) return brc;
  const int variant = level_variant(g, tops);
  const int grid = level_grid(c, g->off64, variant);
  const uint32_t* bounds = nullptr;
  brc = ensure_bounds(c, g, (int64_t)grid * kWavesPerBlock, &bounds);
  if (brc != PBBS_OK) return brc;
  BfsArgs<OffT> a = make_args<OffT>(c, g, own_ws(c, g->n), c->w_parent, c->w_depth, bounds, tops);
  const int init_grid =
      (int)std::max<int64_t>(1, std::min<int64_t>((g->n + kBlock - 1) / kBlock, c->num_cus * 16));
  volatile int32_t* hs = c->host_status;
  hs[0] = 0;
  hs[1] = 0;
  std::vector<hipEvent_t> ev;
  auto cleanup = [&]() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  };
  auto new_event = [&]() -> hipEvent_t {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ev.push_back(e);
    return e;
  };
  hipEvent_t e0 = new_event(), e1 = new_event();
  if (!e0 || !e1) {
    cleanup();
    return fail(PBBS_ERR_HIP, "event creation failed");
  }
  int rc = PBBS_OK;
  std::vector<float> ms;
  do {
    if (hipEventRecord(e0, c->stream) != hipSuccess ||
        launch_bfs_init<OffT>(a, static_cast<BfsArgs<OffT>*>(c->w_args), g->d_iso, s, init_grid,
                              c->stream) != hipSuccess ||
        hipEventRecord(e1, c->stream) != hipSuccess) {
      rc = fail(PBBS_ERR_HIP, "init launch failed");
      break;
    }
    hipEvent_t prev = e1;
    for (int64_t k = 0; k <= g->n + 1; ++k) {
      hipEvent_t e = new_event();
      if (!e || launch_bfs_level<OffT>(static_cast<const BfsArgs<OffT>*>(c->w_args), variant,
                                       grid, c->stream) != hipSuccess ||
          hipEventRecord(e, c->stream) != hipSuccess || hipEventSynchronize(e) != hipSuccess) {
        rc = fail(PBBS_ERR_HIP, "level launch failed");
        break;
      }
      float t = 0.f;
      (void)hipEventElapsedTime(&t, prev, e);
      ms.push_back(t);
      prev = e;
      if (hs[0]) break;
    }
  } while (false);
  if (rc == PBBS_OK) {
    float ti = 0.f;
    (void)hipEventElapsedTime(&ti, e0, e1);
    if (init_ms) *init_ms = ti;
    if (hipMemcpy(c->host_state, c->w_state, sizeof(BfsState), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(PBBS_ERR_HIP, "state read-back failed");
  }
  cleanup();
  if (rc != PBBS_OK) return rc;
  // the last launch saw an empty frontier; the ones before it are the levels
  const int32_t levels = (int32_t)ms.size() - 1;
  const int32_t nrec = std::min<int32_t>(std::min<int32_t>(levels, max_levels), kLevelLogMax);
  for (int32_t k = 0; k < nrec; ++k) {
    out[k].level = k;
    out[k].mode = c->host_state->log[k].mode;
    out[k].frontier = c->host_state->log[k].count;
    out[k].frontier_edges = (int64_t)c->host_state->log[k].edges;
    out[k].ms = ms[k];
  }
  *count = levels;
  return PBBS_OK;
}

extern "C" int pbbs_bfs_profile(pbbs_ctx* c, pbbs_graph* g, int32_t source,
                                pbbs_level_profile* out, int32_t max_levels, int32_t* count,
                                double* init_ms) {
  clear_error();
  if (!c || !g || !count || (max_levels > 0 && !out))
    return fail(PBBS_ERR_INVALID, "NULL argument");
  if (g->device != c->device) return fail(PBBS_ERR_INVALID, "graph lives on another device");
  if (source < 0 || (int64_t)source >= g->n)
    return fail(PBBS_ERR_INVALID, "source %d outside [0, %lld)", source, (long long)g->n);
  HIP_TRY(hipSetDevice(c->device));
  const int rc = ensure_workspace(c, g->n);
  if (rc != PBBS_OK) return rc;
  return g->off64 ? run_profile<int64_t>(c, g, source, out, max_levels, count, init_ms)
                  : run_profile<uint32_t>(c, g, source, out, max_levels, count, init_ms);
}

// ------------------------------------------------------------- preparation
static int batch_grid(pbbs_ctx* c, bool off64, bool sorted);

template <typename OffT>
static int prepare_graph(pbbs_ctx* c, pbbs_graph* g, uint32_t what) {
  if (what & PBBS_PREPARE_SINGLE) {
    const uint32_t* tops = nullptr;
    int rc = ensure_tops<OffT>(c, g, &tops);
    if (rc != PBBS_OK) return rc;
    const int variant = level_variant(g, tops);
    const int grid = level_grid(c, g->off64, variant);
    const uint32_t* bounds = nullptr;
    rc = ensure_bounds(c, g, (int64_t)grid * kWavesPerBlock, &bounds);
    if (rc != PBBS_OK) return rc;
    // and for the grid of a persistent launch
    const int cap = persist_capacity(c, g->off64, variant);
    if (cap >= 1 && persist_grid(c, cap) != grid) {
      rc = ensure_bounds(c, g, (int64_t)persist_grid(c, cap) * kWavesPerBlock, &bounds);
      if (rc != PBBS_OK) return rc;
    }
  }
  if ((what & PBBS_PREPARE_BATCH) && g->symmetric) {
    int rc = ensure_ranks<OffT>(c, g);
    if (rc != PBBS_OK) return rc;
    const int32_t* rows = nullptr;
    rc = ensure_pull_rows<OffT>(c, g, (int64_t)batch_grid(c, g->off64, g->sorted) * kWavesPerBlock,
                                &rows);
    if (rc != PBBS_OK) return rc;
  }
  return PBBS_OK;
}

extern "C" int pbbs_graph_prepare(pbbs_ctx* c, pbbs_graph* g, uint32_t what, double* ms) {
  clear_error();
  if (!c || !g) return fail(PBBS_ERR_INVALID, "NULL context or graph");
  if (g->device != c->device) return fail(PBBS_ERR_INVALID, "graph lives on another device");
  if (what & ~(uint32_t)(PBBS_PREPARE_SINGLE | PBBS_PREPARE_BATCH))
    return fail(PBBS_ERR_INVALID, "unknown prepare flags 0x%x", what);
  HIP_TRY(hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  const int rc = g->off64 ? prepare_graph<int64_t>(c, g, what) : prepare_graph<uint32_t>(c, g, what);
  if (rc != PBBS_OK) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (ms)
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return PBBS_OK;
}

// --------------------------------------------------------------- batched BFS
static int ensure_batch_workspace(pbbs_ctx* c, int64_t n, int64_t out_entries, bool need_out) {
  if (!c->b_state) {
    // all four control buffers or none: b_state marks the set as present
    MbfsState* st = nullptr;
    MbfsState* hst = nullptr;
    void* args = nullptr;
    int32_t* src = nullptr;
    if (hipMalloc(&st, sizeof(MbfsState)) != hipSuccess ||
        hipHostMalloc((void**)&hst, sizeof(MbfsState), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&args, sizeof(MbfsArgs<int64_t>)) != hipSuccess ||
        hipMalloc(&src, kMaxSources * sizeof(int32_t)) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(st);
      if (hst) (void)hipHostFree(hst);
      (void)hipFree(args);
      (void)hipFree(src);
      return fail(PBBS_ERR_NOMEM, "batched control buffers failed");
    }
    c->b_host_state = hst;
    c->b_args = args;
    c->b_sources = src;
    c->b_state = st;
  }
  if (n > c->cap_bn) {
    (void)hipFree(c->b_masks);
    (void)hipFree(c->b_dlev);
    (void)hipFree(c->b_q);
    (void)hipFree(c->b_qe);
    c->b_masks = nullptr;
    c->b_dlev = nullptr;
    c->b_q = nullptr;
    c->b_qe = nullptr;
    c->cap_bn = 0;
    // the union frontier's sub-queues hold n entries each: batched push
    // levels have no bitmap fallback, so a sub-queue must never fill
    if (hipMalloc(&c->b_masks, (5 * n + 4 * ((n + kWave - 1) / kWave)) *
                                   sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&c->b_dlev, n * kMaxSources) != hipSuccess ||
        hipMalloc(&c->b_q, 2 * kSubQueues * n * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&c->b_qe, 2 * kSubQueues * n * sizeof(uint32_t)) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(c->b_masks);
      (void)hipFree(c->b_dlev);
      (void)hipFree(c->b_q);
      (void)hipFree(c->b_qe);
      c->b_masks = nullptr;
      c->b_dlev = nullptr;
      c->b_q = nullptr;
      c->b_qe = nullptr;
      return fail(PBBS_ERR_NOMEM, "batched mask buffers for n=%lld failed", (long long)n);
    }
    c->cap_bn = n;
  }
  if (need_out && out_entries > c->cap_bout) {
    HIP_TRY(hipStreamSynchronize(c->side));   // a fill of the old alt buffer may be running
    (void)hipFree(c->b_parent);
    (void)hipFree(c->b_parent_alt);
    (void)hipFree(c->b_depth);
    c->b_parent = c->b_parent_alt = c->b_depth = nullptr;
    c->alt_cleared = 0;
    c->cap_bout = 0;
    if (hipMalloc(&c->b_parent, out_entries * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&c->b_parent_alt, out_entries * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&c->b_depth, out_entries * sizeof(int32_t)) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(c->b_parent);
      (void)hipFree(c->b_parent_alt);
      (void)hipFree(c->b_depth);
      c->b_parent = c->b_parent_alt = c->b_depth = nullptr;
      return fail(PBBS_ERR_NOMEM, "batched output buffers (%lld entries) failed",
                  (long long)out_entries);
    }
    c->cap_bout = out_entries;
  }
  return PBBS_OK;
}

static int batch_grid(pbbs_ctx* c, bool off64, bool sorted) {
  if (c->tuning.grid_blocks > 0) return std::min(c->tuning.grid_blocks, kMaxGrid);
  int per_cu = 0;
  const void* fn = off64 ? mbfs_level_symbol<int64_t>(sorted) : mbfs_level_symbol<uint32_t>(sorted);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return std::min(per_cu * c->num_cus, kMaxGrid);
}

template <typename OffT>
static int get_batch_graph(pbbs_ctx* c, bool sorted, int grid, int K, hipGraphExec_t* out) {
  const auto key = std::make_tuple((int)sizeof(OffT) * 2 + (sorted ? 1 : 0), grid, K);
  auto it = c->batch_graphs.find(key);
  if (it != c->batch_graphs.end()) {
    *out = it->second;
    return PBBS_OK;
  }
  hipGraph_t graph = nullptr;
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int i = 0; i < K && le == hipSuccess; ++i)
    le = launch_mbfs_level<OffT>(static_cast<const MbfsArgs<OffT>*>(c->b_args), sorted, grid,
                                 c->stream);
  const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
  if (le != hipSuccess || ce != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(PBBS_ERR_HIP, "capturing %d batched level launches failed: %s", K,
                hipGetErrorString(le != hipSuccess ? le : ce));
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) return fail(PBBS_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ie));
  c->batch_graphs[key] = exec;
  *out = exec;
  return PBBS_OK;
}

// own_out: the outputs go to the context's own buffers (d_parent and d_depth
// are ignored), which lets the parents come pre-filled with -1 (see
// b_parent_alt); otherwise the clear kernel fills the caller's parents.
template <typename OffT>
static int begin_batch(pbbs_ctx* c, pbbs_graph* g, int32_t nsrc, int32_t* d_parent,
                       int32_t* d_depth, bool own_out) {
  const int64_t n = g->n;
  const int64_t entries = (int64_t)nsrc * n;
  bool preset = false;
  if (own_out) {
    if (c->b_parent_alt && c->alt_cleared >= entries) {
      std::swap(c->b_parent, c->b_parent_alt);
      c->alt_cleared = 0;   // it now holds the previous batch's parents
      preset = true;
    }
    d_parent = c->b_parent;
    d_depth = c->b_depth;
  }
  MbfsArgs<OffT> a;
  std::memset(&a, 0, sizeof a);
  a.off = static_cast<const OffT*>(g->d_off);
  a.adj = g->d_adj;
  a.parent = d_parent;
  a.depth = d_depth;
  a.vis = c->b_masks;
  a.dlev = c->b_dlev;
  for (int j = 0; j < 4; ++j) a.fr[j] = c->b_masks + (j + 1) * n;
  for (int j = 0; j < 4; ++j) a.fbm[j] = c->b_masks + 5 * n + j * ((n + kWave - 1) / kWave);
  a.q[0] = c->b_q;
  a.q[1] = c->b_q + kSubQueues * c->cap_bn;
  a.qe[0] = c->b_qe;
  a.qe[1] = c->b_qe + kSubQueues * c->cap_bn;
  a.qcap = c->cap_bn;
  a.st = c->b_state;
  a.wg_level = c->w_wg_level;
  a.host_status = c->host_status;
  a.sources = c->b_sources;
  a.n = n;
  a.m = g->m;
  a.nwords = (n + kWave - 1) / kWave;
  a.nsrc = nsrc;
  a.srcmask = nsrc >= 64 ? ~0ull : ((1ull << nsrc) - 1);
  a.td_to_bu_div = c->tuning.td_to_bu_div;
  a.pull_div = c->tuning.batch_pull_div;
  a.num_cus = c->num_cus;
  a.preset_parents = preset ? 1 : 0;
  pbbs_ctx::Run& r = c->run;
  r.kind = 1;
  r.off64 = g->off64;
  r.sorted = g->sorted;
  r.grid = batch_grid(c, g->off64, g->sorted);
  {
    const int32_t* rows = nullptr;
    int rc = ensure_ranks<OffT>(c, g);
    if (rc == PBBS_OK) rc = ensure_pull_rows<OffT>(c, g, (int64_t)r.grid * kWavesPerBlock, &rows);
    if (rc != PBBS_OK) return rc;
    a.pull_rows = rows;
    a.rows = g->d_rows;
    a.nrows = g->nrows;
    a.rank = g->d_rank;
    a.inv_rank = g->d_inv_rank;
    a.adj_rank = g->d_adj_rank;
  }
  r.K = first_chunk(c, g, r.kind);
  r.graph = g;
  r.launched = 0;
  r.finalized = false;
  r.limit = n + 2;
  const int setup_grid =
      (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, c->num_cus * 16));
  if (c->tuning.use_graphs) {
    hipGraphExec_t exec = nullptr;
    const int rc = get_batch_graph<OffT>(c, r.sorted, r.grid, r.K, &exec);
    if (rc != PBBS_OK) return rc;
  }
  volatile int32_t* hs = c->host_status;
  hs[0] = 0;
  hs[1] = 0;
  HIP_TRY(hipEventRecord(c->ev_start, c->stream));
  if (preset) HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_alt, 0));   // inside the timed span
  HIP_TRY(launch_mbfs_setup<OffT>(a, static_cast<MbfsArgs<OffT>*>(c->b_args), setup_grid,
                                  c->stream));
  c->b_last_args_off64 = g->off64;
  std::memcpy(c->b_last_args, &a, sizeof a);
  const int rc = launch_chunk(c);
  if (rc != PBBS_OK) return rc;
  if (own_out && c->b_parent_alt && c->prefill) {
    // the other buffer's contents (the previous batch's outputs) are no
    // longer needed: fill it for the next batch while this one runs
    HIP_TRY(launch_fill_minus1(c->b_parent_alt, entries, std::max(1, c->num_cus / 8), c->side));
    HIP_TRY(hipEventRecord(c->ev_alt, c->side));
    c->alt_cleared = entries;
  }
  return PBBS_OK;
}

static int batch_stats(pbbs_ctx* c, pbbs_bfs_stats* stats) {
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev_start, c->ev_end));
  // the depth pass summed the totals; its copy to pinned memory precedes ev_end
  const MbfsState& hst = *c->b_host_state;
  std::memset(stats, 0, sizeof *stats);
  stats->levels = hst.max_depth + 1;
  stats->td_levels = hst.push_levels;
  stats->bu_levels = hst.pull_levels;
  stats->launches = (int32_t)c->run.launched;
  stats->reached = (int64_t)hst.reached_pairs;
  stats->edges_reached = (int64_t)hst.edges_pairs;
  stats->device_ms = ms;
  return PBBS_OK;
}

template <typename OffT>
static int run_batch(pbbs_ctx* c, pbbs_graph* g, int32_t nsrc, int32_t* d_parent, int32_t* d_depth,
                     bool own_out, pbbs_bfs_stats* stats) {
  int rc = begin_batch<OffT>(c, g, nsrc, d_parent, d_depth, own_out);
  bool done = false;
  while (rc == PBBS_OK && !done) rc = poll_run(c, true, &done);
  if (rc != PBBS_OK) return rc;
  return stats ? batch_stats(c, stats) : PBBS_OK;
}

extern "C" int pbbs_bfs_batch(pbbs_ctx* c, pbbs_graph* g, const int32_t* sources, int32_t nsrc,
                              int32_t* parents, int32_t* depths, uint32_t out_flags,
                              pbbs_bfs_stats* stats) {
  clear_error();
  if (!c || !g || !sources) return fail(PBBS_ERR_INVALID, "NULL argument");
  if (g->device != c->device) return fail(PBBS_ERR_INVALID, "graph lives on another device");
  if (nsrc < 1 || nsrc > kMaxSources)
    return fail(PBBS_ERR_INVALID, "source count %d outside [1, %d]", nsrc, kMaxSources);
  for (int i = 0; i < nsrc; ++i)
    if (sources[i] < 0 || (int64_t)sources[i] >= g->n)
      return fail(PBBS_ERR_INVALID, "source %d (%d) outside [0, %lld)", i, sources[i],
                  (long long)g->n);
  if (!g->symmetric)
    return fail(PBBS_ERR_INVALID,
                "batched BFS needs a symmetric graph (parents are found by scanning each "
                "vertex's row as its in-neighbours)");
  // the deterministic rows also meet the any-parent contract, so that flag
  // is accepted and changes nothing here
  if (out_flags & ~(PBBS_OUT_DEVICE | PBBS_BFS_ANY_PARENT))
    return fail(PBBS_ERR_INVALID, "unknown flags 0x%x", out_flags);
  HIP_TRY(hipSetDevice(c->device));
  const bool dev_out = (out_flags & PBBS_OUT_DEVICE) != 0;
  const int64_t entries = (int64_t)nsrc * g->n;
  int rc = ensure_batch_workspace(c, g->n, entries, !(dev_out && parents && depths));
  if (rc != PBBS_OK) return rc;
  // outputs to the caller's device buffers only when both are given
  const bool own_out = !(dev_out && parents && depths);
  HIP_TRY(hipMemcpyAsync(c->b_sources, sources, nsrc * sizeof(int32_t), hipMemcpyHostToDevice,
                         c->stream));
  rc = g->off64 ? run_batch<int64_t>(c, g, nsrc, parents, depths, own_out, stats)
                : run_batch<uint32_t>(c, g, nsrc, parents, depths, own_out, stats);
  if (rc != PBBS_OK) return rc;
  int32_t* dp = own_out ? c->b_parent : parents;
  int32_t* dd = own_out ? c->b_depth : depths;
  if (!dev_out && (parents || depths)) {
    HIP_TRY(hipEventRecord(c->ev_c0, c->stream));
    if (parents)
      HIP_TRY(hipMemcpyAsync(parents, dp, entries * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
    if (depths)
      HIP_TRY(hipMemcpyAsync(depths, dd, entries * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
    HIP_TRY(hipEventRecord(c->ev_c1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev_c1));
    if (stats) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev_c0, c->ev_c1));
      stats->copy_ms = ms;
    }
  }
  return PBBS_OK;
}

// ------------------------------------------------- entry points for the server
namespace pbbs {

// Device bytes a context allocates for one run on an n-vertex graph: the
// single-source workspace (nsrc <= 1: parent, depth, four bitmaps, two
// frontier queues of queue_cap entries per sub-queue, with their edge
// prefixes) or the batched one (nsrc >= 2: masks, bitmaps, level table,
// union queues of n entries per sub-queue, [nsrc][n] parents and depths).
int64_t bfs_workspace_bytes(int64_t n, int32_t nsrc, int32_t queue_div) {
  const int64_t nwords = (n + kWave - 1) / kWave;
  if (nsrc <= 1)
    return 8 * n + 4 * nwords * 8 + 2 * (int64_t)kSubQueues * queue_cap(n, queue_div) * 8;
  return 5 * n * 8 + 4 * nwords * 8 + n * kMaxSources + 2 * (int64_t)kSubQueues * n * 8 +
         3 * (int64_t)nsrc * n * 4;
}

// Device bytes the context's workspace buffers hold now (they grow with the
// largest run and are kept until ctx_trim or the context's end).
int64_t ctx_workspace_bytes(const pbbs_ctx* c) {
  const int64_t nwords = (c->cap_n + kWave - 1) / kWave;
  const int64_t bwords = (c->cap_bn + kWave - 1) / kWave;
  int64_t b = 0;
  if (c->w_parent) b += 8 * c->cap_n + 4 * nwords * 8 + 2 * (int64_t)kSubQueues * c->cap_q * 8;
  if (c->b_masks)
    b += 5 * c->cap_bn * 8 + 4 * bwords * 8 + c->cap_bn * kMaxSources +
         2 * (int64_t)kSubQueues * c->cap_bn * 8;
  if (c->b_parent) b += 2 * c->cap_bout * 4;
  if (c->b_parent_alt) b += c->cap_bout * 4;
  if (c->t_ctl) b += 12 * c->cap_tn + 21 * c->cap_tr + 8 * c->cap_te + kStMaxGroups * 4;
  if (c->s_flags)
    b += 9 * c->cap_sn + 8 * kMisStripes * mis_region(c->cap_sn) + kMisCtlWords * 8 +
         kMaxGrid * 12;
  return b;
}

// Device bytes the context's workspace will hold after a run of nsrc
// sources on an n-vertex graph (buffers only grow).
int64_t ctx_workspace_bytes_after(const pbbs_ctx* c, int64_t n, int32_t nsrc) {
  int64_t cn = c->w_parent ? c->cap_n : 0, cq = c->w_parent ? c->cap_q : 0;
  int64_t bn = c->b_masks ? c->cap_bn : 0, bo = c->b_parent ? c->cap_bout : 0;
  if (nsrc <= 1) {
    cn = std::max(cn, n);
    cq = std::max(cq, queue_cap(n, c->tuning.queue_div));
  } else {
    bn = std::max(bn, n);
    bo = std::max(bo, (int64_t)nsrc * n);
  }
  const int64_t nwords = (cn + kWave - 1) / kWave, bwords = (bn + kWave - 1) / kWave;
  int64_t b = 0;
  if (cn) b += 8 * cn + 4 * nwords * 8 + 2 * (int64_t)kSubQueues * cq * 8;
  if (bn) b += 5 * bn * 8 + 4 * bwords * 8 + bn * kMaxSources + 2 * (int64_t)kSubQueues * bn * 8;
  if (bo) b += 3 * bo * 4;   // parents, the pre-filled second parent buffer, depths
  return b;
}

// Frees the context's workspace buffers (the stream must be idle).
void ctx_trim(pbbs_ctx* c) {
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->side);
  free_workspace(c);
  free_mis_workspace(c);
  free_st_workspace(c);
  (void)hipFree(c->b_masks);
  (void)hipFree(c->b_dlev);
  (void)hipFree(c->b_q);
  (void)hipFree(c->b_qe);
  (void)hipFree(c->b_parent);
  (void)hipFree(c->b_parent_alt);
  (void)hipFree(c->b_depth);
  c->b_parent_alt = nullptr;
  c->alt_cleared = 0;
  c->b_masks = nullptr;
  c->b_dlev = nullptr;
  c->b_q = nullptr;
  c->b_qe = nullptr;
  c->b_parent = c->b_depth = nullptr;
  c->cap_bn = 0;
  c->cap_bout = 0;
}

int32_t ctx_queue_div(const pbbs_ctx* c) { return c->tuning.queue_div; }
void ctx_no_prefill(pbbs_ctx* c) { c->prefill = false; }

hipStream_t ctx_stream(pbbs_ctx* c) { return c->stream; }
int ctx_device(const pbbs_ctx* c) { return c->device; }
const void* graph_dev_off(const pbbs_graph* g) { return g->d_off; }
const int32_t* graph_dev_adj(const pbbs_graph* g) { return g->d_adj; }
bool graph_off64(const pbbs_graph* g) { return g->off64; }
int64_t graph_m(const pbbs_graph* g) { return g->m; }
int graph_device(const pbbs_graph* g) { return g->device; }
int ctx_num_cus(const pbbs_ctx* c) { return c->num_cus; }
int64_t graph_n(const pbbs_graph* g) { return g->n; }
bool graph_symmetric(const pbbs_graph* g) { return g->symmetric; }

// Starts a single-source BFS (nsrc == 1) or a batched one on c's stream,
// with outputs in c's own buffers.
int async_begin(pbbs_ctx* c, pbbs_graph* g, const int32_t* sources, int32_t nsrc) {
  clear_error();
  if (!c || !g || !sources || nsrc < 1 || nsrc > kMaxSources)
    return fail(PBBS_ERR_INVALID, "bad async BFS arguments");
  for (int i = 0; i < nsrc; ++i)
    if (sources[i] < 0 || (int64_t)sources[i] >= g->n)
      return fail(PBBS_ERR_INVALID, "source %d outside the graph", sources[i]);
  HIP_TRY(hipSetDevice(c->device));
  int rc;
  if (nsrc == 1) {
    rc = ensure_workspace(c, g->n);
    if (rc != PBBS_OK) return rc;
    const SingleWs ws = own_ws(c, g->n);
    rc = g->off64 ? begin_bfs<int64_t>(c, g, sources[0], c->w_parent, c->w_depth, ws, true)
                  : begin_bfs<uint32_t>(c, g, sources[0], c->w_parent, c->w_depth, ws, true);
    if (rc != PBBS_OK) release_persist(c);
    return rc;
  }
  if (!g->symmetric) return fail(PBBS_ERR_INVALID, "batched BFS needs a symmetric graph");
  rc = ensure_batch_workspace(c, g->n, (int64_t)nsrc * g->n, true);
  if (rc != PBBS_OK) return rc;
  HIP_TRY(hipMemcpyAsync(c->b_sources, sources, nsrc * sizeof(int32_t), hipMemcpyHostToDevice,
                         c->stream));
  return g->off64 ? begin_batch<int64_t>(c, g, nsrc, nullptr, nullptr, true)
                  : begin_batch<uint32_t>(c, g, nsrc, nullptr, nullptr, true);
}

int async_poll(pbbs_ctx* c, bool* done) { return poll_run(c, false, done); }

int async_outputs(pbbs_ctx* c, const int32_t** parent, const int32_t** depth) {
  if (c->run.kind == 0) {
    *parent = c->w_parent;
    *depth = c->w_depth;
  } else {
    *parent = c->b_parent;
    *depth = c->b_depth;
  }
  return PBBS_OK;
}

int async_stats(pbbs_ctx* c, pbbs_bfs_stats* stats) {
  return c->run.kind == 0 ? single_stats(c, stats) : batch_stats(c, stats);
}

}  // namespace pbbs

// ------------------------------------------------------ multi-graph launch
namespace pbbs {

constexpr int64_t kJobAlign = 256;
static int64_t job_align(int64_t x) { return (x + kJobAlign - 1) / kJobAlign * kJobAlign; }

// Device memory of one single-source BFS outside the context workspace:
// control block, per-workgroup level counters, argument block, the four
// bitmaps, the two frontier queues and their edge prefixes (queue_cap(n,
// queue_div) entries per sub-queue region), and optionally parent and depth.
struct JobLayout {
  int64_t st, wg, args, bits, q, qe, parent, depth, total, qcap;
};

static JobLayout job_layout_q(int64_t n, int64_t qcap, bool outputs) {
  JobLayout L;
  const int64_t nwords = (n + kWave - 1) / kWave;
  L.qcap = qcap;
  L.st = 0;
  L.wg = job_align(L.st + (int64_t)sizeof(BfsState));
  L.args = job_align(L.wg + kMaxGrid * (int64_t)sizeof(uint32_t));
  L.bits = job_align(L.args + (int64_t)sizeof(BfsArgs<int64_t>));
  L.q = job_align(L.bits + 4 * nwords * (int64_t)sizeof(unsigned long long));
  L.qe = job_align(L.q + 2 * kSubQueues * L.qcap * (int64_t)sizeof(int32_t));
  L.parent = job_align(L.qe + 2 * kSubQueues * L.qcap * (int64_t)sizeof(uint32_t));
  L.depth = outputs ? job_align(L.parent + n * (int64_t)sizeof(int32_t)) : L.parent;
  L.total = outputs ? job_align(L.depth + n * (int64_t)sizeof(int32_t)) : L.parent;
  return L;
}

static JobLayout job_layout(int64_t n, int32_t queue_div, bool outputs) {
  return job_layout_q(n, queue_cap(n, queue_div), outputs);
}

// BfsArgs of a BFS whose workspace is laid out by L at base.
template <typename OffT>
static BfsArgs<OffT> job_args(pbbs_ctx* c, pbbs_graph* g, const JobLayout& L, unsigned char* base,
                              int32_t* parent, int32_t* depth, int32_t* host_status,
                              BfsState* host_state) {
  const int64_t nwords = (g->n + kWave - 1) / kWave;
  unsigned long long* bits = reinterpret_cast<unsigned long long*>(base + L.bits);
  BfsArgs<OffT> a;
  std::memset(&a, 0, sizeof a);
  a.off = static_cast<const OffT*>(g->d_off);
  a.adj = g->d_adj;
  a.parent = parent;
  a.depth = depth;
  a.visited = bits;
  a.fb[0] = bits + nwords;
  a.fb[1] = bits + 2 * nwords;
  a.fb[2] = bits + 3 * nwords;
  a.q[0] = reinterpret_cast<int32_t*>(base + L.q);
  a.q[1] = a.q[0] + kSubQueues * L.qcap;
  a.qe[0] = reinterpret_cast<uint32_t*>(base + L.qe);
  a.qe[1] = a.qe[0] + kSubQueues * L.qcap;
  a.qcap = L.qcap;
  a.st = reinterpret_cast<BfsState*>(base + L.st);
  a.wg_level = reinterpret_cast<uint32_t*>(base + L.wg);
  a.host_status = host_status;
  a.host_state = host_state;
  a.bu_bounds = nullptr;   // equal shares: no per-size bounds are built for a job
  a.bu_tops = nullptr;
  a.n = g->n;
  a.m = g->m;
  a.nwords = nwords;
  a.td_to_bu_div = c->tuning.td_to_bu_div;
  a.bu_to_td_div = c->tuning.bu_to_td_div;
  a.bu_serial_depth = c->tuning.bu_serial_depth;
  a.symmetric = g->symmetric ? 1 : 0;
  a.num_cus = c->num_cus;
  a.small_rows = g->max_degree <= kWave ? 1 : 0;
  a.wait_ticks = (long long)c->tuning.persist_wait_us * c->wall_khz / 1000;
  return a;
}

int64_t multi_workspace_bytes(const int64_t* n, int32_t njobs, int32_t queue_div) {
  int64_t b = 0;
  for (int32_t j = 0; j < njobs; ++j) b += job_layout(n[j], queue_div, true).total;
  return b;
}

// Workgroups per job: proportional to n + m (at least one each, at most one
// per 4096 edges, so a small graph does not spread over the whole device),
// within cap in total.
static void split_grid(const std::vector<pbbs_graph*>& gs, int cap, std::vector<uint32_t>& wg) {
  const size_t J = gs.size();
  double total = 0;
  for (pbbs_graph* g : gs) total += (double)(g->n + g->m);
  std::vector<int64_t> w(J);
  int64_t sum = 0;
  for (size_t j = 0; j < J; ++j) {
    const int64_t share = (int64_t)((double)cap * (double)(gs[j]->n + gs[j]->m) / total);
    const int64_t most = std::max<int64_t>(1, (gs[j]->m + 4095) / 4096);
    w[j] = std::max<int64_t>(1, std::min(share, most));
    sum += w[j];
  }
  while (sum > cap) {   // the floors of one each pushed it over: trim the largest
    size_t big = 0;
    for (size_t j = 1; j < J; ++j)
      if (w[j] > w[big]) big = j;
    --w[big];
    --sum;
  }
  wg.assign(J + 1, 0);
  for (size_t j = 0; j < J; ++j) wg[j + 1] = wg[j] + (uint32_t)w[j];
}

template <typename OffT>
static int run_multi_group(pbbs_ctx* c, const std::vector<int>& ids,
                           const std::vector<pbbs_graph*>& graphs, const int32_t* sources,
                           const std::vector<JobLayout>& L, const std::vector<int64_t>& at,
                           std::vector<int32_t*>& dpar, std::vector<int32_t*>& ddep,
                           pbbs_bfs_stats* stats, bool sorted) {
  const int J = (int)ids.size();
  std::vector<pbbs_graph*> gs(J);
  for (int j = 0; j < J; ++j) gs[j] = graphs[ids[j]];
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, multi_kernel_symbol<OffT>(sorted),
                                                   kBlock, 0) != hipSuccess ||
      per_cu < 1)
    return fail(PBBS_ERR_HIP, "occupancy of the multi-graph kernel unavailable");
  const int cap = std::min(per_cu * c->num_cus, kMaxGrid);
  if (J > cap) return fail(PBBS_ERR_LIMIT, "%d jobs exceed the %d workgroups resident", J, cap);
  std::vector<uint32_t> wg;
  split_grid(gs, cap, wg);
  BfsArgs<OffT>* dargs = static_cast<BfsArgs<OffT>*>(c->m_args);
  for (int j = 0; j < J; ++j) {
    const int i = ids[j];
    unsigned char* base = c->m_ws + at[i];
    int32_t* hs = c->m_host_status + 16 * j;
    hs[0] = hs[1] = hs[2] = 0;
    const BfsArgs<OffT> a =
        job_args<OffT>(c, gs[j], L[i], base, dpar[i], ddep[i], hs, c->m_host_state + j);
    const int init_grid = (int)std::max<int64_t>(
        1, std::min<int64_t>((gs[j]->n + kBlock - 1) / kBlock, c->num_cus * 16));
    HIP_TRY(launch_bfs_init<OffT>(a, dargs + j, gs[j]->d_iso, sources[i], init_grid, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(c->m_wg, wg.data(), (J + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                         c->stream));
  const bool persist = take_persist(c);
  if (persist) {
    HIP_TRY(hipEventRecord(c->ev_start, c->stream));
    const hipError_t le = launch_bfs_multi<OffT>(dargs, c->m_wg, J, sorted, (int)wg[J], c->stream);
    const hipError_t re = le == hipSuccess ? hipEventRecord(c->ev_end, c->stream) : le;
    const hipError_t se = re == hipSuccess ? hipEventSynchronize(c->ev_end) : re;
    release_persist(c);
    if (se != hipSuccess)
      return fail(PBBS_ERR_HIP, "multi-graph launch: %s", hipGetErrorString(se));
  }
  float ms = 0.f;
  if (persist) (void)hipEventElapsedTime(&ms, c->ev_start, c->ev_end);
  for (int j = 0; j < J; ++j) {
    const int i = ids[j];
    volatile int32_t* hs = c->m_host_status + 16 * j;
    if (persist && hs[0]) {
      const BfsState& h = c->m_host_state[j];
      pbbs_bfs_stats& st = stats[i];
      std::memset(&st, 0, sizeof st);
      st.levels = h.max_depth + 1;
      st.td_levels = h.td_levels;
      st.bu_levels = h.bu_levels;
      st.launches = 1;
      st.reached = (int64_t)h.reached;
      st.edges_reached = (int64_t)h.edges_reached;
      st.device_ms = ms;
      continue;
    }
    // the device was busy with another persistent launch, or this job's
    // workgroups could not gather: run it alone
    int rc = ensure_workspace(c, gs[j]->n);
    if (rc == PBBS_OK)
      rc = run_bfs<OffT>(c, gs[j], sources[i], dpar[i], ddep[i], &stats[i], false);
    if (rc != PBBS_OK) return rc;
  }
  return PBBS_OK;
}

}  // namespace pbbs

// Jobs on graphs with more edges than this run alone inside pbbs_bfs_multi.
constexpr int64_t kMultiAloneEdges = int64_t(1) << 23;

extern "C" int pbbs_bfs_multi(pbbs_ctx* c, int32_t njobs, pbbs_graph* const* graphs,
                              const int32_t* sources, int32_t* const* parents,
                              int32_t* const* depths, uint32_t out_flags, pbbs_bfs_stats* stats) {
  clear_error();
  if (!c || !graphs || !sources || !stats) return fail(PBBS_ERR_INVALID, "NULL argument");
  if (njobs < 1 || njobs > kMaxMultiJobs)
    return fail(PBBS_ERR_INVALID, "job count %d outside [1, %d]", njobs, kMaxMultiJobs);
  for (int j = 0; j < njobs; ++j) {
    const pbbs_graph* g = graphs[j];
    if (!g) return fail(PBBS_ERR_INVALID, "graph %d is NULL", j);
    if (g->device != c->device) return fail(PBBS_ERR_INVALID, "graph %d lives on another device", j);
    if (sources[j] < 0 || (int64_t)sources[j] >= g->n)
      return fail(PBBS_ERR_INVALID, "job %d: source %d outside [0, %lld)", j, sources[j],
                  (long long)g->n);
  }
  HIP_TRY(hipSetDevice(c->device));
  if (!c->m_args) {
    void* args = nullptr;
    uint32_t* wg = nullptr;
    int32_t* hs = nullptr;
    BfsState* hst = nullptr;
    if (hipMalloc(&args, kMaxMultiJobs * sizeof(BfsArgs<int64_t>)) != hipSuccess ||
        hipMalloc(&wg, (kMaxMultiJobs + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipHostMalloc((void**)&hs, kMaxMultiJobs * 16 * sizeof(int32_t), hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostMalloc((void**)&hst, kMaxMultiJobs * sizeof(BfsState), hipHostMallocDefault) !=
            hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(args);
      (void)hipFree(wg);
      if (hs) (void)hipHostFree(hs);
      if (hst) (void)hipHostFree(hst);
      return fail(PBBS_ERR_NOMEM, "multi-graph control buffers failed");
    }
    c->m_wg = wg;
    c->m_host_status = hs;
    c->m_host_state = hst;
    c->m_args = args;
  }
  // as in pbbs_bfs_batch: deterministic outputs meet the any-parent contract
  if (out_flags & ~(PBBS_OUT_DEVICE | PBBS_BFS_ANY_PARENT))
    return fail(PBBS_ERR_INVALID, "unknown flags 0x%x", out_flags);
  const bool dev_out = (out_flags & PBBS_OUT_DEVICE) != 0;
  std::vector<pbbs_graph*> gs(graphs, graphs + njobs);
  // every job's workspace, one after another in m_ws, laid out before any
  // launch: the groups below run one after another and each job's outputs
  // (when the caller gave no device pointer) stay in its own region until
  // they are copied out at the end
  std::vector<JobLayout> L(njobs);
  std::vector<int64_t> at(njobs + 1, 0);
  for (int j = 0; j < njobs; ++j) {
    L[j] = job_layout(gs[j]->n, c->tuning.queue_div, true);
    at[j + 1] = at[j] + L[j].total;
  }
  if (at[njobs] > c->m_ws_bytes) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(c->m_ws);
    c->m_ws = nullptr;
    c->m_ws_bytes = 0;
    if (hipMalloc(&c->m_ws, at[njobs]) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PBBS_ERR_NOMEM, "multi-graph workspace of %lld bytes failed",
                  (long long)at[njobs]);
    }
    c->m_ws_bytes = at[njobs];
  }
  std::vector<int32_t*> dpar(njobs, nullptr), ddep(njobs, nullptr);
  for (int j = 0; j < njobs; ++j) {
    unsigned char* base = c->m_ws + at[j];
    dpar[j] = (dev_out && parents && parents[j]) ? parents[j]
                                                 : reinterpret_cast<int32_t*>(base + L[j].parent);
    ddep[j] = (dev_out && depths && depths[j]) ? depths[j]
                                               : reinterpret_cast<int32_t*>(base + L[j].depth);
  }
  // A large job runs alone, one launch sequence on the whole device: on its
  // share of a shared grid it takes longer than alone (8 serving graphs of
  // scales 16-22: 2.03 ms in one launch against 1.34 ms one after another),
  // while small graphs gain by sharing (16 graphs of scales 10-17: 4.8x).
  std::vector<char> alone(njobs, 0);
  for (int j = 0; j < njobs; ++j) {
    if (gs[j]->m <= kMultiAloneEdges) continue;
    alone[j] = 1;
    int rc = ensure_workspace(c, gs[j]->n);
    if (rc == PBBS_OK)
      rc = gs[j]->off64 ? run_bfs<int64_t>(c, gs[j], sources[j], dpar[j], ddep[j], &stats[j], false)
                        : run_bfs<uint32_t>(c, gs[j], sources[j], dpar[j], ddep[j], &stats[j], false);
    if (rc != PBBS_OK) return rc;
  }
  // one launch per (offset width, row order) group: jobs of a launch share
  // the kernel instantiation
  for (int off64 = 0; off64 < 2; ++off64) {
    for (int sorted = 0; sorted < 2; ++sorted) {
      std::vector<int> ids;
      for (int j = 0; j < njobs; ++j)
        if (!alone[j] && (int)gs[j]->off64 == off64 && (int)gs[j]->sorted == sorted)
          ids.push_back(j);
      if (ids.empty()) continue;
      const int rc =
          off64 ? run_multi_group<int64_t>(c, ids, gs, sources, L, at, dpar, ddep, stats, sorted)
                : run_multi_group<uint32_t>(c, ids, gs, sources, L, at, dpar, ddep, stats, sorted);
      if (rc != PBBS_OK) return rc;
    }
  }
  if (!dev_out) {
    for (int j = 0; j < njobs; ++j) {
      const size_t bytes = (size_t)gs[j]->n * sizeof(int32_t);
      if (parents && parents[j])
        HIP_TRY(hipMemcpyAsync(parents[j], dpar[j], bytes, hipMemcpyDeviceToHost, c->stream));
      if (depths && depths[j])
        HIP_TRY(hipMemcpyAsync(depths[j], ddep[j], bytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return PBBS_OK;
}

// ----------------------------------------------- BFS in a caller's memory block
namespace pbbs {

// A single-source BFS whose working memory is the device block [blk, blk +
// bytes), laid out as a multi-graph job's workspace without outputs; the
// frontier queues get what the block leaves (at most queue_cap(n,
// queue_div) entries per sub-queue; a frontier that does not fit is expanded
// from its bitmap, so the outputs do not change).  Nothing is allocated: the
// graph's search data must exist (pbbs_graph_prepare), the levels run in one
// persistent launch or as plain launches (no captured graphs), and the
// outputs go to the caller's device buffers.  The init kernel clears every
// part of the block the levels read before writing it.
int bfs_in_block(pbbs_ctx* c, pbbs_graph* g, int32_t s, int32_t* d_parent, int32_t* d_depth,
                 unsigned char* blk, int64_t bytes, int64_t max_n, pbbs_bfs_stats* stats) {
  clear_error();
  if (!c || !g) return fail(PBBS_ERR_INVALID, "NULL context or graph");
  if (!d_parent || !d_depth)
    return fail(PBBS_ERR_INVALID, "parent and depth must be device buffers of n entries");
  if (g->device != c->device) return fail(PBBS_ERR_INVALID, "graph lives on another device");
  if (g->n > max_n)
    return fail(PBBS_ERR_LIMIT, "n=%lld exceeds the %lld vertices the static scratch serves",
                (long long)g->n, (long long)max_n);
  if (s < 0 || (int64_t)s >= g->n)
    return fail(PBBS_ERR_INVALID, "source %d outside [0, %lld)", s, (long long)g->n);
  if (!blk) return fail(PBBS_ERR_STATE, "no scratch block on device %d", c->device);
  const int64_t fixed = job_layout_q(g->n, 0, false).total;
  const int64_t per_entry = 2 * (int64_t)kSubQueues * (sizeof(int32_t) + sizeof(uint32_t));
  const int64_t room = (bytes - fixed - 2 * kJobAlign) / per_entry;
  const int64_t qcap = std::min(queue_cap(g->n, c->tuning.queue_div), room);
  if (qcap < kWave)
    return fail(PBBS_ERR_LIMIT, "n=%lld leaves no room for the frontier queues in %lld bytes",
                (long long)g->n, (long long)bytes);
  const JobLayout L = job_layout_q(g->n, qcap, false);
  if (L.total > bytes) return fail(PBBS_ERR_LIMIT, "workspace layout exceeds the scratch block");
  SingleWs ws;
  ws.bits = reinterpret_cast<unsigned long long*>(blk + L.bits);
  ws.q = reinterpret_cast<int32_t*>(blk + L.q);
  ws.qe = reinterpret_cast<uint32_t*>(blk + L.qe);
  ws.qcap = qcap;
  ws.st = reinterpret_cast<BfsState*>(blk + L.st);
  ws.wg = reinterpret_cast<uint32_t*>(blk + L.wg);
  ws.args = blk + L.args;
  HIP_TRY(hipSetDevice(c->device));
  int rc = g->off64 ? begin_bfs<int64_t>(c, g, s, d_parent, d_depth, ws, false)
                    : begin_bfs<uint32_t>(c, g, s, d_parent, d_depth, ws, false);
  bool done = false;
  while (rc == PBBS_OK && !done) rc = poll_run(c, true, &done);
  if (rc != PBBS_OK) {
    release_persist(c);
    return rc;
  }
  return stats ? single_stats(c, stats) : PBBS_OK;
}

}  // namespace pbbs

// ------------------------------------------------------------ graph download
extern "C" int pbbs_graph_download(const pbbs_graph* g, pbbs_csr* out) {
  clear_error();
  if (!g || !out) return fail(PBBS_ERR_INVALID, "NULL argument");
  out->n = out->m = 0;
  out->offsets = nullptr;
  out->adj = nullptr;
  HIP_TRY(hipSetDevice(g->device));
  if (g->ctx) HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  int64_t* off = static_cast<int64_t*>(std::malloc((g->n + 1) * sizeof(int64_t)));
  int32_t* adj = static_cast<int32_t*>(std::malloc(std::max<int64_t>(g->m, 1) * sizeof(int32_t)));
  if (!off || !adj) {
    std::free(off);
    std::free(adj);
    return fail(PBBS_ERR_NOMEM, "host CSR allocation failed");
  }
  hipError_t e = hipSuccess;
  if (g->off64) {
    e = hipMemcpy(off, g->d_off, (g->n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost);
  } else {
    std::vector<uint32_t> o32(g->n + 1);
    e = hipMemcpy(o32.data(), g->d_off, (g->n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost);
    for (int64_t i = 0; i <= g->n; ++i) off[i] = o32[i];
  }
  if (e == hipSuccess && g->m > 0)
    e = hipMemcpy(adj, g->d_adj, g->m * sizeof(int32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    std::free(off);
    std::free(adj);
    return fail(PBBS_ERR_HIP, "graph download: %s", hipGetErrorString(e));
  }
  out->n = g->n;
  out->m = g->m;
  out->offsets = off;
  out->adj = adj;
  return PBBS_OK;
}

extern "C" int pbbs_ctx_outputs(const pbbs_ctx* c, int batched, const int32_t** parent,
                                const int32_t** depth) {
  clear_error();
  if (!c || !parent || !depth) return fail(PBBS_ERR_INVALID, "NULL argument");
  *parent = batched ? c->b_parent : c->w_parent;
  *depth = batched ? c->b_depth : c->w_depth;
  if (!*parent) return fail(PBBS_ERR_STATE, "no %s output buffers yet", batched ? "batched" : "BFS");
  return PBBS_OK;
}
