// BFS serving loop: graphs stay resident in HBM; queries arrive on any host
// thread; a worker thread keeps several runs in flight, one per execution
// slot (a context of its own: stream, workspace, captured level graphs).
// Queries are admitted as slots and the HBM budget free up.  When a slot
// frees and several queries for one graph wait, up to max_batch of them are
// merged into one batched traversal.  A run's level loop is advanced by
// polling its chunk event, so runs on different slots overlap on the device;
// the level kernels never wait on other workgroups, so any mix of concurrent
// runs completes.  Results are copied into the caller's buffers through
// pinned staging, and each carries its arrival-to-ready latency.
//
// HBM accounting: the budget covers the resident graphs with their search
// data (built when a graph is added) plus every slot's workspace buffers as
// they stand.  A slot keeps its buffers between runs (they only grow), so a
// run is charged what the slot will hold after it; when that does not fit,
// the slot's buffers, and then idle slots' buffers, are freed first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/pbbs_bfs.h"
#include "engine.hpp"

using pbbs::fail;

// Internal engine entry points shared with the server (engine.cpp).
namespace pbbs {
int64_t bfs_workspace_bytes(int64_t n, int32_t nsrc, int32_t queue_div);
int64_t ctx_workspace_bytes(const pbbs_ctx* c);
int64_t ctx_workspace_bytes_after(const pbbs_ctx* c, int64_t n, int32_t nsrc);
void ctx_trim(pbbs_ctx* c);
int32_t ctx_queue_div(const pbbs_ctx* c);
void ctx_no_prefill(pbbs_ctx* c);
int async_begin(pbbs_ctx* c, pbbs_graph* g, const int32_t* sources, int32_t nsrc);
int async_poll(pbbs_ctx* c, bool* done);
int async_outputs(pbbs_ctx* c, const int32_t** parent, const int32_t** depth);
int async_stats(pbbs_ctx* c, pbbs_bfs_stats* stats);
hipStream_t ctx_stream(pbbs_ctx* c);
int64_t graph_n(const pbbs_graph* g);
bool graph_symmetric(const pbbs_graph* g);
}  // namespace pbbs

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
  uint64_t ticket;
  int32_t graph;
  int32_t source;
  int32_t* parent;
  int32_t* depth;
  Clock::time_point arrival;
  Clock::time_point admitted;
};

struct Slot {
  pbbs_ctx* ctx = nullptr;
  int phase = 0;  // 0 idle, 1 level loop running, 2 outputs being copied
  std::vector<Request> reqs;
  int32_t graph = -1;
  pbbs_graph* g = nullptr;  // captured under the lock: add_graph may grow the table
  int64_t n = 0;
  int64_t retained = 0;  // HBM bytes the slot's context holds (charged to the budget)
  int32_t* pin_parent = nullptr;
  int32_t* pin_depth = nullptr;
  int64_t pin_cap = 0;
  hipEvent_t copied = nullptr;
  double device_ms = 0;
};

}  // namespace

struct pbbs_server {
  int device = 0;
  pbbs_ctx* main = nullptr;
  std::vector<pbbs_graph*> graphs;
  std::vector<Slot> slots;
  int32_t max_batch = 64;
  int64_t budget = 0;
  int64_t graph_bytes = 0;   // resident graphs with their search data
  int64_t max_run = 0;       // the largest single-source workspace of a resident graph
  int64_t in_use = 0;        // graph_bytes + every slot's retained bytes
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request> pending;
  std::vector<pbbs_server_result> completed;
  int64_t running = 0;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;
};

namespace {

void finish(pbbs_server* s, Slot& sl, int status) {
  const auto now = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  for (size_t i = 0; i < sl.reqs.size(); ++i) {
    const Request& r = sl.reqs[i];
    pbbs_server_result res{};
    res.ticket = r.ticket;
    res.graph = r.graph;
    res.source = r.source;
    res.status = status;
    res.batch = (int32_t)sl.reqs.size();
    res.latency_ms = ms_between(r.arrival, now);
    res.queue_ms = ms_between(r.arrival, r.admitted);
    res.device_ms = sl.device_ms;
    s->completed.push_back(res);
  }
  s->running -= (int64_t)sl.reqs.size();
  sl.reqs.clear();
  sl.phase = 0;
  s->cv_done.notify_all();
}

// Frees a slot's device workspace and pinned staging; the caller has taken
// its bytes off in_use.  The slot is idle (worker thread only).
void trim_slot(Slot& sl) {
  pbbs::ctx_trim(sl.ctx);
  if (sl.pin_parent) (void)hipHostFree(sl.pin_parent);
  if (sl.pin_depth) (void)hipHostFree(sl.pin_depth);
  sl.pin_parent = sl.pin_depth = nullptr;
  sl.pin_cap = 0;
}

int ensure_pinned(Slot& sl, int64_t entries) {
  if (entries <= sl.pin_cap) return PBBS_OK;
  if (sl.pin_parent) (void)hipHostFree(sl.pin_parent);
  if (sl.pin_depth) (void)hipHostFree(sl.pin_depth);
  sl.pin_parent = sl.pin_depth = nullptr;
  sl.pin_cap = 0;
  if (hipHostMalloc((void**)&sl.pin_parent, entries * sizeof(int32_t), hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void**)&sl.pin_depth, entries * sizeof(int32_t), hipHostMallocDefault) !=
          hipSuccess)
    return fail(PBBS_ERR_NOMEM, "pinned staging of %lld entries failed", (long long)entries);
  sl.pin_cap = entries;
  return PBBS_OK;
}

// Takes the oldest waiting query and, for a symmetric graph, up to
// max_batch - 1 more for the same graph, if the HBM budget allows the run:
// as the slot's buffers stand, else with them freed first, else with idle
// slots' buffers freed too; the batch shrinks until one of these fits.  A
// lone query always fits an idle server (add_graph checks that).
bool admit(pbbs_server* s, Slot& sl) {
  std::vector<Request> take;
  std::vector<Slot*> trim;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->pending.empty()) return false;
    const int32_t gid = s->pending.front().graph;
    pbbs_graph* g = s->graphs[gid];
    const int64_t n = pbbs::graph_n(g);
    const int cap = pbbs::graph_symmetric(g) ? s->max_batch : 1;
    const int32_t qdiv = pbbs::ctx_queue_div(sl.ctx);
    int want = 0;
    for (const Request& r : s->pending)
      if (r.graph == gid && want < cap) ++want;
    int64_t idle = 0;   // bytes other idle slots hold
    for (Slot& o : s->slots)
      if (&o != &sl && o.phase == 0) idle += o.retained;
    int how = -1;       // 0 keep the buffers, 1 free this slot's, 2 also idle slots'
    int64_t after = 0;
    for (; want >= 1 && how < 0; --want) {
      const int64_t keep = pbbs::ctx_workspace_bytes_after(sl.ctx, n, want);
      const int64_t fresh = pbbs::bfs_workspace_bytes(n, want, qdiv);
      if (s->in_use - sl.retained + keep <= s->budget) {
        how = 0;
        after = keep;
      } else if (s->in_use - sl.retained + fresh <= s->budget) {
        how = 1;
        after = fresh;
      } else if (s->in_use - sl.retained - idle + fresh <= s->budget) {
        how = 2;
        after = fresh;
      }
      if (how >= 0) break;
    }
    if (how < 0) return false;   // wait for running slots to finish
    for (auto it = s->pending.begin(); it != s->pending.end() && (int)take.size() < want;) {
      if (it->graph == gid) {
        take.push_back(*it);
        it = s->pending.erase(it);
      } else {
        ++it;
      }
    }
    if (how >= 1) trim.push_back(&sl);
    if (how == 2)
      for (Slot& o : s->slots)
        if (&o != &sl && o.phase == 0 && o.retained > 0) trim.push_back(&o);
    for (Slot* t : trim) {
      s->in_use -= t->retained;
      t->retained = 0;
    }
    s->in_use += after - sl.retained;
    sl.retained = after;
    s->running += (int64_t)take.size();
    sl.graph = gid;
    sl.g = g;
    sl.n = n;
  }
  for (Slot* t : trim) trim_slot(*t);
  const auto now = Clock::now();
  for (Request& r : take) r.admitted = now;
  sl.reqs = std::move(take);
  std::vector<int32_t> src(sl.reqs.size());
  for (size_t i = 0; i < src.size(); ++i) src[i] = sl.reqs[i].source;
  int rc = ensure_pinned(sl, (int64_t)src.size() * sl.n);
  if (rc == PBBS_OK)
    rc = pbbs::async_begin(sl.ctx, sl.g, src.data(), (int32_t)src.size());
  {
    // charge what the context holds now (equal to the estimate unless an
    // allocation failed)
    std::lock_guard<std::mutex> lk(s->mu);
    const int64_t held = pbbs::ctx_workspace_bytes(sl.ctx);
    s->in_use += held - sl.retained;
    sl.retained = held;
  }
  if (rc != PBBS_OK) {
    finish(s, sl, rc);
    return true;
  }
  sl.phase = 1;
  return true;
}

// Frees idle slots' buffers while the server holds more than its budget
// (a graph added while slots kept buffers from earlier runs).
void trim_over_budget(pbbs_server* s) {
  std::vector<Slot*> trim;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (Slot& o : s->slots) {
      if (s->in_use <= s->budget) break;
      if (o.phase == 0 && o.retained > 0) {
        s->in_use -= o.retained;
        o.retained = 0;
        trim.push_back(&o);
      }
    }
  }
  for (Slot* t : trim) trim_slot(*t);
}

// Level loop finished: stats, then copy outputs to pinned staging.
int start_copy(Slot& sl) {
  pbbs_bfs_stats st;
  int rc = pbbs::async_stats(sl.ctx, &st);
  if (rc != PBBS_OK) return rc;
  sl.device_ms = st.device_ms;
  const int32_t *dp = nullptr, *dd = nullptr;
  rc = pbbs::async_outputs(sl.ctx, &dp, &dd);
  if (rc != PBBS_OK) return rc;
  const int64_t bytes = (int64_t)sl.reqs.size() * sl.n * sizeof(int32_t);
  bool want_parent = false, want_depth = false;
  for (const Request& r : sl.reqs) {
    want_parent |= r.parent != nullptr;
    want_depth |= r.depth != nullptr;
  }
  hipStream_t st_ = pbbs::ctx_stream(sl.ctx);
  if ((want_parent &&
       hipMemcpyAsync(sl.pin_parent, dp, bytes, hipMemcpyDeviceToHost, st_) != hipSuccess) ||
      (want_depth &&
       hipMemcpyAsync(sl.pin_depth, dd, bytes, hipMemcpyDeviceToHost, st_) != hipSuccess) ||
      hipEventRecord(sl.copied, st_) != hipSuccess)
    return fail(PBBS_ERR_HIP, "output copy failed");
  sl.phase = 2;
  return PBBS_OK;
}

// Copies the staged rows into the callers' buffers: pieces of 1M entries,
// spread over host threads (a batch of large graphs is GBs, and first touches
// of fresh caller pages fault one by one).
void deliver(Slot& sl) {
  constexpr int64_t kPiece = 1 << 20;
  const int64_t pieces = (sl.n + kPiece - 1) / kPiece;
  const int64_t tasks = (int64_t)sl.reqs.size() * pieces * 2;
#pragma omp parallel for schedule(dynamic, 1) num_threads(8) if (tasks > 4)
  for (int64_t t = 0; t < tasks; ++t) {
    const int64_t q = t / (pieces * 2), rest = t % (pieces * 2);
    const bool dep = rest >= pieces;
    const int64_t lo = (dep ? rest - pieces : rest) * kPiece;
    const int64_t len = std::min(kPiece, sl.n - lo);
    const Request& r = sl.reqs[q];
    int32_t* dst = dep ? r.depth : r.parent;
    if (!dst) continue;
    const int32_t* src = (dep ? sl.pin_depth : sl.pin_parent) + q * sl.n + lo;
    std::memcpy(dst + lo, src, len * sizeof(int32_t));
  }
}

void worker_loop(pbbs_server* s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    bool active = false;
    for (Slot& sl : s->slots) {
      if (sl.phase == 1) {
        bool done = false;
        int rc = pbbs::async_poll(sl.ctx, &done);
        if (rc == PBBS_OK && done) rc = start_copy(sl);
        if (rc != PBBS_OK) finish(s, sl, rc);
      } else if (sl.phase == 2) {
        const hipError_t e = hipEventQuery(sl.copied);
        if (e == hipSuccess) {
          deliver(sl);
          finish(s, sl, PBBS_OK);
        } else if (e != hipErrorNotReady) {
          finish(s, sl, PBBS_ERR_HIP);
        }
      }
    }
    trim_over_budget(s);
    for (Slot& sl : s->slots)
      if (sl.phase == 0) admit(s, sl);
    for (Slot& sl : s->slots) active |= sl.phase != 0;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->stop && !active && s->pending.empty()) break;
    if (!active && s->pending.empty()) {
      s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
    } else {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
}

}  // namespace

extern "C" {

int pbbs_server_create(int device, const pbbs_server_config* cfg, pbbs_server** out) {
  pbbs::clear_error();
  if (!out) return fail(PBBS_ERR_INVALID, "out is NULL");
  *out = nullptr;
  pbbs_server_config c{};
  if (cfg) c = *cfg;
  if (c.slots <= 0) c.slots = 4;
  if (c.max_batch <= 0) c.max_batch = PBBS_MAX_BATCH;
  if (c.slots > 64 || c.max_batch > PBBS_MAX_BATCH)
    return fail(PBBS_ERR_INVALID, "slots <= 64 and max_batch <= %d", PBBS_MAX_BATCH);
  pbbs_server* s = new pbbs_server();
  s->device = device;
  s->max_batch = c.max_batch;
  int rc = pbbs_ctx_create(device, &s->main);
  if (rc != PBBS_OK) {
    delete s;
    return rc;
  }
  if (c.hbm_budget > 0) {
    s->budget = c.hbm_budget;
  } else {
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    s->budget = (int64_t)(free_b * 0.8);
  }
  s->slots.resize(c.slots);
  for (Slot& sl : s->slots) {
    rc = pbbs_ctx_create(device, &sl.ctx);
    if (rc == PBBS_OK) {
      // slots' runs overlap on the device, so a persistent launch could not
      // count on its grid being resident at once: one level per launch
      pbbs_tuning t;
      rc = pbbs_tuning_get(sl.ctx, &t);
      t.persist = 0;
      if (rc == PBBS_OK) rc = pbbs_tuning_set(sl.ctx, &t);
      // nor does a batched run fill the next run's parents beside itself
      pbbs::ctx_no_prefill(sl.ctx);
    }
    if (rc == PBBS_OK && hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming) != hipSuccess)
      rc = fail(PBBS_ERR_HIP, "event creation failed");
    if (rc != PBBS_OK) {
      pbbs_server_destroy(s);
      return rc;
    }
  }
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return PBBS_OK;
}

int pbbs_server_destroy(pbbs_server* s) {
  if (!s) return PBBS_OK;
  if (s->worker.joinable()) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    s->worker.join();
  }
  for (Slot& sl : s->slots) {
    if (sl.copied) (void)hipEventDestroy(sl.copied);
    if (sl.pin_parent) (void)hipHostFree(sl.pin_parent);
    if (sl.pin_depth) (void)hipHostFree(sl.pin_depth);
    pbbs_ctx_destroy(sl.ctx);
  }
  for (pbbs_graph* g : s->graphs) pbbs_graph_free(g);
  pbbs_ctx_destroy(s->main);
  delete s;
  return PBBS_OK;
}

int pbbs_server_add_graph(pbbs_server* s, int64_t n, int64_t m, const void* offsets,
                          const int32_t* adj, uint32_t flags, int32_t* graph_id) {
  pbbs::clear_error();
  if (!s || !graph_id) return fail(PBBS_ERR_INVALID, "NULL argument");
  pbbs_graph* g = nullptr;
  int rc = pbbs_graph_upload(s->main, n, m, offsets, adj, flags, &g);
  if (rc != PBBS_OK) return rc;
  // search data now, so runs allocate nothing per graph and the bytes are known
  pbbs_graph_info info{};
  rc = pbbs_graph_info_get(g, &info);
  if (rc == PBBS_OK)
    rc = pbbs_graph_prepare(s->main, g,
                            PBBS_PREPARE_SINGLE | (info.symmetric ? PBBS_PREPARE_BATCH : 0u),
                            nullptr);
  if (rc == PBBS_OK) rc = pbbs_graph_info_get(g, &info);
  if (rc != PBBS_OK) {
    pbbs_graph_free(g);
    return rc;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  const int64_t run = pbbs::bfs_workspace_bytes(n, 1, pbbs::ctx_queue_div(s->main));
  if (s->graph_bytes + info.device_bytes + std::max(s->max_run, run) > s->budget) {
    pbbs_graph_free(g);
    return fail(PBBS_ERR_LIMIT,
                "graph of %lld HBM bytes plus a run's workspace does not fit the server budget "
                "(%lld of %lld bytes held by graphs)",
                (long long)info.device_bytes, (long long)s->graph_bytes, (long long)s->budget);
  }
  s->graph_bytes += info.device_bytes;
  s->in_use += info.device_bytes;
  s->max_run = std::max(s->max_run, run);
  s->graphs.push_back(g);
  *graph_id = (int32_t)s->graphs.size() - 1;
  return PBBS_OK;
}

int pbbs_server_submit(pbbs_server* s, int32_t graph_id, int32_t source, int32_t* parent,
                       int32_t* depth, uint64_t* ticket) {
  pbbs::clear_error();
  if (!s) return fail(PBBS_ERR_INVALID, "NULL server");
  Request r{};
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (graph_id < 0 || graph_id >= (int32_t)s->graphs.size())
      return fail(PBBS_ERR_INVALID, "unknown graph %d", graph_id);
    if (source < 0 || (int64_t)source >= pbbs::graph_n(s->graphs[graph_id]))
      return fail(PBBS_ERR_INVALID, "source %d outside graph %d", source, graph_id);
    r.ticket = s->next_ticket++;
    r.graph = graph_id;
    r.source = source;
    r.parent = parent;
    r.depth = depth;
    r.arrival = Clock::now();
    s->pending.push_back(r);
  }
  s->cv_work.notify_one();
  if (ticket) *ticket = r.ticket;
  return PBBS_OK;
}

int pbbs_server_poll(pbbs_server* s, pbbs_server_result* out, int32_t max, int32_t* count,
                     int32_t wait_ms) {
  pbbs::clear_error();
  if (!s || !count || (max > 0 && !out)) return fail(PBBS_ERR_INVALID, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->completed.empty() && wait_ms != 0) {
    auto pred = [&] { return !s->completed.empty(); };
    if (wait_ms < 0)
      s->cv_done.wait(lk, pred);
    else
      s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
  }
  const int32_t k = (int32_t)std::min<size_t>((size_t)std::max(max, 0), s->completed.size());
  for (int32_t i = 0; i < k; ++i) out[i] = s->completed[i];
  s->completed.erase(s->completed.begin(), s->completed.begin() + k);
  *count = k;
  return PBBS_OK;
}

int pbbs_server_load(pbbs_server* s, int64_t* queued, int64_t* running, int64_t* hbm_in_use) {
  pbbs::clear_error();
  if (!s) return fail(PBBS_ERR_INVALID, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = (int64_t)s->pending.size();
  if (running) *running = s->running;
  if (hbm_in_use) *hbm_in_use = s->in_use;
  return PBBS_OK;
}

}  // extern "C"
