// Allocation-free single-source BFS: pbbs_bfs_noalloc runs the same kernels
// as pbbs_bfs, with their working memory (bitmaps, frontier queues, control
// block, per-workgroup level counters, argument block) in a static device
// array of this file instead of buffers the context allocates.  For callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid allocation).
//
// Each GPU the library runs on holds its own copy of the array (kScratchBytes)
// until the process ends.  In host code `scratch` names a host-side stand-in,
// not device memory, so the host never passes it to a kernel: pbbs_ctx_create
// calls scratch_init with the device current, which looks up that device's
// address of the array once (hipGetSymbolAddress) and keeps it per device;
// the kernels receive that address.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's
// contents: calls that overlap on one GPU (two contexts, two threads) race
// and must be serialised by the caller.  The init kernel clears every part
// of the array a BFS reads before it writes it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/pbbs_bfs.h"
#include "engine.hpp"

namespace pbbs {
int bfs_in_block(pbbs_ctx* c, pbbs_graph* g, int32_t s, int32_t* d_parent, int32_t* d_depth,
                 unsigned char* blk, int64_t bytes, int64_t max_n, pbbs_bfs_stats* stats);
int ctx_device(const pbbs_ctx* c);
}

namespace {

constexpr int64_t kScratchBytes = int64_t(32) << 20;   // 32 MiB per GPU
constexpr int64_t kScratchMaxN = int64_t(1) << 25;     // vertices served
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kScratchDevices] = {};

unsigned char* scratch_of(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace

namespace pbbs {

int scratch_init(int device) {
  if (device < 0 || device >= kScratchDevices)
    return fail(PBBS_ERR_LIMIT, "device %d beyond the %d the scratch table serves", device,
                kScratchDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return PBBS_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(PBBS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  return PBBS_OK;
}

}  // namespace pbbs

extern "C" int pbbs_bfs_noalloc(pbbs_ctx* ctx, pbbs_graph* g, int32_t source, int32_t* parent,
                                int32_t* depth, pbbs_bfs_stats* stats) {
  pbbs::clear_error();
  if (!ctx || !g) return pbbs::fail(PBBS_ERR_INVALID, "NULL context or graph");
  // the address scratch_init found on the context's device
  unsigned char* blk = scratch_of(pbbs::ctx_device(ctx));
  return pbbs::bfs_in_block(ctx, g, source, parent, depth, blk, kScratchBytes, kScratchMaxN,
                            stats);
}
