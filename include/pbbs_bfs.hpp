// pbbs_bfs.hpp -- C++ face of the C ABI in pbbs_bfs.h, shaped like the
// reference's benchmark interface (benchmarks/breadthFirstSearch/bench/BFS.h:4-11):
//
//     using vertexId = int;  using edgeId = uint;
//     sequence<vertexId> BFS(vertexId start, const Graph& G, bool verbose);
//
// Here Graph is a CSR resident in HBM (built from a host CSR once, as the
// reference builds graph<int,uint> once in bench/BFSTime.C:58-59), and BFS()
// returns the parent sequence in host memory.  Errors become exceptions on
// this side of the boundary only.
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pbbs_bfs.h"

namespace pbbs_amd {

using vertexId = int;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char* what) {
  if (rc != PBBS_OK) throw Error(rc, std::string(what) + ": " + pbbs_last_error());
}

// Owning host CSR (int64 offsets, int32 ids).
class HostGraph {
 public:
  HostGraph() = default;
  explicit HostGraph(const pbbs_csr& g) : g_(g) {}
  HostGraph(const HostGraph&) = delete;
  HostGraph& operator=(const HostGraph&) = delete;
  HostGraph(HostGraph&& o) noexcept : g_(o.g_) { o.g_ = pbbs_csr{}; }
  HostGraph& operator=(HostGraph&& o) noexcept {
    std::swap(g_, o.g_);
    return *this;
  }
  ~HostGraph() { pbbs_csr_free(&g_); }

  static HostGraph read(const char* path) {  // common/graphIO.h:191-215
    pbbs_csr g{};
    check(pbbs_read_adjacency_graph(path, &g), "readGraphFromFile");
    return HostGraph(g);
  }
  size_t numVertices() const { return (size_t)g_.n; }
  size_t numEdges() const { return (size_t)g_.m; }
  const pbbs_csr& raw() const { return g_; }

 private:
  pbbs_csr g_{};
};

class Context {
 public:
  explicit Context(int device = 0) { check(pbbs_ctx_create(device, &c_), "pbbs_ctx_create"); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  ~Context() { pbbs_ctx_destroy(c_); }
  pbbs_ctx* get() const { return c_; }

 private:
  pbbs_ctx* c_ = nullptr;
};

// Graph resident in HBM.  `wide_offsets` keeps int64 offsets on the device;
// otherwise offsets are narrowed to uint32 (the reference's edgeId).
class Graph {
 public:
  Graph(Context& ctx, const HostGraph& h, bool wide_offsets = false) : n_(h.numVertices()) {
    const pbbs_csr& g = h.raw();
    if (wide_offsets) {
      check(pbbs_graph_upload(ctx.get(), g.n, g.m, g.offsets, g.adj, PBBS_GRAPH_OFFSETS_I64, &g_),
            "pbbs_graph_upload");
    } else {
      std::vector<uint32_t> off32(g.n + 1);
      for (int64_t i = 0; i <= g.n; ++i) off32[i] = (uint32_t)g.offsets[i];
      check(pbbs_graph_upload(ctx.get(), g.n, g.m, off32.data(), g.adj, 0, &g_),
            "pbbs_graph_upload");
    }
    ctx_ = ctx.get();
  }
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;
  ~Graph() { pbbs_graph_free(g_); }
  size_t numVertices() const { return n_; }
  pbbs_graph* get() const { return g_; }
  pbbs_ctx* ctx() const { return ctx_; }

 private:
  pbbs_graph* g_ = nullptr;
  pbbs_ctx* ctx_ = nullptr;
  size_t n_ = 0;
};

// Parent sequence: -1 = unvisited, the start points to itself
// (bench/BFS.h:9-11).  `depth`, if given, receives the level_number depths.
// `flags` may hold PBBS_BFS_ANY_PARENT (backForwardBFS's contract,
// backForwardBFS/BFS.C:38-61).
inline std::vector<vertexId> BFS(vertexId start, const Graph& G, bool verbose = false,
                                 std::vector<vertexId>* depth = nullptr,
                                 pbbs_bfs_stats* stats = nullptr, uint32_t flags = 0) {
  (void)verbose;
  std::vector<vertexId> parent(G.numVertices());
  if (depth) depth->resize(G.numVertices());
  check(pbbs_bfs(G.ctx(), G.get(), start, parent.data(), depth ? depth->data() : nullptr, flags,
                 stats),
        "BFS");
  return parent;
}

// Maximal independent set of G, as maximalIndependentSet(G) returns it
// (maximalIndependentSet/bench/MIS.h, incrementalMIS/MIS.C:62-66): one char
// per vertex, 1 in the set, 2 not.
inline std::vector<char> maximalIndependentSet(const Graph& G, pbbs_mis_stats* stats = nullptr) {
  std::vector<int32_t> f(G.numVertices());
  check(pbbs_mis(G.ctx(), G.get(), f.data(), 0, stats), "maximalIndependentSet");
  return std::vector<char>(f.begin(), f.end());
}

// Edge array (u[i], v[i]) on n vertices, as edgeArray<vertexId> holds it
// (common/graph.h), read with readEdgeArrayFromFile (common/graphIO.h:147-166).
struct EdgeArray {
  int64_t n = 0;
  std::vector<int32_t> u, v;
  static EdgeArray read(const char* path) {
    int64_t n = 0, m = 0;
    int32_t *pu = nullptr, *pv = nullptr;
    check(pbbs_read_edge_array(path, &n, &m, &pu, &pv), "readEdgeArrayFromFile");
    EdgeArray e;
    e.n = n;
    e.u.assign(pu, pu + m);
    e.v.assign(pv, pv + m);
    pbbs_free(pu);
    pbbs_free(pv);
    return e;
  }
};

// Spanning forest of E, as st(E) returns it (spanningForest/bench/ST.h:7,
// incrementalST/ST.C:62-71): the forest's edge ids.  `flags` may hold
// PBBS_ST_ANY_FOREST (ndST's contract).
inline std::vector<uint32_t> st(Context& ctx, const EdgeArray& E, pbbs_st_stats* stats = nullptr,
                                 uint32_t flags = 0) {
  std::vector<int32_t> ids((size_t)std::max<int64_t>(E.n - 1, 1));
  int64_t count = 0;
  check(pbbs_spanning_forest(ctx.get(), E.n, (int64_t)E.u.size(), E.u.data(), E.v.data(), flags,
                             ids.data(), &count, stats),
        "st");
  return std::vector<uint32_t>(ids.begin(), ids.begin() + count);
}

}  // namespace pbbs_amd
