// BFS driver with the command line of the reference's timing driver
// (benchmarks/breadthFirstSearch/bench/BFSTime.C:51-61):
//
//     BFS [-o <outFile>] [-src source] [-r <rounds>] [-v] [-d <depthFile>] [-w] [-any] <inFile>
//
// It reads a PBBS AdjacencyGraph, uploads it to HBM once, runs the timing
// protocol of common/time_loop.h:4-18 (>= 1 s of untimed warm-up runs, then
// `rounds` timed runs, each printed as "Parlay time: <seconds>", the line
// common/runTests.py:36 parses) and writes the parents as sequenceInt.
// `-w` keeps int64 offsets in HBM; `-d` also writes the depth sequence;
// `-any` runs the backForwardBFS contract (backForwardBFS/BFS.C:38-61: any
// neighbour one level up as the parent), which the reference gets by linking
// that implementation into the same driver.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pbbs_bfs.hpp"

namespace {

struct Args {
  int argc;
  char** argv;
  bool has(const char* o) const {
    for (int i = 1; i < argc; ++i)
      if (!std::strcmp(argv[i], o)) return true;
    return false;
  }
  const char* value(const char* o) const {
    for (int i = 1; i < argc - 1; ++i)
      if (!std::strcmp(argv[i], o)) return argv[i + 1];
    return nullptr;
  }
  [[noreturn]] void usage() const {
    std::printf("usage: %s [-o <outFile>] [-src source] [-r <rounds>] [-v] [-d <depthFile>] [-w] "
                "[-any] <inFile>\n",
                argv[0]);
    std::exit(0);
  }
  long nonneg(const char* o, long dflt) const {
    const char* v = value(o);
    if (!v) return dflt;
    const long r = std::atol(v);
    if (r < 0) usage();  // negative values are rejected (common/parse_command_line.h:94-102)
    return r;
  }
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
  Args P{argc, argv};
  if (argc < 2 || P.has("-h") || P.has("-help")) P.usage();
  const char* in = argv[argc - 1];
  const char* out = P.value("-o");
  const char* dout = P.value("-d");
  const int rounds = (int)P.nonneg("-r", 1);
  const long source = P.nonneg("-src", 0);
  const bool verbose = P.has("-v");
  const uint32_t flags = P.has("-any") ? PBBS_BFS_ANY_PARENT : 0u;
  try {
    pbbs_amd::HostGraph hg = pbbs_amd::HostGraph::read(in);
    if (source >= (long)hg.numVertices()) {
      std::fprintf(stderr, "source %ld outside [0, %zu)\n", source, hg.numVertices());
      return 1;
    }
    pbbs_amd::Context ctx(0);
    pbbs_amd::Graph G(ctx, hg, P.has("-w"));
    std::vector<int> parents, depth;
    pbbs_bfs_stats st{};
    // warm-up for at least one second, untimed (common/time_loop.h:8-10)
    const auto t0 = std::chrono::steady_clock::now();
    while (seconds_since(t0) < 1.0)
      parents = pbbs_amd::BFS((int)source, G, verbose, nullptr, nullptr, flags);
    for (int r = 0; r < rounds; ++r) {
      parents.clear();
      const auto t = std::chrono::steady_clock::now();
      parents = pbbs_amd::BFS((int)source, G, verbose, dout ? &depth : nullptr, &st, flags);
      std::printf("Parlay time: %.4f\n", seconds_since(t));
    }
    std::printf("\n");
    if (verbose) {
      size_t visited = 0;
      for (int p : parents) visited += (p != -1);
      std::printf("total visited = %zu\n", visited);
      std::printf("levels = %d (top-down %d, bottom-up %d), device ms = %.3f, m_r = %lld\n",
                  st.levels, st.td_levels, st.bu_levels, st.device_ms, (long long)st.edges_reached);
    }
    if (out) pbbs_amd::check(pbbs_write_sequence_int(out, parents.data(), (int64_t)parents.size()),
                             "writeSequenceToFile");
    if (dout) pbbs_amd::check(pbbs_write_sequence_int(dout, depth.data(), (int64_t)depth.size()),
                              "writeSequenceToFile");
  } catch (const pbbs_amd::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
