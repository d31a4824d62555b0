// Maximal-independent-set driver with the command line of the reference's
// timing driver (benchmarks/maximalIndependentSet/bench/MISTime.C:48-55):
//
//     MIS [-o <outFile>] [-r <rounds>] <inFile>
//
// It reads a PBBS AdjacencyGraph, uploads it to HBM once, runs the timing
// protocol of common/time_loop.h:4-18 (>= 1 s of untimed warm-up runs, then
// `rounds` timed runs, each printed as "Parlay time: <seconds>", the line
// common/runTests.py:36 parses) and writes the flags as sequenceInt (1 in the
// set, 2 not: MISTime.C:38-43).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/pbbs_bfs.hpp"

namespace {

const char* value_of(int argc, char** argv, const char* o) {
  for (int i = 1; i < argc - 1; ++i)
    if (!std::strcmp(argv[i], o)) return argv[i + 1];
  return nullptr;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
  bool help = argc < 2;
  for (int i = 1; i < argc; ++i) help |= !std::strcmp(argv[i], "-h") || !std::strcmp(argv[i], "-help");
  if (help) {
    std::printf("usage: %s [-o <outFile>] [-r <rounds>] <inFile>\n", argv[0]);
    return 0;
  }
  const char* in = argv[argc - 1];
  const char* out = value_of(argc, argv, "-o");
  const char* r = value_of(argc, argv, "-r");
  const int rounds = r ? std::atoi(r) : 1;
  if (rounds < 0) {
    std::printf("usage: %s [-o <outFile>] [-r <rounds>] <inFile>\n", argv[0]);
    return 0;
  }
  std::setvbuf(stdout, nullptr, _IOLBF, 0);
  try {
    pbbs_amd::HostGraph hg = pbbs_amd::HostGraph::read(in);
    pbbs_amd::Context ctx(0);
    pbbs_amd::Graph G(ctx, hg);
    std::vector<int32_t> flags(hg.numVertices());
    pbbs_mis_stats st{};
    const auto t0 = std::chrono::steady_clock::now();
    while (seconds_since(t0) < 1.0)
      pbbs_amd::check(pbbs_mis(ctx.get(), G.get(), flags.data(), 0, &st), "maximalIndependentSet");
    for (int k = 0; k < rounds; ++k) {
      const auto t = std::chrono::steady_clock::now();
      pbbs_amd::check(pbbs_mis(ctx.get(), G.get(), flags.data(), 0, &st), "maximalIndependentSet");
      std::printf("Parlay time: %.4f\n", seconds_since(t));
    }
    std::printf("\n");
    if (out) pbbs_amd::check(pbbs_write_sequence_int(out, flags.data(), (int64_t)flags.size()),
                             "writeIntSeqToFile");
  } catch (const pbbs_amd::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
