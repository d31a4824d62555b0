// Spanning-forest driver with the command line of the reference's timing
// driver (benchmarks/spanningForest/bench/STTime.C:45-52):
//
//     ST [-o <outFile>] [-r <rounds>] <inFile>
//
// It reads a PBBS EdgeArray file, copies the edges to HBM once, runs the
// timing protocol of common/time_loop.h:4-18 (>= 1 s of untimed warm-up runs,
// then `rounds` timed runs printed as "Parlay time: <seconds>") and writes
// the forest's edge ids as sequenceInt (STTime.C:41).
#include <hip/hip_runtime.h>

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

// Device copies of the edges, freed on every exit path.
struct DeviceEdges {
  int32_t* u = nullptr;
  int32_t* v = nullptr;
  int32_t* forest = nullptr;
  ~DeviceEdges() {
    (void)hipFree(u);
    (void)hipFree(v);
    (void)hipFree(forest);
  }
};

}  // namespace

int main(int argc, char** argv) {
  bool help = argc < 2;
  for (int i = 1; i < argc; ++i) help |= !std::strcmp(argv[i], "-h") || !std::strcmp(argv[i], "-help");
  const char* r = value_of(argc, argv, "-r");
  const int rounds = r ? std::atoi(r) : 1;
  if (help || rounds < 0) {
    std::printf("usage: %s [-o <outFile>] [-r <rounds>] <inFile>\n", argv[0]);
    return 0;
  }
  const char* in = argv[argc - 1];
  const char* out = value_of(argc, argv, "-o");
  std::setvbuf(stdout, nullptr, _IOLBF, 0);
  int64_t n = 0, m = 0;
  int32_t *hu = nullptr, *hv = nullptr;
  if (pbbs_read_edge_array(in, &n, &m, &hu, &hv) != PBBS_OK) {
    std::fprintf(stderr, "readEdgeArrayFromFile: %s\n", pbbs_last_error());
    return 1;
  }
  int rc = 0;
  try {
    pbbs_amd::Context ctx(0);
    DeviceEdges d;
    const size_t bytes = (size_t)std::max<int64_t>(m, 1) * sizeof(int32_t);
    if (hipMalloc(&d.u, bytes) != hipSuccess || hipMalloc(&d.v, bytes) != hipSuccess ||
        hipMalloc(&d.forest, (size_t)std::max<int64_t>(n, 1) * sizeof(int32_t)) != hipSuccess ||
        hipMemcpy(d.u, hu, m * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.v, hv, m * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess)
      throw pbbs_amd::Error(PBBS_ERR_HIP, "edge copy into HBM failed");
    int64_t count = 0;
    pbbs_st_stats st{};
    const uint32_t fl = PBBS_EDGES_DEVICE | PBBS_OUT_DEVICE;
    const auto t0 = std::chrono::steady_clock::now();
    while (seconds_since(t0) < 1.0)
      pbbs_amd::check(pbbs_spanning_forest(ctx.get(), n, m, d.u, d.v, fl, d.forest, &count, &st),
                      "st");
    for (int k = 0; k < rounds; ++k) {
      const auto t = std::chrono::steady_clock::now();
      pbbs_amd::check(pbbs_spanning_forest(ctx.get(), n, m, d.u, d.v, fl, d.forest, &count, &st),
                      "st");
      std::printf("Parlay time: %.4f\n", seconds_since(t));
    }
    std::printf("\n");
    if (out) {
      std::vector<int32_t> ids((size_t)count);
      if (count > 0 &&
          hipMemcpy(ids.data(), d.forest, count * sizeof(int32_t), hipMemcpyDeviceToHost) !=
              hipSuccess)
        throw pbbs_amd::Error(PBBS_ERR_HIP, "forest copy from HBM failed");
      pbbs_amd::check(pbbs_write_sequence_int(out, ids.data(), count), "writeIntSeqToFile");
    }
  } catch (const pbbs_amd::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    rc = 1;
  }
  pbbs_free(hu);
  pbbs_free(hv);
  return rc;
}
