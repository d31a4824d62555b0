// Graph ingest: the PBBS AdjacencyGraph text format (docs/fileFormats/
// graph.md:16-39, read as common/graphIO.h:191-215 does) parsed straight
// from a memory-mapped file, and a binary CSR file for fast reloads.
//
// Text: a two-pass parallel parser.  The token range is cut into chunks at
// whitespace; pass 1 counts each chunk's tokens, a prefix gives each chunk
// its first token index, pass 2 parses every token straight into its place
// (offsets, then adjacency).  No per-token table is built (8 bytes per token
// would be ~4 GB for a 2^24-vertex, 2^29-edge graph).
//
// Binary ("PBBSCSR1", this library's own format; the reference has none):
//   bytes 0-7   magic "PBBSCSR1"
//   bytes 8-15  n (uint64, little-endian)
//   bytes 16-23 m (uint64)
//   bytes 24-31 reserved (0)
//   then offsets int64[n+1] (offsets[n] == m), then adjacency int32[m].
#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/pbbs_bfs.h"
#include "engine.hpp"

using pbbs::fail;

namespace {

constexpr char kMagic[8] = {'P', 'B', 'B', 'S', 'C', 'S', 'R', '1'};
constexpr int64_t kBinHeader = 32;

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

struct Mapped {
  const char* p = nullptr;
  int64_t size = 0;
  int fd = -1;
  ~Mapped() {
    if (p && size) munmap(const_cast<char*>(p), (size_t)size);
    if (fd >= 0) close(fd);
  }
};

int map_file(const char* path, Mapped& f) {
  f.fd = open(path, O_RDONLY);
  if (f.fd < 0) return fail(PBBS_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
  struct stat st;
  if (fstat(f.fd, &st) != 0) return fail(PBBS_ERR_IO, "cannot stat %s", path);
  f.size = (int64_t)st.st_size;
  if (f.size == 0) return PBBS_OK;
  void* m = mmap(nullptr, (size_t)f.size, PROT_READ, MAP_PRIVATE, f.fd, 0);
  if (m == MAP_FAILED) return fail(PBBS_ERR_IO, "cannot map %s: %s", path, std::strerror(errno));
  madvise(m, (size_t)f.size, MADV_SEQUENTIAL);
  f.p = static_cast<const char*>(m);
  return PBBS_OK;
}

// Decimal integer with optional '-' at s, ending at whitespace or `end`.
inline const char* parse_tok(const char* s, const char* end, long long* out, bool* ok) {
  bool neg = false;
  if (s < end && *s == '-') {
    neg = true;
    ++s;
  }
  if (!(s < end && *s >= '0' && *s <= '9')) {
    *ok = false;
    while (s < end && !is_space(*s)) ++s;
    return s;
  }
  long long v = 0;
  int digits = 0;
  while (s < end && *s >= '0' && *s <= '9') {
    v = v * 10 + (*s - '0');
    ++s;
    ++digits;
  }
  *ok = digits <= 18 && (s == end || is_space(*s));
  while (s < end && !is_space(*s)) ++s;
  *out = neg ? -v : v;
  return s;
}

inline const char* skip_space(const char* s, const char* end) {
  while (s < end && is_space(*s)) ++s;
  return s;
}

struct Parsed {
  int64_t n = 0, m = 0;
};

// Parses an AdjacencyGraph text into buffers from alloc(n, m) ->
// (int64 offsets[n+1], int32 adj[m]).  Same checks as the reference reader:
// the header, and exactly n + m + 3 tokens (common/graphIO.h:196-206).
template <typename Alloc>
int parse_adjacency(const char* path, const Mapped& f, Alloc alloc, int64_t** off_out,
                    int32_t** adj_out, Parsed* out) {
  const char* b = f.p;
  const char* e = f.p + f.size;
  const char* s = skip_space(b, e);
  const char* h = s;
  while (s < e && !is_space(*s)) ++s;
  if (s - h != 14 || std::memcmp(h, "AdjacencyGraph", 14) != 0)
    return fail(PBBS_ERR_FORMAT, "%s: missing header AdjacencyGraph", path);
  // n and m
  long long n = 0, m = 0;
  bool ok1 = true, ok2 = true;
  s = skip_space(s, e);
  if (s == e) return fail(PBBS_ERR_FORMAT, "%s: missing n and m", path);
  s = parse_tok(s, e, &n, &ok1);
  s = skip_space(s, e);
  if (s == e) return fail(PBBS_ERR_FORMAT, "%s: missing n and m", path);
  s = parse_tok(s, e, &m, &ok2);
  if (!ok1 || !ok2 || n < 0 || m < 0) return fail(PBBS_ERR_FORMAT, "%s: bad n or m", path);
  if (n > INT32_MAX) return fail(PBBS_ERR_LIMIT, "%s: n=%lld >= 2^31", path, n);
  // chunks of the remaining tokens, cut at whitespace
  const int T = std::max(1, omp_get_max_threads());
  const int64_t base = s - b, len = e - s;
  std::vector<int64_t> cut(T + 1);
  cut[0] = base;
  cut[T] = f.size;
  for (int t = 1; t < T; ++t) {
    int64_t p = base + len * t / T;
    while (p < f.size && !is_space(b[p])) ++p;
    cut[t] = std::max(p, cut[t - 1]);
  }
  std::vector<int64_t> cnt(T + 1, 0);
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    int64_t c = 0;
    bool in_tok = false;
    for (int64_t p = cut[t]; p < cut[t + 1]; ++p) {
      const bool sp = is_space(b[p]);
      c += !sp && !in_tok;
      in_tok = !sp;
    }
    cnt[t + 1] = c;
  }
  for (int t = 0; t < T; ++t) cnt[t + 1] += cnt[t];
  if (cnt[T] != n + m)
    return fail(PBBS_ERR_FORMAT, "%s: length = %lld, n+m+3 = %lld", path,
                (long long)cnt[T] + 3, n + m + 3);
  int64_t* off = nullptr;
  int32_t* adj = nullptr;
  const int rc = alloc(n, m, &off, &adj);
  if (rc != PBBS_OK) return rc;
  long long bad = -1;
#pragma omp parallel num_threads(T) reduction(max : bad)
  {
    const int t = omp_get_thread_num();
    int64_t k = cnt[t];
    const char* p = b + cut[t];
    const char* q = b + cut[t + 1];
    while (true) {
      p = skip_space(p, q);
      if (p >= q) break;
      long long x = 0;
      bool ok = true;
      p = parse_tok(p, q, &x, &ok);
      if (k < n) {
        if (!ok) bad = std::max(bad, (long long)k);
        else off[k] = x;
      } else {
        if (!ok || x < INT32_MIN || x > INT32_MAX) bad = std::max(bad, (long long)k);
        else adj[k - n] = (int32_t)x;
      }
      ++k;
    }
  }
  if (bad >= 0)
    return fail(PBBS_ERR_FORMAT, "%s: token %lld is not a valid integer", path, bad + 3);
  off[n] = m;   // tagged on, as common/graphIO.h:209-210 does
  *off_out = off;
  *adj_out = adj;
  out->n = n;
  out->m = m;
  return PBBS_OK;
}

bool is_binary(const Mapped& f) {
  return f.size >= kBinHeader && std::memcmp(f.p, kMagic, 8) == 0;
}

int binary_dims(const char* path, const Mapped& f, int64_t* n, int64_t* m) {
  uint64_t un = 0, um = 0;
  std::memcpy(&un, f.p + 8, 8);
  std::memcpy(&um, f.p + 16, 8);
  if (un > (uint64_t)INT32_MAX || um > (uint64_t)UINT32_MAX)
    return fail(PBBS_ERR_LIMIT, "%s: n=%llu or m=%llu too large", path, (unsigned long long)un,
                (unsigned long long)um);
  const int64_t want = kBinHeader + (int64_t)(un + 1) * 8 + (int64_t)um * 4;
  if (f.size != want)
    return fail(PBBS_ERR_FORMAT, "%s: %lld bytes, a PBBSCSR1 file of n=%llu m=%llu has %lld",
                path, (long long)f.size, (unsigned long long)un, (unsigned long long)um,
                (long long)want);
  *n = (int64_t)un;
  *m = (int64_t)um;
  return PBBS_OK;
}

void parallel_copy(void* dst, const void* src, int64_t bytes) {
  const int64_t piece = 1 << 24;
  const int64_t np = (bytes + piece - 1) / piece;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < np; ++i) {
    const int64_t lo = i * piece, len = std::min(piece, bytes - lo);
    std::memcpy(static_cast<char*>(dst) + lo, static_cast<const char*>(src) + lo, (size_t)len);
  }
}

template <typename Alloc>
int read_binary(const char* path, const Mapped& f, Alloc alloc, int64_t** off_out,
                int32_t** adj_out, Parsed* out) {
  int64_t n = 0, m = 0;
  int rc = binary_dims(path, f, &n, &m);
  if (rc != PBBS_OK) return rc;
  int64_t* off = nullptr;
  int32_t* adj = nullptr;
  rc = alloc(n, m, &off, &adj);
  if (rc != PBBS_OK) return rc;
  parallel_copy(off, f.p + kBinHeader, (n + 1) * 8);
  parallel_copy(adj, f.p + kBinHeader + (n + 1) * 8, m * 4);
  *off_out = off;
  *adj_out = adj;
  out->n = n;
  out->m = m;
  return PBBS_OK;
}

int malloc_csr(int64_t n, int64_t m, int64_t** off, int32_t** adj) {
  *off = static_cast<int64_t*>(std::malloc((n + 1) * sizeof(int64_t)));
  *adj = static_cast<int32_t*>(std::malloc(std::max<int64_t>(m, 1) * sizeof(int32_t)));
  if (!*off || !*adj) {
    std::free(*off);
    std::free(*adj);
    *off = nullptr;
    *adj = nullptr;
    return fail(PBBS_ERR_NOMEM, "host CSR allocation failed");
  }
  return PBBS_OK;
}

// Pinned host buffers (hipHostMalloc), released by the destructor.
struct Pinned {
  int64_t* off = nullptr;
  int32_t* adj = nullptr;
  ~Pinned() {
    if (off) (void)hipHostFree(off);
    if (adj) (void)hipHostFree(adj);
  }
};

}  // namespace

extern "C" {

int pbbs_read_adjacency_graph(const char* path, pbbs_csr* out) {
  pbbs::clear_error();
  if (!path || !out) return fail(PBBS_ERR_INVALID, "NULL argument");
  Mapped f;
  int rc = map_file(path, f);
  if (rc != PBBS_OK) return rc;
  int64_t* off = nullptr;
  int32_t* adj = nullptr;
  Parsed p;
  rc = parse_adjacency(path, f, malloc_csr, &off, &adj, &p);
  if (rc != PBBS_OK) {
    std::free(off);
    std::free(adj);
    return rc;
  }
  out->n = p.n;
  out->m = p.m;
  out->offsets = off;
  out->adj = adj;
  return PBBS_OK;
}

int pbbs_write_graph_binary(const char* path, const pbbs_csr* g) {
  pbbs::clear_error();
  if (!path || !g || g->n < 0 || g->m < 0 || (g->n > 0 && !g->offsets) || (g->m > 0 && !g->adj))
    return fail(PBBS_ERR_INVALID, "bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(PBBS_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
  unsigned char hdr[kBinHeader] = {};
  std::memcpy(hdr, kMagic, 8);
  const uint64_t un = (uint64_t)g->n, um = (uint64_t)g->m;
  std::memcpy(hdr + 8, &un, 8);
  std::memcpy(hdr + 16, &um, 8);
  bool ok = std::fwrite(hdr, 1, sizeof hdr, f) == sizeof hdr;
  ok = ok && std::fwrite(g->offsets, 8, (size_t)g->n + 1, f) == (size_t)g->n + 1;
  ok = ok && (g->m == 0 || std::fwrite(g->adj, 4, (size_t)g->m, f) == (size_t)g->m);
  if (std::fclose(f) != 0) ok = false;
  return ok ? PBBS_OK : fail(PBBS_ERR_IO, "write of %s failed", path);
}

int pbbs_read_graph_binary(const char* path, pbbs_csr* out) {
  pbbs::clear_error();
  if (!path || !out) return fail(PBBS_ERR_INVALID, "NULL argument");
  Mapped f;
  int rc = map_file(path, f);
  if (rc != PBBS_OK) return rc;
  if (!is_binary(f)) return fail(PBBS_ERR_FORMAT, "%s: not a PBBSCSR1 file", path);
  int64_t* off = nullptr;
  int32_t* adj = nullptr;
  Parsed p;
  rc = read_binary(path, f, malloc_csr, &off, &adj, &p);
  if (rc != PBBS_OK) return rc;
  out->n = p.n;
  out->m = p.m;
  out->offsets = off;
  out->adj = adj;
  return PBBS_OK;
}

int pbbs_graph_read(pbbs_ctx* ctx, const char* path, uint32_t flags, pbbs_graph** out) {
  pbbs::clear_error();
  if (!ctx || !path || !out) return fail(PBBS_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (flags & ~(uint32_t)(PBBS_GRAPH_OFFSETS_I64 | PBBS_GRAPH_ASSUME_SYMMETRIC))
    return fail(PBBS_ERR_INVALID, "flags: PBBS_GRAPH_OFFSETS_I64 / PBBS_GRAPH_ASSUME_SYMMETRIC only");
  Mapped f;
  int rc = map_file(path, f);
  if (rc != PBBS_OK) return rc;
  Pinned pin;
  auto pinned = [&](int64_t n, int64_t m, int64_t** off, int32_t** adj) -> int {
    if (hipHostMalloc((void**)&pin.off, (n + 1) * sizeof(int64_t), hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc((void**)&pin.adj, std::max<int64_t>(m, 1) * sizeof(int32_t),
                      hipHostMallocDefault) != hipSuccess)
      return fail(PBBS_ERR_NOMEM, "pinned staging for n=%lld m=%lld failed", (long long)n,
                  (long long)m);
    *off = pin.off;
    *adj = pin.adj;
    return PBBS_OK;
  };
  int64_t* off = nullptr;
  int32_t* adj = nullptr;
  Parsed p;
  rc = is_binary(f) ? read_binary(path, f, pinned, &off, &adj, &p)
                    : parse_adjacency(path, f, pinned, &off, &adj, &p);
  if (rc != PBBS_OK) return rc;
  if (flags & PBBS_GRAPH_OFFSETS_I64)
    return pbbs_graph_upload(ctx, p.n, p.m, off, adj, flags, out);
  if (p.m > (int64_t)UINT32_MAX)
    return fail(PBBS_ERR_LIMIT, "%s: m=%lld needs PBBS_GRAPH_OFFSETS_I64", path, (long long)p.m);
  // Narrow the offsets in place (entry i lands in bytes already read, so the
  // loop runs in ascending order).  Values outside [0, m] would wrap, so they
  // are rejected here; the upload checks the rest.
  uint32_t* o32 = reinterpret_cast<uint32_t*>(off);
  for (int64_t i = 0; i <= p.n; ++i) {
    const int64_t x = off[i];
    if (x < 0 || x > p.m)
      return fail(PBBS_ERR_FORMAT, "%s: offsets[%lld]=%lld outside [0, m]", path, (long long)i,
                  (long long)x);
    o32[i] = (uint32_t)x;
  }
  return pbbs_graph_upload(ctx, p.n, p.m, o32, adj, flags, out);
}

}  // extern "C"
