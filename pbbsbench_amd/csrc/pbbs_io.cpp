// PBBS text formats: AdjacencyGraph (docs/fileFormats/graph.md:16-39),
// EdgeArray (common/graphIO.h:136-166) and sequenceInt
// (docs/fileFormats/sequence.md:10-26).  Tokens are separated by
// any run of space, tab, line feed or carriage return (common/IO.h:41-50).
// Readers parse in parallel chunks split at token boundaries; writers format
// in parallel blocks and emit one value per line, as common/IO.h:113-166 does.
#include <omp.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pbbs_bfs.h"
#include "engine.hpp"

using pbbs::fail;

namespace {

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

int read_file(const char* path, std::vector<char>& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(PBBS_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
  std::fseek(f, 0, SEEK_END);
  const long long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0) {
    std::fclose(f);
    return fail(PBBS_ERR_IO, "cannot size %s", path);
  }
  buf.resize((size_t)sz + 1);
  const size_t got = sz ? std::fread(buf.data(), 1, (size_t)sz, f) : 0;
  std::fclose(f);
  if (got != (size_t)sz) return fail(PBBS_ERR_IO, "short read on %s", path);
  buf[sz] = '\n';
  return PBBS_OK;
}

// Splits buf[begin, end) into tokens; returns the start offset of each token.
// Parallel: chunk boundaries are moved forward to the next whitespace.
std::vector<int64_t> token_starts(const std::vector<char>& buf, int64_t begin, int64_t end) {
  const int T = std::max(1, omp_get_max_threads());
  std::vector<int64_t> cut(T + 1);
  cut[0] = begin;
  cut[T] = end;
  for (int t = 1; t < T; ++t) {
    int64_t p = begin + (end - begin) * t / T;
    while (p < end && !is_space(buf[p])) ++p;
    cut[t] = std::max(p, cut[t - 1]);
  }
  std::vector<std::vector<int64_t>> part(T);
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    auto& out = part[t];
    bool in_tok = false;
    for (int64_t p = cut[t]; p < cut[t + 1]; ++p) {
      const bool sp = is_space(buf[p]);
      if (!sp && !in_tok) out.push_back(p);
      in_tok = !sp;
    }
  }
  size_t total = 0;
  for (auto& p : part) total += p.size();
  std::vector<int64_t> starts;
  starts.reserve(total);
  for (auto& p : part) starts.insert(starts.end(), p.begin(), p.end());
  return starts;
}

// Decimal integer with optional leading '-', ending at whitespace.
inline bool parse_long(const char* s, long long* out) {
  bool neg = false;
  if (*s == '-') {
    neg = true;
    ++s;
  }
  if (!(*s >= '0' && *s <= '9')) return false;
  long long v = 0;
  while (*s >= '0' && *s <= '9') {
    v = v * 10 + (*s - '0');
    ++s;
  }
  if (!is_space(*s)) return false;
  *out = neg ? -v : v;
  return true;
}

std::string header_of(const std::vector<char>& buf, int64_t* after) {
  int64_t p = 0;
  const int64_t n = (int64_t)buf.size();
  while (p < n && is_space(buf[p])) ++p;
  int64_t q = p;
  while (q < n && !is_space(buf[q])) ++q;
  *after = q;
  return std::string(buf.data() + p, buf.data() + q);
}

// Appends "%lld\n" for x at p; returns the end.  Digits are produced in
// reverse into a small buffer, then copied.
inline char* put_line(char* p, long long x) {
  unsigned long long u = x < 0 ? 0ull - (unsigned long long)x : (unsigned long long)x;
  if (x < 0) *p++ = '-';
  char tmp[24];
  int k = 0;
  do {
    tmp[k++] = char('0' + u % 10);
    u /= 10;
  } while (u);
  while (k) *p++ = tmp[--k];
  *p++ = '\n';
  return p;
}

// Formats a[0..n) as "%lld\n" lines in parallel blocks and writes them.
template <typename T>
int write_lines(FILE* f, const T* a, int64_t n) {
  const int64_t block = 1 << 20;
  const int64_t nb = (n + block - 1) / block;
  const int T_ = std::max(1, omp_get_max_threads());
  std::vector<std::vector<char>> out(T_, std::vector<char>(block * 21));
  std::vector<int64_t> len(T_);
  for (int64_t b0 = 0; b0 < nb; b0 += T_) {
    const int64_t b1 = std::min(nb, b0 + T_);
#pragma omp parallel for schedule(static, 1) num_threads(T_)
    for (int64_t b = b0; b < b1; ++b) {
      char* s = out[b - b0].data();
      char* p = s;
      const int64_t lo = b * block, hi = std::min(n, lo + block);
      for (int64_t i = lo; i < hi; ++i) p = put_line(p, (long long)a[i]);
      len[b - b0] = p - s;
    }
    for (int64_t b = b0; b < b1; ++b)
      if (std::fwrite(out[b - b0].data(), 1, (size_t)len[b - b0], f) != (size_t)len[b - b0])
        return fail(PBBS_ERR_IO, "write failed");
  }
  return PBBS_OK;
}

// Formats "u v\n" for each pair in parallel blocks and writes them.
int write_pairs(FILE* f, const int32_t* u, const int32_t* v, int64_t m) {
  const int64_t block = 1 << 20;
  const int64_t nb = (m + block - 1) / block;
  const int T_ = std::max(1, omp_get_max_threads());
  std::vector<std::vector<char>> out(T_, std::vector<char>(block * 24));
  std::vector<int64_t> len(T_);
  for (int64_t b0 = 0; b0 < nb; b0 += T_) {
    const int64_t b1 = std::min(nb, b0 + T_);
#pragma omp parallel for schedule(static, 1) num_threads(T_)
    for (int64_t b = b0; b < b1; ++b) {
      char* s = out[b - b0].data();
      char* p = s;
      const int64_t lo = b * block, hi = std::min(m, lo + block);
      for (int64_t i = lo; i < hi; ++i) {
        p = put_line(p, (long long)u[i]);
        p[-1] = ' ';
        p = put_line(p, (long long)v[i]);
      }
      len[b - b0] = p - s;
    }
    for (int64_t b = b0; b < b1; ++b)
      if (std::fwrite(out[b - b0].data(), 1, (size_t)len[b - b0], f) != (size_t)len[b - b0])
        return fail(PBBS_ERR_IO, "write failed");
  }
  return PBBS_OK;
}

}  // namespace

extern "C" {

int pbbs_write_edge_array(const char* path, int64_t m, const int32_t* u, const int32_t* v) {
  pbbs::clear_error();
  if (!path || m < 0 || (m > 0 && (!u || !v))) return fail(PBBS_ERR_INVALID, "bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(PBBS_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
  std::fputs("EdgeArray\n", f);
  int rc = write_pairs(f, u, v, m);
  if (std::fclose(f) != 0 && rc == PBBS_OK) rc = fail(PBBS_ERR_IO, "close of %s failed", path);
  return rc;
}

int pbbs_read_edge_array(const char* path, int64_t* n_out, int64_t* m_out, int32_t** u_out,
                         int32_t** v_out) {
  pbbs::clear_error();
  if (!path || !n_out || !m_out || !u_out || !v_out) return fail(PBBS_ERR_INVALID, "NULL argument");
  *u_out = *v_out = nullptr;
  *n_out = *m_out = 0;
  std::vector<char> buf;
  int rc = read_file(path, buf);
  if (rc != PBBS_OK) return rc;
  int64_t after = 0;
  if (header_of(buf, &after) != "EdgeArray")
    return fail(PBBS_ERR_FORMAT, "%s: missing header EdgeArray", path);
  const std::vector<int64_t> tok = token_starts(buf, after, (int64_t)buf.size());
  // (tokens - 1) / 2 pairs with the header counted (common/graphIO.h:154): an
  // odd last token is ignored, as there
  const int64_t m = (int64_t)tok.size() / 2;
  int32_t* u = static_cast<int32_t*>(std::malloc(std::max<int64_t>(m, 1) * sizeof(int32_t)));
  int32_t* v = static_cast<int32_t*>(std::malloc(std::max<int64_t>(m, 1) * sizeof(int32_t)));
  if (!u || !v) {
    std::free(u);
    std::free(v);
    return fail(PBBS_ERR_NOMEM, "allocation failed");
  }
  long long bad = -1, mx = 0;
#pragma omp parallel for schedule(static) reduction(max : bad, mx)
  for (int64_t i = 0; i < m; ++i) {
    long long a = 0, b = 0;
    if (!parse_long(&buf[tok[2 * i]], &a) || !parse_long(&buf[tok[2 * i + 1]], &b) || a < 0 ||
        b < 0 || a >= INT32_MAX || b >= INT32_MAX) {
      bad = std::max(bad, (long long)i);
    } else {
      u[i] = (int32_t)a;
      v[i] = (int32_t)b;
      mx = std::max(mx, std::max(a, b));
    }
  }
  if (bad >= 0) {
    std::free(u);
    std::free(v);
    return fail(PBBS_ERR_FORMAT, "%s: edge %lld is not two vertex ids in [0, 2^31 - 1)", path,
                bad);
  }
  *u_out = u;
  *v_out = v;
  *m_out = m;
  *n_out = mx + 1;   // the largest id + 1 (common/graphIO.h:159-164)
  return PBBS_OK;
}

void pbbs_free(void* p) { std::free(p); }

int pbbs_write_adjacency_graph(const char* path, const pbbs_csr* g) {
  pbbs::clear_error();
  if (!path || !g || (g->n > 0 && !g->offsets) || (g->m > 0 && !g->adj))
    return fail(PBBS_ERR_INVALID, "NULL argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(PBBS_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
  std::fprintf(f, "AdjacencyGraph\n%lld\n%lld\n", (long long)g->n, (long long)g->m);
  int rc = write_lines(f, g->offsets, g->n);
  if (rc == PBBS_OK) rc = write_lines(f, g->adj, g->m);
  if (std::fclose(f) != 0 && rc == PBBS_OK) rc = fail(PBBS_ERR_IO, "close of %s failed", path);
  return rc;
}

int pbbs_write_sequence_int(const char* path, const int32_t* a, int64_t n) {
  pbbs::clear_error();
  if (!path || (n > 0 && !a) || n < 0) return fail(PBBS_ERR_INVALID, "bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(PBBS_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
  std::fputs("sequenceInt\n", f);
  int rc = write_lines(f, a, n);
  if (std::fclose(f) != 0 && rc == PBBS_OK) rc = fail(PBBS_ERR_IO, "close of %s failed", path);
  return rc;
}

int pbbs_read_sequence_int(const char* path, int32_t** out, int64_t* n_out) {
  pbbs::clear_error();
  if (!path || !out || !n_out) return fail(PBBS_ERR_INVALID, "NULL argument");
  *out = nullptr;
  *n_out = 0;
  std::vector<char> buf;
  int rc = read_file(path, buf);
  if (rc != PBBS_OK) return rc;
  int64_t after = 0;
  // common/IO.h:217-229 accepts only the sequenceInt header
  if (header_of(buf, &after) != "sequenceInt")
    return fail(PBBS_ERR_FORMAT, "%s: missing header sequenceInt", path);
  const std::vector<int64_t> tok = token_starts(buf, after, (int64_t)buf.size());
  const int64_t n = (int64_t)tok.size();
  int32_t* a = static_cast<int32_t*>(std::malloc(std::max<int64_t>(n, 1) * sizeof(int32_t)));
  if (!a) return fail(PBBS_ERR_NOMEM, "allocation failed");
  long long bad = -1;
#pragma omp parallel for schedule(static) reduction(max : bad)
  for (int64_t i = 0; i < n; ++i) {
    long long x = 0;
    if (!parse_long(&buf[tok[i]], &x) || x < INT32_MIN || x > INT32_MAX)
      bad = std::max(bad, (long long)i);
    else
      a[i] = (int32_t)x;
  }
  if (bad >= 0) {
    std::free(a);
    return fail(PBBS_ERR_FORMAT, "%s: element %lld is not a valid int", path, bad);
  }
  *out = a;
  *n_out = n;
  return PBBS_OK;
}

}  // extern "C"
