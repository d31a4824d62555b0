// Host CSR construction and the seeded graph generators of the benchmark
// inputs.  Construction follows common/graphUtils.h:54-94: drop self-loops and
// add reverse edges when symmetrising (makeSymmetric, :63-71), remove
// duplicates after sorting by (u,v) (remDuplicates, :54-61), then bucket by u
// into offsets (graphFromEdges, :73-94).  The sort here is an LSD radix sort
// over packed (u,v) keys, parallelised with OpenMP.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/pbbs_bfs.h"
#include "engine.hpp"

using pbbs::fail;

namespace {

// SplitMix64 finaliser (Steele, Lea, Flood 2014), used as a counter-based
// hash; it stands in for parlay::hash64, which is not available here.
inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

inline double unit32(uint32_t x) { return (x + 0.5) * (1.0 / 4294967296.0); }

int bits_for(int64_t n) {
  int b = 0;
  while ((int64_t(1) << b) < n) ++b;
  return std::max(b, 1);
}

// Uninitialised 64-bit key buffer: the parallel loops that fill it are the
// first to touch its pages, so page faults spread over all threads.
struct Keys {
  uint64_t* p = nullptr;
  int64_t n = 0;
  explicit Keys(int64_t count) : p(static_cast<uint64_t*>(std::malloc(std::max<int64_t>(count, 1) * 8))), n(count) {
    if (!p) throw std::bad_alloc();
  }
  Keys(const Keys&) = delete;
  Keys& operator=(const Keys&) = delete;
  ~Keys() { std::free(p); }
  uint64_t& operator[](int64_t i) { return p[i]; }
  const uint64_t& operator[](int64_t i) const { return p[i]; }
  int64_t size() const { return n; }
};

// Stable parallel LSD radix sort of keys on their low `key_bits` bits.
void radix_sort(Keys& a, int key_bits) {
  const int64_t N = a.size();
  if (N < 2) return;
  constexpr int D = 11;
  constexpr int B = 1 << D;
  Keys tmp(N);
  const int T = std::max(1, omp_get_max_threads());
  std::vector<int64_t> hist((size_t)T * B);
  for (int shift = 0; shift < key_bits; shift += D) {
    std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      const int64_t lo = N * t / T, hi = N * (t + 1) / T;
      int64_t* h = &hist[(size_t)t * B];
      for (int64_t i = lo; i < hi; ++i) h[(a[i] >> shift) & (B - 1)]++;
    }
    int64_t run = 0;
    for (int d = 0; d < B; ++d)
      for (int t = 0; t < T; ++t) {
        const int64_t c = hist[(size_t)t * B + d];
        hist[(size_t)t * B + d] = run;
        run += c;
      }
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      const int64_t lo = N * t / T, hi = N * (t + 1) / T;
      int64_t* h = &hist[(size_t)t * B];
      for (int64_t i = lo; i < hi; ++i) tmp[h[(a[i] >> shift) & (B - 1)]++] = a[i];
    }
    std::swap(a.p, tmp.p);
  }
}

// keys[i] = (u << vb) | v, sorted and deduplicated into a CSR.  Keys equal to
// UINT64_MAX are dropped (they mark removed self-loops).
int csr_from_keys(int64_t n, Keys& keys, int vb, pbbs_csr* out) {
  radix_sort(keys, 2 * vb);
  int64_t N = (int64_t)keys.size();
  while (N > 0 && keys[N - 1] == ~0ull) --N;
  // unique, in parallel: per-thread counts, prefix, compaction into place order
  const int T = std::max(1, omp_get_max_threads());
  std::vector<int64_t> cnt(T + 1, 0);
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    const int64_t lo = N * t / T, hi = N * (t + 1) / T;
    int64_t c = 0;
    for (int64_t i = lo; i < hi; ++i) c += (i == 0 || keys[i] != keys[i - 1]);
    cnt[t + 1] = c;
  }
  for (int t = 0; t < T; ++t) cnt[t + 1] += cnt[t];
  const int64_t m = cnt[T];
  if (m > (int64_t)UINT32_MAX) return fail(PBBS_ERR_LIMIT, "graph has %lld >= 2^32 edges", (long long)m);
  int64_t* offsets = static_cast<int64_t*>(std::malloc((n + 1) * sizeof(int64_t)));
  int32_t* adj = static_cast<int32_t*>(std::malloc(std::max<int64_t>(m, 1) * sizeof(int32_t)));
  if (!offsets || !adj) {
    std::free(offsets);
    std::free(adj);
    return fail(PBBS_ERR_NOMEM, "host CSR allocation failed");
  }
  const uint64_t vmask = (uint64_t(1) << vb) - 1;
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    const int64_t lo = N * t / T, hi = N * (t + 1) / T;
    int64_t w = cnt[t];
    for (int64_t i = lo; i < hi; ++i) {
      if (i != 0 && keys[i] == keys[i - 1]) continue;
      // keys are sorted, so keys[i-1] equals the previous distinct key; when
      // its row differs, rows prev_u+1 .. u (the empty ones too) start here.
      const int64_t u = (int64_t)(keys[i] >> vb);
      const int64_t prev_u = i == 0 ? -1 : (int64_t)(keys[i - 1] >> vb);
      for (int64_t x = prev_u + 1; x <= u; ++x) offsets[x] = w;
      adj[w] = (int32_t)(keys[i] & vmask);
      ++w;
    }
  }
  // empty rows after the last source
  const int64_t last_u = m > 0 ? (int64_t)(keys[N - 1] >> vb) : -1;
  for (int64_t x = last_u + 1; x <= n; ++x) offsets[x] = m;
  out->n = n;
  out->m = m;
  out->offsets = offsets;
  out->adj = adj;
  return PBBS_OK;
}

// Relabels both endpoints of every key with a seeded random permutation
// (Fisher-Yates), as graphReorder does before writing an unordered graph
// (common/graphUtils.h:166-189).
void permute_keys(Keys& keys, int64_t n, int vb, uint64_t seed) {
  std::vector<int32_t> perm(n);
  pbbs::make_permutation(n, seed, perm.data());
  const uint64_t vmask = (uint64_t(1) << vb) - 1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < keys.size(); ++i) {
    const uint64_t k = keys[i];
    if (k == ~0ull) continue;
    keys[i] = ((uint64_t)perm[k >> vb] << vb) | (uint64_t)perm[k & vmask];
  }
}

}  // namespace

namespace pbbs {

void make_permutation(int64_t n, uint64_t seed, int32_t* perm) {
  for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
  uint64_t s = mix64(seed ^ 0x5DEECE66Dull);
  for (int64_t i = n - 1; i > 0; --i) {
    s = mix64(s);
    const int64_t j = (int64_t)(s % (uint64_t)(i + 1));
    std::swap(perm[i], perm[j]);
  }
}

}  // namespace pbbs

extern "C" {

void pbbs_csr_free(pbbs_csr* g) {
  if (!g) return;
  std::free(g->offsets);
  std::free(g->adj);
  g->offsets = nullptr;
  g->adj = nullptr;
  g->n = g->m = 0;
}

int pbbs_csr_from_edges(int64_t n, int64_t num_edges, const int32_t* u, const int32_t* v,
                        int symmetrise, pbbs_csr* out) {
  pbbs::clear_error();
  if (!out || (num_edges > 0 && (!u || !v)) || num_edges < 0)
    return fail(PBBS_ERR_INVALID, "bad edge-list arguments");
  if (n <= 0 || n > INT32_MAX) return fail(PBBS_ERR_LIMIT, "n=%lld outside [1, 2^31)", (long long)n);
  for (int64_t i = 0; i < num_edges; ++i)
    if (u[i] < 0 || u[i] >= n || v[i] < 0 || v[i] >= n)
      return fail(PBBS_ERR_INVALID, "edge %lld = (%d,%d) outside [0,%lld)", (long long)i, u[i],
                  v[i], (long long)n);
  const int vb = bits_for(n);
  try {
    Keys keys(num_edges * (symmetrise ? 2 : 1));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < num_edges; ++i) {
      const uint64_t a = (uint64_t)u[i], b = (uint64_t)v[i];
      if (symmetrise) {
        const bool loop = a == b;
        keys[2 * i] = loop ? ~0ull : (a << vb) | b;
        keys[2 * i + 1] = loop ? ~0ull : (b << vb) | a;
      } else {
        keys[i] = (a << vb) | b;
      }
    }
    return csr_from_keys(n, keys, vb, out);
  } catch (const std::bad_alloc&) {
    return fail(PBBS_ERR_NOMEM, "edge buffer allocation failed");
  }
}

int pbbs_csr_permute(const pbbs_csr* g, uint64_t seed, pbbs_csr* out) {
  pbbs::clear_error();
  if (!g || !out || !g->offsets || (g->m > 0 && !g->adj))
    return fail(PBBS_ERR_INVALID, "bad graph arguments");
  if (seed == 0) return fail(PBBS_ERR_INVALID, "seed must be nonzero");
  const int64_t n = g->n;
  if (n <= 0 || n > INT32_MAX) return fail(PBBS_ERR_LIMIT, "n=%lld outside [1, 2^31)", (long long)n);
  if (g->offsets[0] != 0 || g->offsets[n] != g->m)
    return fail(PBBS_ERR_INVALID, "offsets do not span [0, m)");
  for (int64_t u = 0; u < n; ++u)
    if (g->offsets[u + 1] < g->offsets[u])
      return fail(PBBS_ERR_INVALID, "offsets decrease at row %lld", (long long)u);
  for (int64_t e = 0; e < g->m; ++e)
    if (g->adj[e] < 0 || g->adj[e] >= n)
      return fail(PBBS_ERR_INVALID, "edge %lld names vertex %d outside [0,%lld)", (long long)e,
                  g->adj[e], (long long)n);
  const int vb = bits_for(n);
  try {
    Keys keys(g->m);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t u = 0; u < n; ++u)
      for (int64_t e = g->offsets[u]; e < g->offsets[u + 1]; ++e)
        keys[e] = ((uint64_t)u << vb) | (uint64_t)(uint32_t)g->adj[e];
    permute_keys(keys, n, vb, seed);
    return csr_from_keys(n, keys, vb, out);
  } catch (const std::bad_alloc&) {
    return fail(PBBS_ERR_NOMEM, "edge buffer allocation failed");
  }
}

int pbbs_gen_rmat(int32_t scale, int64_t draws, double a, double b, double c, uint64_t seed,
                  uint64_t permute_seed, pbbs_csr* out) {
  pbbs::clear_error();
  if (!out) return fail(PBBS_ERR_INVALID, "out is NULL");
  if (scale < 1 || scale > 30) return fail(PBBS_ERR_LIMIT, "scale %d outside [1, 30]", scale);
  if (draws < 0) return fail(PBBS_ERR_INVALID, "negative edge count");
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
    return fail(PBBS_ERR_INVALID, "quadrant probabilities must be >= 0 with a+b+c <= 1");
  const int64_t n = int64_t(1) << scale;
  const int vb = scale;
  const double ab = a + b, abc = a + b + c;
  const uint64_t hseed = mix64(seed);
  try {
    Keys keys(draws * 2);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < draws; ++i) {
      uint64_t st = mix64(hseed ^ ((uint64_t)i * 0xD1B54A32D192ED03ull));
      uint64_t u = 0, v = 0;
      for (int l = 0; l < scale; l += 2) {
        st = mix64(st);
        for (int h = 0; h < 2 && l + h < scale; ++h) {
          const double r = unit32((uint32_t)(h ? (st >> 32) : st));
          const uint64_t ub = r >= ab, vbit = (r >= a && r < ab) || r >= abc;
          u = (u << 1) | ub;
          v = (v << 1) | vbit;
        }
      }
      const bool loop = u == v;
      keys[2 * i] = loop ? ~0ull : (u << vb) | v;
      keys[2 * i + 1] = loop ? ~0ull : (v << vb) | u;
    }
    if (permute_seed) permute_keys(keys, n, vb, permute_seed);
    return csr_from_keys(n, keys, vb, out);
  } catch (const std::bad_alloc&) {
    return fail(PBBS_ERR_NOMEM, "edge buffer allocation failed");
  }
}

int pbbs_gen_grid3d(int32_t d, pbbs_csr* out) {
  pbbs::clear_error();
  if (!out) return fail(PBBS_ERR_INVALID, "out is NULL");
  if (d < 1 || (int64_t)d * d * d > INT32_MAX) return fail(PBBS_ERR_LIMIT, "bad grid side %d", d);
  const int64_t D = d, n = D * D * D;
  const int vb = bits_for(n);
  auto loc = [D](int64_t i, int64_t j, int64_t k) {
    return ((i + D) % D) * D * D + ((j + D) % D) * D + (k + D) % D;
  };
  try {
    // three forward torus edges per vertex (gridGraph.C:62-70), symmetrised
    Keys keys(n * 6);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < D; ++i)
      for (int64_t j = 0; j < D; ++j)
        for (int64_t k = 0; k < D; ++k) {
          const uint64_t l = (uint64_t)loc(i, j, k);
          const uint64_t nb[3] = {(uint64_t)loc(i + 1, j, k), (uint64_t)loc(i, j + 1, k),
                                  (uint64_t)loc(i, j, k + 1)};
          for (int e = 0; e < 3; ++e) {
            const bool loop = nb[e] == l;
            keys[6 * l + 2 * e] = loop ? ~0ull : (l << vb) | nb[e];
            keys[6 * l + 2 * e + 1] = loop ? ~0ull : (nb[e] << vb) | l;
          }
        }
    return csr_from_keys(n, keys, vb, out);
  } catch (const std::bad_alloc&) {
    return fail(PBBS_ERR_NOMEM, "edge buffer allocation failed");
  }
}

int pbbs_gen_randlocal(int64_t n, int32_t k, uint64_t seed, pbbs_csr* out) {
  pbbs::clear_error();
  if (!out) return fail(PBBS_ERR_INVALID, "out is NULL");
  if (n < 2 || n > INT32_MAX) return fail(PBBS_ERR_LIMIT, "n=%lld outside [2, 2^31)", (long long)n);
  if (k < 0) return fail(PBBS_ERR_INVALID, "negative degree");
  const int vb = bits_for(n);
  const uint64_t hseed = mix64(seed ^ 0xA0761D6478BD642Full);
  const double dn = (double)n;
  try {
    Keys keys(n * k * 2);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      uint64_t st = mix64(hseed ^ ((uint64_t)i * 0xD1B54A32D192ED03ull));
      for (int r = 0; r < k; ++r) {
        st = mix64(st);
        const double u = unit32((uint32_t)st);
        int64_t dist = (int64_t)std::floor(std::pow(dn, u));
        dist = std::min<int64_t>(std::max<int64_t>(dist, 1), n - 1);
        const bool neg = (st >> 63) != 0;
        const int64_t j = neg ? (i - dist + n) % n : (i + dist) % n;
        const size_t p = (size_t)(i * k + r) * 2;
        keys[p] = ((uint64_t)i << vb) | (uint64_t)j;
        keys[p + 1] = ((uint64_t)j << vb) | (uint64_t)i;
      }
    }
    return csr_from_keys(n, keys, vb, out);
  } catch (const std::bad_alloc&) {
    return fail(PBBS_ERR_NOMEM, "edge buffer allocation failed");
  }
}

}  // extern "C"
