// spm_host.cpp -- host data layer: seeded graph generators, graph.hpp-semantics
// CSR construction and text parsing, and the algorithmic-byte models.
//
// CSR semantics follow StarPlat's graphcode/graph.hpp:
//   parseEdges (:464-515)        max id -> num_nodes, duplicate pairs summed
//   parseEdgesContent (:720-859) per-row ascending destinations, reverse CSR
//                                 with per-row ascending sources
// The reference builds these through a std::map of every pair and an
// O(indeg^2) edgeMap loop; here the same arrays come from a parallel sort of
// packed (src,dst) keys, which produces identical offsets and row contents.
#include "../../include/starplat_host.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  // splitmix64 finaliser (Steele, Lea, Flood 2014).
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Counter-based stream: value depends only on (seed, stream, counter).
inline uint64_t rng_base(uint64_t seed, uint64_t stream) { return mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull)); }
inline uint64_t rng_at(uint64_t base, uint64_t counter) { return mix64(base + counter); }
inline uint64_t rng(uint64_t seed, uint64_t stream, uint64_t counter) { return rng_at(rng_base(seed, stream), counter); }

inline uint64_t key_of(uint32_t s, uint32_t d) { return (uint64_t(s) << 32) | d; }
inline int32_t key_src(uint64_t k) { return int32_t(k >> 32); }
inline int32_t key_dst(uint64_t k) { return int32_t(k & 0xffffffffu); }

int32_t* alloc_i32(int64_t n) {
  int32_t* p = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (n > 0 ? n : 1)));
  if (!p) { std::fprintf(stderr, "spm_host: out of memory (%lld ints)\n", (long long)n); std::abort(); }
  return p;
}

// Exclusive prefix sum of n counts: out[i] = cnt[0] + ... + cnt[i-1], out[n]
// the total (two passes over per-thread slices).
int64_t scan_counts(const int32_t* cnt, int64_t n, int64_t* out) {
  const int T = omp_get_max_threads();
  std::vector<int64_t> part(size_t(T) + 1, 0);
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    int64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += cnt[i];
    part[size_t(t) + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int i = 0; i < nt; ++i) part[size_t(i) + 1] += part[size_t(i)];
    int64_t acc = part[size_t(t)];
    for (int64_t i = lo; i < hi; ++i) {
      out[i] = acc;
      acc += cnt[i];
    }
  }
  int64_t total = 0;
  for (int t = 0; t < T; ++t) total += part[size_t(t) + 1] - part[size_t(t)];
  out[n] = total;
  return total;
}

// Ascending sort of one adjacency row (rows are mostly short).
inline void sort_row(int32_t* a, int64_t n) {
  if (n > 32) {
    std::sort(a, a + n);
    return;
  }
  for (int64_t i = 1; i < n; ++i) {
    const int32_t x = a[i];
    int64_t j = i - 1;
    while (j >= 0 && a[j] > x) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = x;
  }
}

spmh_graph* alloc_graph(int32_t V, int64_t E) {
  spmh_graph* g = static_cast<spmh_graph*>(std::calloc(1, sizeof(spmh_graph)));
  g->V = V;
  g->E = E;
  g->indexofNodes = alloc_i32(int64_t(V) + 1);
  g->edgeList = alloc_i32(E);
  g->edgeLen = alloc_i32(E);
  g->rev_indexofNodes = alloc_i32(int64_t(V) + 1);
  g->srcList = alloc_i32(E);
  return g;
}

// Stable bucketing of n items (row(i) in [0, V), val(i)) into CSR rows:
// out[start[r] .. start[r+1]) receives the values of row r in input order.
// Two counting passes, no atomics: items are first partitioned by the high
// bits of their row (per-thread histograms over contiguous input slices, so
// each thread writes its items in input order), then each partition is
// counted and placed by the full row id within its small row range.
template <class Row, class Val>
void stable_rows(int64_t n, int32_t V, Row row, Val val, int64_t* start, int32_t* out) {
  int shift = 0;
  while ((int64_t(V - 1) >> shift) >= 1024) ++shift;  // at most 1024 partitions
  const int32_t B = int32_t((int64_t(V - 1) >> shift) + 1);
  const int T = omp_get_max_threads();
  std::vector<int64_t> hist(size_t(T) * size_t(B), 0);
  std::vector<uint64_t> tmp(size_t(std::max<int64_t>(n, 1)));
  std::vector<int64_t> pstart(size_t(B) + 1, 0);
  int used = T;
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
#pragma omp single
    used = nt;
    const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    int64_t* h = hist.data() + size_t(t) * B;
    for (int64_t i = lo; i < hi; ++i) {
      const int32_t r = row(i);
      if (r >= 0) ++h[r >> shift];
    }
#pragma omp barrier
#pragma omp single
    {
      int64_t acc = 0;
      for (int32_t b = 0; b < B; ++b) {
        pstart[size_t(b)] = acc;
        for (int u = 0; u < nt; ++u) {
          const int64_t c = hist[size_t(u) * B + b];
          hist[size_t(u) * B + b] = acc;
          acc += c;
        }
      }
      pstart[size_t(B)] = acc;
    }
    for (int64_t i = lo; i < hi; ++i) {
      const int32_t r = row(i);
      if (r >= 0) tmp[size_t(h[r >> shift]++)] = (uint64_t(uint32_t(r)) << 32) | uint32_t(val(i));
    }
  }
  (void)used;
  const int32_t span = int32_t(1) << shift;
#pragma omp parallel
  {
    std::vector<int64_t> local(size_t(span) + 1);
#pragma omp for schedule(dynamic, 1)
    for (int32_t b = 0; b < B; ++b) {
      const int32_t r0 = b << shift;
      const int32_t r1 = int32_t(std::min<int64_t>(int64_t(r0) + span, V));
      std::fill(local.begin(), local.end(), 0);
      const int64_t a = pstart[size_t(b)], z = pstart[size_t(b) + 1];
      for (int64_t i = a; i < z; ++i) ++local[size_t(int32_t(tmp[size_t(i)] >> 32) - r0) + 1];
      for (int32_t r = r0; r < r1; ++r) {
        local[size_t(r - r0) + 1] += local[size_t(r - r0)];
        start[r] = a + local[size_t(r - r0)];
      }
      for (int64_t i = a; i < z; ++i) {
        const uint64_t x = tmp[size_t(i)];
        out[a + local[size_t(int32_t(x >> 32) - r0)]++] = int32_t(uint32_t(x));
      }
    }
  }
  start[V] = pstart[size_t(B)];
}

// Reverse CSR of g's forward CSR: the edges, visited in forward order (sources
// ascending), are bucketed stably by destination, so every srcList row comes
// out ascending (graph.hpp:810-832 sorts each row the same way).
void build_reverse(spmh_graph* g) {
  const int32_t V = g->V;
  const int64_t E = g->E;
  std::vector<int32_t> src_of(size_t(std::max<int64_t>(E, 1)));
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t u = 0; u < V; ++u)
    for (int32_t e = g->indexofNodes[u]; e < g->indexofNodes[u + 1]; ++e) src_of[size_t(e)] = u;
  std::vector<int64_t> start(size_t(V) + 1);
  stable_rows(
      E, V, [&](int64_t e) { return g->edgeList[e]; }, [&](int64_t e) { return src_of[size_t(e)]; }, start.data(),
      g->srcList);
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v <= V; ++v) g->rev_indexofNodes[v] = int32_t(start[size_t(v)]);
}

// Builds the forward and reverse CSR from keys sorted ascending and unique,
// with weights aligned to the keys (or NULL for zero weights).
spmh_graph* build_csr(const std::vector<uint64_t>& keys, const int32_t* w, int32_t V) {
  const int64_t E = (int64_t)keys.size();
  spmh_graph* g = alloc_graph(V, E);
  std::vector<int32_t> cnt(size_t(V), 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < E; ++i) {
    __atomic_fetch_add(&cnt[size_t(key_src(keys[i]))], 1, __ATOMIC_RELAXED);
    g->edgeList[i] = key_dst(keys[i]);
    g->edgeLen[i] = w ? w[i] : 0;
  }
  std::vector<int64_t> off(size_t(V) + 1);
  scan_counts(cnt.data(), V, off.data());
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v <= V; ++v) g->indexofNodes[v] = int32_t(off[size_t(v)]);
  build_reverse(g);
  return g;
}

// CSR of the distinct pairs among n raw (src, dst) samples in any order, with
// duplicates: self-loops dropped; for undirected graphs every kept pair also
// stands for its mirror. num_nodes = largest id of a kept pair + 1 (graph.hpp
// starts nodesTotal at 0, :73). Weights, when asked for, are hashed from the
// pair (undirected: from the ordered pair, so both directions agree). The
// samples are bucketed by source (stable_rows) and each row sorted and
// deduplicated on its own, instead of sorting all samples globally.
spmh_graph* finish(const uint64_t* keys, int64_t n, uint64_t seed, int undirected, int wmin, int wmax) {
  int32_t m = 0;
#pragma omp parallel for reduction(max : m) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const int32_t s = key_src(keys[i]), d = key_dst(keys[i]);
    if (s != d) m = std::max(m, std::max(s, d));
  }
  const int32_t V = m + 1;
  // Item j < n is sample j as (src, dst); with undirected, item n + j its mirror.
  const int64_t items = undirected ? 2 * n : n;
  auto row = [&](int64_t j) {
    const uint64_t k = keys[j < n ? j : j - n];
    const int32_t s = key_src(k), d = key_dst(k);
    return s == d ? -1 : (j < n ? s : d);
  };
  auto val = [&](int64_t j) {
    const uint64_t k = keys[j < n ? j : j - n];
    return j < n ? key_dst(k) : key_src(k);
  };
  std::vector<int64_t> start(size_t(V) + 1);
  int64_t kept = 0;
#pragma omp parallel for reduction(+ : kept) schedule(static)
  for (int64_t i = 0; i < n; ++i) kept += key_src(keys[i]) != key_dst(keys[i]);
  std::vector<int32_t> rows(size_t(std::max<int64_t>(undirected ? 2 * kept : kept, 1)));
  stable_rows(items, V, row, val, start.data(), rows.data());
  // Sort and deduplicate each row in place; cnt[v] becomes its distinct count.
  std::vector<int32_t> cnt(static_cast<size_t>(V));
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t v = 0; v < V; ++v) {
    int32_t* a = rows.data() + start[size_t(v)];
    const int64_t len = start[size_t(v) + 1] - start[size_t(v)];
    sort_row(a, len);
    cnt[size_t(v)] = len ? int32_t(std::unique(a, a + len) - a) : 0;
  }
  std::vector<int64_t> off(size_t(V) + 1);
  const int64_t E = scan_counts(cnt.data(), V, off.data());
  if (E > 0x7fffffff) {
    std::fprintf(stderr, "spm_host: %lld edges exceed the int32 CSR\n", (long long)E);
    std::abort();
  }
  spmh_graph* g = alloc_graph(V, E);
  const uint64_t span = wmax > 0 ? uint64_t(wmax - wmin + 1) : 1;
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t v = 0; v < V; ++v) {
    const int32_t* a = rows.data() + start[size_t(v)];
    const int64_t o = off[size_t(v)];
    g->indexofNodes[v] = int32_t(o);
    for (int32_t j = 0; j < cnt[size_t(v)]; ++j) {
      const int32_t d = a[j];
      g->edgeList[o + j] = d;
      uint32_t s0 = uint32_t(v), d0 = uint32_t(d);
      if (undirected && s0 > d0) std::swap(s0, d0);
      g->edgeLen[o + j] = wmax > 0 ? wmin + int32_t(rng(seed, 0x57, key_of(s0, d0)) % span) : 0;
    }
  }
  g->indexofNodes[V] = int32_t(E);
  if (undirected) {  // symmetric: the in-rows are the out-rows
    std::memcpy(g->rev_indexofNodes, g->indexofNodes, sizeof(int32_t) * (size_t(V) + 1));
    std::memcpy(g->srcList, g->edgeList, sizeof(int32_t) * size_t(E));
  } else {
    build_reverse(g);
  }
  return g;
}

// Uninitialised sample buffer (the sampling loop writes every element, so
// its pages are first touched by the threads that fill them).
struct Samples {
  uint64_t* p;
  explicit Samples(int64_t n) : p(static_cast<uint64_t*>(std::malloc(sizeof(uint64_t) * size_t(n > 0 ? n : 1)))) {
    if (!p) { std::fprintf(stderr, "spm_host: out of memory (%lld samples)\n", (long long)n); std::abort(); }
  }
  ~Samples() { std::free(p); }
};

}  // namespace

extern "C" {

spmh_graph* spmh_rmat(int scale, int edge_factor, double a, double b, double c,
                      uint64_t seed, int undirected, int wmin, int wmax) {
  if (scale < 1 || scale > 30 || edge_factor < 1) return nullptr;
  const int64_t n = int64_t(edge_factor) << scale;
  // Integer thresholds on 32-bit uniforms: quadrant a | b | c | d.
  const uint64_t ta = uint64_t(a * 4294967296.0);
  const uint64_t tb = uint64_t((a + b) * 4294967296.0);
  const uint64_t tc = uint64_t((a + b + c) * 4294967296.0);
  Samples keys(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    uint32_t s = 0, d = 0;
    const uint64_t base = rng_base(seed, uint64_t(i));
    for (int l = 0; l < scale; l += 2) {
      const uint64_t h = rng_at(base, uint64_t(l));
      for (int k = 0; k < 2 && l + k < scale; ++k) {
        const uint64_t r = k ? (h >> 32) : (h & 0xffffffffu);
        // Quadrant a | b | c | d of [0, 2^32) without branches (their
        // outcomes are random, so a branchy form mispredicts every level):
        // bs = r >= tb; bd = (ta <= r < tb) or (r >= tc).
        const uint32_t bs = uint32_t(r >= tb);
        const uint32_t bd = uint32_t(r >= ta) ^ bs ^ uint32_t(r >= tc);
        s = (s << 1) | bs;
        d = (d << 1) | bd;
      }
    }
    keys.p[i] = key_of(s, d);
  }
  return finish(keys.p, n, seed, undirected, wmin, wmax);
}

spmh_graph* spmh_er(int64_t nv, int64_t m, uint64_t seed, int undirected, int wmin, int wmax) {
  if (nv < 1 || nv > 0x7fffffff || m < 0) return nullptr;
  Samples keys(m);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    const uint64_t base = rng_base(seed, uint64_t(i));
    const uint64_t h1 = rng_at(base, 1), h2 = rng_at(base, 2);
    // Multiply-shift maps a 64-bit uniform onto [0, nv).
    const uint32_t s = uint32_t((unsigned __int128)h1 * (uint64_t)nv >> 64);
    const uint32_t d = uint32_t((unsigned __int128)h2 * (uint64_t)nv >> 64);
    keys.p[i] = key_of(s, d);
  }
  return finish(keys.p, m, seed, undirected, wmin, wmax);
}

spmh_graph* spmh_grid(int rows, int cols, uint64_t seed, int wmin, int wmax) {
  if (rows < 1 || cols < 1 || int64_t(rows) * cols > 0x7fffffff) return nullptr;
  std::vector<uint64_t> keys;
  keys.reserve(size_t(4) * rows * cols);
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) {
      const uint32_t v = uint32_t(r) * cols + c;
      if (c + 1 < cols) keys.push_back(key_of(v, v + 1));
      if (r + 1 < rows) keys.push_back(key_of(v, v + cols));
    }
  return finish(keys.data(), int64_t(keys.size()), seed, 1, wmin, wmax);
}

spmh_graph* spmh_from_edges(int64_t n, const int32_t* src, const int32_t* dst, const int32_t* w) {
  struct KW { uint64_t k; int32_t w; };
  std::vector<KW> kw(n);
  int32_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (src[i] < 0 || dst[i] < 0) return nullptr;
    kw[i].k = key_of((uint32_t)src[i], (uint32_t)dst[i]);
    kw[i].w = w ? w[i] : 0;
    m = std::max(m, std::max(src[i], dst[i]));
  }
  std::stable_sort(kw.begin(), kw.end(), [](const KW& x, const KW& y) { return x.k < y.k; });
  std::vector<uint64_t> keys;
  std::vector<int32_t> ws;
  for (int64_t i = 0; i < n; ++i) {
    if (!keys.empty() && keys.back() == kw[i].k) ws.back() += kw[i].w;  // mpp[p] += w
    else { keys.push_back(kw[i].k); ws.push_back(kw[i].w); }
  }
  return build_csr(keys, ws.data(), m + 1);
}

spmh_graph* spmh_parse_text(const char* path) {
  std::ifstream in(path);
  if (!in.is_open()) return nullptr;
  std::vector<int32_t> s, d, w;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] < '0' || line[0] > '9') continue;
    // strtol mirrors "ss >> x": leading whitespace skipped, a missing field
    // leaves the value 0.
    const char* p = line.c_str();
    char* end;
    long a = std::strtol(p, &end, 10); p = end;
    long bb = std::strtol(p, &end, 10);
    const bool have_b = end != p; p = end;
    long ww = std::strtol(p, &end, 10);
    const bool have_w = end != p;
    s.push_back((int32_t)a);
    d.push_back(have_b ? (int32_t)bb : 0);
    w.push_back(have_w ? (int32_t)ww : 0);
  }
  return spmh_from_edges((int64_t)s.size(), s.data(), d.data(), w.data());
}

spmh_graph* spmh_union(const spmh_graph* const* gs, int32_t n, int64_t* bases) {
  if (n < 1 || !gs) return nullptr;
  int64_t V = 0, E = 0;
  std::vector<int64_t> vb(n + 1), eb(n + 1);
  for (int32_t i = 0; i < n; ++i) {
    vb[i] = V;
    eb[i] = E;
    V += gs[i]->V;
    E += gs[i]->E;
  }
  vb[n] = V;
  eb[n] = E;
  if (V > 0x7fffffff || E > 0x7fffffff) return nullptr;
  spmh_graph* g = static_cast<spmh_graph*>(std::calloc(1, sizeof(spmh_graph)));
  g->V = int32_t(V);
  g->E = E;
  g->indexofNodes = alloc_i32(V + 1);
  g->edgeList = alloc_i32(E);
  g->edgeLen = alloc_i32(E);
  g->rev_indexofNodes = alloc_i32(V + 1);
  g->srcList = alloc_i32(E);
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t i = 0; i < n; ++i) {
    const spmh_graph* s = gs[i];
    const int32_t vo = int32_t(vb[i]), eo = int32_t(eb[i]);
    for (int32_t v = 0; v < s->V; ++v) {
      g->indexofNodes[vo + v] = s->indexofNodes[v] + eo;
      g->rev_indexofNodes[vo + v] = s->rev_indexofNodes[v] + eo;
    }
    for (int64_t e = 0; e < s->E; ++e) {
      g->edgeList[eo + e] = s->edgeList[e] + vo;
      g->srcList[eo + e] = s->srcList[e] + vo;
      g->edgeLen[eo + e] = s->edgeLen[e];
    }
  }
  g->indexofNodes[V] = int32_t(E);
  g->rev_indexofNodes[V] = int32_t(E);
  if (bases) std::copy(vb.begin(), vb.end(), bases);
  return g;
}

int spmh_write_text(const spmh_graph* g, const char* path, int weights) {
  FILE* f = std::fopen(path, "w");
  if (!f) return -1;
  for (int32_t v = 0; v < g->V; ++v)
    for (int32_t e = g->indexofNodes[v]; e < g->indexofNodes[v + 1]; ++e) {
      if (weights) std::fprintf(f, "%d %d %d\n", v, g->edgeList[e], g->edgeLen[e]);
      else std::fprintf(f, "%d %d\n", v, g->edgeList[e]);
    }
  return std::fclose(f) == 0 ? 0 : -1;
}

namespace {
struct BinHeader {
  char magic[8];
  int64_t V;
  int64_t E;
  int64_t flags;
  uint64_t checksum;  // FNV-1a over the five arrays
  int64_t pad[3];
};
static_assert(sizeof(BinHeader) == 64, "header is 64 bytes");

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001B3ull;
  return h;
}

uint64_t csr_checksum(const spmh_graph* g) {
  uint64_t h = 0xCBF29CE484222325ull;
  const size_t vb = sizeof(int32_t) * (size_t(g->V) + 1), eb = sizeof(int32_t) * size_t(g->E);
  h = fnv1a(h, g->indexofNodes, vb);
  h = fnv1a(h, g->edgeList, eb);
  h = fnv1a(h, g->edgeLen, eb);
  h = fnv1a(h, g->rev_indexofNodes, vb);
  return fnv1a(h, g->srcList, eb);
}
}  // namespace

int spmh_write_binary(const spmh_graph* g, const char* path) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return -1;
  BinHeader h{};
  std::memcpy(h.magic, "SPMCSR01", 8);
  h.V = g->V;
  h.E = g->E;
  h.checksum = csr_checksum(g);
  const size_t vb = size_t(g->V) + 1, eb = size_t(g->E);
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  ok = ok && std::fwrite(g->indexofNodes, sizeof(int32_t), vb, f) == vb;
  ok = ok && std::fwrite(g->edgeList, sizeof(int32_t), eb, f) == eb;
  ok = ok && std::fwrite(g->edgeLen, sizeof(int32_t), eb, f) == eb;
  ok = ok && std::fwrite(g->rev_indexofNodes, sizeof(int32_t), vb, f) == vb;
  ok = ok && std::fwrite(g->srcList, sizeof(int32_t), eb, f) == eb;
  return (std::fclose(f) == 0 && ok) ? 0 : -1;
}

spmh_graph* spmh_read_binary(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return nullptr;
  BinHeader h{};
  if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, "SPMCSR01", 8) != 0 || h.V < 1 ||
      h.V > 0x7fffffff || h.E < 0 || h.E > 0x7fffffff) {
    std::fclose(f);
    return nullptr;
  }
  spmh_graph* g = static_cast<spmh_graph*>(std::calloc(1, sizeof(spmh_graph)));
  g->V = int32_t(h.V);
  g->E = h.E;
  const size_t vb = size_t(h.V) + 1, eb = size_t(h.E);
  g->indexofNodes = alloc_i32(int64_t(vb));
  g->edgeList = alloc_i32(h.E);
  g->edgeLen = alloc_i32(h.E);
  g->rev_indexofNodes = alloc_i32(int64_t(vb));
  g->srcList = alloc_i32(h.E);
  bool ok = std::fread(g->indexofNodes, sizeof(int32_t), vb, f) == vb;
  ok = ok && std::fread(g->edgeList, sizeof(int32_t), eb, f) == eb;
  ok = ok && std::fread(g->edgeLen, sizeof(int32_t), eb, f) == eb;
  ok = ok && std::fread(g->rev_indexofNodes, sizeof(int32_t), vb, f) == vb;
  ok = ok && std::fread(g->srcList, sizeof(int32_t), eb, f) == eb;
  std::fclose(f);
  if (!ok || csr_checksum(g) != h.checksum || g->indexofNodes[0] != 0 || g->indexofNodes[g->V] != g->E) {
    spmh_free(g);
    return nullptr;
  }
  return g;
}

void spmh_free(spmh_graph* g) {
  if (!g) return;
  std::free(g->indexofNodes);
  std::free(g->edgeList);
  std::free(g->edgeLen);
  std::free(g->rev_indexofNodes);
  std::free(g->srcList);
  std::free(g);
}

int spmh_sources(int32_t V, int32_t n, uint64_t seed, int32_t* out) {
  if (n < 0 || n > V) return -1;
  std::vector<int32_t> picked;
  std::vector<uint8_t> used(V, 0);
  for (uint64_t i = 0; (int32_t)picked.size() < n; ++i) {
    const int32_t v = int32_t((unsigned __int128)rng(seed, 0x5eed, i) * (uint64_t)V >> 64);
    if (!used[v]) { used[v] = 1; picked.push_back(v); }
  }
  std::sort(picked.begin(), picked.end());
  std::copy(picked.begin(), picked.end(), out);
  return 0;
}

int spmh_edge_map(int32_t V, const int32_t* off, const int32_t* adj, const int32_t* roff, const int32_t* radj,
                  int32_t* emap) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(| : bad)
  for (int32_t u = 0; u < V; ++u) {
    for (int32_t j = off[u]; j < off[u + 1]; ++j) {
      const int32_t v = adj[j];
      int32_t k = 0;  // copies of (u, v) before position j
      while (j - 1 - k >= off[u] && adj[j - 1 - k] == v) ++k;
      const int32_t* lo = std::lower_bound(radj + roff[v], radj + roff[v + 1], u);
      const int64_t p = (lo - radj) + k;
      if (p >= roff[v + 1] || radj[p] != u) {
        bad = 1;
        continue;
      }
      emap[p] = j;
    }
  }
  return bad ? -1 : 0;
}

void spmh_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

int spmh_features(const spmh_graph* g, float job, float* out) {
  const int32_t V = g->V;
  const double E = double(g->E);
  const int32_t* off = g->indexofNodes;
  const int32_t* roff = g->rev_indexofNodes;
  double s1 = 0, s2 = 0, s3 = 0;
  int64_t maxo = 0, maxi = 0, zero_out = 0, low_in = 0, hub_edges = 0;
#pragma omp parallel for reduction(+ : s1, s2, zero_out, low_in, hub_edges) reduction(max : maxo, maxi) schedule(static)
  for (int32_t v = 0; v < V; ++v) {
    const int64_t d = off[v + 1] - off[v], di = roff[v + 1] - roff[v];
    s1 += double(d);
    s2 += double(d) * double(d);
    maxo = std::max(maxo, d);
    maxi = std::max(maxi, di);
    zero_out += d == 0;
    low_in += di <= 4;
    if (di > 64) hub_edges += di;
  }
  const double mean = s1 / V, var = std::max(0.0, s2 / V - mean * mean);
#pragma omp parallel for reduction(+ : s3) schedule(static)
  for (int32_t v = 0; v < V; ++v) {
    const double x = double(off[v + 1] - off[v]) - mean;
    s3 += x * x * x;
  }
  const double skew = var > 0 ? (s3 / V) / std::pow(var, 1.5) : 0.0;
  // Double-sweep BFS over out-edges: from the highest-degree vertex, then
  // from the deepest vertex of that sweep (the smallest id of its last level,
  // so the estimate does not depend on the thread schedule). Level-synchronous;
  // levels with many vertices are expanded by all threads (CAS on the level).
  auto bfs = [&](int32_t root, int32_t* far, int64_t* reached, int64_t* widest) {
    std::vector<int32_t> level(size_t(V), -1), cur{root}, nxt;
    level[size_t(root)] = 0;
    int32_t depth = 0;
    int64_t seen = 1, wide = 1;
    int32_t last = root;
    const int T = omp_get_max_threads();
    std::vector<std::vector<int32_t>> loc(static_cast<size_t>(T));
    for (;;) {
      nxt.clear();
      const int32_t nl = depth + 1;
      if (cur.size() < 1024) {
        for (const int32_t v : cur)
          for (int32_t e = off[v]; e < off[v + 1]; ++e) {
            const int32_t w = g->edgeList[e];
            if (level[size_t(w)] < 0) {
              level[size_t(w)] = nl;
              nxt.push_back(w);
            }
          }
      } else {
#pragma omp parallel num_threads(T)
        {
          std::vector<int32_t>& L = loc[size_t(omp_get_thread_num())];
          L.clear();
#pragma omp for schedule(dynamic, 64)
          for (int64_t i = 0; i < int64_t(cur.size()); ++i) {
            const int32_t v = cur[size_t(i)];
            for (int32_t e = off[v]; e < off[v + 1]; ++e) {
              const int32_t w = g->edgeList[e];
              int32_t expect = -1;
              if (__atomic_load_n(&level[size_t(w)], __ATOMIC_RELAXED) < 0 &&
                  __atomic_compare_exchange_n(&level[size_t(w)], &expect, nl, false, __ATOMIC_RELAXED,
                                              __ATOMIC_RELAXED))
                L.push_back(w);
            }
          }
        }
        for (const auto& L : loc) nxt.insert(nxt.end(), L.begin(), L.end());
      }
      if (nxt.empty()) break;
      depth = nl;
      seen += int64_t(nxt.size());
      wide = std::max<int64_t>(wide, int64_t(nxt.size()));
      last = *std::min_element(nxt.begin(), nxt.end());
      cur.swap(nxt);
    }
    *far = last;
    *reached = seen;
    *widest = wide;
    return depth;
  };
  int32_t root = 0;
  for (int32_t v = 1; v < V; ++v)
    if (off[v + 1] - off[v] > off[root + 1] - off[root]) root = v;
  int32_t far = root;
  int64_t reached = 1, widest = 1, r2 = 1, w2 = 1;
  int32_t d1 = bfs(root, &far, &reached, &widest);
  int32_t far2 = far;
  int32_t d2 = bfs(far, &far2, &r2, &w2);
  const int32_t diam = std::max(d1, d2);
  // Reciprocated edges: (u, w) with (w, u) also present, i.e. w in both the
  // out-row and the in-row of u; both rows are sorted, so one merge per vertex.
  int64_t recip = 0;
#pragma omp parallel for reduction(+ : recip) schedule(dynamic, 1024)
  for (int32_t u = 0; u < V; ++u) {
    int32_t j = roff[u];
    const int32_t jend = roff[u + 1];
    for (int32_t e = off[u]; e < off[u + 1]; ++e) {
      const int32_t w = g->edgeList[e];
      while (j < jend && g->srcList[j] < w) ++j;
      recip += (j < jend && g->srcList[j] == w) ? 1 : 0;
    }
  }
  out[0] = float(std::log2(double(V)));
  out[1] = float(std::log2(E + 1));
  out[2] = float(mean);
  out[3] = float(std::log2(1 + var));
  out[4] = float(std::log2(1.0 + double(maxo)));
  out[5] = float(std::log2(1.0 + double(maxi)));
  out[6] = float((skew >= 0 ? 1 : -1) * std::log1p(std::fabs(skew)));
  out[7] = float(double(zero_out) / V);
  out[8] = float(double(low_in) / V);
  out[9] = E > 0 ? float(double(hub_edges) / E) : 0.f;
  out[10] = float(std::log2(1.0 + diam));
  out[11] = float(double(reached) / V);
  out[12] = float(std::log2(1.0 + 1000.0 * (d1 + 1) / double(reached)));
  out[13] = float(double(widest) / V);
  out[14] = job;
  out[15] = E > 0 ? float(double(recip) / E) : 0.f;
  return 0;
}

double spmh_model_pr_bytes(const spmh_graph* g) {
  return 8.0 * double(g->E) + 20.0 * double(g->V) + 4.0;
}

double spmh_model_sssp(const spmh_graph* g, const int32_t* weight, int32_t src,
                       int64_t* rounds_out, int64_t* m_reach_out, int64_t* relax_out) {
  // Synchronous (Jacobi) frontier trace: round k relaxes the out-edges of
  // F_k with the distances of round k-1; F_{k+1} is the set whose distance
  // dropped.
  const int32_t V = g->V;
  if (src < 0 || src >= V) return -1.0;
  std::vector<int32_t> dist(V, INT32_MAX), nxt(V, INT32_MAX);
  std::vector<uint8_t> inq(V, 0);
  std::vector<int32_t> front{src};
  dist[src] = nxt[src] = 0;
  double bytes = 0.0;
  int64_t rounds = 0, relax = 0;
  while (!front.empty()) {
    int64_t mk = 0;
#pragma omp parallel for reduction(+ : mk) schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)front.size(); ++i) {
      const int32_t v = front[i];
      const int32_t dv = dist[v];
      for (int32_t e = g->indexofNodes[v]; e < g->indexofNodes[v + 1]; ++e) {
        const int32_t n = g->edgeList[e];
        const int32_t nd = dv + weight[e];
        int32_t cur = __atomic_load_n(&nxt[n], __ATOMIC_RELAXED);
        while (nd < cur && !__atomic_compare_exchange_n(&nxt[n], &cur, nd, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
      }
      mk += g->indexofNodes[v + 1] - g->indexofNodes[v];
    }
    std::vector<int32_t> next;
#pragma omp parallel
    {
      std::vector<int32_t> loc;
#pragma omp for schedule(static) nowait
      for (int32_t v = 0; v < V; ++v)
        if (nxt[v] < dist[v]) loc.push_back(v);
#pragma omp critical
      next.insert(next.end(), loc.begin(), loc.end());
    }
    for (int32_t v : next) dist[v] = nxt[v];
    bytes += 12.0 * double(front.size()) + 12.0 * double(mk) + 4.0 * double(next.size());
    relax += mk;
    ++rounds;
    front.swap(next);
  }
  int64_t mr = 0;
#pragma omp parallel for reduction(+ : mr) schedule(static)
  for (int32_t v = 0; v < V; ++v)
    if (dist[v] != INT32_MAX) mr += g->indexofNodes[v + 1] - g->indexofNodes[v];
  if (rounds_out) *rounds_out = rounds;
  if (m_reach_out) *m_reach_out = mr;
  if (relax_out) *relax_out = relax;
  return bytes;
}

double spmh_model_tc_bytes(const spmh_graph* g, int64_t* m_und_out, int64_t* work_out) {
  const int32_t V = g->V;
  std::vector<int32_t> dplus(V, 0);
  auto deg = [&](int32_t v) { return g->indexofNodes[v + 1] - g->indexofNodes[v]; };
  auto above = [&](int32_t u, int32_t w) {
    const int32_t du = deg(u), dw = deg(w);
    return dw > du || (dw == du && w > u);
  };
  int64_t m = 0;
#pragma omp parallel for reduction(+ : m) schedule(dynamic, 1024)
  for (int32_t u = 0; u < V; ++u) {
    int32_t c = 0;
    for (int32_t e = g->indexofNodes[u]; e < g->indexofNodes[u + 1]; ++e)
      c += above(u, g->edgeList[e]) ? 1 : 0;
    dplus[u] = c;
    m += c;
  }
  int64_t work = 0;
#pragma omp parallel for reduction(+ : work) schedule(dynamic, 1024)
  for (int32_t u = 0; u < V; ++u)
    for (int32_t e = g->indexofNodes[u]; e < g->indexofNodes[u + 1]; ++e) {
      const int32_t w = g->edgeList[e];
      if (above(u, w)) work += int64_t(dplus[u]) + dplus[w];
    }
  if (m_und_out) *m_und_out = m;
  if (work_out) *work_out = work;
  return 8.0 * double(V + 1) + 4.0 * double(m) + 4.0 * double(work);
}

double spmh_model_bc_bytes(const spmh_graph* g, const int32_t* sources, int32_t n,
                           int64_t* edges_out) {
  const int32_t V = g->V;
  double total = 0.0;
  int64_t edges = 0;
#pragma omp parallel reduction(+ : total, edges)
  {
    std::vector<int32_t> level(V, -1), q;
    q.reserve(V);
#pragma omp for schedule(dynamic, 1)
    for (int32_t i = 0; i < n; ++i) {
      std::fill(level.begin(), level.end(), -1);
      q.clear();
      const int32_t s = sources[i];
      level[s] = 0;
      q.push_back(s);
      int64_t Er = 0, D = 0;
      for (size_t h = 0; h < q.size(); ++h) {
        const int32_t v = q[h];
        for (int32_t e = g->indexofNodes[v]; e < g->indexofNodes[v + 1]; ++e) {
          const int32_t w = g->edgeList[e];
          if (level[w] < 0) { level[w] = level[v] + 1; q.push_back(w); }
          if (level[w] == level[v] + 1) ++D;
        }
        Er += g->indexofNodes[v + 1] - g->indexofNodes[v];
      }
      total += 32.0 * double(q.size()) + 16.0 * double(Er) + 20.0 * double(D);
      edges += Er;
    }
  }
  if (edges_out) *edges_out = edges;
  return total;
}

int spmh_bc_gather_profile(const spmh_graph* g, const int32_t* sources, int32_t n, double* out) {
  if (!g || !sources || !out || n < 1 || n > 64) return -1;
  const int32_t V = g->V;
  const int nb = n;
  // lev[v * nb + s]: BFS level of v from source s (-1 unreached).
  std::vector<int16_t> lev(size_t(V) * nb, int16_t(-1));
  int32_t depth = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : depth)
  for (int s = 0; s < nb; ++s) {
    std::vector<int32_t> q;
    q.reserve(V);
    const int32_t src = sources[s];
    lev[size_t(src) * nb + s] = 0;
    q.push_back(src);
    for (size_t h = 0; h < q.size(); ++h) {
      const int32_t v = q[h];
      const int16_t lv = lev[size_t(v) * nb + s];
      depth = std::max(depth, int32_t(lv));
      for (int32_t e = g->indexofNodes[v]; e < g->indexofNodes[v + 1]; ++e) {
        const int32_t w = g->edgeList[e];
        if (lev[size_t(w) * nb + s] < 0) {
          lev[size_t(w) * nb + s] = int16_t(lv + 1);
          q.push_back(w);
        }
      }
    }
  }
  double dag = 0, gathers = 0, lines64 = 0, lines32 = 0, listed = 0, row_reads = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : dag, gathers, lines64, lines32, listed, row_reads)
  for (int32_t w = 0; w < V; ++w) {
    const int16_t* lw = &lev[size_t(w) * nb];
    // Levels >= 1 at which w is listed (some source reaches w there).
    int16_t lv[64];
    int nl = 0;
    for (int s = 0; s < nb; ++s) {
      if (lw[s] < 1) continue;
      bool seen = false;
      for (int k = 0; k < nl && !seen; ++k) seen = lv[k] == lw[s];
      if (!seen) lv[nl++] = lw[s];
    }
    const int32_t indeg = g->rev_indexofNodes[w + 1] - g->rev_indexofNodes[w];
    listed += nl;
    row_reads += double(nl) * indeg;
    for (int32_t e = g->rev_indexofNodes[w]; e < g->rev_indexofNodes[w + 1]; ++e) {
      const int16_t* lu = &lev[size_t(g->srcList[e]) * nb];
      for (int k = 0; k < nl; ++k) {
        uint64_t m = 0;
        for (int s = 0; s < nb; ++s)
          if (lw[s] == lv[k] && lu[s] == lv[k] - 1) m |= 1ull << s;
        if (!m) continue;
        gathers += 1;
        dag += __builtin_popcountll(m);
        for (int l = 0; l < 4; ++l) lines64 += ((m >> (16 * l)) & 0xffffull) != 0;
        for (int l = 0; l < 2; ++l) lines32 += ((m >> (32 * l)) & 0xffffffffull) != 0;
      }
    }
  }
  out[0] = dag;
  out[1] = gathers;
  out[2] = lines64;
  out[3] = lines32;
  out[4] = listed;
  out[5] = row_reads;
  out[6] = double(depth) + 1.0;
  return 0;
}

}  // extern "C"
