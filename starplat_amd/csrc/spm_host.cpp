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
#include <parallel/algorithm>
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
inline uint64_t rng(uint64_t seed, uint64_t stream, uint64_t counter) {
  return mix64(mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull)) + counter);
}

inline uint64_t key_of(uint32_t s, uint32_t d) { return (uint64_t(s) << 32) | d; }
inline int32_t key_src(uint64_t k) { return int32_t(k >> 32); }
inline int32_t key_dst(uint64_t k) { return int32_t(k & 0xffffffffu); }

int32_t* alloc_i32(int64_t n) {
  int32_t* p = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (n > 0 ? n : 1)));
  if (!p) { std::fprintf(stderr, "spm_host: out of memory (%lld ints)\n", (long long)n); std::abort(); }
  return p;
}

// Builds the forward and reverse CSR from keys sorted ascending and unique,
// with weights aligned to the keys (or NULL for zero weights).
spmh_graph* build_csr(const std::vector<uint64_t>& keys, const int32_t* w, int32_t V) {
  const int64_t E = (int64_t)keys.size();
  spmh_graph* g = static_cast<spmh_graph*>(std::calloc(1, sizeof(spmh_graph)));
  g->V = V;
  g->E = E;
  g->indexofNodes = alloc_i32(V + 1);
  g->edgeList = alloc_i32(E);
  g->edgeLen = alloc_i32(E);
  g->rev_indexofNodes = alloc_i32(V + 1);
  g->srcList = alloc_i32(E);

  std::vector<std::atomic<int32_t>> outc(V), inc(V);
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < V; ++v) { outc[v].store(0, std::memory_order_relaxed); inc[v].store(0, std::memory_order_relaxed); }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < E; ++i) {
    outc[key_src(keys[i])].fetch_add(1, std::memory_order_relaxed);
    inc[key_dst(keys[i])].fetch_add(1, std::memory_order_relaxed);
    g->edgeList[i] = key_dst(keys[i]);
    g->edgeLen[i] = w ? w[i] : 0;
  }
  int64_t acc = 0, racc = 0;
  for (int32_t v = 0; v < V; ++v) {
    g->indexofNodes[v] = (int32_t)acc;
    g->rev_indexofNodes[v] = (int32_t)racc;
    acc += outc[v].load(std::memory_order_relaxed);
    racc += inc[v].load(std::memory_order_relaxed);
  }
  g->indexofNodes[V] = (int32_t)acc;
  g->rev_indexofNodes[V] = (int32_t)racc;

  // Reverse CSR: scatter sources with per-row cursors, then sort each row
  // ascending (graph.hpp:810-832 sorts each srcList row the same way).
  std::vector<std::atomic<int32_t>>& cur = inc;
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < V; ++v) cur[v].store(g->rev_indexofNodes[v], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < E; ++i) {
    int32_t d = key_dst(keys[i]);
    int32_t pos = cur[d].fetch_add(1, std::memory_order_relaxed);
    g->srcList[pos] = key_src(keys[i]);
  }
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < V; ++v)
    std::sort(g->srcList + g->rev_indexofNodes[v], g->srcList + g->rev_indexofNodes[v + 1]);
  return g;
}

int32_t max_id(const std::vector<uint64_t>& keys) {
  int32_t m = 0;  // graph.hpp starts nodesTotal at 0 (graph.hpp:73)
#pragma omp parallel for reduction(max : m) schedule(static)
  for (int64_t i = 0; i < (int64_t)keys.size(); ++i) {
    m = std::max(m, key_src(keys[i]));
    m = std::max(m, key_dst(keys[i]));
  }
  return m;
}

// Sort, drop duplicates, optionally assign hashed weights, build the CSR.
spmh_graph* finish(std::vector<uint64_t>& keys, uint64_t seed, int undirected, int wmin, int wmax) {
  __gnu_parallel::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const int32_t V = max_id(keys) + 1;
  std::vector<int32_t> w;
  if (wmax > 0) {
    w.resize(keys.size());
    const uint64_t span = uint64_t(wmax - wmin + 1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)keys.size(); ++i) {
      uint32_t s = (uint32_t)key_src(keys[i]), d = (uint32_t)key_dst(keys[i]);
      if (undirected && s > d) std::swap(s, d);
      w[i] = wmin + int32_t(rng(seed, 0x57, key_of(s, d)) % span);
    }
  }
  return build_csr(keys, w.empty() ? nullptr : w.data(), V);
}

void drop_loops_and_symmetrise(std::vector<uint64_t>& keys, int undirected) {
  // Self-loops removed; for undirected graphs each kept pair is mirrored.
  const int64_t n = (int64_t)keys.size();
  std::vector<uint8_t> keep(n);
  int64_t cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    keep[i] = key_src(keys[i]) != key_dst(keys[i]);
    cnt += keep[i];
  }
  std::vector<uint64_t> out;
  out.reserve(undirected ? 2 * cnt : cnt);
  for (int64_t i = 0; i < n; ++i)
    if (keep[i]) out.push_back(keys[i]);
  if (undirected) {
    const int64_t m = (int64_t)out.size();
    out.resize(2 * m);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) out[m + i] = key_of((uint32_t)key_dst(out[i]), (uint32_t)key_src(out[i]));
  }
  keys.swap(out);
}

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
  std::vector<uint64_t> keys(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    uint32_t s = 0, d = 0;
    for (int l = 0; l < scale; l += 2) {
      const uint64_t h = rng(seed, uint64_t(i), uint64_t(l));
      for (int k = 0; k < 2 && l + k < scale; ++k) {
        const uint64_t r = k ? (h >> 32) : (h & 0xffffffffu);
        uint32_t bs, bd;
        if (r < ta) { bs = 0; bd = 0; }
        else if (r < tb) { bs = 0; bd = 1; }
        else if (r < tc) { bs = 1; bd = 0; }
        else { bs = 1; bd = 1; }
        s = (s << 1) | bs;
        d = (d << 1) | bd;
      }
    }
    keys[i] = key_of(s, d);
  }
  drop_loops_and_symmetrise(keys, undirected);
  return finish(keys, seed, undirected, wmin, wmax);
}

spmh_graph* spmh_er(int64_t nv, int64_t m, uint64_t seed, int undirected, int wmin, int wmax) {
  if (nv < 1 || nv > 0x7fffffff || m < 0) return nullptr;
  std::vector<uint64_t> keys(m);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    const uint64_t h1 = rng(seed, uint64_t(i), 1), h2 = rng(seed, uint64_t(i), 2);
    // Multiply-shift maps a 64-bit uniform onto [0, nv).
    const uint32_t s = uint32_t((unsigned __int128)h1 * (uint64_t)nv >> 64);
    const uint32_t d = uint32_t((unsigned __int128)h2 * (uint64_t)nv >> 64);
    keys[i] = key_of(s, d);
  }
  drop_loops_and_symmetrise(keys, undirected);
  return finish(keys, seed, undirected, wmin, wmax);
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
  drop_loops_and_symmetrise(keys, 1);
  spmh_graph* g = finish(keys, seed, 1, wmin, wmax);
  return g;
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
  // from the deepest vertex of that sweep.
  auto bfs = [&](int32_t root, int32_t* far, int64_t* reached, int64_t* widest) {
    std::vector<int32_t> level(V, -1), q;
    q.reserve(V);
    level[root] = 0;
    q.push_back(root);
    int32_t depth = 0;
    std::vector<int64_t> width(1, 1);
    for (size_t h = 0; h < q.size(); ++h) {
      const int32_t v = q[h];
      for (int32_t e = off[v]; e < off[v + 1]; ++e) {
        const int32_t w = g->edgeList[e];
        if (level[w] < 0) {
          level[w] = level[v] + 1;
          q.push_back(w);
          if (level[w] > depth) { depth = level[w]; width.push_back(0); }
          ++width[level[w]];
        }
      }
    }
    *far = q.back();
    *reached = int64_t(q.size());
    *widest = *std::max_element(width.begin(), width.end());
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
  int64_t recip = 0;
#pragma omp parallel for reduction(+ : recip) schedule(dynamic, 1024)
  for (int32_t u = 0; u < V; ++u)
    for (int32_t e = off[u]; e < off[u + 1]; ++e) {
      const int32_t w = g->edgeList[e];
      recip += std::binary_search(g->edgeList + off[w], g->edgeList + off[w + 1], u) ? 1 : 0;
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

}  // extern "C"
