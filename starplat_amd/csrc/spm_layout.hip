// spm_layout.hip -- CSR layouts built on the device: sorted CSR (and reverse
// CSR) from (row << 32 | column, weight) pairs, used by the fold of edge
// updates; and the degree-ordered relabeled copy of the CSR that PageRank and
// SSSP run on when the schedule asks for it.
//
// Relabeling. Vertices are renumbered by descending out-degree (ties by
// original id). On power-law graphs the few vertices that most in-rows and
// frontiers touch then share cache lines, so the random 4-byte gathers of
// PageRank's contribution array and SSSP's distance array hit L2 more often
// (RMAT-22: PageRank 48.2 -> 41.9 ms per 100 iterations, SSSP 1.85 -> 1.69 ms).
// The copy is built once per handle (and again after edge updates or new
// weights); sources are mapped in and results gathered back, so callers only
// ever see original ids. The in-row summation order of PageRank follows the
// new ids (same sums within round-off).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "spm_internal.hpp"

namespace spm {

constexpr int kLayBlock = 256;
constexpr int kLayWaves = kLayBlock / kWave;

struct Relabel {
  int32_t* perm = nullptr;  // [V] original id -> new id
  int32_t* eidx = nullptr;  // [E] original CSR position of each new position
  int32_t *off = nullptr, *adj = nullptr, *w = nullptr, *roff = nullptr, *radj = nullptr;
  int64_t E = 0;
};

static int lay_grid(const spm_graph* g, int64_t n) {
  return int(std::max<int64_t>(1, std::min<int64_t>((n + kLayBlock - 1) / kLayBlock, int64_t(g->num_cus) * 16)));
}

// adj[i] = low half of keys[i]; w[i] = vals[i]; and the swapped pairs.
__global__ void lay_unpack(int64_t n, const unsigned long long* __restrict__ keys, const int32_t* __restrict__ vals,
                           int32_t* __restrict__ adj, int32_t* __restrict__ w, unsigned long long* __restrict__ rkeys,
                           int32_t* __restrict__ rvals) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const unsigned long long k = keys[i];
    const uint32_t r = uint32_t(k >> 32), c = uint32_t(k);
    adj[i] = int32_t(c);
    if (w) w[i] = vals[i];
    if (rkeys) {
      rkeys[i] = (static_cast<unsigned long long>(c) << 32) | r;
      rvals[i] = vals[i];
    }
  }
}

// Row offsets of sorted keys (row in the high half): off[x] = first i whose
// row is >= x, for x in [0, V] (off[V] = n), one binary search per row. (Writing
// the offsets from the row boundaries instead leaves a run of empty rows to
// one thread: the ~2M rows without out-edges that a degree-ordered relabel
// puts last took 70 ms that way at RMAT-22.)
__global__ void lay_segoff(int64_t n, const unsigned long long* __restrict__ keys, int32_t V,
                           int32_t* __restrict__ off) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t x = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; x <= V; x += stride) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (int64_t(keys[mid] >> 32) < x) lo = mid + 1;
      else hi = mid;
    }
    off[x] = int32_t(lo);
  }
}

__global__ void lay_fill(int64_t n, int32_t* __restrict__ a, int32_t v) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = v;
}

int csr_from_pairs(spm_graph* g, int64_t n, unsigned long long* k0, unsigned long long* k1, int32_t* v0, int32_t* v1,
                   int32_t* off, int32_t* adj, int32_t* w, int32_t* roff, int32_t* radj, int32_t* rw) {
  const int32_t V = g->V;
  if (n == 0) {
    SPM_TRY(launch(g, "lay_fill", lay_fill, dim3(lay_grid(g, int64_t(V) + 1)), dim3(kLayBlock), 0, int64_t(V) + 1,
                   off, 0));
    if (roff)
      SPM_TRY(launch(g, "lay_fill", lay_fill, dim3(lay_grid(g, int64_t(V) + 1)), dim3(kLayBlock), 0, int64_t(V) + 1,
                     roff, 0));
    return SPM_OK;
  }
  int bits = 1;
  while ((int64_t(1) << bits) <= int64_t(V)) ++bits;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = SPM_OK;
  auto sort = [&](hipcub::DoubleBuffer<unsigned long long>& kb, hipcub::DoubleBuffer<int32_t>& vb) -> int {
    size_t need = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, kb, vb, int(n), 0, 32 + bits, g->stream);
    if (e != hipSuccess) return set_error(SPM_EHIP, std::string("csr sort: ") + hipGetErrorString(e));
    if (need > tmp_bytes) {
      dfree(g, tmp, tmp_bytes);
      tmp = nullptr;
      tmp_bytes = 0;
      SPM_TRY(dmalloc(g, &tmp, need, "csr sort scratch"));
      tmp_bytes = need;
    }
    e = hipcub::DeviceRadixSort::SortPairs(tmp, need, kb, vb, int(n), 0, 32 + bits, g->stream);
    if (e != hipSuccess) return set_error(SPM_EHIP, std::string("csr sort: ") + hipGetErrorString(e));
    return SPM_OK;
  };
  hipcub::DoubleBuffer<unsigned long long> kb(k0, k1);
  hipcub::DoubleBuffer<int32_t> vb(v0, v1);
  rc = sort(kb, vb);
  unsigned long long* rk = kb.Current() == k0 ? k1 : k0;
  int32_t* rv = vb.Current() == v0 ? v1 : v0;
  if (rc == SPM_OK)
    rc = launch(g, "lay_unpack", lay_unpack, dim3(lay_grid(g, n)), dim3(kLayBlock), 0, n,
                (const unsigned long long*)kb.Current(), (const int32_t*)vb.Current(), adj, w,
                roff ? rk : (unsigned long long*)nullptr, roff ? rv : (int32_t*)nullptr);
  if (rc == SPM_OK)
    rc = launch(g, "lay_segoff", lay_segoff, dim3(lay_grid(g, int64_t(V) + 1)), dim3(kLayBlock), 0, n,
                (const unsigned long long*)kb.Current(), V, off);
  if (rc == SPM_OK && roff) {
    hipcub::DoubleBuffer<unsigned long long> rkb(rk, kb.Current());
    hipcub::DoubleBuffer<int32_t> rvb(rv, vb.Current());
    rc = sort(rkb, rvb);
    if (rc == SPM_OK)
      rc = launch(g, "lay_unpack", lay_unpack, dim3(lay_grid(g, n)), dim3(kLayBlock), 0, n,
                  (const unsigned long long*)rkb.Current(), (const int32_t*)rvb.Current(), radj, rw,
                  (unsigned long long*)nullptr, (int32_t*)nullptr);
    if (rc == SPM_OK)
      rc = launch(g, "lay_segoff", lay_segoff, dim3(lay_grid(g, int64_t(V) + 1)), dim3(kLayBlock), 0, n,
                  (const unsigned long long*)rkb.Current(), V, roff);
  }
  if (rc == SPM_OK) {
    hipError_t e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = set_error(SPM_EHIP, std::string("csr build: ") + hipGetErrorString(e));
  }
  dfree(g, tmp, tmp_bytes);
  return rc;
}

// ---- relabeling ---------------------------------------------------------------

// key[v] = (INT_MAX - outdeg(v)) << 32 | v: ascending keys = descending degree, then id.
__global__ void lay_degkeys(int32_t V, const int32_t* __restrict__ off, unsigned long long* __restrict__ keys) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride)
    keys[v] = (static_cast<unsigned long long>(uint32_t(0x7fffffff - (off[v + 1] - off[v]))) << 32) | uint32_t(v);
}

__global__ void lay_perm(int32_t V, const unsigned long long* __restrict__ sorted, int32_t* __restrict__ perm) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < V; i += stride)
    perm[uint32_t(sorted[i])] = int32_t(i);
}

// Pairs (perm[u] << 32 | perm[x], j) of every edge u -> x at CSR position j,
// one wavefront per row.
__global__ __launch_bounds__(kLayBlock) void lay_relabel_pairs(int32_t V, const int32_t* __restrict__ off,
                                                               const int32_t* __restrict__ adj,
                                                               const int32_t* __restrict__ perm,
                                                               unsigned long long* __restrict__ keys,
                                                               int32_t* __restrict__ vals) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * kLayWaves;
  for (int64_t x_ = int64_t(blockIdx.x) * kLayWaves + threadIdx.x / kWave; x_ < V; x_ += nw) {
    const int64_t u = spread_vertex(x_, V);
    const unsigned long long hi = static_cast<unsigned long long>(uint32_t(perm[u])) << 32;
    for (int32_t j = off[u] + lane; j < off[u + 1]; j += kWave) {
      keys[j] = hi | uint32_t(perm[adj[j]]);
      vals[j] = j;
    }
  }
}

// out[i] = in[idx[i]] (4-byte elements), i < n.
__global__ void lay_take(int64_t n, const int32_t* __restrict__ idx, const uint32_t* __restrict__ in,
                         uint32_t* __restrict__ out) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[idx[i]];
}

// out[v] = in[perm[v]] (4-byte elements).
__global__ void lay_gather(int32_t V, const int32_t* __restrict__ perm, const uint32_t* __restrict__ in,
                           uint32_t* __restrict__ out) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) out[v] = in[perm[v]];
}

int relabel_release(spm_graph* g) {
  Relabel* r = g->rl;
  if (!r) return SPM_OK;
  const size_t vb = sizeof(int32_t) * (size_t(g->V) + 1), eb = sizeof(int32_t) * size_t(r->E);
  dfree(g, r->perm, sizeof(int32_t) * size_t(g->V));
  dfree(g, r->eidx, eb);
  dfree(g, r->off, vb);
  dfree(g, r->adj, eb);
  dfree(g, r->w, eb);
  dfree(g, r->roff, vb);
  dfree(g, r->radj, eb);
  delete r;
  g->rl = nullptr;
  // PR and SSSP captured graphs and bins refer to the arrays just freed
  pr_release(g);
  sssp_release(g);
  return SPM_OK;
}

// Weights of the relabeled CSR from the handle's: w'[i] = w[eidx[i]].
int relabel_weights(spm_graph* g) {
  Relabel* r = g->rl;
  if (!r || !g->w) return SPM_OK;
  if (!r->w) SPM_TRY(dmalloc(g, (void**)&r->w, sizeof(int32_t) * size_t(r->E), "relabeled edgeLen"));
  SPM_TRY(launch(g, "lay_take", lay_take, dim3(lay_grid(g, r->E)), dim3(kLayBlock), 0, r->E, (const int32_t*)r->eidx,
                 (const uint32_t*)g->w, (uint32_t*)r->w));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  return SPM_OK;
}

static int relabel_build(spm_graph* g) {
  const int32_t V = g->V;
  const int64_t E = g->E;
  Relabel* r = new Relabel();
  g->rl = r;
  r->E = E;
  const size_t vb = sizeof(int32_t) * (size_t(V) + 1), eb = sizeof(int32_t) * size_t(E);
  SPM_TRY(dmalloc(g, (void**)&r->perm, sizeof(int32_t) * size_t(V), "relabel map"));
  SPM_TRY(dmalloc(g, (void**)&r->off, vb, "relabeled indexofNodes"));
  SPM_TRY(dmalloc(g, (void**)&r->adj, eb, "relabeled edgeList"));
  SPM_TRY(dmalloc(g, (void**)&r->eidx, eb, "relabeled edge positions"));
  if (g->roff) {
    SPM_TRY(dmalloc(g, (void**)&r->roff, vb, "relabeled rev_indexofNodes"));
    SPM_TRY(dmalloc(g, (void**)&r->radj, eb, "relabeled srcList"));
  }
  const size_t n = size_t(std::max<int64_t>(E, V));
  unsigned long long *k0 = nullptr, *k1 = nullptr;
  int32_t *v0 = nullptr, *v1 = nullptr;
  int rc = dmalloc(g, (void**)&k0, sizeof(unsigned long long) * n, "relabel keys");
  if (rc == SPM_OK) rc = dmalloc(g, (void**)&k1, sizeof(unsigned long long) * n, "relabel keys");
  if (rc == SPM_OK) rc = dmalloc(g, (void**)&v0, sizeof(int32_t) * n, "relabel weights");
  if (rc == SPM_OK) rc = dmalloc(g, (void**)&v1, sizeof(int32_t) * n, "relabel weights");
  // vertex order
  if (rc == SPM_OK)
    rc = launch(g, "lay_degkeys", lay_degkeys, dim3(lay_grid(g, V)), dim3(kLayBlock), 0, V, (const int32_t*)g->off, k0);
  if (rc == SPM_OK) {
    hipcub::DoubleBuffer<unsigned long long> kb(k0, k1);
    size_t need = 0;
    void* tmp = nullptr;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(nullptr, need, kb, int(V), 0, 64, g->stream);
    if (e == hipSuccess) rc = dmalloc(g, &tmp, need, "relabel sort scratch");
    else rc = set_error(SPM_EHIP, std::string("relabel sort: ") + hipGetErrorString(e));
    if (rc == SPM_OK) {
      e = hipcub::DeviceRadixSort::SortKeys(tmp, need, kb, int(V), 0, 64, g->stream);
      if (e != hipSuccess) rc = set_error(SPM_EHIP, std::string("relabel sort: ") + hipGetErrorString(e));
    }
    if (rc == SPM_OK)
      rc = launch(g, "lay_perm", lay_perm, dim3(lay_grid(g, V)), dim3(kLayBlock), 0, V,
                  (const unsigned long long*)kb.Current(), r->perm);
    if (rc == SPM_OK) {
      hipError_t e2 = hipStreamSynchronize(g->stream);
      if (e2 != hipSuccess) rc = set_error(SPM_EHIP, std::string("relabel: ") + hipGetErrorString(e2));
    }
    dfree(g, tmp, need);
  }
  // edges in new ids
  if (rc == SPM_OK && E > 0)
    rc = launch(g, "lay_relabel_pairs", lay_relabel_pairs, dim3(lay_grid(g, int64_t(V) * kWave)), dim3(kLayBlock), 0,
                V, (const int32_t*)g->off, (const int32_t*)g->adj, (const int32_t*)r->perm, k0, v0);
  if (rc == SPM_OK) rc = csr_from_pairs(g, E, k0, k1, v0, v1, r->off, r->adj, r->eidx, r->roff, r->radj, nullptr);
  if (rc == SPM_OK && g->w) rc = relabel_weights(g);
  dfree(g, k0, sizeof(unsigned long long) * n);
  dfree(g, k1, sizeof(unsigned long long) * n);
  dfree(g, v0, sizeof(int32_t) * n);
  dfree(g, v1, sizeof(int32_t) * n);
  if (rc != SPM_OK) {
    relabel_release(g);
    return rc;
  }
  return SPM_OK;
}

int csr_view(spm_graph* g, CsrView* v) {
  if (g->sched.relabel && !g->rl) {
    // PR and SSSP state built for the original ids must not mix with the new ones
    pr_release(g);
    sssp_release(g);
    SPM_TRY(relabel_build(g));
  }
  if (!g->sched.relabel && g->rl) SPM_TRY(relabel_release(g));
  const Relabel* r = g->rl;
  v->off = r ? r->off : g->off;
  v->adj = r ? r->adj : g->adj;
  v->w = r ? r->w : g->w;
  v->roff = r ? r->roff : g->roff;
  v->radj = r ? r->radj : g->radj;
  v->perm = r ? r->perm : nullptr;
  return SPM_OK;
}

int gather_back(spm_graph* g, const CsrView& v, const void* in, void* out) {
  if (!v.perm) {
    if (in != out)
      SPM_CHECK_HIP(hipMemcpyAsync(out, in, sizeof(int32_t) * size_t(g->V), hipMemcpyDeviceToDevice, g->stream));
    return SPM_OK;
  }
  return launch(g, "lay_gather", lay_gather, dim3(lay_grid(g, g->V)), dim3(kLayBlock), 0, g->V, v.perm,
                (const uint32_t*)in, (uint32_t*)out);
}

}  // namespace spm
