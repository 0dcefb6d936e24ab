// spm_dyn.hip -- batched edge updates of the device graph and dynamic SSSP
// for gfx950: the replacement of graph::updateCSRDel / graph::updateCSRAdd
// (graphcode/graph.hpp:869-975, over graph_ompv2.hpp:13-454) and of the code
// StarPlat generates for the Dynamic / Decremental / Incremental functions of
// graphcode/dynamicDSLCodes/dynamicBatchSSSP, dynamicBatchTCV2 and
// dynamicBatchPR (batch loop, OnDelete, OnAdd:
// backend_omp/dsl_dyn_cpp_generator.cpp:44-160, update calls :301-316).
//
// Layout. The handle's CSR and reverse CSR stay where they are and stay
// sorted. A deletion sets a bit in a removed-slot bitset per direction (one
// bit per CSR position, found by binary search in the sorted row) and, for
// entries of the overflow rows, in their own bitset. An addition goes to the
// per-vertex overflow ("diff") rows of each direction; an add batch rebuilds
// those rows in O(V + batch): live old entries are copied and the new ones
// placed behind them. Neighbour iteration is therefore "row minus removed
// slots, then overflow row minus removed entries", as graph::getNeighbors
// (graph.hpp:982-1023) iterates the CSR and then diff_edgeList. Unlike
// graph_ompv2.hpp:378-400, additions never refill removed CSR slots, so rows
// stay sorted and deletions stay binary searches; the edge multiset is the
// same. A deletion removes every copy of the edge (the reference removes the
// CSR copies, and the diff copies only when the CSR held none, which depends
// on its slot layout).
//
// The static kernels assume a clean CSR, so any static algorithm on an
// updated handle first folds the updates into a fresh sorted CSR
// (dyn_settle: live entries of both parts per row, a device radix sort by
// (vertex, neighbour), the reverse CSR by a second sort).
//
// Dynamic SSSP keeps (distance, parent) packed into one 64-bit word per
// vertex, distance in the high half, so a 64-bit atomicMin settles both at
// once and the parent chosen among equally short predecessors is the
// smallest id -- a deterministic rule (the reference's parent depends on the
// thread schedule). Distances >= INT_MAX/2 are infinite: the reference's
// generated code adds weights to them (INT_MAX + w overflows), here they are
// never relaxed from.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "spm_internal.hpp"

namespace spm {

constexpr int kDynBlock = 256;
constexpr int kDynWaves = kDynBlock / kWave;
constexpr int32_t kInf2 = INT_MAX / 2;             // the reference's "removed" distance
constexpr unsigned long long kNoParent = 0xFFFFFFFFull;

__host__ __device__ __forceinline__ unsigned long long dpack(int32_t d, unsigned long long p) {
  return (static_cast<unsigned long long>(static_cast<uint32_t>(d)) << 32) | p;
}
__device__ __forceinline__ int32_t ddist(unsigned long long k) { return int32_t(k >> 32); }

__device__ __forceinline__ bool bit_of(const uint32_t* __restrict__ bits, int64_t j) {
  return (bits[j >> 5] >> (j & 31)) & 1u;
}

// Overflow rows of one direction.
struct DiffRows {
  int32_t* off = nullptr;   // [V+1]
  int32_t* adj = nullptr;   // [cap]
  int32_t* w = nullptr;     // [cap]
  uint32_t* dead = nullptr; // [cap/32 + 1] removed entries
};

struct DiffStore {
  DiffRows r;
  size_t cap = 0;   // entries allocated
  int64_t n = 0;    // bound on the entries in use (off[V] <= n; not read back)
};

struct DynWork {
  uint32_t* dead = nullptr;     // [E/32+1] removed CSR slots
  uint32_t* rdead = nullptr;    // [E/32+1] removed reverse-CSR slots
  int32_t* dead_row = nullptr;  // [V] removed CSR slots per vertex
  int32_t* rdead_row = nullptr; // [V]
  size_t bits_words = 0;
  int32_t* rw = nullptr;        // [E] weight of each reverse-CSR slot
  bool rw_ok = false;
  DiffStore out[2], in[2];      // current and rebuild target
  int cur = 0;
  int64_t removed = 0, rremoved = 0;  // CSR slots removed since the last fold
  int32_t* ubuf = nullptr;      // staged batch: src, dst, weight
  size_t ucap = 0;
  int32_t* h_stage = nullptr;   // pinned host copy of the staged batch
  hipEvent_t stage_done = nullptr;  // its copies to ubuf have completed
  bool stage_pending = false;
  int32_t* cnt = nullptr;       // [V+1] per-vertex counts / lengths
  int32_t* cursor = nullptr;    // [V]
  int32_t* live = nullptr;      // [V]
  unsigned long long* ctr = nullptr;  // [8]
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  // dynamic SSSP
  unsigned long long* dp = nullptr;  // [V] (distance << 32 | parent)
  uint8_t* mod = nullptr;            // [V] Decremental's modified
  uint8_t* fl[2] = {nullptr, nullptr};  // [V] Incremental frontier flags
  int32_t* list = nullptr;           // [V]
  int32_t* nsel = nullptr;           // [1]
  int32_t src = -1;
  // dynamic TC
  unsigned long long* tset = nullptr;  // [tset_cap] the phase's (s << 32 | t) pairs, open addressing
  size_t tset_cap = 0;
  int32_t* tchunk = nullptr;           // [tchunk_cap] work-item prefix over the phase's updates
  size_t tchunk_cap = 0;
  unsigned long long* tcv = nullptr;   // [4] class counts 1..3 of the phase, triangle count
  bool tc_ready = false;
  // dynamic PageRank
  float* pr = nullptr;                 // [V] ranks
  float* pr_nxt = nullptr;             // [V]
  float* pr_c = nullptr;               // [V] rank / base out-degree
  float* pr_deg = nullptr;             // [V] out-degree in the CSR the run started from
  uint8_t* pr_flag = nullptr;          // [V] the phase's modified flags
  float* pr_part = nullptr;            // [pr_part_n + 1] per-workgroup signed diffs, then the sum
  int pr_part_n = 0;
  int32_t* pr_heavy = nullptr;         // [pr_nheavy] vertices with long in-rows (starting CSR)
  int32_t pr_nheavy = 0;
  uint8_t* pr_isheavy = nullptr;       // [V] 1 for the listed vertices
  float pr_beta = 0.f, pr_delta = 0.f;
  int32_t pr_maxIter = 0;
  std::vector<int32_t> pr_iters;       // sweeps of staticPR, then Decremental / Incremental per batch
  bool pr_ready = false;
};

static bool dyn_dirty(const spm_graph* g) {
  const DynWork* w = g->dyn;
  return w && (w->removed > 0 || w->rremoved > 0 || w->out[w->cur].n > 0 || w->in[w->cur].n > 0);
}

// ---------------------------------------------------------------- kernels

// One lane per update (s, t): removes every copy of t from s's row (binary
// search in the sorted row, then the run of equal entries) and from s's
// overflow row. The removal totals are summed per wavefront before the one
// atomic on each shared counter.
__global__ void dyn_delete(int32_t n, const int32_t* __restrict__ us, const int32_t* __restrict__ ut,
                           const int32_t* __restrict__ off, const int32_t* __restrict__ adj,
                           uint32_t* __restrict__ dead, int32_t* __restrict__ dead_row, DiffRows d,
                           unsigned long long* __restrict__ ctr) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long rb = 0, ro = 0;
  if (i < n) {
    const int32_t s = us[i], t = ut[i];
    int32_t lo = off[s], hi = off[s + 1];
    const int32_t end = hi;
    while (lo < hi) {
      const int32_t mid = lo + ((hi - lo) >> 1);
      if (adj[mid] < t) lo = mid + 1;
      else hi = mid;
    }
    for (int32_t j = lo; j < end && adj[j] == t; ++j) {
      const uint32_t b = 1u << (j & 31);
      if (!(atomicOr(&dead[j >> 5], b) & b)) {
        atomicAdd(&dead_row[s], 1);
        ++rb;
      }
    }
    for (int32_t j = d.off[s]; j < d.off[s + 1]; ++j) {
      if (d.adj[j] != t) continue;
      const uint32_t b = 1u << (j & 31);
      if (!(atomicOr(&d.dead[j >> 5], b) & b)) ++ro;
    }
  }
  rb = wave_sum(rb);
  ro = wave_sum(ro);
  if (__lane_id() == 0) {
    if (rb) atomicAdd(&ctr[0], rb);
    if (ro) atomicAdd(&ctr[1], ro);
  }
}

__global__ void dyn_count(int32_t n, const int32_t* __restrict__ key, int32_t* __restrict__ cnt) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[key[i]], 1);
}

// live[v] = entries of v's overflow row not removed; len[v] = live + adds.
__global__ void dyn_difflen(int32_t V, DiffRows d, int32_t* __restrict__ len, int32_t* __restrict__ live) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    int32_t L = 0;
    for (int32_t j = d.off[v]; j < d.off[v + 1]; ++j) L += bit_of(d.dead, j) ? 0 : 1;
    live[v] = L;
    len[v] += L;
  }
}

// Copies each row's live entries to the new rows; cursor[v] = first free slot.
__global__ void dyn_diffcopy(int32_t V, DiffRows src, DiffRows dst, int32_t* __restrict__ cursor) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    int32_t p = dst.off[v];
    for (int32_t j = src.off[v]; j < src.off[v + 1]; ++j) {
      if (bit_of(src.dead, j)) continue;
      dst.adj[p] = src.adj[j];
      dst.w[p] = src.w[j];
      ++p;
    }
    cursor[v] = p;
  }
}

// Places update i (key -> val, weight) behind the copied entries of key's row.
__global__ void dyn_place(int32_t n, const int32_t* __restrict__ key, const int32_t* __restrict__ val,
                          const int32_t* __restrict__ wt, DiffRows dst, int32_t* __restrict__ cursor) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t p = atomicAdd(&cursor[key[i]], 1);
  dst.adj[p] = val[i];
  dst.w[p] = wt ? wt[i] : 1;
}

// Live entries per row: CSR row minus removed slots plus live overflow entries.
__global__ void dyn_rowlen(int32_t V, const int32_t* __restrict__ off, const int32_t* __restrict__ dead_row,
                           DiffRows d, int32_t* __restrict__ len) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    int32_t L = off[v + 1] - off[v] - dead_row[v];
    for (int32_t j = d.off[v]; j < d.off[v + 1]; ++j) L += bit_of(d.dead, j) ? 0 : 1;
    len[v] = L;
  }
}

// Writes the live entries of every row as (vertex << 32 | neighbour, weight)
// pairs at the row's new position, CSR part first (ballot compaction per 64
// slots), then the overflow part. One wavefront per vertex.
__global__ __launch_bounds__(kDynBlock) void dyn_emit(int32_t V, const int32_t* __restrict__ off,
                                                      const int32_t* __restrict__ adj, const int32_t* __restrict__ w,
                                                      const uint32_t* __restrict__ dead, DiffRows d,
                                                      const int32_t* __restrict__ noff,
                                                      unsigned long long* __restrict__ keys, int32_t* __restrict__ vals) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * kDynWaves;
  for (int64_t x_ = int64_t(blockIdx.x) * kDynWaves + threadIdx.x / kWave; x_ < V; x_ += nw) {
    const int64_t v = spread_vertex(x_, V);
    int64_t pos = noff[v];
    const unsigned long long hi = static_cast<unsigned long long>(v) << 32;
    for (int part = 0; part < 2; ++part) {
      const int32_t a = part == 0 ? off[v] : d.off[v];
      const int32_t b = part == 0 ? off[v + 1] : d.off[v + 1];
      const int32_t* A = part == 0 ? adj : d.adj;
      const int32_t* W = part == 0 ? w : d.w;
      const uint32_t* D = part == 0 ? dead : d.dead;
      for (int32_t j0 = a; j0 < b; j0 += kWave) {
        const int32_t j = j0 + lane;
        const bool keep = j < b && !bit_of(D, j);
        const unsigned long long bl = __ballot(keep);
        if (keep) {
          const int64_t p = pos + __popcll(bl & ((1ull << lane) - 1ull));
          keys[p] = hi | static_cast<uint32_t>(A[j]);
          vals[p] = W ? W[j] : 0;
        }
        pos += __popcll(bl);
      }
    }
  }
}

// Weight of every reverse-CSR slot, from the CSR: the k-th copy of u in v's
// in-row gets the weight of the k-th copy of v in u's out-row. One wavefront
// per vertex u, lanes over its out-edges.
__global__ __launch_bounds__(kDynBlock) void dyn_build_rw(int32_t V, const int32_t* __restrict__ off,
                                                          const int32_t* __restrict__ adj,
                                                          const int32_t* __restrict__ w,
                                                          const int32_t* __restrict__ roff,
                                                          const int32_t* __restrict__ radj, int32_t* __restrict__ rw) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * kDynWaves;
  for (int64_t x_ = int64_t(blockIdx.x) * kDynWaves + threadIdx.x / kWave; x_ < V; x_ += nw) {
    const int64_t u = spread_vertex(x_, V);
    const int32_t a = off[u], b = off[u + 1];
    for (int32_t j = a + lane; j < b; j += kWave) {
      const int32_t v = adj[j];
      int32_t k = 0;
      while (j - 1 - k >= a && adj[j - 1 - k] == v) ++k;
      int32_t lo = roff[v], hi = roff[v + 1];
      while (lo < hi) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        if (radj[mid] < int32_t(u)) lo = mid + 1;
        else hi = mid;
      }
      if (lo + k < roff[v + 1] && radj[lo + k] == int32_t(u)) rw[lo + k] = w[j];
    }
  }
}

// ---- dynamic TC ----------------------------------------------------------------
// The Decremental / Incremental functions of dynamicBatchTCV2 (:23-135): for
// every update (v1, v2) of the phase with v1 != v2, every neighbour
// occurrence v3 of v1 (CSR row minus removed slots, then the live overflow
// entries, as graph::getNeighbors lists them) with v3 != v1, v2 that is a
// neighbour of v2 adds one to class 1 + [(v1,v3) in the phase] +
// [(v2,v3) in the phase]; the count moves by class1/2 + class2/4 + class3/6.
// (The DSL marks modified_del / modified_add through graph::getEdge, which
// returns the first copy of a pair, so a pair is marked or not as a whole;
// the counting loop only asks about pairs that are edges, so "marked" is
// "in the phase's updates".) Work items are 256 neighbour positions of one
// update, so a hub's row is spread over many wavefronts.
constexpr int kDtcChunk = 256;
constexpr unsigned long long kSetEmpty = ~0ull;

__device__ __forceinline__ uint32_t dtc_hash(unsigned long long k) {
  k ^= k >> 31;
  k *= 0x9E3779B97F4A7C15ull;
  return uint32_t(k >> 32);
}

__global__ void dtc_set_insert(int32_t n, const int32_t* __restrict__ us, const int32_t* __restrict__ ut,
                               unsigned long long* __restrict__ set, uint32_t mask) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long key = (static_cast<unsigned long long>(uint32_t(us[i])) << 32) | uint32_t(ut[i]);
  uint32_t h = dtc_hash(key) & mask;
  for (;;) {
    const unsigned long long prev = atomicCAS(&set[h], kSetEmpty, key);
    if (prev == kSetEmpty || prev == key) return;
    h = (h + 1) & mask;
  }
}

__device__ __forceinline__ bool dtc_set_has(const unsigned long long* __restrict__ set, uint32_t mask, int32_t a,
                                            int32_t b) {
  const unsigned long long key = (static_cast<unsigned long long>(uint32_t(a)) << 32) | uint32_t(b);
  uint32_t h = dtc_hash(key) & mask;
  for (;;) {
    const unsigned long long k = set[h];
    if (k == key) return true;
    if (k == kSetEmpty) return false;
    h = (h + 1) & mask;
  }
}

// Work items per update: ceil(neighbour positions of v1 / kDtcChunk), none
// for v1 == v2 (the DSL's filter).
__global__ void dtc_chunks(int32_t n, const int32_t* __restrict__ us, const int32_t* __restrict__ ut,
                           const int32_t* __restrict__ off, DiffRows d, int32_t* __restrict__ cnt) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    cnt[n] = 0;
    return;
  }
  const int32_t v1 = us[i];
  const int32_t len = v1 == ut[i] ? 0 : (off[v1 + 1] - off[v1]) + (d.off[v1 + 1] - d.off[v1]);
  cnt[i] = (len + kDtcChunk - 1) / kDtcChunk;
}

__global__ __launch_bounds__(kDynBlock) void dtc_count(int32_t n, const int32_t* __restrict__ us,
                                                       const int32_t* __restrict__ ut,
                                                       const int32_t* __restrict__ ioff,
                                                       const int32_t* __restrict__ off,
                                                       const int32_t* __restrict__ adj,
                                                       const uint32_t* __restrict__ dead, DiffRows d,
                                                       const unsigned long long* __restrict__ set, uint32_t mask,
                                                       unsigned long long* __restrict__ cls) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * kDynWaves;
  const int64_t total = ioff[n];
  uint32_t c1 = 0, c2 = 0, c3 = 0;
  for (int64_t x = int64_t(blockIdx.x) * kDynWaves + threadIdx.x / kWave; x < total; x += nw) {
    int32_t lo = 0, hi = n - 1;  // the update owning item x: last i with ioff[i] <= x
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (ioff[mid] <= x) lo = mid;
      else hi = mid - 1;
    }
    const int32_t v1 = us[lo], v2 = ut[lo];
    const int64_t base = (x - ioff[lo]) * int64_t(kDtcChunk);
    const int32_t a = off[v1], la = off[v1 + 1] - a;
    const int32_t da = d.off[v1], len = la + (d.off[v1 + 1] - da);
    // A lane's kDtcChunk / 64 candidates v3 are looked up in v2's sorted row
    // together: the binary searches advance in lockstep, so their loads are in
    // flight at once instead of one dependent chain after another.
    constexpr int Q = kDtcChunk / kWave;
    const int32_t b2 = off[v2], e2 = off[v2 + 1];
    int32_t v3[Q], slo[Q], shi[Q];
    // The candidates load first, all of them, and are tested afterwards
    // (testing each as it arrives put every load's wait before the next).
    int32_t xs[Q];
    bool lv[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int64_t p = base + q * kWave + lane;
      xs[q] = -1;
      lv[q] = false;
      if (p >= len) continue;
      if (p < la) {
        const int32_t j = a + int32_t(p);
        lv[q] = !bit_of(dead, j);
        xs[q] = adj[j];
      } else {
        const int32_t j = da + int32_t(p - la);
        lv[q] = !bit_of(d.dead, j);
        xs[q] = d.adj[j];
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int32_t x = xs[q];
      const bool take = lv[q] && x >= 0 && x != v1 && x != v2;
      v3[q] = take ? x : -1;
      slo[q] = b2;
      shi[q] = take ? e2 : b2;
    }
    for (bool more = true; more;) {  // lower bound of each v3 in v2's row
      int32_t mid[Q], val[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        mid[q] = slo[q] + ((shi[q] - slo[q]) >> 1);
        val[q] = slo[q] < shi[q] ? adj[mid[q]] : 0;
      }
      more = false;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if (slo[q] >= shi[q]) continue;
        if (val[q] < v3[q]) slo[q] = mid[q] + 1;
        else shi[q] = mid[q];
        more |= slo[q] < shi[q];
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (v3[q] < 0) continue;
      // every copy of a pair is removed together, so one live copy decides;
      // then the overflow row
      bool nb = false;
      for (int32_t j = slo[q]; j < e2 && adj[j] == v3[q] && !nb; ++j) nb = !bit_of(dead, j);
      for (int32_t j = d.off[v2]; j < d.off[v2 + 1] && !nb; ++j) nb = d.adj[j] == v3[q] && !bit_of(d.dead, j);
      if (!nb) continue;
      const int k = 1 + int(dtc_set_has(set, mask, v1, v3[q])) + int(dtc_set_has(set, mask, v2, v3[q]));
      c1 += k == 1;
      c2 += k == 2;
      c3 += k == 3;
    }
  }
  const unsigned long long s1 = wave_sum(static_cast<unsigned long long>(c1));
  const unsigned long long s2 = wave_sum(static_cast<unsigned long long>(c2));
  const unsigned long long s3 = wave_sum(static_cast<unsigned long long>(c3));
  if (lane == 0) {
    if (s1) atomicAdd(&cls[0], s1);
    if (s2) atomicAdd(&cls[1], s2);
    if (s3) atomicAdd(&cls[2], s3);
  }
}

// count += sign * (class1/2 + class2/4 + class3/6); clears the classes.
__global__ void dtc_apply(unsigned long long* tcv, int sign) {
  const long long delta = static_cast<long long>(tcv[0] / 2 + tcv[1] / 4 + tcv[2] / 6);
  tcv[3] = static_cast<unsigned long long>(static_cast<long long>(tcv[3]) + sign * delta);
  tcv[0] = tcv[1] = tcv[2] = 0ull;
}

// ---- dynamic PageRank -----------------------------------------------------------
// dynamicBatchPR as the dynamic OMP generator emits it: count_outNbrs is the
// row length of the CSR the run starts from (dsl_dyn_cpp_generator.cpp:279-288;
// updates never move indexofNodes), nodes_to iterates the live in-edges (CSR
// part, then overflow part), propagateNodeFlags flags the destination of every
// live edge (graph.hpp:444-462), and the diff is signed (val - pr[v]).

// c[u] = pr[u] / base out-degree (IEEE division, as the emitted expression).
__global__ void dpr_contrib(int32_t V, const float* __restrict__ pr, const float* __restrict__ deg,
                            float* __restrict__ c) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) c[v] = pr[v] / deg[v];
}

// flag[v] |= v has a live in-edge (propagateNodeFlags): its reverse row minus
// removed slots, or a live entry of its reverse overflow row.
__global__ void dpr_propagate(int32_t V, const int32_t* __restrict__ roff, const int32_t* __restrict__ rdead_row,
                              DiffRows in, uint8_t* __restrict__ flag) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    bool live = roff[v + 1] - roff[v] - rdead_row[v] > 0;
    for (int32_t j = in.off[v]; !live && j < in.off[v + 1]; ++j) live = !bit_of(in.dead, j);
    if (live) flag[v] = 1;
  }
}

__global__ void dpr_flag_dests(int32_t n, const int32_t* __restrict__ ut, uint8_t* __restrict__ flag) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) flag[ut[i]] = 1;
}

// One sweep over the flagged vertices (all when flag is null): sum of c over
// live in-edges, val = (1 - delta)/N + delta*sum without contraction, the
// signed diff val - pr[v] summed per workgroup in a fixed order into part[].
// Vertices whose in-row in the starting CSR holds at most kDprHeavy entries
// take kDprGroup lanes each, in id order (dpr_sweep_light); the others, listed
// and marked once at the start (so a later fold cannot move a vertex between
// the two kernels), take a workgroup each (dpr_sweep_heavy).
constexpr int kDprGroup = 4;
constexpr int32_t kDprHeavy = 32;

// Live entries of adj[a, b) taken by one lane of `lanes`, four at a time into
// independent sums so the lane keeps four gathers in flight.
__device__ __forceinline__ float dpr_part_sum(int32_t a, int32_t b, int lane, int lanes,
                                              const int32_t* __restrict__ adj, const uint32_t* __restrict__ dead,
                                              const float* __restrict__ c) {
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  int32_t j = a + lane;
  // Every gather is issued before its term is added (a removed slot still
  // names a vertex; its term adds +0, exactly), so no branch puts one
  // gather's wait in front of the next.
  for (; j + 3 * lanes < b; j += 4 * lanes) {
    const int32_t j1 = j + lanes, j2 = j + 2 * lanes, j3 = j + 3 * lanes;
    const int32_t u0 = adj[j], u1 = adj[j1], u2 = adj[j2], u3 = adj[j3];
    const float t0 = c[u0], t1 = c[u1], t2 = c[u2], t3 = c[u3];
    s0 += bit_of(dead, j) ? 0.f : t0;
    s1 += bit_of(dead, j1) ? 0.f : t1;
    s2 += bit_of(dead, j2) ? 0.f : t2;
    s3 += bit_of(dead, j3) ? 0.f : t3;
  }
  // The last 0..3 terms of the lane, loaded together, added in order.
  float t[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int32_t jk = j + k * lanes;
    t[k] = 0.f;
    if (jk < b && !bit_of(dead, jk)) t[k] = c[adj[jk]];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) s0 += t[k];
  return (s0 + s1) + (s2 + s3);
}

__device__ __forceinline__ float dpr_row_sum(int64_t v, int lane, int lanes, const int32_t* __restrict__ roff,
                                             const int32_t* __restrict__ radj, const uint32_t* __restrict__ rdead,
                                             const DiffRows& in, const float* __restrict__ c) {
  return dpr_part_sum(roff[v], roff[v + 1], lane, lanes, radj, rdead, c) +
         dpr_part_sum(in.off[v], in.off[v + 1], lane, lanes, in.adj, in.dead, c);
}

__device__ __forceinline__ void dpr_part_store(float mydiff, float* __restrict__ part, float* s_red) {
  mydiff = wave_sum(mydiff);
  if (__lane_id() == 0) s_red[threadIdx.x / kWave] = mydiff;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < kDynWaves; ++w) t += s_red[w];
    part[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kDynBlock) void dpr_sweep_light(int32_t V, const uint8_t* __restrict__ is_heavy,
                                                             const int32_t* __restrict__ roff,
                                                             const int32_t* __restrict__ radj,
                                                             const uint32_t* __restrict__ rdead, DiffRows in,
                                                             const float* __restrict__ c, const float* __restrict__ pr,
                                                             const uint8_t* __restrict__ flag, float* __restrict__ nxt,
                                                             float base, float delta, float* __restrict__ part) {
  __shared__ float s_red[kDynWaves];
  const int g = threadIdx.x % kDprGroup;
  const int64_t ng = int64_t(gridDim.x) * (kDynBlock / kDprGroup);
  float mydiff = 0.f;
  for (int64_t v = int64_t(blockIdx.x) * (kDynBlock / kDprGroup) + threadIdx.x / kDprGroup; v < V; v += ng) {
    const bool mine = !is_heavy[v] && (!flag || flag[v]);  // uniform in the group
    float sum = mine ? dpr_row_sum(v, g, kDprGroup, roff, radj, rdead, in, c) : 0.f;
    sum = group_sum<kDprGroup>(sum);
    if (mine && g == 0) {
      const float val = __fadd_rn(base, __fmul_rn(delta, sum));
      mydiff += val - pr[v];
      nxt[v] = val;
    }
  }
  dpr_part_store(mydiff, part, s_red);
}

// Live in-edge sum of v over the lanes of a workgroup, four independent
// gathers in flight per lane.
__device__ __forceinline__ float dpr_row_sum4(int32_t a, int32_t b, const int32_t* __restrict__ adj,
                                              const uint32_t* __restrict__ dead, const float* __restrict__ c) {
  return dpr_part_sum(a, b, int(threadIdx.x), kDynBlock, adj, dead, c);
}

// One workgroup per listed vertex (the list is ordered by in-row length,
// longest first, so the largest hubs start first).
__global__ __launch_bounds__(kDynBlock) void dpr_sweep_heavy(int32_t nh, const int32_t* __restrict__ heavy,
                                                             const int32_t* __restrict__ roff,
                                                             const int32_t* __restrict__ radj,
                                                             const uint32_t* __restrict__ rdead, DiffRows in,
                                                             const float* __restrict__ c, const float* __restrict__ pr,
                                                             const uint8_t* __restrict__ flag, float* __restrict__ nxt,
                                                             float base, float delta, float* __restrict__ part) {
  __shared__ float s_red[kDynWaves];
  __shared__ float s_sum[kDynWaves];
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  float mydiff = 0.f;
  for (int32_t i = blockIdx.x; i < nh; i += gridDim.x) {
    const int32_t v = heavy[i];
    if (flag && !flag[v]) continue;  // uniform across the workgroup
    float sum = dpr_row_sum4(roff[v], roff[v + 1], radj, rdead, c) +
                dpr_row_sum4(in.off[v], in.off[v + 1], in.adj, in.dead, c);
    sum = wave_sum(sum);
    if (lane == 0) s_sum[wid] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
      float t = 0.f;
      for (int w = 0; w < kDynWaves; ++w) t += s_sum[w];
      const float val = __fadd_rn(base, __fmul_rn(delta, t));
      mydiff += val - pr[v];
      nxt[v] = val;
    }
    __syncthreads();  // s_sum reuse
  }
  dpr_part_store(mydiff, part, s_red);
}

// staticPR's start: every rank 1/N; the base out-degrees of the run.
__global__ void dpr_init(int32_t V, const int32_t* __restrict__ off, float init, float* __restrict__ pr,
                         float* __restrict__ deg) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    pr[v] = init;
    deg[v] = float(off[v + 1] - off[v]);
  }
}

__global__ void dpr_apply(int32_t V, const uint8_t* __restrict__ flag, const float* __restrict__ nxt,
                          float* __restrict__ pr) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride)
    if (!flag || flag[v]) pr[v] = nxt[v];
}

// The signed diff of one sweep: the per-workgroup parts summed in order (one
// workgroup), written to out[0].
__global__ __launch_bounds__(kDynBlock) void dpr_diff(const float* __restrict__ part, int n, float* __restrict__ out) {
  __shared__ float s_red[kDynWaves];
  float t = 0.f;
  for (int i = threadIdx.x; i < n; i += kDynBlock) t += part[i];
  t = wave_sum(t);
  if (__lane_id() == 0) s_red[threadIdx.x / kWave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    float d = 0.f;
    for (int w = 0; w < kDynWaves; ++w) d += s_red[w];
    out[0] = d;
  }
}

// ---- dynamic SSSP ------------------------------------------------------------

// Smallest-id predecessor u with dist[u] + w(u,v) == dist[v], packed with
// dist[v]; the source and unreachable vertices get no parent.
__global__ __launch_bounds__(kDynBlock) void dsp_parent(int32_t V, int32_t src, const int32_t* __restrict__ roff,
                                                        const int32_t* __restrict__ radj,
                                                        const uint32_t* __restrict__ rdead,
                                                        const int32_t* __restrict__ rw, DiffRows in,
                                                        const int32_t* __restrict__ dist,
                                                        unsigned long long* __restrict__ dp) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * kDynWaves;
  for (int64_t x_ = int64_t(blockIdx.x) * kDynWaves + threadIdx.x / kWave; x_ < V; x_ += nw) {
    const int64_t v = spread_vertex(x_, V);
    const int32_t dv = dist[v];
    unsigned long long best = kNoParent;
    if (v != src && dv < kInf2) {
      for (int part = 0; part < 2; ++part) {
        const int32_t a = part == 0 ? roff[v] : in.off[v];
        const int32_t b = part == 0 ? roff[v + 1] : in.off[v + 1];
        for (int32_t j = a + lane; j < b; j += kWave) {
          // The slot's removed bit, neighbour and weight load together (a
          // removed slot still names a vertex); only then is it skipped.
          const bool gone = bit_of(part == 0 ? rdead : in.dead, j);
          const int32_t u = part == 0 ? radj[j] : in.adj[j];
          const int32_t wt = part == 0 ? rw[j] : in.w[j];
          if (gone) continue;
          const int32_t du = dist[u];
          if (du < kInf2 && du + wt == dv && static_cast<unsigned long long>(u) < best) best = u;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(best, o, kWave);
        best = x < best ? x : best;
      }
    }
    if (lane == 0) dp[v] = dpack(dv, best);
  }
}

// OnDelete: a deleted edge (s, t) that was t's parent edge cuts t off.
__global__ void dsp_on_delete(int32_t n, const int32_t* __restrict__ us, const int32_t* __restrict__ ut,
                              unsigned long long* __restrict__ dp, uint8_t* __restrict__ mod) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t s = us[i], t = ut[i];
  if (uint32_t(dp[t]) == uint32_t(s)) {
    dp[t] = dpack(kInf2, kNoParent);
    mod[t] = 1;
  }
}

// Fixed-point loops run as groups of launches with one host check per group.
// Iteration k reads f[k % 3] (did iteration k-1 change anything?), records its
// own changes in f[(k+1) % 3] and clears f[(k+2) % 3] for iteration k+1; an
// iteration whose predecessor changed nothing returns at once, so launches
// past convergence cost only their launch.
struct IterFlags {
  const unsigned long long* in;
  unsigned long long* out;
  unsigned long long* clear;
};

__device__ __forceinline__ bool iter_begin(const IterFlags& f) {
  const bool go = __hip_atomic_load(f.in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0) *f.clear = 0ull;
  return go;
}

__global__ void dyn_flags_init(unsigned long long* f) {
  if (threadIdx.x < 3) f[threadIdx.x] = threadIdx.x == 0 ? 1ull : 0ull;
}

// Decremental, first loop: a vertex whose parent was cut off is cut off too.
__global__ void dsp_invalidate(int32_t V, unsigned long long* __restrict__ dp, uint8_t* __restrict__ mod,
                               IterFlags fl) {
  if (!iter_begin(fl)) return;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  bool any = false;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    if (mod[v]) continue;
    const uint32_t p = uint32_t(dp[v]);
    if (p != uint32_t(kNoParent) && __hip_atomic_load(&mod[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      mod[v] = 1;
      dp[v] = dpack(kInf2, kNoParent);
      any = true;
    }
  }
  if (__any(any) && __lane_id() == 0) atomicOr(fl.out, 1ull);
}

// Decremental, second loop: each cut-off vertex takes the best finite
// in-neighbour (distance + weight, then smallest id). One wavefront per
// listed vertex.
__global__ __launch_bounds__(kDynBlock) void dsp_pull(const int32_t* __restrict__ list, const int32_t* __restrict__ nsel,
                                                      const int32_t* __restrict__ roff, const int32_t* __restrict__ radj,
                                                      const uint32_t* __restrict__ rdead, const int32_t* __restrict__ rw,
                                                      DiffRows in, unsigned long long* __restrict__ dp,
                                                      IterFlags fl) {
  if (!iter_begin(fl)) return;
  const int lane = __lane_id();
  const int64_t n = *nsel;
  const int64_t nw = int64_t(gridDim.x) * kDynWaves;
  bool any = false;
  for (int64_t i = int64_t(blockIdx.x) * kDynWaves + threadIdx.x / kWave; i < n; i += nw) {
    const int32_t v = list[i];
    unsigned long long best = ~0ull;
    for (int part = 0; part < 2; ++part) {
      const int32_t a = part == 0 ? roff[v] : in.off[v];
      const int32_t b = part == 0 ? roff[v + 1] : in.off[v + 1];
      for (int32_t j = a + lane; j < b; j += kWave) {
        const bool gone = bit_of(part == 0 ? rdead : in.dead, j);
        const int32_t u = part == 0 ? radj[j] : in.adj[j];
        const int32_t wt = part == 0 ? rw[j] : in.w[j];
        if (gone) continue;
        const int32_t du = ddist(__hip_atomic_load(&dp[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (du >= kInf2) continue;
        const unsigned long long k = dpack(du + wt, uint32_t(u));
        best = k < best ? k : best;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long x = __shfl_xor(best, o, kWave);
      best = x < best ? x : best;
    }
    if (lane == 0 && best < dp[v]) {
      __hip_atomic_store(&dp[v], best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      any = true;
    }
  }
  if (__any(any) && lane == 0) atomicOr(fl.out, 1ull);
}

// OnAdd: an added edge (s, t, w) that shortens t marks both ends.
__global__ void dsp_on_add(int32_t n, const int32_t* __restrict__ us, const int32_t* __restrict__ ut,
                           const int32_t* __restrict__ uw, const unsigned long long* __restrict__ dp,
                           uint8_t* __restrict__ madd) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t s = us[i], t = ut[i], w = uw ? uw[i] : 1;
  const int32_t ds = ddist(dp[s]);
  if (ds < kInf2 && ddist(dp[t]) > ds + w) {
    madd[s] = 1;
    madd[t] = 1;
  }
}

// Incremental: one push round over the flagged vertices. A wavefront's 64
// vertices have their out-edges (CSR row, then overflow row) flattened across
// the lanes; each edge tries a 64-bit atomicMin of (distance, parent).
__global__ __launch_bounds__(kDynBlock) void dsp_push(int32_t V, const int32_t* __restrict__ off,
                                                      const int32_t* __restrict__ adj, const int32_t* __restrict__ w,
                                                      const uint32_t* __restrict__ dead, DiffRows out,
                                                      unsigned long long* __restrict__ dp, uint8_t* __restrict__ cur,
                                                      uint8_t* __restrict__ nxt, IterFlags fl) {
  if (!iter_begin(fl)) return;
  __shared__ uint32_t s_pre[kDynWaves][kWave];
  __shared__ int32_t s_a[kDynWaves][kWave];
  __shared__ int32_t s_da[kDynWaves][kWave];
  __shared__ int32_t s_nb[kDynWaves][kWave];
  __shared__ int32_t s_dv[kDynWaves][kWave];
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  bool found = false;
  const int64_t nwv = int64_t(gridDim.x) * kDynWaves;
  for (int64_t base = (int64_t(blockIdx.x) * kDynWaves + wid) * kWave; base < V; base += nwv * kWave) {
    const int64_t v = base + lane;
    uint32_t deg = 0;
    int32_t a = 0, da = 0, nb = 0, dv = 0;
    if (v < V && cur[v]) {
      cur[v] = 0;
      dv = ddist(__hip_atomic_load(&dp[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (dv < kInf2) {
        a = off[v];
        nb = off[v + 1] - a;
        da = out.off[v];
        deg = uint32_t(nb + out.off[v + 1] - da);
      }
    }
    const uint32_t incl = wave_incl_scan(deg);
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    s_pre[wid][lane] = incl - deg;
    s_a[wid][lane] = a;
    s_da[wid][lane] = da;
    s_nb[wid][lane] = nb;
    s_dv[wid][lane] = dv;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t x = lane; x < total; x += kWave) {
      int lo = 0, hi = kWave - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[wid][mid] <= x) lo = mid;
        else hi = mid - 1;
      }
      const int32_t k = int32_t(x - s_pre[wid][lo]);
      int32_t t, wt;
      if (k < s_nb[wid][lo]) {
        const int32_t j = s_a[wid][lo] + k;
        if (bit_of(dead, j)) continue;
        t = adj[j];
        wt = w[j];
      } else {
        const int32_t j = s_da[wid][lo] + (k - s_nb[wid][lo]);
        if (bit_of(out.dead, j)) continue;
        t = out.adj[j];
        wt = out.w[j];
      }
      const int32_t nd = s_dv[wid][lo] + wt;
      const unsigned long long key = dpack(nd, uint32_t(base + lo));
      if (key < __hip_atomic_load(&dp[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
        const unsigned long long old = atomicMin(&dp[t], key);
        if (nd < ddist(old)) {
          nxt[t] = 1;
          found = true;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (__any(found) && lane == 0 && __hip_atomic_load(fl.out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull)
    atomicOr(fl.out, 1ull);
}

__global__ void dsp_unpack(int32_t V, const unsigned long long* __restrict__ dp, int32_t* __restrict__ dist,
                           int32_t* __restrict__ parent) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    const unsigned long long k = dp[v];
    if (dist) dist[v] = int32_t(k >> 32);
    if (parent) parent[v] = uint32_t(k) == uint32_t(kNoParent) ? -1 : int32_t(uint32_t(k));
  }
}

__global__ void dsp_min_weight(int64_t n, const int32_t* __restrict__ w, unsigned long long* __restrict__ flag) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  bool bad = false;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) bad |= w[i] < 1;
  if (__any(bad) && __lane_id() == 0) atomicOr(flag, 1ull);
}

// ---------------------------------------------------------------- host side

static int grid_for(const spm_graph* g, int64_t n, int block = kDynBlock) {
  return int(std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, int64_t(g->num_cus) * 16)));
}

static int tmp_reserve(spm_graph* g, size_t need) {
  DynWork* w = g->dyn;
  if (need <= w->tmp_bytes) return SPM_OK;
  dfree(g, w->tmp, w->tmp_bytes);
  w->tmp = nullptr;
  w->tmp_bytes = 0;
  SPM_TRY(dmalloc(g, &w->tmp, need, "update scratch"));
  w->tmp_bytes = need;
  return SPM_OK;
}

static void diff_free(spm_graph* g, DiffStore& d) {
  const size_t V = size_t(g->V);
  dfree(g, d.r.off, sizeof(int32_t) * (V + 1));
  dfree(g, d.r.adj, sizeof(int32_t) * d.cap);
  dfree(g, d.r.w, sizeof(int32_t) * d.cap);
  dfree(g, d.r.dead, sizeof(uint32_t) * (d.cap / 32 + 1));
  d = DiffStore();
}

static int diff_init(spm_graph* g, DiffStore& d) {
  SPM_TRY(dmalloc(g, (void**)&d.r.off, sizeof(int32_t) * (size_t(g->V) + 1), "overflow row offsets"));
  SPM_CHECK_HIP(hipMemsetAsync(d.r.off, 0, sizeof(int32_t) * (size_t(g->V) + 1), g->stream));
  d.cap = 0;
  d.n = 0;
  return SPM_OK;
}

static int diff_reserve(spm_graph* g, DiffStore& d, size_t need) {
  if (need <= d.cap && d.r.adj) return SPM_OK;
  const size_t cap = std::max<size_t>(need + need / 2, 1024);
  dfree(g, d.r.adj, sizeof(int32_t) * d.cap);
  dfree(g, d.r.w, sizeof(int32_t) * d.cap);
  dfree(g, d.r.dead, sizeof(uint32_t) * (d.cap / 32 + 1));
  d.r.adj = d.r.w = nullptr;
  d.r.dead = nullptr;
  d.cap = 0;
  SPM_TRY(dmalloc(g, (void**)&d.r.adj, sizeof(int32_t) * cap, "overflow rows"));
  SPM_TRY(dmalloc(g, (void**)&d.r.w, sizeof(int32_t) * cap, "overflow weights"));
  SPM_TRY(dmalloc(g, (void**)&d.r.dead, sizeof(uint32_t) * (cap / 32 + 1), "overflow removed bits"));
  d.cap = cap;
  return SPM_OK;
}

// Removed-slot state sized for the current CSR (all clear).
static int dyn_reset_base(spm_graph* g) {
  DynWork* w = g->dyn;
  const size_t V = size_t(g->V);
  dfree(g, w->dead, sizeof(uint32_t) * w->bits_words);
  dfree(g, w->rdead, sizeof(uint32_t) * w->bits_words);
  w->dead = w->rdead = nullptr;
  w->bits_words = size_t(g->E) / 32 + 1;
  SPM_TRY(dmalloc(g, (void**)&w->dead, sizeof(uint32_t) * w->bits_words, "removed slots"));
  SPM_CHECK_HIP(hipMemsetAsync(w->dead, 0, sizeof(uint32_t) * w->bits_words, g->stream));
  if (g->roff) {
    SPM_TRY(dmalloc(g, (void**)&w->rdead, sizeof(uint32_t) * w->bits_words, "removed reverse slots"));
    SPM_CHECK_HIP(hipMemsetAsync(w->rdead, 0, sizeof(uint32_t) * w->bits_words, g->stream));
  }
  SPM_CHECK_HIP(hipMemsetAsync(w->dead_row, 0, sizeof(int32_t) * V, g->stream));
  SPM_CHECK_HIP(hipMemsetAsync(w->rdead_row, 0, sizeof(int32_t) * V, g->stream));
  w->removed = w->rremoved = 0;
  for (int i = 0; i < 2; ++i) {
    SPM_CHECK_HIP(hipMemsetAsync(w->out[i].r.off, 0, sizeof(int32_t) * (V + 1), g->stream));
    SPM_CHECK_HIP(hipMemsetAsync(w->in[i].r.off, 0, sizeof(int32_t) * (V + 1), g->stream));
    w->out[i].n = w->in[i].n = 0;
  }
  return SPM_OK;
}

static int dyn_alloc(spm_graph* g) {
  if (g->dyn) return SPM_OK;
  DynWork* w = new DynWork();
  g->dyn = w;
  const size_t V = size_t(g->V);
  SPM_TRY(dmalloc(g, (void**)&w->dead_row, sizeof(int32_t) * V, "removed per row"));
  SPM_TRY(dmalloc(g, (void**)&w->rdead_row, sizeof(int32_t) * V, "removed per reverse row"));
  SPM_TRY(dmalloc(g, (void**)&w->cnt, sizeof(int32_t) * (V + 1), "update counts"));
  SPM_TRY(dmalloc(g, (void**)&w->cursor, sizeof(int32_t) * V, "update cursors"));
  SPM_TRY(dmalloc(g, (void**)&w->live, sizeof(int32_t) * V, "overflow live counts"));
  SPM_TRY(dmalloc(g, (void**)&w->ctr, sizeof(unsigned long long) * 8, "update counters"));
  for (int i = 0; i < 2; ++i) {
    SPM_TRY(diff_init(g, w->out[i]));
    SPM_TRY(diff_init(g, w->in[i]));
  }
  SPM_TRY(dyn_reset_base(g));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  return SPM_OK;
}

int dyn_release(spm_graph* g) {
  DynWork* w = g->dyn;
  if (!w) return SPM_OK;
  const size_t V = size_t(g->V);
  dfree(g, w->dead, sizeof(uint32_t) * w->bits_words);
  dfree(g, w->rdead, sizeof(uint32_t) * w->bits_words);
  dfree(g, w->dead_row, sizeof(int32_t) * V);
  dfree(g, w->rdead_row, sizeof(int32_t) * V);
  dfree(g, w->rw, sizeof(int32_t) * size_t(g->E));
  for (int i = 0; i < 2; ++i) {
    diff_free(g, w->out[i]);
    diff_free(g, w->in[i]);
  }
  dfree(g, w->ubuf, sizeof(int32_t) * 3 * w->ucap);
  if (w->stage_pending) (void)hipEventSynchronize(w->stage_done);
  if (w->h_stage) (void)hipHostFree(w->h_stage);
  if (w->stage_done) (void)hipEventDestroy(w->stage_done);
  dfree(g, w->cnt, sizeof(int32_t) * (V + 1));
  dfree(g, w->cursor, sizeof(int32_t) * V);
  dfree(g, w->live, sizeof(int32_t) * V);
  dfree(g, w->ctr, sizeof(unsigned long long) * 8);
  dfree(g, w->tmp, w->tmp_bytes);
  dfree(g, w->dp, sizeof(unsigned long long) * V);
  dfree(g, w->mod, V);
  dfree(g, w->fl[0], V);
  dfree(g, w->fl[1], V);
  dfree(g, w->list, sizeof(int32_t) * V);
  dfree(g, w->nsel, sizeof(int32_t));
  dfree(g, w->tset, sizeof(unsigned long long) * w->tset_cap);
  dfree(g, w->tchunk, sizeof(int32_t) * w->tchunk_cap);
  dfree(g, w->tcv, sizeof(unsigned long long) * 4);
  dfree(g, w->pr, sizeof(float) * V);
  dfree(g, w->pr_nxt, sizeof(float) * V);
  dfree(g, w->pr_c, sizeof(float) * V);
  dfree(g, w->pr_deg, sizeof(float) * V);
  dfree(g, w->pr_flag, V);
  dfree(g, w->pr_part, sizeof(float) * (size_t(w->pr_part_n) + 1));
  dfree(g, w->pr_heavy, sizeof(int32_t) * size_t(w->pr_nheavy));
  dfree(g, w->pr_isheavy, V);
  delete w;
  g->dyn = nullptr;
  return SPM_OK;
}

// Copies a batch (src, dst, optional weight) to the device; returns device
// pointers to the three arrays (weight null when not given).
static int stage(spm_graph* g, const int32_t* s, const int32_t* t, const int32_t* wt, int32_t n,
                 int32_t** ds, int32_t** dt, int32_t** dw) {
  DynWork* w = g->dyn;
  for (int32_t i = 0; i < n; ++i)
    if (s[i] < 0 || s[i] >= g->V || t[i] < 0 || t[i] >= g->V)
      return set_error(SPM_EINVAL, "edge update names a vertex outside 0..num_nodes()-1");
  // the pinned buffer may still feed the previous batch's copies
  if (w->stage_pending) {
    SPM_CHECK_HIP(hipEventSynchronize(w->stage_done));
    w->stage_pending = false;
  }
  if (!w->stage_done) SPM_CHECK_HIP(hipEventCreateWithFlags(&w->stage_done, hipEventDisableTiming));
  if (size_t(n) > w->ucap) {
    dfree(g, w->ubuf, sizeof(int32_t) * 3 * w->ucap);
    if (w->h_stage) (void)hipHostFree(w->h_stage);
    w->ubuf = nullptr;
    w->h_stage = nullptr;
    w->ucap = 0;
    const size_t cap = std::max<size_t>(size_t(n), 4096);
    SPM_TRY(dmalloc(g, (void**)&w->ubuf, sizeof(int32_t) * 3 * cap, "update batch"));
    SPM_CHECK_HIP(hipHostMalloc((void**)&w->h_stage, sizeof(int32_t) * 3 * cap, hipHostMallocDefault));
    w->ucap = cap;
  }
  std::memcpy(w->h_stage, s, sizeof(int32_t) * n);
  std::memcpy(w->h_stage + w->ucap, t, sizeof(int32_t) * n);
  if (wt) std::memcpy(w->h_stage + 2 * w->ucap, wt, sizeof(int32_t) * n);
  *ds = w->ubuf;
  *dt = w->ubuf + w->ucap;
  *dw = wt ? w->ubuf + 2 * w->ucap : nullptr;
  SPM_CHECK_HIP(hipMemcpyAsync(w->ubuf, w->h_stage, sizeof(int32_t) * (wt ? 3 : 2) * w->ucap, hipMemcpyHostToDevice,
                               g->stream));
  SPM_CHECK_HIP(hipEventRecord(w->stage_done, g->stream));
  w->stage_pending = true;
  return SPM_OK;
}

static int delete_staged(spm_graph* g, const int32_t* ds, const int32_t* dt, int32_t n) {
  DynWork* w = g->dyn;
  if (n == 0) return SPM_OK;
  SPM_CHECK_HIP(hipMemsetAsync(w->ctr, 0, sizeof(unsigned long long) * 4, g->stream));
  const int grid = int((n + kDynBlock - 1) / kDynBlock);
  SPM_TRY(launch(g, "dyn_delete", dyn_delete, dim3(grid), dim3(kDynBlock), 0, n, ds, dt, (const int32_t*)g->off,
                 (const int32_t*)g->adj, w->dead, w->dead_row, w->out[w->cur].r, w->ctr));
  if (g->roff)
    SPM_TRY(launch(g, "dyn_delete", dyn_delete, dim3(grid), dim3(kDynBlock), 0, n, dt, ds, (const int32_t*)g->roff,
                   (const int32_t*)g->radj, w->rdead, w->rdead_row, w->in[w->cur].r, w->ctr + 2));
  // counted as pending without reading the device counters back: an upper
  // bound is enough to know a fold is due
  w->removed += n;
  if (g->roff) w->rremoved += n;
  return SPM_OK;
}

// Rebuilds one direction's overflow rows with the batch (key -> val) added.
static int add_one_side(spm_graph* g, DiffStore* st, const int32_t* key, const int32_t* val, const int32_t* wt,
                        int32_t n) {
  DynWork* w = g->dyn;
  const int32_t V = g->V;
  DiffStore& src = st[w->cur];
  DiffStore& dst = st[1 - w->cur];
  SPM_CHECK_HIP(hipMemsetAsync(w->cnt, 0, sizeof(int32_t) * (size_t(V) + 1), g->stream));
  SPM_TRY(launch(g, "dyn_count", dyn_count, dim3((n + kDynBlock - 1) / kDynBlock), dim3(kDynBlock), 0, n, key, w->cnt));
  SPM_TRY(launch(g, "dyn_difflen", dyn_difflen, dim3(grid_for(g, V)), dim3(kDynBlock), 0, V, src.r, w->cnt, w->live));
  size_t need = 0;
  SPM_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, w->cnt, dst.r.off, V + 1, g->stream));
  SPM_TRY(tmp_reserve(g, need));
  SPM_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(w->tmp, need, w->cnt, dst.r.off, V + 1, g->stream));
  // live old entries + adds <= old entries + adds: sized without reading the
  // scan's total back
  const int64_t bound = src.n + n;
  SPM_TRY(diff_reserve(g, dst, size_t(bound)));
  SPM_CHECK_HIP(hipMemsetAsync(dst.r.dead, 0, sizeof(uint32_t) * (dst.cap / 32 + 1), g->stream));
  SPM_TRY(launch(g, "dyn_diffcopy", dyn_diffcopy, dim3(grid_for(g, V)), dim3(kDynBlock), 0, V, src.r, dst.r, w->cursor));
  SPM_TRY(launch(g, "dyn_place", dyn_place, dim3((n + kDynBlock - 1) / kDynBlock), dim3(kDynBlock), 0, n, key, val, wt,
                 dst.r, w->cursor));
  dst.n = bound;
  return SPM_OK;
}

static int add_staged(spm_graph* g, const int32_t* ds, const int32_t* dt, const int32_t* dw, int32_t n) {
  DynWork* w = g->dyn;
  if (n == 0) return SPM_OK;
  if (g->E + int64_t(w->out[w->cur].n) + n > 0x7fffffffLL)
    return set_error(SPM_EINVAL, "edge additions would exceed 2^31 - 1 edges");
  SPM_TRY(add_one_side(g, w->out, ds, dt, dw, n));
  if (g->roff) SPM_TRY(add_one_side(g, w->in, dt, ds, dw, n));
  w->cur = 1 - w->cur;
  return SPM_OK;
}

// Folds removed slots and overflow rows into a fresh sorted CSR (and reverse
// CSR with per-slot weights); algorithm workspaces sized for the old CSR are
// dropped.
static int dyn_fold(spm_graph* g) {
  DynWork* w = g->dyn;
  const int32_t V = g->V;
  const size_t Vs = size_t(V);
  DiffRows out = w->out[w->cur].r;
  // new row lengths and offsets
  SPM_TRY(launch(g, "dyn_rowlen", dyn_rowlen, dim3(grid_for(g, V)), dim3(kDynBlock), 0, V, (const int32_t*)g->off,
                 (const int32_t*)w->dead_row, out, w->cnt));
  SPM_CHECK_HIP(hipMemsetAsync(w->cnt + V, 0, sizeof(int32_t), g->stream));
  int32_t* noff = nullptr;
  SPM_TRY(dmalloc(g, (void**)&noff, sizeof(int32_t) * (Vs + 1), "indexofNodes"));
  size_t need = 0;
  SPM_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, w->cnt, noff, V + 1, g->stream));
  SPM_TRY(tmp_reserve(g, need));
  SPM_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(w->tmp, need, w->cnt, noff, V + 1, g->stream));
  // the largest new row length into cnt[V] (the scan has consumed it)
  SPM_CHECK_HIP(hipcub::DeviceReduce::Max(nullptr, need, w->cnt, w->cnt + V, V, g->stream));
  SPM_TRY(tmp_reserve(g, need));
  SPM_CHECK_HIP(hipcub::DeviceReduce::Max(w->tmp, need, w->cnt, w->cnt + V, V, g->stream));
  int32_t* h = g->h_pinned;
  SPM_CHECK_HIP(hipMemcpyAsync(h, noff + V, sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipMemcpyAsync(h + 1, w->cnt + V, sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  const int32_t E2 = h[0], maxdeg = h[1];
  const size_t En = size_t(E2);
  // pairs, sorted by (vertex, neighbour); the sort is stable, so copies of
  // an edge keep CSR-before-overflow order
  unsigned long long *k0 = nullptr, *k1 = nullptr;
  int32_t *v0 = nullptr, *v1 = nullptr;
  int rc = dmalloc(g, (void**)&k0, sizeof(unsigned long long) * En, "fold keys");
  if (rc == SPM_OK) rc = dmalloc(g, (void**)&k1, sizeof(unsigned long long) * En, "fold keys");
  if (rc == SPM_OK) rc = dmalloc(g, (void**)&v0, sizeof(int32_t) * En, "fold weights");
  if (rc == SPM_OK) rc = dmalloc(g, (void**)&v1, sizeof(int32_t) * En, "fold weights");
  int32_t *nadj = nullptr, *nw = nullptr, *nroff = nullptr, *nradj = nullptr, *nrw = nullptr;
  if (rc == SPM_OK) rc = dmalloc(g, (void**)&nadj, sizeof(int32_t) * En, "edgeList");
  if (rc == SPM_OK && g->w) rc = dmalloc(g, (void**)&nw, sizeof(int32_t) * En, "edgeLen");
  if (rc == SPM_OK && g->roff) rc = dmalloc(g, (void**)&nroff, sizeof(int32_t) * (Vs + 1), "rev_indexofNodes");
  if (rc == SPM_OK && g->roff) rc = dmalloc(g, (void**)&nradj, sizeof(int32_t) * En, "srcList");
  if (rc == SPM_OK && g->roff && g->w) rc = dmalloc(g, (void**)&nrw, sizeof(int32_t) * En, "reverse weights");
  // live entries as (vertex << 32 | neighbour, weight) pairs, then the same
  // sorted CSR / reverse CSR build the relabeled layout uses (the sort is
  // stable, so copies of an edge keep CSR-before-overflow order)
  if (rc == SPM_OK && En > 0)
    rc = launch(g, "dyn_emit", dyn_emit, dim3(grid_for(g, int64_t(V) * kWave)), dim3(kDynBlock), 0, V,
                (const int32_t*)g->off, (const int32_t*)g->adj, (const int32_t*)g->w, (const uint32_t*)w->dead, out,
                (const int32_t*)noff, k0, v0);
  if (rc == SPM_OK) rc = csr_from_pairs(g, int64_t(En), k0, k1, v0, v1, noff, nadj, nw, nroff, nradj, nrw);
  if (rc == SPM_OK) {
    hipError_t e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = set_error(SPM_EHIP, std::string("fold: ") + hipGetErrorString(e));
  }
  dfree(g, k0, sizeof(unsigned long long) * En);
  dfree(g, k1, sizeof(unsigned long long) * En);
  dfree(g, v0, sizeof(int32_t) * En);
  dfree(g, v1, sizeof(int32_t) * En);
  if (rc != SPM_OK) {
    dfree(g, noff, sizeof(int32_t) * (Vs + 1));
    dfree(g, nadj, sizeof(int32_t) * En);
    dfree(g, nw, sizeof(int32_t) * En);
    dfree(g, nroff, sizeof(int32_t) * (Vs + 1));
    dfree(g, nradj, sizeof(int32_t) * En);
    dfree(g, nrw, sizeof(int32_t) * En);
    return rc;
  }
  // swap in the new arrays
  pr_release(g);
  sssp_release(g);
  tc_release(g);
  bc_release(g);
  relabel_release(g);
  const size_t eb = sizeof(int32_t) * size_t(g->E), vb = sizeof(int32_t) * (Vs + 1);
  dfree(g, g->off, vb);
  dfree(g, g->adj, eb);
  dfree(g, g->w, eb);
  dfree(g, g->roff, vb);
  dfree(g, g->radj, eb);
  dfree(g, w->rw, eb);
  g->off = noff;
  g->adj = nadj;
  g->w = nw;
  g->roff = nroff;
  g->radj = nradj;
  w->rw = nrw;
  w->rw_ok = nrw != nullptr;
  g->E = int64_t(E2);
  g->max_outdeg = maxdeg;
  schedule_after_fold(g);
  SPM_TRY(dyn_reset_base(g));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  return SPM_OK;
}

// Reverse-slot weights for the current (clean) CSR.
static int dyn_ensure_rw(spm_graph* g) {
  DynWork* w = g->dyn;
  if (w->rw_ok) return SPM_OK;
  if (!g->w || !g->roff) return set_error(SPM_ENOGRAPH, "dynamic SSSP needs weights and the reverse CSR");
  if (!w->rw) SPM_TRY(dmalloc(g, (void**)&w->rw, sizeof(int32_t) * size_t(g->E), "reverse weights"));
  SPM_TRY(launch(g, "dyn_build_rw", dyn_build_rw, dim3(grid_for(g, int64_t(g->V) * kWave)), dim3(kDynBlock), 0, g->V,
                 (const int32_t*)g->off, (const int32_t*)g->adj, (const int32_t*)g->w, (const int32_t*)g->roff,
                 (const int32_t*)g->radj, w->rw));
  w->rw_ok = true;
  return SPM_OK;
}

int dyn_settle(spm_graph* g) {
  if (!dyn_dirty(g)) return SPM_OK;
  return dyn_fold(g);
}

bool dyn_pending(const spm_graph* g) { return dyn_dirty(g); }

void dyn_weights_changed(spm_graph* g) {
  if (g->dyn) g->dyn->rw_ok = false;
}

// ---- dynamic SSSP host steps ------------------------------------------------

static int dsp_alloc(spm_graph* g) {
  DynWork* w = g->dyn;
  if (w->dp) return SPM_OK;
  const size_t V = size_t(g->V);
  SPM_TRY(dmalloc(g, (void**)&w->dp, sizeof(unsigned long long) * V, "dynamic SSSP state"));
  SPM_TRY(dmalloc(g, (void**)&w->mod, V, "dynamic SSSP flags"));
  SPM_TRY(dmalloc(g, (void**)&w->fl[0], V, "dynamic SSSP frontier"));
  SPM_TRY(dmalloc(g, (void**)&w->fl[1], V, "dynamic SSSP frontier"));
  SPM_TRY(dmalloc(g, (void**)&w->list, sizeof(int32_t) * V, "dynamic SSSP list"));
  SPM_TRY(dmalloc(g, (void**)&w->nsel, sizeof(int32_t), "dynamic SSSP list size"));
  return SPM_OK;
}

static int dsp_read_flag(spm_graph* g, unsigned long long* dflag, bool* set) {
  unsigned long long* h = reinterpret_cast<unsigned long long*>(g->h_pinned);
  SPM_CHECK_HIP(hipMemcpyAsync(h, dflag, sizeof(unsigned long long), hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  *set = h[0] != 0ull;
  return SPM_OK;
}

static int dsp_init(spm_graph* g, int32_t src) {
  if (src < 0 || src >= g->V) return set_error(SPM_EINVAL, "dynamic SSSP: source out of range");
  if (!g->w) return set_error(SPM_ENOGRAPH, "dynamic SSSP needs edge weights (edgeLen or spm_graph_set_weights)");
  if (!g->roff) return set_error(SPM_ENOGRAPH, "dynamic SSSP needs the reverse CSR (rev_indexofNodes, srcList)");
  SPM_TRY(dyn_alloc(g));
  SPM_TRY(dyn_settle(g));
  DynWork* w = g->dyn;
  SPM_CHECK_HIP(hipMemsetAsync(w->ctr, 0, sizeof(unsigned long long), g->stream));
  SPM_TRY(launch(g, "dsp_min_weight", dsp_min_weight, dim3(grid_for(g, g->E)), dim3(kDynBlock), 0, g->E,
                 (const int32_t*)g->w, w->ctr));
  bool bad = false;
  SPM_TRY(dsp_read_flag(g, w->ctr, &bad));
  if (bad) return set_error(SPM_EINVAL, "dynamic SSSP needs weights >= 1 (parent pointers must strictly shorten)");
  SPM_TRY(dsp_alloc(g));
  SPM_TRY(dyn_ensure_rw(g));
  SPM_TRY(sssp_run(g, &src, 1, nullptr, nullptr));
  SPM_TRY(launch(g, "dsp_parent", dsp_parent, dim3(grid_for(g, int64_t(g->V) * kWave)), dim3(kDynBlock), 0, g->V, src,
                 (const int32_t*)g->roff, (const int32_t*)g->radj, (const uint32_t*)w->rdead, (const int32_t*)w->rw,
                 w->in[w->cur].r, (const int32_t*)sssp_dist(g), w->dp));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  w->src = src;
  return prof_flush(g);
}

// Runs iteration(k, flags) in groups of kIterGroup launches, reading one flag
// back per group, until an iteration changes nothing.
constexpr int kIterGroup = 8;
template <typename Iter>
static int fixed_point(spm_graph* g, const char* what, Iter iteration) {
  unsigned long long* f = g->dyn->ctr + 4;  // three rotating flags
  SPM_TRY(launch(g, "dyn_flags_init", dyn_flags_init, dim3(1), dim3(64), 0, f));
  for (int64_t base = 0;; base += kIterGroup) {
    for (int64_t k = base; k < base + kIterGroup; ++k)
      SPM_TRY(iteration(k, IterFlags{f + k % 3, f + (k + 1) % 3, f + (k + 2) % 3}));
    bool changed = false;
    SPM_TRY(dsp_read_flag(g, f + (base + kIterGroup) % 3, &changed));
    if (!changed) return SPM_OK;
    if (base > int64_t(g->V)) return set_error(SPM_EDIVERGE, std::string("dynamic SSSP: ") + what + " did not converge");
  }
}

// One batch of the Dynamic function: OnDelete, updateCSRDel, Decremental,
// OnAdd, updateCSRAdd, Incremental (added edges weigh 1, update.hpp:56-60).
static int dsp_batch(spm_graph* g, const char* type, const int32_t* us, const int32_t* ut, int32_t n) {
  DynWork* w = g->dyn;
  if (!w || w->src < 0) return set_error(SPM_EINVAL, "dynamic SSSP: call spm_dynsssp_init first");
  const int32_t V = g->V;
  std::vector<int32_t> ds, dt, as, at;
  for (int32_t i = 0; i < n; ++i) {
    if (type[i] == 'd') { ds.push_back(us[i]); dt.push_back(ut[i]); }
    else if (type[i] == 'a') { as.push_back(us[i]); at.push_back(ut[i]); }
    else return set_error(SPM_EINVAL, "dynamic SSSP: update type must be 'a' or 'd'");
  }
  const int vgrid = grid_for(g, V);
  const int wgrid = g->num_cus * 8;
  // OnDelete + updateCSRDel + Decremental
  SPM_CHECK_HIP(hipMemsetAsync(w->mod, 0, size_t(V), g->stream));
  if (!ds.empty()) {
    int32_t *s = nullptr, *t = nullptr, *x = nullptr;
    const int32_t nd = int32_t(ds.size());
    SPM_TRY(stage(g, ds.data(), dt.data(), nullptr, nd, &s, &t, &x));
    SPM_TRY(launch(g, "dsp_on_delete", dsp_on_delete, dim3((nd + kDynBlock - 1) / kDynBlock), dim3(kDynBlock), 0, nd,
                   (const int32_t*)s, (const int32_t*)t, w->dp, w->mod));
    SPM_TRY(delete_staged(g, s, t, nd));
    SPM_TRY(fixed_point(g, "the cut-off loop", [&](int64_t, IterFlags fl) {
      return launch(g, "dsp_invalidate", dsp_invalidate, dim3(vgrid), dim3(kDynBlock), 0, V, w->dp, w->mod, fl);
    }));
    size_t need = 0;
    hipcub::CountingInputIterator<int32_t> ids(0);
    SPM_CHECK_HIP(hipcub::DeviceSelect::Flagged(nullptr, need, ids, w->mod, w->list, w->nsel, V, g->stream));
    SPM_TRY(tmp_reserve(g, need));
    SPM_CHECK_HIP(hipcub::DeviceSelect::Flagged(w->tmp, need, ids, w->mod, w->list, w->nsel, V, g->stream));
    SPM_TRY(fixed_point(g, "Decremental", [&](int64_t, IterFlags fl) {
      return launch(g, "dsp_pull", dsp_pull, dim3(wgrid), dim3(kDynBlock), 0, (const int32_t*)w->list,
                    (const int32_t*)w->nsel, (const int32_t*)g->roff, (const int32_t*)g->radj,
                    (const uint32_t*)w->rdead, (const int32_t*)w->rw, w->in[w->cur].r, w->dp, fl);
    }));
  }
  // OnAdd + updateCSRAdd + Incremental
  if (!as.empty()) {
    int32_t *s = nullptr, *t = nullptr, *x = nullptr;
    const int32_t na = int32_t(as.size());
    SPM_TRY(stage(g, as.data(), at.data(), nullptr, na, &s, &t, &x));
    SPM_CHECK_HIP(hipMemsetAsync(w->fl[0], 0, size_t(V), g->stream));
    SPM_CHECK_HIP(hipMemsetAsync(w->fl[1], 0, size_t(V), g->stream));
    SPM_TRY(launch(g, "dsp_on_add", dsp_on_add, dim3((na + kDynBlock - 1) / kDynBlock), dim3(kDynBlock), 0, na,
                   (const int32_t*)s, (const int32_t*)t, (const int32_t*)nullptr, (const unsigned long long*)w->dp,
                   w->fl[0]));
    SPM_TRY(add_staged(g, s, t, nullptr, na));
    SPM_TRY(fixed_point(g, "Incremental", [&](int64_t k, IterFlags fl) {
      return launch(g, "dsp_push", dsp_push, dim3(wgrid), dim3(kDynBlock), 0, V, (const int32_t*)g->off,
                    (const int32_t*)g->adj, (const int32_t*)g->w, (const uint32_t*)w->dead, w->out[w->cur].r, w->dp,
                    w->fl[k & 1], w->fl[(k + 1) & 1], fl);
    }));
  }
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  return prof_flush(g);
}

// ---- dynamic TC, host side ---------------------------------------------------

// One counting phase (Decremental with sign -1, Incremental with +1) over the
// staged updates (s, t), on the graph as it stands.
static int dtc_phase(spm_graph* g, const int32_t* s, const int32_t* t, int32_t n, int sign) {
  DynWork* w = g->dyn;
  if (n == 0) return SPM_OK;
  size_t cap = 64;
  while (cap < 2 * size_t(n)) cap <<= 1;
  if (cap > w->tset_cap) {
    dfree(g, w->tset, sizeof(unsigned long long) * w->tset_cap);
    w->tset = nullptr;
    w->tset_cap = 0;
    SPM_TRY(dmalloc(g, (void**)&w->tset, sizeof(unsigned long long) * cap, "dynamic TC pair set"));
    w->tset_cap = cap;
  }
  if (size_t(n) + 1 > w->tchunk_cap) {
    dfree(g, w->tchunk, sizeof(int32_t) * w->tchunk_cap);
    w->tchunk = nullptr;
    w->tchunk_cap = 0;
    const size_t c = std::max<size_t>(size_t(n) + 1, 4097);
    SPM_TRY(dmalloc(g, (void**)&w->tchunk, sizeof(int32_t) * c, "dynamic TC work items"));
    w->tchunk_cap = c;
  }
  const uint32_t mask = uint32_t(cap - 1);
  const DiffRows d = w->out[w->cur].r;
  SPM_CHECK_HIP(hipMemsetAsync(w->tset, 0xFF, sizeof(unsigned long long) * cap, g->stream));
  const int ugrid = int((int64_t(n) + kDynBlock) / kDynBlock);  // n + 1 threads for dtc_chunks
  SPM_TRY(launch(g, "dtc_set_insert", dtc_set_insert, dim3(ugrid), dim3(kDynBlock), 0, n, s, t, w->tset, mask));
  SPM_TRY(launch(g, "dtc_chunks", dtc_chunks, dim3(ugrid), dim3(kDynBlock), 0, n, s, t, (const int32_t*)g->off, d,
                 w->tchunk));
  size_t need = 0;
  SPM_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, w->tchunk, w->tchunk, n + 1, g->stream));
  SPM_TRY(tmp_reserve(g, need));
  SPM_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(w->tmp, need, w->tchunk, w->tchunk, n + 1, g->stream));
  SPM_TRY(launch(g, "dtc_count", dtc_count, dim3(g->num_cus * 8), dim3(kDynBlock), 0, n, s, t,
                 (const int32_t*)w->tchunk, (const int32_t*)g->off, (const int32_t*)g->adj,
                 (const uint32_t*)w->dead, d, (const unsigned long long*)w->tset, mask, w->tcv));
  return launch(g, "dtc_apply", dtc_apply, dim3(1), dim3(1), 0, w->tcv, sign);
}

static int dtc_init(spm_graph* g) {
  uint64_t count = 0;
  SPM_TRY(tc_run(g, &count));  // staticTC, on the folded graph
  SPM_TRY(dyn_alloc(g));
  DynWork* w = g->dyn;
  if (!w->tcv) SPM_TRY(dmalloc(g, (void**)&w->tcv, sizeof(unsigned long long) * 4, "dynamic TC counters"));
  const unsigned long long h[4] = {0ull, 0ull, 0ull, static_cast<unsigned long long>(count)};
  SPM_CHECK_HIP(hipMemcpyAsync(w->tcv, h, sizeof h, hipMemcpyHostToDevice, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  w->tc_ready = true;
  return SPM_OK;
}

// One batch of DynTC (dynamicBatchTCV2:137-187): OnDelete + Decremental on the
// graph before the deletions, updateCSRDel, updateCSRAdd, OnAdd + Incremental.
static int dtc_batch(spm_graph* g, const char* type, const int32_t* us, const int32_t* ut, int32_t n) {
  DynWork* w = g->dyn;
  if (!w || !w->tc_ready) return set_error(SPM_EINVAL, "dynamic TC: call spm_dyntc_init first");
  std::vector<int32_t> ds, dt, as, at;
  for (int32_t i = 0; i < n; ++i) {
    if (type[i] == 'd') { ds.push_back(us[i]); dt.push_back(ut[i]); }
    else if (type[i] == 'a') { as.push_back(us[i]); at.push_back(ut[i]); }
    else return set_error(SPM_EINVAL, "dynamic TC: update type must be 'a' or 'd'");
  }
  if (!ds.empty()) {
    int32_t *s = nullptr, *t = nullptr, *x = nullptr;
    const int32_t nd = int32_t(ds.size());
    SPM_TRY(stage(g, ds.data(), dt.data(), nullptr, nd, &s, &t, &x));
    SPM_TRY(dtc_phase(g, s, t, nd, -1));
    SPM_TRY(delete_staged(g, s, t, nd));
  }
  if (!as.empty()) {
    int32_t *s = nullptr, *t = nullptr, *x = nullptr;
    const int32_t na = int32_t(as.size());
    SPM_TRY(stage(g, as.data(), at.data(), nullptr, na, &s, &t, &x));
    SPM_TRY(add_staged(g, s, t, nullptr, na));
    SPM_TRY(dtc_phase(g, s, t, na, +1));
  }
  return prof_flush(g);
}

// ---- dynamic PageRank, host side ----------------------------------------------

// One do-while of sweeps (staticPR with flag null, Decremental or
// Incremental): stops when !(diff > beta && iterCount < maxIter).
static int dpr_sweeps(spm_graph* g, const uint8_t* flag) {
  DynWork* w = g->dyn;
  const int32_t V = g->V;
  const float base = (1 - w->pr_delta) / float(V);  // (1 - delta) / g.num_nodes()
  const int vgrid = grid_for(g, V);
  const int lgrid = g->num_cus * 8, hgrid = w->pr_part_n - lgrid;
  const int sgrid = w->pr_part_n;
  float* h = reinterpret_cast<float*>(g->h_pinned);
  int32_t it = 0;
  for (;;) {
    SPM_TRY(launch(g, "dpr_contrib", dpr_contrib, dim3(vgrid), dim3(kDynBlock), 0, V, (const float*)w->pr,
                   (const float*)w->pr_deg, w->pr_c));
    SPM_TRY(launch(g, "dpr_sweep_light", dpr_sweep_light, dim3(lgrid), dim3(kDynBlock), 0, V,
                   (const uint8_t*)w->pr_isheavy, (const int32_t*)g->roff,
                   (const int32_t*)g->radj, (const uint32_t*)w->rdead, w->in[w->cur].r, (const float*)w->pr_c,
                   (const float*)w->pr, flag, w->pr_nxt, base, w->pr_delta, w->pr_part));
    SPM_TRY(launch(g, "dpr_sweep_heavy", dpr_sweep_heavy, dim3(hgrid), dim3(kDynBlock), 0, w->pr_nheavy,
                   (const int32_t*)w->pr_heavy, (const int32_t*)g->roff, (const int32_t*)g->radj,
                   (const uint32_t*)w->rdead, w->in[w->cur].r, (const float*)w->pr_c, (const float*)w->pr, flag,
                   w->pr_nxt, base, w->pr_delta, w->pr_part + lgrid));
    SPM_TRY(launch(g, "dpr_apply", dpr_apply, dim3(vgrid), dim3(kDynBlock), 0, V, flag, (const float*)w->pr_nxt,
                   w->pr));
    SPM_TRY(launch(g, "dpr_diff", dpr_diff, dim3(1), dim3(kDynBlock), 0, (const float*)w->pr_part, sgrid,
                   w->pr_part + sgrid));
    SPM_CHECK_HIP(hipMemcpyAsync(h, w->pr_part + sgrid, sizeof(float), hipMemcpyDeviceToHost, g->stream));
    SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    ++it;
    if (!(h[0] > w->pr_beta && it < w->pr_maxIter)) break;
  }
  w->pr_iters.push_back(it);
  return SPM_OK;
}

static int dpr_init_run(spm_graph* g, float beta, float delta, int32_t maxIter) {
  if (!g->roff) return set_error(SPM_ENOGRAPH, "dynamic PageRank needs the reverse CSR (rev_indexofNodes, srcList)");
  SPM_TRY(dyn_settle(g));  // the run starts from a clean CSR: its row lengths are count_outNbrs
  SPM_TRY(dyn_alloc(g));
  DynWork* w = g->dyn;
  const size_t V = size_t(g->V);
  if (!w->pr) {
    SPM_TRY(dmalloc(g, (void**)&w->pr, sizeof(float) * V, "dynamic PR ranks"));
    SPM_TRY(dmalloc(g, (void**)&w->pr_nxt, sizeof(float) * V, "dynamic PR next ranks"));
    SPM_TRY(dmalloc(g, (void**)&w->pr_c, sizeof(float) * V, "dynamic PR contributions"));
    SPM_TRY(dmalloc(g, (void**)&w->pr_deg, sizeof(float) * V, "dynamic PR base degrees"));
    SPM_TRY(dmalloc(g, (void**)&w->pr_flag, V, "dynamic PR flags"));
    w->pr_part_n = g->num_cus * 8 + g->num_cus * 8;  // light sweep workgroups, then heavy ones
    SPM_TRY(dmalloc(g, (void**)&w->pr_part, sizeof(float) * (size_t(w->pr_part_n) + 1), "dynamic PR diffs"));
  }
  {  // vertices whose in-row (starting CSR) is long take a wavefront each
    std::vector<int32_t> roff(V + 1), heavy;
    std::vector<uint8_t> mark(V, 0);
    SPM_CHECK_HIP(hipMemcpyAsync(roff.data(), g->roff, sizeof(int32_t) * (V + 1), hipMemcpyDeviceToHost, g->stream));
    SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    for (size_t v = 0; v < V; ++v)
      if (roff[v + 1] - roff[v] > kDprHeavy) {
        heavy.push_back(int32_t(v));
        mark[v] = 1;
      }
    std::stable_sort(heavy.begin(), heavy.end(), [&](int32_t a, int32_t b) {
      return roff[a + 1] - roff[a] > roff[b + 1] - roff[b];
    });
    if (!w->pr_isheavy) SPM_TRY(dmalloc(g, (void**)&w->pr_isheavy, V, "dynamic PR heavy marks"));
    SPM_CHECK_HIP(hipMemcpyAsync(w->pr_isheavy, mark.data(), V, hipMemcpyHostToDevice, g->stream));
    dfree(g, w->pr_heavy, sizeof(int32_t) * size_t(w->pr_nheavy));
    w->pr_heavy = nullptr;
    w->pr_nheavy = int32_t(heavy.size());
    if (!heavy.empty()) {
      SPM_TRY(dmalloc(g, (void**)&w->pr_heavy, sizeof(int32_t) * heavy.size(), "dynamic PR heavy rows"));
      SPM_CHECK_HIP(hipMemcpyAsync(w->pr_heavy, heavy.data(), sizeof(int32_t) * heavy.size(), hipMemcpyHostToDevice,
                                   g->stream));
    }
    SPM_CHECK_HIP(hipStreamSynchronize(g->stream));  // the host vectors end with this block
  }
  w->pr_beta = beta;
  w->pr_delta = delta;
  w->pr_maxIter = maxIter;
  w->pr_iters.clear();
  const float num_nodes = float(g->V);
  SPM_TRY(launch(g, "dpr_init", dpr_init, dim3(grid_for(g, g->V)), dim3(kDynBlock), 0, g->V, (const int32_t*)g->off,
                 1 / num_nodes, w->pr, w->pr_deg));
  SPM_TRY(dpr_sweeps(g, nullptr));  // staticPR
  w->pr_ready = true;
  return prof_flush(g);
}

// One batch of DynPR (dynamicBatchPR:97-135): OnDelete + propagateNodeFlags
// on the graph before the deletions, updateCSRDel, Decremental; OnAdd +
// propagateNodeFlags on the graph after them, updateCSRAdd, Incremental.
static int dpr_batch(spm_graph* g, const char* type, const int32_t* us, const int32_t* ut, int32_t n) {
  DynWork* w = g->dyn;
  if (!w || !w->pr_ready) return set_error(SPM_EINVAL, "dynamic PageRank: call spm_dynpr_init first");
  std::vector<int32_t> ds, dt, as, at;
  for (int32_t i = 0; i < n; ++i) {
    if (type[i] == 'd') { ds.push_back(us[i]); dt.push_back(ut[i]); }
    else if (type[i] == 'a') { as.push_back(us[i]); at.push_back(ut[i]); }
    else return set_error(SPM_EINVAL, "dynamic PageRank: update type must be 'a' or 'd'");
  }
  const int32_t V = g->V;
  const int vgrid = grid_for(g, V);
  for (int phase = 0; phase < 2; ++phase) {
    const std::vector<int32_t>& s = phase == 0 ? ds : as;
    const std::vector<int32_t>& t = phase == 0 ? dt : at;
    const int32_t m = int32_t(s.size());
    SPM_CHECK_HIP(hipMemsetAsync(w->pr_flag, 0, size_t(V), g->stream));
    int32_t *ps = nullptr, *pt = nullptr, *px = nullptr;
    if (m > 0) {
      SPM_TRY(stage(g, s.data(), t.data(), nullptr, m, &ps, &pt, &px));
      SPM_TRY(launch(g, "dpr_flag_dests", dpr_flag_dests, dim3((m + kDynBlock - 1) / kDynBlock), dim3(kDynBlock), 0,
                     m, (const int32_t*)pt, w->pr_flag));
    }
    SPM_TRY(launch(g, "dpr_propagate", dpr_propagate, dim3(vgrid), dim3(kDynBlock), 0, V, (const int32_t*)g->roff,
                   (const int32_t*)w->rdead_row, w->in[w->cur].r, w->pr_flag));
    if (m > 0) SPM_TRY(phase == 0 ? delete_staged(g, ps, pt, m) : add_staged(g, ps, pt, nullptr, m));
    SPM_TRY(dpr_sweeps(g, w->pr_flag));
  }
  return prof_flush(g);
}

}  // namespace spm

using namespace spm;

extern "C" {

int spm_graph_delete_edges(spm_graph* g, const int32_t* src, const int32_t* dst, int32_t n) {
  if (!g || n < 0 || (n > 0 && (!src || !dst))) return set_error(SPM_EINVAL, "spm_graph_delete_edges: bad argument");
  SPM_TRY(dyn_alloc(g));
  int32_t *s = nullptr, *t = nullptr, *x = nullptr;
  if (n == 0) return SPM_OK;
  SPM_TRY(stage(g, src, dst, nullptr, n, &s, &t, &x));
  return delete_staged(g, s, t, n);
}

int spm_graph_add_edges(spm_graph* g, const int32_t* src, const int32_t* dst, const int32_t* weight, int32_t n) {
  if (!g || n < 0 || (n > 0 && (!src || !dst))) return set_error(SPM_EINVAL, "spm_graph_add_edges: bad argument");
  SPM_TRY(dyn_alloc(g));
  int32_t *s = nullptr, *t = nullptr, *x = nullptr;
  if (n == 0) return SPM_OK;
  SPM_TRY(stage(g, src, dst, weight, n, &s, &t, &x));
  return add_staged(g, s, t, x, n);
}

int spm_graph_compact(spm_graph* g) {
  if (!g) return set_error(SPM_EINVAL, "spm_graph_compact: NULL handle");
  return dyn_settle(g);
}

int spm_graph_num_edges(spm_graph* g, int64_t* num_edges) {
  if (!g || !num_edges) return set_error(SPM_EINVAL, "spm_graph_num_edges: NULL argument");
  SPM_TRY(dyn_settle(g));
  *num_edges = g->E;
  return SPM_OK;
}

int spm_graph_export(spm_graph* g, int32_t* indexofNodes, int32_t* edgeList, int32_t* edgeLen,
                     int32_t* rev_indexofNodes, int32_t* srcList) {
  if (!g) return set_error(SPM_EINVAL, "spm_graph_export: NULL handle");
  SPM_TRY(dyn_settle(g));
  const size_t vb = sizeof(int32_t) * (size_t(g->V) + 1), eb = sizeof(int32_t) * size_t(g->E);
  if ((edgeLen && !g->w) || ((rev_indexofNodes || srcList) && !g->roff))
    return set_error(SPM_ENOGRAPH, "spm_graph_export: the handle has no such array");
  if (indexofNodes) SPM_CHECK_HIP(hipMemcpyAsync(indexofNodes, g->off, vb, hipMemcpyDeviceToHost, g->stream));
  if (edgeList && eb) SPM_CHECK_HIP(hipMemcpyAsync(edgeList, g->adj, eb, hipMemcpyDeviceToHost, g->stream));
  if (edgeLen && eb) SPM_CHECK_HIP(hipMemcpyAsync(edgeLen, g->w, eb, hipMemcpyDeviceToHost, g->stream));
  if (rev_indexofNodes) SPM_CHECK_HIP(hipMemcpyAsync(rev_indexofNodes, g->roff, vb, hipMemcpyDeviceToHost, g->stream));
  if (srcList && eb) SPM_CHECK_HIP(hipMemcpyAsync(srcList, g->radj, eb, hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  return SPM_OK;
}

int spm_dynsssp_init(spm_graph* g, int32_t src) {
  if (!g) return set_error(SPM_EINVAL, "spm_dynsssp_init: NULL handle");
  return dsp_init(g, src);
}

int spm_dynsssp_batch(spm_graph* g, const char* type, const int32_t* src, const int32_t* dst, int32_t n) {
  if (!g || n < 0 || (n > 0 && (!type || !src || !dst))) return set_error(SPM_EINVAL, "spm_dynsssp_batch: bad argument");
  for (int32_t i = 0; i < n; ++i)
    if (src[i] < 0 || src[i] >= g->V || dst[i] < 0 || dst[i] >= g->V)
      return set_error(SPM_EINVAL, "edge update names a vertex outside 0..num_nodes()-1");
  return dsp_batch(g, type, src, dst, n);
}

int spm_dynsssp_device(spm_graph* g, int32_t* d_dist, int32_t* d_parent) {
  if (!g || !g->dyn || g->dyn->src < 0) return set_error(SPM_EINVAL, "spm_dynsssp_device: call spm_dynsssp_init first");
  SPM_TRY(launch(g, "dsp_unpack", dsp_unpack, dim3(grid_for(g, g->V)), dim3(kDynBlock), 0, g->V,
                 (const unsigned long long*)g->dyn->dp, d_dist, d_parent));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  return SPM_OK;
}

int spm_dynsssp_result(spm_graph* g, int32_t* dist, int32_t* parent) {
  if (!g || !g->dyn || g->dyn->src < 0) return set_error(SPM_EINVAL, "spm_dynsssp_result: call spm_dynsssp_init first");
  int32_t* d = nullptr;
  SPM_TRY(dmalloc(g, (void**)&d, sizeof(int32_t) * 2 * size_t(g->V), "dynamic SSSP result"));
  int rc = launch(g, "dsp_unpack", dsp_unpack, dim3(grid_for(g, g->V)), dim3(kDynBlock), 0, g->V,
                  (const unsigned long long*)g->dyn->dp, d, d + g->V);
  if (rc == SPM_OK) {
    hipError_t e = hipSuccess;
    if (dist) e = hipMemcpyAsync(dist, d, sizeof(int32_t) * g->V, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess && parent) e = hipMemcpyAsync(parent, d + g->V, sizeof(int32_t) * g->V, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = set_error(SPM_EHIP, std::string("spm_dynsssp_result copy: ") + hipGetErrorString(e));
  }
  dfree(g, d, sizeof(int32_t) * 2 * size_t(g->V));
  return rc;
}

int spm_dyntc_init(spm_graph* g) {
  if (!g) return set_error(SPM_EINVAL, "spm_dyntc_init: NULL handle");
  return dtc_init(g);
}

int spm_dyntc_batch(spm_graph* g, const char* type, const int32_t* src, const int32_t* dst, int32_t n) {
  if (!g || n < 0 || (n > 0 && (!type || !src || !dst))) return set_error(SPM_EINVAL, "spm_dyntc_batch: bad argument");
  return dtc_batch(g, type, src, dst, n);
}

int spm_dyntc_count(spm_graph* g, int64_t* triangleCount) {
  if (!g || !triangleCount) return set_error(SPM_EINVAL, "spm_dyntc_count: NULL argument");
  DynWork* w = g->dyn;
  if (!w || !w->tc_ready) return set_error(SPM_EINVAL, "spm_dyntc_count: call spm_dyntc_init first");
  unsigned long long c = 0;
  SPM_CHECK_HIP(hipMemcpyAsync(&c, w->tcv + 3, sizeof c, hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  *triangleCount = static_cast<int64_t>(c);
  return SPM_OK;
}

int spm_dyn_tc(spm_graph* g, const char* type, const int32_t* usrc, const int32_t* udst, int32_t n, int32_t batchSize,
               int64_t* triangleCount) {
  if (!g || n < 0 || batchSize < 1 || !triangleCount) return set_error(SPM_EINVAL, "spm_dyn_tc: bad argument");
  SPM_TRY(spm_dyntc_init(g));
  for (int32_t k = 0; k < n; k += batchSize) {
    const int32_t m = std::min(batchSize, n - k);
    SPM_TRY(spm_dyntc_batch(g, type + k, usrc + k, udst + k, m));
  }
  return spm_dyntc_count(g, triangleCount);
}

int spm_dynpr_init(spm_graph* g, float beta, float delta, int32_t maxIter) {
  if (!g) return set_error(SPM_EINVAL, "spm_dynpr_init: NULL handle");
  return dpr_init_run(g, beta, delta, maxIter);
}

int spm_dynpr_batch(spm_graph* g, const char* type, const int32_t* src, const int32_t* dst, int32_t n) {
  if (!g || n < 0 || (n > 0 && (!type || !src || !dst))) return set_error(SPM_EINVAL, "spm_dynpr_batch: bad argument");
  return dpr_batch(g, type, src, dst, n);
}

int spm_dynpr_result(spm_graph* g, float* pageRank) {
  if (!g || !pageRank) return set_error(SPM_EINVAL, "spm_dynpr_result: NULL argument");
  DynWork* w = g->dyn;
  if (!w || !w->pr_ready) return set_error(SPM_EINVAL, "spm_dynpr_result: call spm_dynpr_init first");
  SPM_CHECK_HIP(hipMemcpyAsync(pageRank, w->pr, sizeof(float) * size_t(g->V), hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  return SPM_OK;
}

int spm_dynpr_iterations(spm_graph* g, int32_t* iterCounts, int32_t cap, int32_t* count) {
  if (!g || !count || (cap > 0 && !iterCounts)) return set_error(SPM_EINVAL, "spm_dynpr_iterations: bad argument");
  DynWork* w = g->dyn;
  if (!w || !w->pr_ready) return set_error(SPM_EINVAL, "spm_dynpr_iterations: call spm_dynpr_init first");
  *count = int32_t(w->pr_iters.size());
  for (int32_t i = 0; i < cap && i < *count; ++i) iterCounts[i] = w->pr_iters[size_t(i)];
  return SPM_OK;
}

int spm_dyn_pr(spm_graph* g, float beta, float delta, int32_t maxIter, float* pageRank, const char* type,
               const int32_t* usrc, const int32_t* udst, int32_t n, int32_t batchSize) {
  if (!g || n < 0 || batchSize < 1 || !pageRank) return set_error(SPM_EINVAL, "spm_dyn_pr: bad argument");
  SPM_TRY(spm_dynpr_init(g, beta, delta, maxIter));
  for (int32_t k = 0; k < n; k += batchSize) {
    const int32_t m = std::min(batchSize, n - k);
    SPM_TRY(spm_dynpr_batch(g, type + k, usrc + k, udst + k, m));
  }
  return spm_dynpr_result(g, pageRank);
}

int spm_dyn_sssp(spm_graph* g, int32_t src, const char* type, const int32_t* usrc, const int32_t* udst, int32_t n,
                 int32_t batchSize, int32_t* dist, int32_t* parent) {
  if (!g || n < 0 || batchSize < 1) return set_error(SPM_EINVAL, "spm_dyn_sssp: bad argument");
  SPM_TRY(spm_dynsssp_init(g, src));
  for (int32_t k = 0; k < n; k += batchSize) {
    const int32_t m = std::min(batchSize, n - k);
    SPM_TRY(spm_dynsssp_batch(g, type + k, usrc + k, udst + k, m));
  }
  return spm_dynsssp_result(g, dist, parent);
}

}  // extern "C"
