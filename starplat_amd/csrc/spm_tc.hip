// spm_tc.hip -- triangle counting for gfx950, the replacement of the code
// StarPlat generates for Compute_TC (graphcode/staticDSLCodes/
// triangle_counting_dsl:1-13):
//   forall v: forall u in nbrs(v), u < v: forall w in nbrs(v), w > v:
//     if (is_an_edge(u, w)) triangle_count += 1
// with is_an_edge -> graph::check_if_nbr, a linear scan per call
// (graph.hpp:233-287; OMP emission dsl_cpp_generator.cpp:2485-2495).
//
// Two exact forms of the same sum:
//  * symmetric graphs (every edge stored both ways, the usual TC input): each
//    triangle is counted exactly once by the DSL (at its middle vertex), so
//    the count equals the number of triangles, which is computed by forward
//    counting over a degree-ordered orientation: u -> w when
//    (deg u, u) < (deg w, w). Out-rows of the orientation are at most
//    sqrt(2m) long, which bounds the work of every hub.
//  * any other graph: the DSL's own sum, item = CSR edge (v,u) with u < v,
//    count |nbrs(u) intersect {w in nbrs(v) : w > v}|.
// Both reduce to "size of the intersection of two ascending rows" per item.
// Symmetric graphs are counted per oriented source vertex with an LDS hash
// set of its row (tc_count_hash, the default). The search form
// (tc_count, schedule tc_mode 0, and the general sum) has a wavefront take 64
// items, flatten the elements of each item's shorter row across its lanes
// (wave prefix sum, LDS lookup) and binary-search each element in the longer
// row; hits are counted with ballot + popcount. The integer total is exact,
// so results are bit-identical to the reference.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "spm_internal.hpp"

namespace spm {

constexpr int kTcBlock = 256;
constexpr int kTcWaves = kTcBlock / kWave;

struct TcWork {
  int checked = 0;       // rows verified sorted; symmetric flag valid
  int symmetric = 0;
  int oriented = 0;      // oriented CSR built
  int32_t* dplus = nullptr;   // [V+1] oriented out-degree, then offsets
  int32_t* padj = nullptr;    // [m] oriented rows
  int32_t* psrc = nullptr;    // [m] oriented edge source
  int32_t* csrc = nullptr;    // [E] CSR edge source (general form)
  int64_t m = 0;
  int32_t* maxd = nullptr;    // [1] longest oriented row (device)
  int32_t max_dplus = 0;
  unsigned long long* acc = nullptr;  // [3] total, flags, item counter
  int2* items = nullptr;      // [nitems] hashed-count work items (u, first neighbour)
  uint32_t nitems = 0;
  int items_built = 0;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
};

__device__ __forceinline__ bool tc_above(int32_t du, int32_t u, int32_t dw, int32_t w) {
  return dw > du || (dw == du && w > u);
}

// First index in a[0..n) whose value is >= x.
__device__ __forceinline__ int32_t lower_bound_dev(const int32_t* __restrict__ a, int32_t n, int32_t x) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Checks ascending rows and symmetry; flags[0] |= 1 on an unsorted row,
// flags[0] |= 2 when some edge (u,w) has no (w,u).
// Flags: 1 = a row is not strictly ascending; 2 = some edge has no reverse
// edge (not symmetric); 4 = the rows differ from the reverse CSR's (symmetry
// undecided by the streaming test). With the reverse CSR at hand (kRev), a
// graph is symmetric exactly when every out-row equals the in-row of the same
// vertex, which is a streaming comparison; without it, or when the rows
// differ, each edge's reverse is looked up by binary search.
template <bool kRev>
__global__ __launch_bounds__(kTcBlock) void tc_check(int32_t V, const int32_t* __restrict__ off,
                                                      const int32_t* __restrict__ adj,
                                                      const int32_t* __restrict__ roff,
                                                      const int32_t* __restrict__ radj,
                                                      unsigned long long* flags) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * kTcWaves;
  unsigned long long f = 0;
  for (int64_t x_ = int64_t(blockIdx.x) * kTcWaves + threadIdx.x / kWave; x_ < V; x_ += nw) {
    const int64_t u = spread_vertex(x_, V);
    const int32_t b = off[u], e = off[u + 1];
    if (kRev && (roff[u] != b || roff[u + 1] != e)) f |= 4;
    for (int32_t i = b + lane; i < e; i += kWave) {
      const int32_t w = adj[i];
      if (i > b && adj[i - 1] >= w) f |= 1;
      if (kRev) {
        if (!(f & 4) && radj[i] != w) f |= 4;
      } else {
        const int32_t wb = off[w], wl = off[w + 1] - wb;
        const int32_t p = lower_bound_dev(adj + wb, wl, int32_t(u));
        if (p >= wl || adj[wb + p] != u) f |= 2;
      }
    }
  }
  // One atomic per wavefront with any flag.
  unsigned long long any = 0;
  for (int k = 0; k < 3; ++k)
    if (__ballot((f >> k) & 1ull)) any |= 1ull << k;
  if (lane == 0 && any) atomicOr(flags, any);
}

__global__ __launch_bounds__(kTcBlock) void tc_orient_count(int32_t V, const int32_t* __restrict__ off,
                                                             const int32_t* __restrict__ adj,
                                                             int32_t* __restrict__ dplus) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * kTcWaves;
  for (int64_t x_ = int64_t(blockIdx.x) * kTcWaves + threadIdx.x / kWave; x_ < V; x_ += nw) {
    const int64_t u = spread_vertex(x_, V);
    const int32_t b = off[u], e = off[u + 1], du = e - b;
    int32_t c = 0;
    for (int32_t i = b + lane; i < e; i += kWave) {
      const int32_t w = adj[i];
      c += tc_above(du, int32_t(u), off[w + 1] - off[w], w) ? 1 : 0;
    }
    c = wave_sum(c);
    if (lane == 0) dplus[u] = c;
  }
}

// Oriented rows keep ascending order: lanes write their kept elements at the
// ballot prefix of the lanes below them.
__global__ __launch_bounds__(kTcBlock) void tc_orient_fill(int32_t V, const int32_t* __restrict__ off,
                                                            const int32_t* __restrict__ adj,
                                                            const int32_t* __restrict__ poff,
                                                            int32_t* __restrict__ padj,
                                                            int32_t* __restrict__ psrc) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * kTcWaves;
  for (int64_t x_ = int64_t(blockIdx.x) * kTcWaves + threadIdx.x / kWave; x_ < V; x_ += nw) {
    const int64_t u = spread_vertex(x_, V);
    const int32_t b = off[u], e = off[u + 1], du = e - b;
    int32_t out = poff[u];
    for (int32_t i0 = b; i0 < e; i0 += kWave) {
      const int32_t i = i0 + lane;
      bool keep = false;
      int32_t w = 0;
      if (i < e) {
        w = adj[i];
        keep = tc_above(du, int32_t(u), off[w + 1] - off[w], w);
      }
      const unsigned long long m = __ballot(keep);
      if (keep) {
        const int32_t pos = out + __popcll(m & ((1ull << lane) - 1ull));
        padj[pos] = w;
        psrc[pos] = int32_t(u);
      }
      out += __popcll(m);
    }
  }
}

__global__ void tc_expand_src(int32_t V, const int32_t* __restrict__ off, int32_t* __restrict__ csrc) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * kTcWaves;
  for (int64_t x_ = int64_t(blockIdx.x) * kTcWaves + threadIdx.x / kWave; x_ < V; x_ += nw) {
    const int64_t u = spread_vertex(x_, V);
    for (int32_t i = off[u] + lane; i < off[u + 1]; i += kWave) csrc[i] = int32_t(u);
  }
}

// Intersection counting. ORIENTED: item e is oriented edge (src[e], adj[e])
// and both rows are oriented rows. Otherwise item e is CSR edge (v=src[e],
// u=adj[e]); it counts only when u < v, against row(u) and the part of
// row(v) above v.
template <bool ORIENTED>
__global__ __launch_bounds__(kTcBlock) void tc_count(int64_t nitems, const int32_t* __restrict__ roff,
                                                      const int32_t* __restrict__ radj,
                                                      const int32_t* __restrict__ src,
                                                      unsigned long long* __restrict__ total) {
  __shared__ uint32_t s_pre[kTcWaves][kWave];
  __shared__ int32_t s_sp[kTcWaves][kWave];
  __shared__ int32_t s_bp[kTcWaves][kWave];
  __shared__ int32_t s_bl[kTcWaves][kWave];
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  unsigned long long cnt = 0;
  const int64_t nw = int64_t(gridDim.x) * kTcWaves;
  for (int64_t base = (int64_t(blockIdx.x) * kTcWaves + wid) * kWave; base < nitems; base += nw * kWave) {
    const int64_t e = base + lane;
    uint32_t work = 0;
    int32_t sp = 0, bp = 0, bl = 0;
    if (e < nitems) {
      const int32_t x = src[e], y = radj[e];
      int32_t ap, al, cp, cl;
      bool live = true;
      if (ORIENTED) {
        ap = roff[x];
        al = roff[x + 1] - ap;
        cp = roff[y];
        cl = roff[y + 1] - cp;
      } else {
        live = y < x;  // u = y < v = x
        ap = roff[y];
        al = roff[y + 1] - ap;
        const int32_t vb = roff[x], vl = roff[x + 1] - vb;
        const int32_t s0 = live ? lower_bound_dev(radj + vb, vl, x + 1) : vl;
        cp = vb + s0;
        cl = vl - s0;
      }
      if (live && al > 0 && cl > 0) {
        if (al <= cl) { sp = ap; work = uint32_t(al); bp = cp; bl = cl; }
        else { sp = cp; work = uint32_t(cl); bp = ap; bl = al; }
      }
    }
    const uint32_t incl = wave_incl_scan(work);
    const uint32_t tot = __shfl(incl, kWave - 1, kWave);
    s_pre[wid][lane] = incl - work;
    s_sp[wid][lane] = sp;
    s_bp[wid][lane] = bp;
    s_bl[wid][lane] = bl;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t t0 = 0; t0 < tot; t0 += kWave) {
      const uint32_t t = t0 + lane;
      bool hit = false;
      if (t < tot) {
        int lo = 0, hi = kWave - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_pre[wid][mid] <= t) lo = mid;
          else hi = mid - 1;
        }
        const int32_t xv = radj[s_sp[wid][lo] + int32_t(t - s_pre[wid][lo])];
        const int32_t* big = radj + s_bp[wid][lo];
        const int32_t n = s_bl[wid][lo];
        const int32_t p = lower_bound_dev(big, n, xv);
        hit = p < n && big[p] == xv;
      }
      cnt += __popcll(__ballot(hit));
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0 && cnt) atomicAdd(total, cnt);
}

// Oriented rows, hashed per source vertex. Work items are (u, c): the
// oriented neighbours N+(u)[c, c+64) of a vertex u with at least two (a
// triangle needs two). A wavefront takes items in batches from a counter and
// keeps a hash set of N+(u) in its slice of LDS, rebuilt only when the item's
// vertex changes (items of one vertex are adjacent). It then walks the rows
// N+(v) of the item's v's one at a time, all 64 lanes striding over a row
// (coalesced loads, kTcUnroll in flight per lane), and probes each element:
// every hit is one triangle u < v < w in the orientation. At RMAT-21, 99% of
// the probed elements lie in rows of 64 or more, so striding rows keeps the
// lanes busy without locating each element's row.
// The set is bucketed: 4-slot buckets (one 16-byte LDS read), at least d+(u)
// buckets (load <= 1/4 while 4 d+(u) <= kTcHash slots, <= 1/2 at the size
// limit), keys placed in the first free slot of the first non-full bucket
// from the hashed one. A bucket with a free slot ends a search, so a miss
// almost always costs one read.
// The host launches this kernel only when the longest oriented row fits the
// set (2 d+ <= kTcHash; the orientation bounds rows by sqrt(2m)), otherwise
// it runs the search; a row that does not fit is counted in `overflow`. Taking
// items from a counter matters: with a static vertex-to-wavefront map, RMAT
// ids with many zero low bits (the heavy vertices) land on the same
// wavefronts (24x the mean load at RMAT-21).
constexpr int kTcHash = 2048;  // slots per wavefront (int32 keys, -1 empty)
constexpr int kTcItem = kWave;  // oriented neighbours per work item (one per lane)
// Items per counter bump: kTcGrab once fewer than kTcGrabDiv per wavefront
// are left (the last items set the tail), up to kTcGrabMax before that (one
// counter shared by every wavefront; RMAT-21 17.34 -> 17.15 ms with 32 / 16,
// 8 / 4 17.73, 128 / 64 22.1, a fixed 1 19.6 and 8 18.0).
constexpr uint32_t kTcGrab = 2;
constexpr uint32_t kTcGrabMax = 32;
constexpr uint32_t kTcGrabDiv = 16;
constexpr int kTcUnroll = 8;     // row elements per lane per step

__global__ void tc_item_count(int32_t V, const int32_t* __restrict__ poff, int32_t* __restrict__ nitem) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t u = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; u < V; u += stride) {
    const int32_t du = poff[u + 1] - poff[u];
    nitem[u] = du >= 2 ? (du + kTcItem - 1) / kTcItem : 0;
  }
}

__global__ void tc_item_fill(int32_t V, const int32_t* __restrict__ ioff, int2* __restrict__ items) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t u = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; u < V; u += stride)
    for (int32_t k = ioff[u]; k < ioff[u + 1]; ++k) items[k] = make_int2(int32_t(u), (k - ioff[u]) * kTcItem);
}

__global__ __launch_bounds__(kTcBlock) void tc_count_hash(uint32_t nitems, const int2* __restrict__ items,
                                                          const int32_t* __restrict__ poff,
                                                          const int32_t* __restrict__ padj,
                                                          unsigned long long* __restrict__ total,
                                                          unsigned long long* __restrict__ overflow,
                                                          uint32_t* __restrict__ next) {
  __shared__ __align__(16) int32_t s_tab[kTcWaves][kTcHash];
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  int32_t* tab = s_tab[wid];
  const int4* bkt = reinterpret_cast<const int4*>(tab);
  unsigned long long cnt = 0;  // per lane
  int32_t cur = -1, ub = 0, du = 0, shift = 0;
  uint32_t bmask = 0;
  bool fits = false;
  const uint32_t nwaves = gridDim.x * (blockDim.x / kWave);
  uint32_t grab = kTcGrabMax;
  for (;;) {
    uint32_t b = 0;
    if (lane == 0) b = atomicAdd(next, grab);
    b = __shfl(b, 0, kWave);
    if (b >= nitems) break;
    const uint32_t e = min(b + grab, nitems);
    grab = max(kTcGrab, min(kTcGrabMax, (nitems - e) / (nwaves * kTcGrabDiv)));
    for (uint32_t it = b; it < e; ++it) {
      const int2 item = items[it];
      if (item.x != cur) {
        cur = item.x;
        ub = poff[cur];
        du = poff[cur + 1] - ub;
        fits = 2 * du <= kTcHash;
        if (!fits) {  // excluded by the host's check of the longest row
          if (lane == 0) atomicAdd(overflow, 1ull);
          continue;
        }
        int lg = 4;  // log2 of the bucket count
        while ((1 << lg) < du && (4 << lg) < kTcHash) ++lg;
        bmask = (1u << lg) - 1u;
        shift = 32 - lg;  // Fibonacci hashing: the top lg bits of x * 2^32/phi
        __builtin_amdgcn_wave_barrier();  // the previous vertex's probes are done
        for (int i = lane; i < (4 << lg); i += kWave) tab[i] = -1;
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < du; i += kWave) {
          const int32_t x = padj[ub + i];
          uint32_t h = (uint32_t(x) * 2654435769u) >> shift;
          for (;;) {
            int s = 0;
            while (s < 4 && atomicCAS(&tab[4 * h + s], -1, x) != -1) ++s;
            if (s < 4) break;
            h = (h + 1) & bmask;
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
      if (!fits) continue;
      const int32_t i = item.y + lane;
      int32_t rb = 0, len = 0;
      if (i < du) {
        const int32_t v = padj[ub + i];
        rb = poff[v];
        len = poff[v + 1] - rb;
      }
      const int nv = min(kTcItem, du - item.y);
      for (int k = 0; k < nv; ++k) {
        const int32_t r0 = __builtin_amdgcn_readlane(rb, k);
        const int32_t n = __builtin_amdgcn_readlane(len, k);
        for (int32_t j0 = 0; j0 < n; j0 += kTcUnroll * kWave) {
          int32_t wq[kTcUnroll];
#pragma unroll
          for (int q = 0; q < kTcUnroll; ++q) {
            const int32_t j = j0 + q * kWave + lane;
            wq[q] = j < n ? padj[r0 + j] : -1;
          }
#pragma unroll
          for (int q = 0; q < kTcUnroll; ++q) {
            const int32_t w = wq[q];
            if (w < 0) continue;
            uint32_t h = (uint32_t(w) * 2654435769u) >> shift;
            for (;;) {
              const int4 c = bkt[h];
              if (c.x == w || c.y == w || c.z == w || c.w == w) {
                ++cnt;
                break;
              }
              if (min(min(c.x, c.y), min(c.z, c.w)) < 0) break;  // a free slot: w is absent
              h = (h + 1) & bmask;
            }
          }
        }
      }
    }
  }
  cnt = wave_sum(cnt);
  if (lane == 0 && cnt) atomicAdd(total, cnt);
}

static int tc_prepare(spm_graph* g) {
  if (!g->tc) g->tc = new TcWork();
  TcWork* w = g->tc;
  const int32_t V = g->V;
  const int grid = g->num_cus * 8;
  if (!w->acc) SPM_TRY(dmalloc(g, (void**)&w->acc, 3 * sizeof(unsigned long long), "tc counters"));
  if (!w->maxd) SPM_TRY(dmalloc(g, (void**)&w->maxd, sizeof(int32_t), "tc max row"));
  if (!w->checked) {
    SPM_CHECK_HIP(hipMemsetAsync(w->acc, 0, 2 * sizeof(unsigned long long), g->stream));
    unsigned long long f = 4;
    if (g->roff) {
      SPM_TRY(launch(g, "tc_check", tc_check<true>, dim3(grid), dim3(kTcBlock), 0, V, (const int32_t*)g->off,
                     (const int32_t*)g->adj, (const int32_t*)g->roff, (const int32_t*)g->radj, w->acc + 1));
      SPM_CHECK_HIP(hipMemcpyAsync(&f, w->acc + 1, sizeof f, hipMemcpyDeviceToHost, g->stream));
      SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    }
    if (f & 4) {
      SPM_CHECK_HIP(hipMemsetAsync(w->acc + 1, 0, sizeof(unsigned long long), g->stream));
      SPM_TRY(launch(g, "tc_check", tc_check<false>, dim3(grid), dim3(kTcBlock), 0, V, (const int32_t*)g->off,
                     (const int32_t*)g->adj, (const int32_t*)nullptr, (const int32_t*)nullptr, w->acc + 1));
      SPM_CHECK_HIP(hipMemcpyAsync(&f, w->acc + 1, sizeof f, hipMemcpyDeviceToHost, g->stream));
      SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    }
    if (f & 1) return set_error(SPM_EINVAL, "TC needs ascending adjacency rows (graph.hpp sorts them)");
    w->symmetric = (f & 2) == 0;
    w->checked = 1;
  }
  if (w->symmetric && !w->oriented) {
    SPM_TRY(dmalloc(g, (void**)&w->dplus, sizeof(int32_t) * (size_t(V) + 1), "tc oriented offsets"));
    SPM_TRY(launch(g, "tc_orient_count", tc_orient_count, dim3(grid), dim3(kTcBlock), 0, V,
                   (const int32_t*)g->off, (const int32_t*)g->adj, w->dplus));
    SPM_CHECK_HIP(hipMemsetAsync(w->dplus + V, 0, sizeof(int32_t), g->stream));
    {  // longest oriented row, to choose the hashed count kernel
      size_t need_max = 0;
      SPM_CHECK_HIP(hipcub::DeviceReduce::Max(nullptr, need_max, w->dplus, w->maxd, V, g->stream));
      if (need_max > w->tmp_bytes) {
        dfree(g, w->tmp, w->tmp_bytes);
        w->tmp = nullptr;
        w->tmp_bytes = 0;
        SPM_TRY(dmalloc(g, &w->tmp, need_max, "tc reduce scratch"));
        w->tmp_bytes = need_max;
      }
      SPM_CHECK_HIP(hipcub::DeviceReduce::Max(w->tmp, need_max, w->dplus, w->maxd, V, g->stream));
      SPM_CHECK_HIP(hipMemcpyAsync(&w->max_dplus, w->maxd, sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
      SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    }
    size_t need = 0;
    SPM_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, w->dplus, w->dplus, V + 1, g->stream));
    if (need > w->tmp_bytes) {
      dfree(g, w->tmp, w->tmp_bytes);
      w->tmp = nullptr;
      w->tmp_bytes = 0;
      SPM_TRY(dmalloc(g, &w->tmp, need, "tc scan scratch"));
      w->tmp_bytes = need;
    }
    SPM_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(w->tmp, need, w->dplus, w->dplus, V + 1, g->stream));
    int32_t m = 0;
    SPM_CHECK_HIP(hipMemcpyAsync(&m, w->dplus + V, sizeof m, hipMemcpyDeviceToHost, g->stream));
    SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    w->m = m;
    SPM_TRY(dmalloc(g, (void**)&w->padj, sizeof(int32_t) * size_t(m), "tc oriented rows"));
    SPM_TRY(dmalloc(g, (void**)&w->psrc, sizeof(int32_t) * size_t(m), "tc oriented sources"));
    SPM_TRY(launch(g, "tc_orient_fill", tc_orient_fill, dim3(grid), dim3(kTcBlock), 0, V,
                   (const int32_t*)g->off, (const int32_t*)g->adj, (const int32_t*)w->dplus, w->padj, w->psrc));
    w->oriented = 1;
  }
  if (!w->symmetric && !w->csrc) {
    SPM_TRY(dmalloc(g, (void**)&w->csrc, sizeof(int32_t) * size_t(g->E), "tc edge sources"));
    SPM_TRY(launch(g, "tc_expand_src", tc_expand_src, dim3(grid), dim3(kTcBlock), 0, V, (const int32_t*)g->off, w->csrc));
  }
  return SPM_OK;
}

// Work items of the hashed count, built once per orientation.
static int tc_build_items(spm_graph* g) {
  TcWork* w = g->tc;
  if (w->items_built) return SPM_OK;
  const int32_t V = g->V;
  int32_t* ioff = nullptr;
  SPM_TRY(dmalloc(g, (void**)&ioff, sizeof(int32_t) * (size_t(V) + 1), "tc item offsets"));
  const int igrid = int(std::min<int64_t>((int64_t(V) + 255) / 256, int64_t(g->num_cus) * 16));
  int rc = launch(g, "tc_item_count", tc_item_count, dim3(igrid), dim3(256), 0, V, (const int32_t*)w->dplus, ioff);
  size_t need = 0;
  if (rc == SPM_OK && hipMemsetAsync(ioff + V, 0, sizeof(int32_t), g->stream) != hipSuccess)
    rc = set_error(SPM_EHIP, "tc items: memset");
  if (rc == SPM_OK && hipcub::DeviceScan::ExclusiveSum(nullptr, need, ioff, ioff, V + 1, g->stream) != hipSuccess)
    rc = set_error(SPM_EHIP, "tc items: scan size");
  if (rc == SPM_OK && need > w->tmp_bytes) {
    dfree(g, w->tmp, w->tmp_bytes);
    w->tmp = nullptr;
    w->tmp_bytes = 0;
    rc = dmalloc(g, &w->tmp, need, "tc scan scratch");
    if (rc == SPM_OK) w->tmp_bytes = need;
  }
  if (rc == SPM_OK && hipcub::DeviceScan::ExclusiveSum(w->tmp, need, ioff, ioff, V + 1, g->stream) != hipSuccess)
    rc = set_error(SPM_EHIP, "tc items: scan");
  int32_t n = 0;
  if (rc == SPM_OK && (hipMemcpyAsync(&n, ioff + V, sizeof n, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
                       hipStreamSynchronize(g->stream) != hipSuccess))
    rc = set_error(SPM_EHIP, "tc items: count read-back");
  if (rc == SPM_OK && n > 0) {
    rc = dmalloc(g, (void**)&w->items, sizeof(int2) * size_t(n), "tc work items");
    if (rc == SPM_OK) {
      w->nitems = uint32_t(n);
      rc = launch(g, "tc_item_fill", tc_item_fill, dim3(igrid), dim3(256), 0, V, (const int32_t*)ioff, w->items);
    }
  }
  if (rc == SPM_OK && hipStreamSynchronize(g->stream) != hipSuccess) rc = set_error(SPM_EHIP, "tc items: sync");
  dfree(g, ioff, sizeof(int32_t) * (size_t(V) + 1));
  if (rc == SPM_OK) w->items_built = 1;
  return rc;
}

int tc_release(spm_graph* g) {
  TcWork* w = g->tc;
  if (!w) return SPM_OK;
  dfree(g, w->dplus, sizeof(int32_t) * (size_t(g->V) + 1));
  dfree(g, w->padj, sizeof(int32_t) * size_t(w->m));
  dfree(g, w->psrc, sizeof(int32_t) * size_t(w->m));
  dfree(g, w->csrc, sizeof(int32_t) * size_t(g->E));
  dfree(g, w->acc, 3 * sizeof(unsigned long long));
  dfree(g, w->items, sizeof(int2) * w->nitems);
  dfree(g, w->maxd, sizeof(int32_t));
  dfree(g, w->tmp, w->tmp_bytes);
  delete w;
  g->tc = nullptr;
  return SPM_OK;
}

int tc_run(spm_graph* g, uint64_t* out) {
  SPM_TRY(dyn_settle(g));
  SPM_TRY(tc_prepare(g));
  TcWork* w = g->tc;
  SPM_CHECK_HIP(hipMemsetAsync(w->acc, 0, sizeof(unsigned long long), g->stream));
  const int grid = g->num_cus * 8;
  // tc_mode 1 (default): hashed per-vertex count, 17.4 ms at RMAT-21. It
  // visits sum d+(v) = 11.3G row elements, each one coalesced load and about
  // one LDS read, against sum min(d+(u), d+(v)) = 4.8G binary searches of
  // ~10 dependent loads each in mode 0 (41.6 ms). Mode 0 also serves rows
  // too long for the LDS set and non-symmetric graphs.
  const bool hashed = w->symmetric && g->sched.tc_mode == 1 && 2 * w->max_dplus <= kTcHash;
  if (hashed) {
    SPM_TRY(tc_build_items(g));
    SPM_CHECK_HIP(hipMemsetAsync(w->acc + 1, 0, 2 * sizeof(unsigned long long), g->stream));
    // one resident wave of workgroups: work is taken from the item counter
    if (w->nitems)
      SPM_TRY(launch(g, "tc_count_hash", tc_count_hash, dim3(g->num_cus * 4), dim3(kTcBlock), 0, w->nitems,
                     (const int2*)w->items, (const int32_t*)w->dplus, (const int32_t*)w->padj, w->acc, w->acc + 1,
                     reinterpret_cast<uint32_t*>(w->acc + 2)));
  } else if (w->symmetric) {
    if (w->m) SPM_TRY(launch(g, "tc_count", tc_count<true>, dim3(grid), dim3(kTcBlock), 0, w->m,
                            (const int32_t*)w->dplus, (const int32_t*)w->padj, (const int32_t*)w->psrc, w->acc));
  } else {
    if (g->E) SPM_TRY(launch(g, "tc_count_general", tc_count<false>, dim3(grid), dim3(kTcBlock), 0, g->E,
                            (const int32_t*)g->off, (const int32_t*)g->adj, (const int32_t*)w->csrc, w->acc));
  }
  unsigned long long t[2] = {0, 0};
  SPM_CHECK_HIP(hipMemcpyAsync(t, w->acc, sizeof t, hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  if (hashed && t[1] != 0)
    return set_error(SPM_EHIP, "tc: oriented row larger than the LDS hash (inconsistent orientation)");
  *out = t[0];
  SPM_TRY(prof_flush(g));
  return SPM_OK;
}

}  // namespace spm

using namespace spm;

extern "C" {

int spm_compute_tc(spm_graph* g, uint64_t* triangle_count) {
  if (!g || !triangle_count) return set_error(SPM_EINVAL, "spm_compute_tc: NULL argument");
  return tc_run(g, triangle_count);
}

}  // extern "C"
