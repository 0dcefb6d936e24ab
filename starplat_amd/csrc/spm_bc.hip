// spm_bc.hip -- batched Brandes betweenness centrality for gfx950, the
// replacement of the code StarPlat generates for Compute_BC
// (graphcode/staticDSLCodes/bc_dslV2:1-21):
//   for src in sourceSet: delta = 0; sigma = 0; sigma[src] = 1;
//     iterateInBFS(v from src): for w in nbrs(v), bfsDist[w]==bfsDist[v]+1:
//       sigma[w] += sigma[v]
//     iterateInReverse(v != src): for w in nbrs(v) on the BFS DAG:
//       delta[v] += sigma[v]/sigma[w] * (1 + delta[w]);  BC[v] += delta[v]
// (OMP emission: backend_omp/dsl_cpp_generator.cpp:66-193; CUDA kernels with
//  a host round trip per level: backend_cuda/...cpp:275-570).
//
// Design. Sources are processed 64 at a time, one per lane of a wavefront:
// sigma is FP64 in vertex-major rows of 64 entries (512 B), so the lanes of a
// wavefront working on one vertex touch one contiguous row. BFS levels are kept as 64-bit masks per vertex and level
// (bit s set = vertex at this level for source s), plus a visited mask.
// Each level runs:
//   scan   one lane per vertex: next = OR of the in-neighbours' current
//          masks, minus visited (bottom-up, no atomics); appends the vertex
//          to the level's list when next != 0; or, while the frontier is
//          small, push: the frontier's out-edges OR their sources into the
//          targets' next masks (top-down, atomics), with the same result;
//   sigma  one wavefront per listed vertex: for every in-neighbour whose
//          current mask meets next, lanes of those sources add the
//          neighbour's sigma (integers below 2^53, so exact in any order).
// The reverse pass walks the level lists backwards, one wavefront per vertex:
// lane s sums the FP32 terms sigma_v * coef_w over out-neighbours w at the
// next level for s (ascending neighbour order, as the reference), where
// coef_w = (1 + delta_w) / sigma_w was left by w's own step in an FP32 row
// [V][64] -- one multiply and one 4-byte read per DAG edge instead of a
// division and two reads (a batch whose path counts reach 2^96, where the
// reciprocal would leave FP32's normal range, keeps FP64 coefficients in
// place of sigma instead). A gather touches on average ~2.6 of a 512-byte
// FP64 row's four 128-byte lines for ~6 sources at RMAT/ER sizes; the FP32
// row halves the lines it can touch (the coefficient's rounding, 2^-24
// relative, is below the FP32 accumulation's own). The 64 per-source
// dependencies are reduced into BC[v] (pairwise over the batch, then once
// per batch).
#include <algorithm>
#include <cstring>
#include <vector>

#include "spm_internal.hpp"

namespace spm {

constexpr int kBcBlock = 256;
constexpr int kBcWaves = kBcBlock / kWave;
constexpr int kBatch = 64;
// A level whose frontier holds fewer than V / kPushRatio vertices is expanded
// top-down (its out-edges) rather than bottom-up (every vertex's in-edges).
constexpr uint64_t kPushRatio = 14;
// Active neighbours whose sigma rows one wavefront loads per step.
constexpr int kGather = 4;
// In-edges one lane of the bottom-up scan reads per step.
constexpr int kScanStep = 4;

// The forward pass runs its levels from device-side state, so the host reads
// it once per window of levels rather than once per level.
// The cursor sits on a cache line of its own: the flags are read by every
// workgroup of every level's launches, the cursor is hit by atomics.
struct BcCtl {
  unsigned int cnt;  // list append cursor
  unsigned int pad_[31];
  int done;          // no vertex reached at the next level
  int levels;        // deepest non-empty level, once done
  int push;          // the current level expands top-down
  int wide;          // some sigma reached 2^96: the reverse pass keeps FP64 coefficients
};

struct BcWork {
  double* sigma = nullptr;          // [V][64]
  float* coef = nullptr;            // [V][64] (1 + delta) / sigma of the reverse pass
  unsigned long long* visited = nullptr;  // [V]
  unsigned long long* masks = nullptr;    // [mask_levels][V]
  int32_t* lists = nullptr;         // level lists, concatenated
  BcCtl* ctl = nullptr;             // the forward pass's device-side control
  unsigned int* lstart = nullptr;   // [lstart_cap] level d's list is lstart[d] .. lstart[d+1]
  size_t lstart_cap = 0;
  int32_t* srcbuf = nullptr;        // [64]
  size_t mask_levels = 0;
  size_t list_cap = 0;
};

// Clears the visited and level-0 masks; the sources' bits are set by bc_seed,
// a separate launch, so no workgroup can clear a bit another one has set.
__global__ void bc_init(int32_t V, unsigned long long* __restrict__ visited,
                        unsigned long long* __restrict__ mask0, const int32_t* __restrict__ src, int nb,
                        int32_t* __restrict__ list0, BcCtl* ctl, unsigned int* __restrict__ lstart) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    visited[v] = 0ull;
    mask0[v] = 0ull;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->cnt = unsigned(nb);
    ctl->done = 0;
    ctl->levels = 0;
    ctl->push = uint64_t(nb) * kPushRatio < uint64_t(V);
    ctl->wide = 0;
    lstart[0] = 0;
    lstart[1] = unsigned(nb);
  }
  if (blockIdx.x == 0 && int(threadIdx.x) < nb) list0[threadIdx.x] = src[threadIdx.x];
}

// Flag reads of the control block: written by one thread of an earlier
// launch (visible after the launch boundary), fixed while a kernel runs.
__device__ __forceinline__ int ctl_load(const int* p) { return *p; }

// End of level d's expansion (one thread): the list of level d+1 ends at the
// cursor; an empty one ends the pass, else the size picks how level d+1
// expands.
__global__ void bc_level_end(int32_t V, int d, BcCtl* ctl, unsigned int* __restrict__ lstart) {
  if (ctl->done) return;
  const unsigned int used = lstart[d + 1], total = ctl->cnt;
  if (total == used) {
    ctl->levels = d;
    ctl->done = 1;
    return;
  }
  lstart[d + 2] = total;
  ctl->push = uint64_t(total - used) * kPushRatio < uint64_t(V);
}

// The next level's masks start at zero when the level expands top-down.
__global__ void bc_clear_if_push(int32_t V, unsigned long long* __restrict__ nxt, const BcCtl* ctl) {
  if (ctl_load(&ctl->done) || !ctl_load(&ctl->push)) return;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) nxt[v] = 0ull;
}

__global__ void bc_seed(const int32_t* __restrict__ src, int nb, unsigned long long* __restrict__ visited,
                        unsigned long long* __restrict__ mask0, double* __restrict__ sigma) {
  const int s = threadIdx.x;
  if (s < nb) {
    const int32_t v = src[s];
    atomicOr(&visited[v], 1ull << s);
    atomicOr(&mask0[v], 1ull << s);
    sigma[size_t(v) * kBatch + s] = 1.0;
  }
}

__global__ __launch_bounds__(kBcBlock) void bc_scan(
    int32_t V, const int32_t* __restrict__ roff, const int32_t* __restrict__ radj,
    const unsigned long long* __restrict__ cur, unsigned long long* __restrict__ nxt,
    unsigned long long* __restrict__ visited, unsigned long long batchmask,
    int32_t* __restrict__ list, BcCtl* ctl) {
  __shared__ unsigned int s_cnt[kBcWaves];
  __shared__ unsigned int s_base;
  if (ctl_load(&ctl->done) || ctl_load(&ctl->push)) return;
  unsigned int* cnt = &ctl->cnt;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  // Rounded so whole workgroups take part in the aggregated append.
  const int64_t vend = (int64_t(V) + stride - 1) / stride * stride;
  for (int64_t w = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; w < vend; w += stride) {
    unsigned long long nm = 0;
    if (w < V) {
      const unsigned long long vis = visited[w];
      const unsigned long long unv = batchmask & ~vis;
      if (unv) {
        // Four in-edges per step (their loads issued together); stops once
        // every source still missing at w has been found. Edges read past
        // that point only add bits that are set already or masked off below.
        const int32_t e = roff[w + 1];
        for (int32_t i = roff[w]; i < e && (nm & unv) != unv; i += kScanStep) {
          int32_t u[kScanStep];
#pragma unroll
          for (int k = 0; k < kScanStep; ++k) u[k] = i + k < e ? radj[i + k] : -1;
          unsigned long long c[kScanStep];
#pragma unroll
          for (int k = 0; k < kScanStep; ++k) c[k] = u[k] >= 0 ? cur[u[k]] : 0ull;
#pragma unroll
          for (int k = 0; k < kScanStep; ++k) nm |= c[k];
        }
        nm &= unv;
        if (nm) visited[w] = vis | nm;
      }
      nxt[w] = nm;
    }
    // One list reservation per workgroup and step: a shared counter hit once
    // per wavefront serialises (~65K same-address atomics a level at 4M
    // vertices).
    const unsigned long long b = __ballot(nm != 0ull);
    if (lane == 0) s_cnt[wid] = unsigned(__popcll(b));
    __syncthreads();
    unsigned int before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < kBcWaves; ++k) {
      const unsigned int x = s_cnt[k];
      before += k < wid ? x : 0u;
      all += x;
    }
    if (threadIdx.x == 0) s_base = all ? atomicAdd(cnt, all) : 0u;
    __syncthreads();
    if (nm) list[s_base + before + __popcll(b & ((1ull << lane) - 1ull))] = int32_t(w);
    __syncthreads();  // s_cnt and s_base are reused by the next step
  }
}

// Top-down step for a small frontier. A workgroup takes kPushTile frontier
// vertices per round, each wavefront one vertex at a time with its lanes over
// the out-edges: the sources of v not yet at w are ORed into nxt[w] and
// visited[w], and the lane whose OR made nxt[w] non-zero appends w. Appends
// are staged in LDS and reserved with one global atomic per workgroup round
// (a shared counter hit once per vertex would serialise the launch); a round
// that overflows the stage reserves the excess per wavefront.
// nxt must be zero on entry. A stale read of visited[w] is harmless: it holds
// at least the bits of earlier levels (written by earlier launches), and any
// bit of this level it already holds is in nxt[w] too, so nxt[w] ends as the
// OR over in-neighbours at this level minus the earlier levels -- the
// bottom-up scan's value.
constexpr int kPushTile = 64;
constexpr int kPushStage = 4096;
__global__ __launch_bounds__(kBcBlock) void bc_push(
    const unsigned int* __restrict__ lstart, int d, const int32_t* __restrict__ off,
    const int32_t* __restrict__ adj, const unsigned long long* __restrict__ cur,
    unsigned long long* __restrict__ nxt, unsigned long long* __restrict__ visited,
    int32_t* __restrict__ list, BcCtl* ctl) {
  __shared__ int32_t s_buf[kPushStage];
  __shared__ unsigned int s_n, s_base;
  if (ctl_load(&ctl->done) || !ctl_load(&ctl->push)) return;
  unsigned int* cnt = &ctl->cnt;
  const int32_t* front = list + lstart[d];
  const unsigned int n = lstart[d + 1] - lstart[d];
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (int64_t t0 = int64_t(blockIdx.x) * kPushTile; t0 < n; t0 += int64_t(gridDim.x) * kPushTile) {
    const int64_t t1 = std::min<int64_t>(t0 + kPushTile, n);
    for (int64_t i = t0 + wid; i < t1; i += kBcWaves) {
      const int32_t v = front[i];
      const unsigned long long cv = cur[v];
      const int32_t b = off[v], e = off[v + 1];
      for (int32_t i0 = b; i0 < e; i0 += kWave) {
        const int32_t j = i0 + lane;
        bool add = false;
        int32_t w = 0;
        if (j < e) {
          w = adj[j];
          const unsigned long long bits =
              cv & ~__hip_atomic_load(&visited[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (bits) {
            add = atomicOr(&nxt[w], bits) == 0ull;
            atomicOr(&visited[w], bits);
          }
        }
        const unsigned long long bl = __ballot(add);
        if (bl) {
          const int leader = __ffsll((long long)bl) - 1;
          const unsigned int k = unsigned(__popcll(bl));
          unsigned int p = 0;
          if (lane == leader) p = atomicAdd(&s_n, k);
          p = __shfl(p, leader, kWave);
          // Stage slots p .. p+k-1; the ones at or past kPushStage go to the
          // list directly, reserved by the group's leader.
          const unsigned int cap = unsigned(kPushStage);
          const unsigned int over = p + k > cap ? std::min(k, p + k - cap) : 0u;
          unsigned int g = 0;
          if (over && lane == leader) g = atomicAdd(cnt, over);
          g = __shfl(g, leader, kWave);
          const unsigned int q = p + unsigned(__popcll(bl & ((1ull << lane) - 1ull)));
          if (add) {
            if (q < cap) s_buf[q] = w;
            else list[g + (q - std::max(p, cap))] = w;
          }
        }
      }
    }
    __syncthreads();
    const unsigned int m = std::min(s_n, unsigned(kPushStage));
    if (threadIdx.x == 0 && m) s_base = atomicAdd(cnt, m);
    __syncthreads();
    for (unsigned int k = threadIdx.x; k < m; k += kBcBlock) list[s_base + k] = s_buf[k];
    __syncthreads();
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBcBlock) void bc_sigma(
    const int32_t* __restrict__ lists, const unsigned int* __restrict__ lstart, int d, BcCtl* ctl,
    const int32_t* __restrict__ roff, const int32_t* __restrict__ radj, const unsigned long long* __restrict__ cur,
    const unsigned long long* __restrict__ nxt, double* __restrict__ sigma) {
  const int lane = __lane_id();
  if (ctl_load(&ctl->done)) return;
  // level d+1's list
  const int32_t* list = lists + lstart[d + 1];
  const unsigned int n = lstart[d + 2] - lstart[d + 1];
  const int64_t nw = int64_t(gridDim.x) * kBcWaves;
  for (int64_t i = int64_t(blockIdx.x) * kBcWaves + threadIdx.x / kWave; i < n; i += nw) {
    const int32_t w = list[i];
    const unsigned long long nm = nxt[w];
    double acc = 0.0;
    const int32_t b = roff[w], e = roff[w + 1];
    for (int32_t i0 = b; i0 < e; i0 += kWave) {
      const int32_t j = i0 + lane;
      int32_t v = 0;
      unsigned long long m = 0;
      if (j < e) {
        v = radj[j];
        m = cur[v] & nm;
      }
      unsigned long long act = __ballot(m != 0ull);
      while (act) {
        // Up to kGather active in-neighbours per step: their row loads are
        // issued together (sums of integers below 2^53 are exact in any order).
        double x[kGather];
#pragma unroll
        for (int k = 0; k < kGather; ++k) {
          x[k] = 0.0;
          if (act) {
            const int src_lane = __ffsll((long long)act) - 1;
            act &= act - 1;
            const int32_t vv = __shfl(v, src_lane, kWave);
            const unsigned long long mm = __shfl(m, src_lane, kWave);
            if ((mm >> lane) & 1ull) x[k] = sigma[size_t(vv) * kBatch + lane];
          }
        }
#pragma unroll
        for (int k = 0; k < kGather; ++k) acc += x[k];
      }
    }
    if ((nm >> lane) & 1ull) {
      sigma[size_t(w) * kBatch + lane] = acc;
      // Path counts from 2^96 on would leave (1 + delta) / sigma near FP32's
      // smallest normal numbers (2^-126): the batch's reverse pass then
      // keeps FP64 coefficients (a grid's counts are binomials, ~10^59 at
      // 200 levels).
      if (acc >= 0x1p96) ctl->wide = 1;
    }
  }
}

// CT = float: coefficients in their own FP32 rows (coef); CT = double: in
// place of sigma_v (coef == sigma), for batches whose path counts are too
// large for FP32 reciprocals.
template <typename CT>
__global__ __launch_bounds__(kBcBlock) void bc_back(
    const int32_t* __restrict__ list, unsigned int n, const int32_t* __restrict__ off,
    const int32_t* __restrict__ adj, const unsigned long long* __restrict__ md,
    const unsigned long long* __restrict__ md1, const double* sigma, CT* coef, float* __restrict__ BC) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * kBcWaves;
  for (int64_t i = int64_t(blockIdx.x) * kBcWaves + threadIdx.x / kWave; i < n; i += nw) {
    const int32_t v = list[i];
    const unsigned long long mv = md[v];
    const bool mine = (mv >> lane) & 1ull;
    const double sv = mine ? sigma[size_t(v) * kBatch + lane] : 1.0;
    float acc = 0.f;
    const int32_t b = off[v], e = off[v + 1];
    for (int32_t i0 = b; i0 < e; i0 += kWave) {
      const int32_t j = i0 + lane;
      int32_t w = 0;
      unsigned long long m = 0;
      if (j < e) {
        w = adj[j];
        m = md1[w] & mv;
      }
      unsigned long long act = __ballot(m != 0ull);
      while (act) {
        // Up to kGather active successors per step: loads issued together,
        // terms added in ascending neighbour order. coef[w][s] of a successor
        // holds (1 + delta_w) / sigma_w, written by w's own step.
        CT c[kGather];
        bool on[kGather];
#pragma unroll
        for (int k = 0; k < kGather; ++k) {
          on[k] = false;
          c[k] = CT(0);
          if (act) {
            const int src_lane = __ffsll((long long)act) - 1;
            act &= act - 1;
            const int32_t ww = __shfl(w, src_lane, kWave);
            const unsigned long long mm = __shfl(m, src_lane, kWave);
            on[k] = (mm >> lane) & 1ull;
            if (on[k]) c[k] = coef[size_t(ww) * kBatch + lane];
          }
        }
#pragma unroll
        for (int k = 0; k < kGather; ++k)
          if (on[k]) acc += float(sv * double(c[k]));
      }
    }
    float part = 0.f;
    if (mine) {
      // delta_v is final: leave coef_v for the predecessors at level d-1
      // (only this source's lane).
      coef[size_t(v) * kBatch + lane] = CT((1.0 + double(acc)) / sv);
      part = acc;
    }
    part = wave_sum(part);
    if (lane == 0 && mv) BC[v] += part;
  }
}

__global__ void bc_extract(int32_t V, const unsigned long long* __restrict__ masks, int levels,
                           const double* __restrict__ sigma, int32_t* __restrict__ lev,
                           double* __restrict__ sig) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    int32_t l = -1;
    for (int d = 0; d < levels && l < 0; ++d)
      if (masks[size_t(d) * V + v] & 1ull) l = d;
    lev[v] = l;
    sig[v] = l >= 0 ? sigma[size_t(v) * kBatch] : 0.0;
  }
}

__global__ void bc_zero(int32_t V, float* __restrict__ BC) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) BC[v] = 0.f;
}

static int bc_alloc(spm_graph* g) {
  if (!g->roff) return set_error(SPM_ENOGRAPH, "BC needs the reverse CSR (rev_indexofNodes, srcList)");
  if (g->bc) return SPM_OK;
  BcWork* w = new BcWork();
  g->bc = w;
  const size_t V = size_t(g->V);
  SPM_TRY(dmalloc(g, (void**)&w->sigma, sizeof(double) * V * kBatch, "bc sigma"));
  SPM_TRY(dmalloc(g, (void**)&w->coef, sizeof(float) * V * kBatch, "bc coefficients"));
  SPM_TRY(dmalloc(g, (void**)&w->visited, sizeof(unsigned long long) * V, "bc visited"));
  SPM_TRY(dmalloc(g, (void**)&w->ctl, sizeof(BcCtl), "bc control"));
  SPM_TRY(dmalloc(g, (void**)&w->srcbuf, sizeof(int32_t) * kBatch, "bc sources"));
  return SPM_OK;
}

// Makes room for `levels` mask levels, `entries` list entries and the list
// bounds of `levels` levels (grow only; contents are kept).
static int bc_reserve(spm_graph* g, size_t levels, size_t entries) {
  BcWork* w = g->bc;
  const size_t V = size_t(g->V);
  if (levels + 2 > w->lstart_cap) {
    const size_t nc = std::max(levels + 2, w->lstart_cap * 2);
    unsigned int* p = nullptr;
    SPM_TRY(dmalloc(g, (void**)&p, sizeof(unsigned int) * nc, "bc level bounds"));
    if (w->lstart) {
      SPM_CHECK_HIP(hipMemcpyAsync(p, w->lstart, sizeof(unsigned int) * w->lstart_cap, hipMemcpyDeviceToDevice,
                                   g->stream));
      SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
      dfree(g, w->lstart, sizeof(unsigned int) * w->lstart_cap);
    }
    w->lstart = p;
    w->lstart_cap = nc;
  }
  if (levels > w->mask_levels) {
    const size_t nl = std::max(levels, w->mask_levels * 2);
    unsigned long long* p = nullptr;
    SPM_TRY(dmalloc(g, (void**)&p, sizeof(unsigned long long) * V * nl, "bc level masks"));
    if (w->masks) {
      SPM_CHECK_HIP(hipMemcpyAsync(p, w->masks, sizeof(unsigned long long) * V * w->mask_levels, hipMemcpyDeviceToDevice, g->stream));
      SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
      dfree(g, w->masks, sizeof(unsigned long long) * V * w->mask_levels);
    }
    w->masks = p;
    w->mask_levels = nl;
  }
  if (entries > w->list_cap) {
    const size_t nc = std::max(entries, w->list_cap * 2);
    int32_t* p = nullptr;
    SPM_TRY(dmalloc(g, (void**)&p, sizeof(int32_t) * nc, "bc level lists"));
    if (w->lists) {
      SPM_CHECK_HIP(hipMemcpyAsync(p, w->lists, sizeof(int32_t) * w->list_cap, hipMemcpyDeviceToDevice, g->stream));
      SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
      dfree(g, w->lists, sizeof(int32_t) * w->list_cap);
    }
    w->lists = p;
    w->list_cap = nc;
  }
  return SPM_OK;
}

int bc_release(spm_graph* g) {
  BcWork* w = g->bc;
  if (!w) return SPM_OK;
  const size_t V = size_t(g->V);
  dfree(g, w->sigma, sizeof(double) * V * kBatch);
  dfree(g, w->coef, sizeof(float) * V * kBatch);
  dfree(g, w->visited, sizeof(unsigned long long) * V);
  dfree(g, w->masks, sizeof(unsigned long long) * V * w->mask_levels);
  dfree(g, w->lists, sizeof(int32_t) * w->list_cap);
  dfree(g, w->ctl, sizeof(BcCtl));
  dfree(g, w->lstart, sizeof(unsigned int) * w->lstart_cap);
  dfree(g, w->srcbuf, sizeof(int32_t) * kBatch);
  delete w;
  g->bc = nullptr;
  return SPM_OK;
}

// Forward pass of one batch; returns the deepest non-empty level and the list
// start of every level (lstart[d] .. lstart[d+1]). Levels run in windows of
// K launches-worth (4, doubling to 32): each level is five launches -- clear
// (top-down levels), push or scan (the one not chosen returns at once), the
// level's end (one thread: next list bound, next mode, or done), sigma --
// reading their list bounds and mode from device memory, and the host reads
// the control block once per window. Levels after the last one return at
// once. (One read-back per level costs a round trip per level on
// high-diameter graphs: ~2,000 levels on a 1024^2 grid.)
static int bc_forward(spm_graph* g, const int32_t* h_src, int nb, int* levels_out,
                      std::vector<unsigned int>& lstart, bool* wide = nullptr) {
  BcWork* w = g->bc;
  const int32_t V = g->V;
  const size_t Vs = size_t(V);
  SPM_TRY(bc_reserve(g, 2, Vs + kBatch));
  SPM_CHECK_HIP(hipMemcpyAsync(w->srcbuf, h_src, sizeof(int32_t) * nb, hipMemcpyHostToDevice, g->stream));
  const int igrid = int(std::min<int64_t>((V + 255) / 256, 4096));
  SPM_TRY(launch(g, "bc_init", bc_init, dim3(igrid), dim3(256), 0, V, w->visited, w->masks,
                 (const int32_t*)w->srcbuf, nb, w->lists, w->ctl, w->lstart));
  SPM_TRY(launch(g, "bc_seed", bc_seed, dim3(1), dim3(kBatch), 0, (const int32_t*)w->srcbuf, nb,
                 w->visited, w->masks, w->sigma));
  const unsigned long long batchmask = nb == 64 ? ~0ull : ((1ull << nb) - 1ull);
  // pinned read-back slots: the cursor, then done, levels, push, wide
  unsigned int* hc = reinterpret_cast<unsigned int*>(g->h_pinned);
  int* hf = g->h_pinned + 4;
  const int sgrid = int(std::min<int64_t>((V + kBcBlock - 1) / kBcBlock, int64_t(g->num_cus) * 16));
  const int wgrid = g->num_cus * 8;
  unsigned int used = unsigned(nb);  // list entries appended so far (as of the last read)
  int d = 0, K = 4;
  for (;;) {
    // Room for the window: K more levels, each appending at most V entries.
    SPM_TRY(bc_reserve(g, size_t(d) + size_t(K) + 2, size_t(used) + size_t(K) * Vs + kBatch));
    for (int k = 0; k < K; ++k, ++d) {
      unsigned long long* cur = w->masks + size_t(d) * Vs;
      unsigned long long* nxt = w->masks + size_t(d + 1) * Vs;
      SPM_TRY(launch(g, "bc_clear", bc_clear_if_push, dim3(igrid), dim3(256), 0, V, nxt, (const BcCtl*)w->ctl));
      SPM_TRY(launch(g, "bc_push", bc_push, dim3(wgrid), dim3(kBcBlock), 0, (const unsigned int*)w->lstart, d,
                     (const int32_t*)g->off, (const int32_t*)g->adj, (const unsigned long long*)cur, nxt,
                     w->visited, w->lists, w->ctl));
      SPM_TRY(launch(g, "bc_scan", bc_scan, dim3(sgrid), dim3(kBcBlock), 0, V, (const int32_t*)g->roff,
                     (const int32_t*)g->radj, (const unsigned long long*)cur, nxt, w->visited, batchmask,
                     w->lists, w->ctl));
      SPM_TRY(launch(g, "bc_level_end", bc_level_end, dim3(1), dim3(1), 0, V, d, w->ctl, w->lstart));
      SPM_TRY(launch(g, "bc_sigma", bc_sigma, dim3(wgrid), dim3(kBcBlock), 0, (const int32_t*)w->lists,
                     (const unsigned int*)w->lstart, d, w->ctl, (const int32_t*)g->roff,
                     (const int32_t*)g->radj, (const unsigned long long*)cur, (const unsigned long long*)nxt,
                     w->sigma));
    }
    SPM_CHECK_HIP(hipMemcpyAsync(hc, &w->ctl->cnt, sizeof(unsigned int), hipMemcpyDeviceToHost, g->stream));
    SPM_CHECK_HIP(hipMemcpyAsync(hf, &w->ctl->done, 4 * sizeof(int), hipMemcpyDeviceToHost, g->stream));
    SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    used = *hc;
    if (hf[0]) break;
    K = std::min(2 * K, 32);
  }
  const int L = hf[1];  // levels 0..L are non-empty; masks[L+1] is all zero
  if (wide) *wide = hf[3] != 0;
  lstart.resize(size_t(L) + 2);
  SPM_CHECK_HIP(hipMemcpyAsync(lstart.data(), w->lstart, sizeof(unsigned int) * (size_t(L) + 2),
                               hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  *levels_out = L;
  return SPM_OK;
}

static int bc_backward(spm_graph* g, int L, const std::vector<unsigned int>& lstart, bool wide, float* d_BC) {
  BcWork* w = g->bc;
  const size_t Vs = size_t(g->V);
  const int wgrid = g->num_cus * 8;
  for (int d = L; d >= 1; --d) {
    const unsigned int b = lstart[d], e = lstart[d + 1];
    if (e == b) continue;
    const unsigned long long* md = w->masks + size_t(d) * Vs;
    const unsigned long long* md1 = w->masks + size_t(d + 1) * Vs;
    if (wide)
      SPM_TRY(launch(g, "bc_back", bc_back<double>, dim3(wgrid), dim3(kBcBlock), 0, (const int32_t*)(w->lists + b),
                     e - b, (const int32_t*)g->off, (const int32_t*)g->adj, md, md1, (const double*)w->sigma,
                     w->sigma, d_BC));
    else
      SPM_TRY(launch(g, "bc_back", bc_back<float>, dim3(wgrid), dim3(kBcBlock), 0, (const int32_t*)(w->lists + b),
                     e - b, (const int32_t*)g->off, (const int32_t*)g->adj, md, md1, (const double*)w->sigma,
                     w->coef, d_BC));
  }
  return SPM_OK;
}

static int bc_check_sources(const spm_graph* g, const int32_t* sources, int32_t nsrc, std::vector<int32_t>& s) {
  if (nsrc < 0 || (nsrc > 0 && !sources)) return set_error(SPM_EINVAL, "BC: bad source list");
  s.assign(sources, sources + nsrc);
  std::sort(s.begin(), s.end());
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] < 0 || s[i] >= g->V) return set_error(SPM_EINVAL, "BC: source out of range");
    if (i && s[i] == s[i - 1]) return set_error(SPM_EINVAL, "BC: duplicate source (sourceSet is a set)");
  }
  return SPM_OK;
}

int bc_run(spm_graph* g, const int32_t* sources, int32_t nsrc, float* d_BC) {
  std::vector<int32_t> s;
  SPM_TRY(bc_check_sources(g, sources, nsrc, s));
  SPM_TRY(dyn_settle(g));
  SPM_TRY(bc_alloc(g));
  const int igrid = int(std::min<int64_t>((g->V + 255) / 256, 4096));
  SPM_TRY(launch(g, "bc_zero", bc_zero, dim3(igrid), dim3(256), 0, g->V, d_BC));
  std::vector<unsigned int> lstart;
  for (size_t b = 0; b < s.size(); b += kBatch) {
    const int nb = int(std::min<size_t>(kBatch, s.size() - b));
    int L = 0;
    bool wide = false;
    SPM_TRY(bc_forward(g, s.data() + b, nb, &L, lstart, &wide));
    SPM_TRY(bc_backward(g, L, lstart, wide, d_BC));
  }
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  SPM_TRY(prof_flush(g));
  return SPM_OK;
}

}  // namespace spm

using namespace spm;

extern "C" {

int spm_bc_device(spm_graph* g, const int32_t* sources, int32_t nsrc, float* d_BC) {
  if (!g || !d_BC) return set_error(SPM_EINVAL, "spm_bc_device: NULL argument");
  return bc_run(g, sources, nsrc, d_BC);
}

int spm_compute_bc(spm_graph* g, float* BC, const int32_t* sources, int32_t nsrc) {
  if (!g || !BC) return set_error(SPM_EINVAL, "spm_compute_bc: NULL argument");
  float* d = nullptr;
  SPM_TRY(dmalloc(g, (void**)&d, sizeof(float) * g->V, "bc output"));
  int rc = bc_run(g, sources, nsrc, d);
  if (rc == SPM_OK) {
    hipError_t e = hipMemcpyAsync(BC, d, sizeof(float) * g->V, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = set_error(SPM_EHIP, std::string("spm_compute_bc copy: ") + hipGetErrorString(e));
  }
  dfree(g, d, sizeof(float) * g->V);
  return rc;
}

int spm_bc_paths(spm_graph* g, int32_t src, int32_t* bfsDist, double* sigma) {
  if (!g || !bfsDist || !sigma) return set_error(SPM_EINVAL, "spm_bc_paths: NULL argument");
  if (src < 0 || src >= g->V) return set_error(SPM_EINVAL, "spm_bc_paths: source out of range");
  SPM_TRY(dyn_settle(g));
  SPM_TRY(bc_alloc(g));
  std::vector<unsigned int> lstart;
  int L = 0;
  SPM_TRY(bc_forward(g, &src, 1, &L, lstart));
  int32_t* dl = nullptr;
  double* ds = nullptr;
  SPM_TRY(dmalloc(g, (void**)&dl, sizeof(int32_t) * g->V, "bc paths levels"));
  int rc = dmalloc(g, (void**)&ds, sizeof(double) * g->V, "bc paths sigma");
  if (rc == SPM_OK) {
    const int igrid = int(std::min<int64_t>((g->V + 255) / 256, 4096));
    rc = launch(g, "bc_extract", bc_extract, dim3(igrid), dim3(256), 0, g->V, (const unsigned long long*)g->bc->masks,
                L + 1, (const double*)g->bc->sigma, dl, ds);
  }
  if (rc == SPM_OK) {
    hipError_t e = hipMemcpyAsync(bfsDist, dl, sizeof(int32_t) * g->V, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sigma, ds, sizeof(double) * g->V, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = set_error(SPM_EHIP, std::string("spm_bc_paths copy: ") + hipGetErrorString(e));
  }
  dfree(g, dl, sizeof(int32_t) * g->V);
  dfree(g, ds, sizeof(double) * g->V);
  if (rc == SPM_OK) rc = prof_flush(g);
  return rc;
}

}  // extern "C"
