// spm_pr.hip -- pull PageRank for gfx950, the replacement of the code StarPlat
// generates for Compute_PR (graphcode/staticDSLCodes/PageRankDSLV2:1-19, with
// the diff update of PageRankDSLV3:15-19):
//   do { diff = 0; forall v: sum = sum_{u in nodes_to(v)} pr[u]/outdeg(u);
//        pr_nxt[v] = (1-delta)/N + delta*sum; diff += |pr_nxt[v]-pr[v]|;
//        pr = pr_nxt; iter++ } while (diff > beta && iter < maxIter)
// (OMP emission: backend_omp/dsl_cpp_generator.cpp:988-1015, :1094-1113,
//  :2475-2484; do-while :568-575).
//
// Design. Each iteration is one kernel. A vertex's in-row (srcList) is read
// once; the per-source term pr[u]/outdeg(u) is produced by the previous
// iteration (or the init kernel) into a contribution array, so the gather
// reads one float per in-edge. Vertices are binned once by in-degree and the
// workgroups of one launch are split over the bins: one lane per vertex for
// short rows, a group of G lanes, a 64-lane wavefront, or a 256-lane
// workgroup for hubs. Each workgroup stores its share of the diff; a
// one-workgroup check kernel sums the shares, evaluates the loop condition and
// raises a device flag; later iterations see the flag and return at once, so
// a whole run is captured into one hipGraph with no host round trip. (No
// workgroup updates a shared counter: same-address device-scope atomics from
// 10^5 workgroups cost milliseconds per iteration.)
//
// On the degree-ordered copy of a large graph a persistent variant of the
// sweep (pr_iter_hot, schedule pr_mode 3) first copies the contributions of
// the lowest ids -- the sources of most in-edges -- into each CU's LDS and
// reads those terms there; it also evaluates the loop test of the sweep
// before it, so a captured batch has one check kernel in all.
//
// Arithmetic follows the emitted code: the per-term division pr[u]/outdeg(u)
// is the same IEEE single division, and the update is evaluated without FMA
// contraction. Only the order of the in-row summation differs (tree vs
// sequential), which keeps ranks within a few ulps of the restatement.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "spm_internal.hpp"

namespace spm {

constexpr int kPrBlock = 256;

struct PrState {
  int done;
  int iter;
  int maxIter;
  float beta;
  float delta;
  float base;  // (1 - delta) / num_nodes
};

struct PrBins {
  int32_t nb[4];   // vertices per bin
  int32_t blk[5];  // first workgroup of each bin; blk[4] = grid size
  int32_t group;   // lanes per vertex in bin 1
  int32_t zb;      // bin 0 starts with the in-degree-0 vertices: its first zb
                   // workgroups hold only those (0 for batched jobs)
};

// Batched jobs (spm_pr_multi_device): vertex v belongs to job job[v], each
// job a contiguous range of the union graph's ids with no edge leaving it.
// Jobs iterate with their own (1 - delta) / N_k and stopping test; a job that
// has stopped keeps its ranks and is skipped by later iterations.
struct PrJobs {
  const int32_t* job;   // [V]
  const int32_t* live;  // [n] 1 while the job iterates
  const float* base;    // [n] (1 - delta) / N_k
  float* dv;            // [V] |new - old| of the vertex's last iteration
};

struct PrWork {
  float* pr[2] = {nullptr, nullptr};
  float* c[2] = {nullptr, nullptr};
  float* degf = nullptr;
  float* acc = nullptr;      // push-mode accumulators
  float* partial = nullptr;  // per-workgroup diff sums of one iteration
  int partial_n = 0;
  int32_t* sp = nullptr;     // segment split points [V][S-1]
  size_t sp_n = 0;
  int S = 1;                 // source segments per pull sweep
  bool hot = false;          // layout binned for the staged sweep (pr_iter_hot)
  int4* rows = nullptr;      // {v, in-row start, in-row end, 0} grouped by bin
  float* out = nullptr;      // [V] result in new ids before the gather back (relabeled runs)
  CsrView view{};            // the CSR the kernels run on
  PrState* st = nullptr;
  PrBins bins{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int R = 0;
};

__global__ void pr_init(int32_t V, const int32_t* __restrict__ off, float* __restrict__ degf,
                        float* __restrict__ pr, float* __restrict__ c, float* __restrict__ acc, PrState* st,
                        float init, int maxIter, float beta, float delta, float base) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    const float d = float(off[v + 1] - off[v]);
    degf[v] = d;
    pr[v] = init;
    c[v] = d > 0.f ? init / d : 0.f;
    if (acc) acc[v] = 0.f;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->done = 0;
    st->iter = 0;
    st->maxIter = maxIter;
    st->beta = beta;
    st->delta = delta;
    st->base = base;
  }
}

// v's new rank from its in-row sum; old = pr_in[v] and d = degf[v] are
// loaded by the caller beside the gather rather than after it.
__device__ __forceinline__ void pr_store(int32_t v, float sum, const PrState& s, float old, float d,
                                         float* __restrict__ pr_out, float* __restrict__ c_out, float& mydiff) {
  const float nv = __fadd_rn(s.base, __fmul_rn(s.delta, sum));
  pr_out[v] = nv;
  c_out[v] = d > 0.f ? nv / d : 0.f;
  mydiff += nv - old >= 0.f ? nv - old : old - nv;
}

// End of an iteration inside a workgroup: the sum of |new - old| over its
// vertices goes to partial[blockIdx.x] with a plain store. No workgroup
// touches a shared counter: with ~10^5 workgroups per iteration at RMAT-22,
// same-address device-scope atomics would serialise the whole iteration.
template <int B = kPrBlock>
__device__ __forceinline__ void pr_finish(float mydiff, float* __restrict__ partial, float* s_red, int b) {
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  mydiff = wave_sum(mydiff);
  if (lane == 0) s_red[wid] = mydiff;
  __syncthreads();
  if (tid == 0) {
    float t = 0.f;
    for (int w = 0; w < B / kWave; ++w) t += s_red[w];
    partial[b] = t;
  }
}

// The loop test (diff > beta && iter < maxIter), one workgroup: sums the
// per-workgroup partials in a fixed order, advances the iteration count by
// `step` (1 after each sweep; the batch length where the sweeps test their
// predecessors themselves) and raises the done flag that makes later
// iterations return at once.
constexpr int kCheckBlock = 1024;

// The sum of partial[0..n) in pr_check's fixed order, computed by a
// kCheckBlock-lane workgroup; every thread returns the same value.
__device__ __forceinline__ float check_sum(const float* __restrict__ partial, int n, float* s_red) {
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  // 16-byte loads into four independent sums, two loads per step, so the one
  // workgroup keeps several requests in flight (partial is 256-byte aligned).
  const float4* p4 = reinterpret_cast<const float4*>(partial);
  const int n4 = n >> 2;
  float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
  int i = tid;
  for (; i + kCheckBlock < n4; i += 2 * kCheckBlock) {
    const float4 x = p4[i], y = p4[i + kCheckBlock];
    a.x += x.x; a.y += x.y; a.z += x.z; a.w += x.w;
    b.x += y.x; b.y += y.y; b.z += y.z; b.w += y.w;
  }
  if (i < n4) {
    const float4 x = p4[i];
    a.x += x.x; a.y += x.y; a.z += x.z; a.w += x.w;
  }
  float t = ((a.x + b.x) + (a.y + b.y)) + ((a.z + b.z) + (a.w + b.w));
  if (tid < (n & 3)) t += partial[(n4 << 2) + tid];
  t = wave_sum(t);
  if (lane == 0) s_red[wid] = t;
  __syncthreads();
  float d = 0.f;
  for (int w = 0; w < kCheckBlock / kWave; ++w) d += s_red[w];
  return d;
}

__global__ __launch_bounds__(kCheckBlock) void pr_check(const float* __restrict__ partial, int n,
                                                        PrState* __restrict__ st, int step) {
  __shared__ float s_red[kCheckBlock / kWave];
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const float d = check_sum(partial, n, s_red);
  if (threadIdx.x == 0) {
    const int it = st->iter + step;
    st->iter = it;
    if (!(d > st->beta && it < st->maxIter)) st->done = 1;
  }
}

// Source segmentation (cache blocking). With S > 1 segments, pass `seg` sums
// only the in-edges whose source lies in [seg*2^k, (seg+1)*2^k); the split
// points of every in-row (rows are ascending) are precomputed in sp[v][S-1].
// Passes 0..S-2 accumulate into acc[v]; the last pass adds its own part and
// writes the rank. Each pass gathers from a 2^k-entry slice of the
// contribution array that stays resident in L2.
struct PrSeg {
  const int32_t* sp;  // [V][S-1] split points, or null when S == 1
  float* acc;         // [V] partial sums across passes
  int S;
  int seg;
};

__device__ __forceinline__ void seg_range(int32_t v, const int32_t* __restrict__ roff, const PrSeg& sg,
                                          int32_t& a, int32_t& b) {
  if (sg.S == 1) {
    a = roff[v];
    b = roff[v + 1];
    return;
  }
  const int32_t* p = sg.sp + int64_t(v) * (sg.S - 1);
  a = sg.seg == 0 ? roff[v] : p[sg.seg - 1];
  b = sg.seg == sg.S - 1 ? roff[v + 1] : p[sg.seg];
}

__device__ __forceinline__ void seg_emit(int32_t v, float sum, const PrSeg& sg, const PrState& s, float old,
                                         float dg, float* __restrict__ pr_out, float* __restrict__ c_out,
                                         float& mydiff) {
  if (sg.S > 1) {
    if (sg.seg < sg.S - 1) {
      sg.acc[v] = sg.seg == 0 ? sum : sg.acc[v] + sum;
      return;
    }
    sum = sg.acc[v] + sum;
  }
  pr_store(v, sum, s, old, dg, pr_out, c_out, mydiff);
}

// A vertex's new rank: the single-job path through seg_emit; batched jobs use
// the job's base term and leave |new - old| in dv for the per-job test.
template <bool MULTI>
__device__ __forceinline__ void pr_emit(int32_t v, float sum, const PrSeg& sg, const PrState& s, float old, float dg,
                                        float* __restrict__ pr_out, float* __restrict__ c_out, float& mydiff,
                                        const PrJobs& jb) {
  if (!MULTI) {
    seg_emit(v, sum, sg, s, old, dg, pr_out, c_out, mydiff);
    return;
  }
  const float nv = __fadd_rn(jb.base[jb.job[v]], __fmul_rn(s.delta, sum));
  pr_out[v] = nv;
  c_out[v] = dg > 0.f ? nv / dg : 0.f;
  jb.dv[v] = nv - old >= 0.f ? nv - old : old - nv;
}

// Batched jobs: a vertex of a stopped job is skipped (no gather, no writes).
template <bool MULTI>
__device__ __forceinline__ bool pr_live(int32_t v, const PrJobs& jb) {
  return !MULTI || jb.live[jb.job[v]] != 0;
}

// Sum of c[radj[e]] for e = ea+start, ea+start+stride, ... < eb, eight row
// elements at a time: eight independent index loads, then eight independent
// gathers, so a lane keeps several memory requests in flight instead of a
// chain of dependent ones (a lane with up to 8 terms waits for two round
// trips in all).
//
// Staged variant (HOT): the contributions of sources [0, H) -- after the
// degree-ordered relabeling, the sources with the most out-edges -- were
// copied into LDS by the workgroup, and those terms are read there. The
// values, and so the sums, are the same as the plain gather's.
struct PrHot {
  const float* lds;  // [H] contributions of the sources below H (LDS)
  int32_t H;
};

template <bool HOT>
__device__ __forceinline__ float pr_term(const float* __restrict__ c, const PrHot& hot, int32_t u) {
  if (HOT && u < hot.H) return hot.lds[u];
  return c[u];
}

// W partial sums (and W loads in flight per lane); 16 on the staged sweep
// measured slower than 8 (RMAT-22 27.92 against 27.22 ms per run).
template <bool HOT = false, int W = 8>
__device__ __forceinline__ float gather_sum(const int32_t* __restrict__ radj, const float* __restrict__ c,
                                            int32_t ea, int32_t eb, int32_t start, int32_t stride,
                                            const PrHot& hot = PrHot{nullptr, 0}) {
  float s[W];
#pragma unroll
  for (int k = 0; k < W; ++k) s[k] = 0.f;
  int32_t e = ea + start;
  for (; e + (W - 1) * stride < eb; e += W * stride) {
    int32_t a[W];
#pragma unroll
    for (int k = 0; k < W; ++k) a[k] = radj[e + k * stride];
#pragma unroll
    for (int k = 0; k < W; ++k) s[k] += pr_term<HOT>(c, hot, a[k]);
  }
  // The remaining 0..W-1 terms, loaded together: every gather is issued
  // before any is added (a missing term stays +0, which adds exactly), so
  // no addition puts one gather's wait in front of the next.
  int32_t a[W - 1];
#pragma unroll
  for (int k = 0; k < W - 1; ++k) a[k] = e + k * stride < eb ? radj[e + k * stride] : -1;
  float t[W - 1];
#pragma unroll
  for (int k = 0; k < W - 1; ++k) t[k] = 0.f;
#pragma unroll
  for (int k = 0; k < W - 1; ++k)
    if (a[k] >= 0) t[k] = pr_term<HOT>(c, hot, a[k]);
#pragma unroll
  for (int k = 0; k < W - 1; ++k) s[k] += t[k];
  // pairwise: ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7)) for W = 8
#pragma unroll
  for (int h = 1; h < W; h <<= 1)
#pragma unroll
    for (int k = 0; k + h < W; k += 2 * h) s[k] += s[k + h];
  return s[0];
}

// rows[i] = {v, first in-edge, end of in-edges, 0} for the i-th vertex in bin
// order: one 16-byte load replaces the order -> offsets chain. pr_slot gives
// the i that thread threadIdx.x of workgroup b reads (-1: none): bin 0 one
// lane per vertex, bin 1 G lanes, bin 2 a wavefront, bin 3 the workgroup.
// B is the workgroup size the layout was binned for (kPrBlock, or kHotBlock
// for the staged sweep).
template <int G, int B = kPrBlock>
__device__ __forceinline__ int64_t pr_slot(int b, const PrBins& bins) {
  const int tid = threadIdx.x;
  if (b < bins.blk[1]) {
    const int64_t i = int64_t(b) * B + tid;
    return i < bins.nb[0] ? i : -1;
  }
  if (b < bins.blk[2]) {
    const int64_t i = int64_t(b - bins.blk[1]) * (B / G) + tid / G;
    return i < bins.nb[1] ? bins.nb[0] + i : -1;
  }
  if (b < bins.blk[3]) {
    const int64_t i = int64_t(b - bins.blk[2]) * (B / kWave) + tid / kWave;
    return i < bins.nb[2] ? int64_t(bins.nb[0]) + bins.nb[1] + i : -1;
  }
  return int64_t(bins.nb[0]) + bins.nb[1] + bins.nb[2] + (b - bins.blk[3]);
}

template <int G, int B = kPrBlock>
__device__ __forceinline__ int4 pr_row(const int4* __restrict__ rows, int b, const PrBins& bins) {
  const int64_t i = pr_slot<G, B>(b, bins);
  return i >= 0 ? rows[i] : make_int4(-1, 0, 0, 0);
}

// The sweep of one (virtual) workgroup b of the binned launch layout, given
// the thread's row (pr_row): the rows of bin 0 one lane each, of bin 1 G
// lanes each, of bin 2 a wavefront each, of bin 3 the whole workgroup. Adds
// the workgroup's |new - old| to the calling thread's mydiff. Control flow
// is uniform across the workgroup.
// One block's in-row sums, without writing: the lane that writes vertex ev
// gets its sum, old rank and out-degree (ev < 0 in the other lanes).
template <int G, bool MULTI, int B = kPrBlock, bool HOT = false>
__device__ __forceinline__ void pr_block_sums(int b, int4 row, const int32_t* __restrict__ roff,
                                              const int32_t* __restrict__ radj, const float* __restrict__ degf,
                                              const PrBins& bins, const float* __restrict__ pr_in,
                                              const float* __restrict__ c_in, const PrSeg& sg, const PrJobs& jb,
                                              float* s_red, int32_t& ev, float& esum, float& eold, float& edg,
                                              const PrHot& hot = PrHot{nullptr, 0}) {
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  int32_t v = row.x, ea = row.y, eb = row.z;
  if (v >= 0 && !pr_live<MULTI>(v, jb)) v = -1;
  if (v >= 0 && sg.S > 1) seg_range(v, roff, sg, ea, eb);
  // The old rank and the out-degree are in flight during the gather.
  float old = 0.f, dg = 0.f;
  if (v >= 0) {
    old = pr_in[v];
    dg = degf[v];
  }
  ev = -1;
  esum = 0.f;
  eold = old;
  edg = dg;
  if (b < bins.blk[1]) {
    // bin 0: one lane per vertex
    if (v >= 0) {
      esum = gather_sum<HOT>(radj, c_in, ea, eb, 0, 1, hot);
      ev = v;
    }
  } else if (b < bins.blk[2]) {
    // bin 1: G lanes per vertex
    const int g = tid % G;
    float sum = v >= 0 ? gather_sum<HOT>(radj, c_in, ea, eb, g, G, hot) : 0.f;
    esum = group_sum<G>(sum);
    if (v >= 0 && g == 0) ev = v;
  } else if (b < bins.blk[3]) {
    // bin 2: one wavefront per vertex
    float sum = v >= 0 ? gather_sum<HOT>(radj, c_in, ea, eb, lane, kWave, hot) : 0.f;
    esum = wave_sum(sum);
    if (v >= 0 && lane == 0) ev = v;
  } else if (v >= 0) {
    // bin 3: one workgroup per vertex (v is the same in every lane)
    float sum = gather_sum<HOT>(radj, c_in, ea, eb, tid, B, hot);
    sum = wave_sum(sum);
    if (lane == 0) s_red[wid] = sum;
    __syncthreads();
    if (tid == 0) {
      float t = 0.f;
      for (int w = 0; w < B / kWave; ++w) t += s_red[w];
      esum = t;
      ev = v;
    }
    __syncthreads();
  }
}

// One block's sweep: the sums, then the writes.
template <int G, bool MULTI, int B = kPrBlock, bool HOT = false>
__device__ __forceinline__ void pr_sweep_block(int b, int4 row, const int32_t* __restrict__ roff,
                                               const int32_t* __restrict__ radj, const float* __restrict__ degf,
                                               const PrBins& bins, const float* __restrict__ pr_in,
                                               const float* __restrict__ c_in, float* __restrict__ pr_out,
                                               float* __restrict__ c_out, const PrState& s, const PrSeg& sg,
                                               const PrJobs& jb, float* s_red, float& mydiff,
                                               const PrHot& hot = PrHot{nullptr, 0}) {
  int32_t ev;
  float esum, eold, edg;
  pr_block_sums<G, MULTI, B, HOT>(b, row, roff, radj, degf, bins, pr_in, c_in, sg, jb, s_red, ev, esum, eold, edg,
                                  hot);
  if (ev >= 0) pr_emit<MULTI>(ev, esum, sg, s, eold, edg, pr_out, c_out, mydiff, jb);
}

template <int G, bool MULTI>
__global__ __launch_bounds__(kPrBlock) void pr_iter(
    const int32_t* __restrict__ roff, const int32_t* __restrict__ radj,
    const int4* __restrict__ rows, const float* __restrict__ degf, PrBins bins,
    const float* __restrict__ pr_in, const float* __restrict__ c_in, float* __restrict__ pr_out,
    float* __restrict__ c_out, const PrState* __restrict__ st, float* __restrict__ partial, PrSeg sg,
    PrJobs jb) {
  __shared__ float s_red[kPrBlock / kWave];
  // The row, the run's constants and the flag are loaded together; the flag
  // is tested before anything that depends on the row.
  const int4 row = pr_row<G>(rows, blockIdx.x, bins);
  const PrState s = *st;
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  if (!MULTI && int(blockIdx.x) < bins.zb && s.iter >= 2) {
    // In-degree-0 vertices only: from the third sweep on their rank is base
    // in both buffers and |new - old| = 0, so the workgroup's share is 0.
    if (sg.seg == sg.S - 1 && threadIdx.x == 0) partial[blockIdx.x] = 0.f;
    return;
  }
  float mydiff = 0.f;
  pr_sweep_block<G, MULTI>(blockIdx.x, row, roff, radj, degf, bins, pr_in, c_in, pr_out, c_out, s, sg, jb, s_red,
                           mydiff);
  if (!MULTI && sg.seg == sg.S - 1) pr_finish(mydiff, partial, s_red, blockIdx.x);
}

// The loop test folded into the next sweep (schedule pr_mode 2, grids of at
// most kLagMaxGrid workgroups). Sweep r of a captured batch (r >= 1) starts
// by summing sweep r-1's shares -- every workgroup the same shares in the
// same order as pr_check, so all of them reach the same verdict -- and
// evaluates (diff > beta && iter < maxIter) for iteration st->iter + r. On
// "stop", workgroup 0 records the count and raises the flag, and the sweep
// returns without writing. The shares alternate between two buffers, so
// sweep r reads r-1's while writing its own. Sweep 0 of a batch only reads
// the flag: the pr_check that ends every batch has judged the batch before.
// One launch per iteration instead of two; the extra read is at most 8 KB
// per workgroup.
constexpr int kLagMaxGrid = 2048;

// pr_check's sum of partial[0..n), reproduced by a 256-lane workgroup in two
// steps (load, then reduce), so the loads can be issued before a sweep's
// gathers and summed after them. Lane tid stands for pr_check's threads tid,
// tid+256, tid+512 and tid+768, whose lanes sit in waves wid, wid+4, wid+8
// and wid+12 of the 1024-lane check. With n <= kLagMaxGrid (n4 <= 512) each
// of those threads reads at most one 16-byte word (pr_check's second
// accumulator stays +0) and thread tid < (n & 3) one scalar.
struct LagShares {
  float4 x[4];
  float tail;
};

__device__ __forceinline__ LagShares lag_load(const float* __restrict__ partial, int n) {
  const int tid = threadIdx.x;
  const float4* p4 = reinterpret_cast<const float4*>(partial);
  const int n4 = n >> 2;
  LagShares l;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int vt = tid + k * kPrBlock;
    l.x[k] = vt < n4 ? p4[vt] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  l.tail = tid < (n & 3) ? partial[(n4 << 2) + tid] : 0.f;
  return l;
}

__device__ __forceinline__ float lag_reduce(const LagShares& l, int n, float* s_red16) {
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  float t[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 b = a;
    a.x += l.x[k].x; a.y += l.x[k].y; a.z += l.x[k].z; a.w += l.x[k].w;
    t[k] = ((a.x + b.x) + (a.y + b.y)) + ((a.z + b.z) + (a.w + b.w));
    if (k == 0 && tid < (n & 3)) t[k] += l.tail;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float w = wave_sum(t[k]);
    if (lane == 0) s_red16[wid + 4 * k] = w;
  }
  __syncthreads();
  float d = 0.f;
  for (int w = 0; w < kCheckBlock / kWave; ++w) d += s_red16[w];
  return d;
}

// OVERLAP: the sums are formed before the verdict and written after it (the
// shares' loads overlap the gathers); otherwise the verdict comes first and
// the sweep writes as it goes. The first pays on skewed graphs, whose long
// rows set the sweep's length (RMAT-16 0.917 -> 0.836 ms a run), the second
// on even ones, where the barrier between the sums and the writes costs
// more (ER 2^16 0.826 against 0.859 ms).
template <int G, bool OVERLAP>
__global__ __launch_bounds__(kPrBlock) void pr_iter_lag(
    const int32_t* __restrict__ roff, const int32_t* __restrict__ radj, const int4* __restrict__ rows,
    const float* __restrict__ degf, PrBins bins, const float* __restrict__ pr_in, const float* __restrict__ c_in,
    float* __restrict__ pr_out, float* __restrict__ c_out, PrState* __restrict__ st,
    const float* __restrict__ prev, float* __restrict__ partial, int r) {
  __shared__ float s_red[kCheckBlock / kWave];
  const int4 row = pr_row<G>(rows, blockIdx.x, bins);
  const PrState s = *st;
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  // The previous sweep's shares load first and the sweep's sums are formed
  // while they are in flight; the verdict then decides whether the sums are
  // written (a stopping sweep discards them), so the loop test costs no
  // round trip of its own.
  LagShares shares;
  if (r > 0) shares = lag_load(prev, bins.blk[4]);
  // In-degree-0 vertices only, from the third sweep on (see pr_iter).
  const bool zero = int(blockIdx.x) < bins.zb && s.iter + r >= 2;
  int32_t ev = -1;
  float esum = 0.f, eold = 0.f, edg = 0.f;
  if (OVERLAP && !zero)
    pr_block_sums<G, false>(blockIdx.x, row, roff, radj, degf, bins, pr_in, c_in, PrSeg{nullptr, nullptr, 1, 0},
                            PrJobs{}, s_red, ev, esum, eold, edg);
  if (r > 0) {
    // Every workgroup computes the same d and it. Workgroup 0 alone writes;
    // a workgroup that reads its new iter or flag also stops (a larger
    // count fails the test too), so the verdict is the same everywhere.
    if (OVERLAP) __syncthreads();  // s_red: the sums' reduction is done with it
    const float d = lag_reduce(shares, bins.blk[4], s_red);
    const int it = s.iter + r;
    if (!(d > s.beta && it < s.maxIter)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->iter = it;
        __hip_atomic_store(&st->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      return;
    }
    __syncthreads();  // s_red is reused below
  }
  if (zero) {
    if (threadIdx.x == 0) partial[blockIdx.x] = 0.f;
    return;
  }
  const PrSeg sg{nullptr, nullptr, 1, 0};
  const PrJobs jb{};
  float mydiff = 0.f;
  if (OVERLAP) {
    if (ev >= 0) pr_emit<false>(ev, esum, sg, s, eold, edg, pr_out, c_out, mydiff, jb);
  } else {
    pr_sweep_block<G, false>(blockIdx.x, row, roff, radj, degf, bins, pr_in, c_in, pr_out, c_out, s, sg, jb, s_red,
                             mydiff);
  }
  pr_finish(mydiff, partial, s_red, blockIdx.x);
}

// ---- staged sweep (schedule pr_mode 3) ---------------------------------------
// On the degree-ordered copy of a large skewed graph the sources with the most
// out-edges have the smallest ids, and a few of them carry much of every
// sweep's gathers (at RMAT-22 the first 40,832 sources hold 56% of the
// in-edges). Their contributions are L2 hits that still cost an L1 miss each,
// and L1 misses are what the gather-bound sweep waits on. The staged sweep runs
// one persistent workgroup of kHotBlock lanes per CU; each first copies the
// contributions of sources [0, H) into LDS (159.5 KiB of the CU's 160 KiB) and
// then reads those terms there, so only the other terms leave the CU. The
// layout is binned for kHotBlock-lane blocks (a "virtual block" per step); a
// workgroup walks virtual blocks blockIdx.x, blockIdx.x + gridDim.x, ...
// without a barrier between them (bin 3, a whole block per vertex, reduces
// through LDS, so its threshold is set high), and the diff is summed per
// workgroup at the end. Measured at RMAT-22 (100 iterations): 33.5 ms for the
// plain sweep, 30.0 ms staging 20,416 sources in two workgroups per CU,
// 28.4 ms staging 40,832 in one (bin 3 above 4096 in-edges).
//
// The loop test is folded in as in pr_iter_lag: sweep r >= 1 of a captured
// batch first sums sweep r-1's shares (prev[0..nprev), pr_check's order, so
// every workgroup reaches pr_check's verdict) and stops the run itself; the
// batch's last sweep is judged by one pr_check.
constexpr int kHotBlock = 1024;
constexpr int kHotPerCu = 1;
constexpr int kHotMax = 40832;
static_assert(kHotBlock == kCheckBlock, "the staged sweep reproduces pr_check's sum with its own lanes");

template <int G>
__global__ __launch_bounds__(kHotBlock) void pr_iter_hot(
    const int32_t* __restrict__ roff, const int32_t* __restrict__ radj, const int4* __restrict__ rows,
    const float* __restrict__ degf, PrBins bins, const float* __restrict__ pr_in, const float* __restrict__ c_in,
    float* __restrict__ pr_out, float* __restrict__ c_out, PrState* __restrict__ st,
    const float* __restrict__ prev, int nprev, float* __restrict__ partial, int32_t H, int r) {
  __shared__ float4 s_hot4[kHotMax / 4];
  __shared__ float s_red[kHotBlock / kWave];
  const PrState s = *st;
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  if (r > 0) {
    // Same verdict in every workgroup; workgroup 0 alone writes (see pr_iter_lag).
    const float d = check_sum(prev, nprev, s_red);
    const int it = s.iter + r;
    if (!(d > s.beta && it < s.maxIter)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->iter = it;
        __hip_atomic_store(&st->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      return;
    }
    __syncthreads();  // s_red is reused by the sweep
  }
  const int tid = threadIdx.x;
  // H is a multiple of 4 and at most kHotMax; c_in is 256-byte aligned.
  const float4* c4 = reinterpret_cast<const float4*>(c_in);
  for (int i = tid; i < H / 4; i += kHotBlock) s_hot4[i] = c4[i];
  __syncthreads();
  const PrHot hot{reinterpret_cast<const float*>(s_hot4), H};
  const PrSeg sg{nullptr, nullptr, 1, 0};
  const PrJobs jb{};
  float mydiff = 0.f;
  // This sweep's iteration: st->iter advances once per batch.
  const bool skip0 = s.iter + r >= 2;
  for (int b = blockIdx.x; b < bins.blk[4]; b += gridDim.x) {
    // In-degree-0 vertices only: from the third sweep on |new - old| = 0.
    if (b < bins.zb && skip0) continue;
    const int4 row = pr_row<G, kHotBlock>(rows, b, bins);
    pr_sweep_block<G, false, kHotBlock, true>(b, row, roff, radj, degf, bins, pr_in, c_in, pr_out, c_out, s, sg, jb,
                                              s_red, mydiff, hot);
  }
  pr_finish<kHotBlock>(mydiff, partial, s_red, blockIdx.x);
}

// ---- batched jobs -------------------------------------------------------------

// job[v] = k for v in [bases[k], bases[k+1]) (binary search over n+1 bases).
__global__ void pr_job_ids(int32_t V, const int32_t* __restrict__ bases, int32_t n, int32_t* __restrict__ job) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    int32_t lo = 0, hi = n - 1;  // last k with bases[k] <= v
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (bases[mid] <= v) lo = mid;
      else hi = mid - 1;
    }
    job[v] = lo;
  }
}

// flag |= 1 if some edge u -> x joins two jobs (one wavefront per row).
__global__ __launch_bounds__(kPrBlock) void pr_job_edges(int32_t V, const int32_t* __restrict__ off,
                                                         const int32_t* __restrict__ adj,
                                                         const int32_t* __restrict__ job,
                                                         unsigned int* __restrict__ flag) {
  const int lane = __lane_id();
  const int64_t nw = int64_t(gridDim.x) * (kPrBlock / kWave);
  bool cross = false;
  for (int64_t x_ = int64_t(blockIdx.x) * (kPrBlock / kWave) + threadIdx.x / kWave; x_ < V; x_ += nw) {
    const int64_t u = spread_vertex(x_, V);
    const int32_t k = job[u];
    for (int32_t e = off[u] + lane; e < off[u + 1]; e += kWave) cross |= job[adj[e]] != k;
  }
  if (__any(cross) && lane == 0) atomicOr(flag, 1u);
}

__global__ void pr_init_multi(int32_t V, const int32_t* __restrict__ off, const int32_t* __restrict__ job,
                              const float* __restrict__ init, float* __restrict__ degf, float* __restrict__ pr,
                              float* __restrict__ c, PrState* st, int maxIter, float beta, float delta) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->done = 0;
    st->iter = 0;
    st->maxIter = maxIter;
    st->beta = beta;
    st->delta = delta;
    st->base = 0.f;  // per job
  }
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    const float d = float(off[v + 1] - off[v]);
    const float x = init[job[v]];
    degf[v] = d;
    pr[v] = x;
    c[v] = d > 0.f ? x / d : 0.f;
  }
}

// One workgroup per job: the loop test of job k on the sum of dv over its
// vertices, in a fixed order (so a batched job stops where the same sum
// tells it to, run after run).
__global__ __launch_bounds__(kPrBlock) void pr_job_check(const int32_t* __restrict__ bases, const float* __restrict__ dv,
                                                         int32_t* __restrict__ live, int32_t* __restrict__ iters,
                                                         const PrState* __restrict__ st) {
  __shared__ float s_red[kPrBlock / kWave];
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const int k = blockIdx.x;
  if (!live[k]) return;
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  float t = 0.f;
  for (int32_t v = bases[k] + tid; v < bases[k + 1]; v += kPrBlock) t += dv[v];
  t = wave_sum(t);
  if (lane == 0) s_red[wid] = t;
  __syncthreads();
  if (tid == 0) {
    float d = 0.f;
    for (int w = 0; w < kPrBlock / kWave; ++w) d += s_red[w];
    const int it = iters[k] + 1;
    iters[k] = it;
    if (!(d > st->beta && it < st->maxIter)) live[k] = 0;
  }
}

// One workgroup: the run is done when no job is live.
__global__ __launch_bounds__(kPrBlock) void pr_jobs_done(const int32_t* __restrict__ live, int32_t n,
                                                         PrState* __restrict__ st) {
  __shared__ int s_any;
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  bool any = false;
  for (int32_t k = threadIdx.x; k < n; k += kPrBlock) any |= live[k] != 0;
  if (__any(any) && __lane_id() == 0) s_any = 1;
  __syncthreads();
  if (threadIdx.x == 0) {
    st->iter += 1;
    if (!s_any) st->done = 1;
  }
}

// out[v] = the rank buffer of job[v]'s last iteration.
__global__ void pr_job_select(int32_t V, const int32_t* __restrict__ job, const int32_t* __restrict__ iters,
                              const float* __restrict__ pr0, const float* __restrict__ pr1, float* __restrict__ out) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride)
    out[v] = (iters[job[v]] & 1) ? pr1[v] : pr0[v];
}

// Split points of every in-row at the segment boundaries (one lane per vertex,
// binary search in the ascending row).
__global__ void pr_split(int32_t V, const int32_t* __restrict__ roff, const int32_t* __restrict__ radj, int S,
                         int seg_log2, int32_t* __restrict__ sp) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += stride) {
    const int32_t a = roff[v], b = roff[v + 1];
    for (int s = 1; s < S; ++s) {
      const int32_t x = int32_t(int64_t(s) << seg_log2);
      int32_t lo = a, hi = b;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (radj[mid] < x) lo = mid + 1;
        else hi = mid;
      }
      sp[v * (S - 1) + (s - 1)] = lo;
    }
  }
}

// Push iteration, step 1: every out-edge (u,v) adds u's contribution to
// acc[v] (hardware FP32 atomics; the order of additions is unspecified, so
// sums differ from the pull order by round-off only). A wavefront flattens
// the out-edges of 64 consecutive vertices across its lanes.
__global__ __launch_bounds__(kPrBlock) void pr_push_scatter(int32_t V, const int32_t* __restrict__ off,
                                                            const int32_t* __restrict__ adj,
                                                            const float* __restrict__ c_in, float* __restrict__ acc,
                                                            const PrState* __restrict__ st) {
  __shared__ uint32_t s_pre[kPrBlock / kWave][kWave];
  __shared__ int32_t s_base[kPrBlock / kWave][kWave];
  __shared__ float s_c[kPrBlock / kWave][kWave];
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  const int64_t nw = int64_t(gridDim.x) * (kPrBlock / kWave);
  for (int64_t base = (int64_t(blockIdx.x) * (kPrBlock / kWave) + wid) * kWave; base < V; base += nw * kWave) {
    const int64_t u = base + lane;
    uint32_t deg = 0;
    int32_t b = 0;
    float cu = 0.f;
    if (u < V) {
      b = off[u];
      deg = uint32_t(off[u + 1] - b);
      cu = c_in[u];
    }
    const uint32_t incl = wave_incl_scan(deg);
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    s_pre[wid][lane] = incl - deg;
    s_base[wid][lane] = b;
    s_c[wid][lane] = cu;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t x = lane; x < total; x += kWave) {
      int lo = 0, hi = kWave - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[wid][mid] <= x) lo = mid;
        else hi = mid - 1;
      }
      const int32_t v = adj[s_base[wid][lo] + int32_t(x - s_pre[wid][lo])];
      unsafeAtomicAdd(&acc[v], s_c[wid][lo]);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Push iteration, step 2: one lane per vertex applies the update, clears its
// accumulator for the next iteration and joins the convergence test.
__global__ __launch_bounds__(kPrBlock) void pr_push_apply(int32_t V, const float* __restrict__ degf,
                                                          float* __restrict__ acc, const float* __restrict__ pr_in,
                                                          float* __restrict__ pr_out, float* __restrict__ c_out,
                                                          const PrState* __restrict__ st,
                                                          float* __restrict__ partial) {
  __shared__ float s_red[kPrBlock / kWave];
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const PrState s = *st;
  float mydiff = 0.f;
  const int64_t stride = int64_t(gridDim.x) * kPrBlock;
  for (int64_t v = int64_t(blockIdx.x) * kPrBlock + threadIdx.x; v < V; v += stride) {
    const float sum = acc[v];
    acc[v] = 0.f;
    pr_store(int32_t(v), sum, s, pr_in[v], degf[v], pr_out, c_out, mydiff);
  }
  pr_finish(mydiff, partial, s_red, blockIdx.x);
}

// ---- in-degree bins on the device -------------------------------------------
// Vertices are grouped by in-degree into the four bins of the lane mapping,
// in id order within each bin (the order the kernels' per-workgroup diff sums
// follow, so it is fixed). Three launches: per-chunk bin counts, one scan,
// and a stable placement (ballot ranks inside each wavefront).
struct PrThresh {
  int32_t t1, t2, t3;
};
constexpr int kBinBlock = 256;
constexpr int kBinChunk = 4 * kBinBlock;  // vertices per binning workgroup

// Placement categories: in-degree 0 first, then the rest of bin 0, bins 1-3;
// category c > 0 is bin c - 1, and category 0 heads bin 0.
constexpr int kCat = 5;
__device__ __forceinline__ int pr_bin_of(int32_t d, const PrThresh& t) {
  return d == 0 ? 0 : d <= t.t1 ? 1 : d <= t.t2 ? 2 : d <= t.t3 ? 3 : 4;
}

// cnt[c * nblk + blk] = vertices of category c among chunk blk's vertices.
__global__ __launch_bounds__(kBinBlock) void pr_bin_count(int32_t V, const int32_t* __restrict__ roff, PrThresh t,
                                                          int32_t nblk, int32_t* __restrict__ cnt) {
  __shared__ int32_t s_cnt[kCat];
  if (threadIdx.x < kCat) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  int32_t mine[kCat] = {0, 0, 0, 0, 0};
  const int64_t base = int64_t(blockIdx.x) * kBinChunk;
  for (int k = 0; k < kBinChunk / kBinBlock; ++k) {
    const int64_t v = base + k * kBinBlock + threadIdx.x;
    if (v < V) ++mine[pr_bin_of(roff[v + 1] - roff[v], t)];
  }
  for (int b = 0; b < kCat; ++b) {
    const int32_t w = wave_sum(mine[b]);
    if (__lane_id() == 0 && w) atomicAdd(&s_cnt[b], w);
  }
  __syncthreads();
  if (threadIdx.x < kCat) cnt[int64_t(threadIdx.x) * nblk + blockIdx.x] = s_cnt[threadIdx.x];
}

// One workgroup: cnt (category-major) becomes the exclusive prefix over all
// (category, chunk) pairs, i.e. each chunk's first position within the
// category order; nb[c] receives the size of category c.
__global__ __launch_bounds__(1024) void pr_bin_scan(int32_t nblk, int32_t* __restrict__ cnt, int32_t* __restrict__ nb) {
  __shared__ int32_t s_wave[1024 / kWave];
  __shared__ int32_t s_carry;
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  const int64_t n = int64_t(kCat) * nblk;
  for (int64_t base = 0; base < n; base += 1024) {
    const int64_t i = base + tid;
    const int32_t x = i < n ? cnt[i] : 0;
    const int32_t incl = wave_incl_scan(x);
    if (lane == kWave - 1) s_wave[wid] = incl;
    __syncthreads();
    int32_t before = s_carry;
    for (int w = 0; w < wid; ++w) before += s_wave[w];
    if (i < n) cnt[i] = before + incl - x;
    __syncthreads();
    if (tid == 0) {
      int32_t t = 0;
      for (int w = 0; w < 1024 / kWave; ++w) t += s_wave[w];
      s_carry += t;
    }
    __syncthreads();
  }
  // category sizes: start of the next category minus start of this one
  if (tid < kCat) {
    const int32_t a = cnt[int64_t(tid) * nblk];
    const int32_t b = tid < kCat - 1 ? cnt[int64_t(tid + 1) * nblk] : s_carry;
    nb[tid] = b - a;
  }
}

// rows[pos] = {v, first in-edge, end, 0} with pos = the chunk's start in v's
// category plus v's rank among the chunk's earlier vertices of the same
// category (so bin 0's rows are its in-degree-0 vertices, then the rest).
__global__ __launch_bounds__(kBinBlock) void pr_bin_place(int32_t V, const int32_t* __restrict__ roff, PrThresh t,
                                                          int32_t nblk, const int32_t* __restrict__ start,
                                                          int4* __restrict__ rows) {
  __shared__ int32_t s_next[kCat];
  __shared__ int32_t s_wave[kCat][kBinBlock / kWave];
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  if (tid < kCat) s_next[tid] = start[int64_t(tid) * nblk + blockIdx.x];
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * kBinChunk;
  const uint64_t below = (uint64_t(1) << lane) - 1;
  for (int k = 0; k < kBinChunk / kBinBlock; ++k) {
    const int64_t v = base + k * kBinBlock + tid;
    int b = -1;
    int32_t a = 0, z = 0;
    if (v < V) {
      a = roff[v];
      z = roff[v + 1];
      b = pr_bin_of(z - a, t);
    }
    int32_t rank = 0;
    for (int bb = 0; bb < kCat; ++bb) {
      const uint64_t m = __ballot(b == bb);
      if (b == bb) rank = __popcll(m & below);
      if (lane == 0) s_wave[bb][wid] = __popcll(m);
    }
    __syncthreads();
    if (b >= 0) {
      int32_t pos = s_next[b] + rank;
      for (int w = 0; w < wid; ++w) pos += s_wave[b][w];
      rows[pos] = make_int4(int32_t(v), a, z, 0);
    }
    __syncthreads();
    if (tid < kCat) {
      int32_t tot = 0;
      for (int w = 0; w < kBinBlock / kWave; ++w) tot += s_wave[tid][w];
      s_next[tid] += tot;
    }
    __syncthreads();
  }
}

// Bins the in-rows of `roff` into `rows` (int4[V]) with the schedule's
// thresholds; cnt is scratch of kCat * ceil(V / kBinChunk) + kCat ints. Reads
// the category sizes back (one synchronisation) and fills the launch layout
// for workgroups of B lanes.
static int pr_bin_device(spm_graph* g, const int32_t* roff, int4* rows, int32_t* cnt, int32_t* h_nb, PrBins* out,
                         int B = kPrBlock) {
  const spm_schedule& s = g->sched;
  const int32_t V = g->V;
  const int32_t nblk = int32_t((int64_t(V) + kBinChunk - 1) / kBinChunk);
  const PrThresh t{s.pr_t1, s.pr_t2, s.pr_t3};
  int32_t* nb = cnt + int64_t(kCat) * nblk;
  SPM_TRY(launch(g, "pr_bin_count", pr_bin_count, dim3(nblk), dim3(kBinBlock), 0, V, roff, t, nblk, cnt));
  SPM_TRY(launch(g, "pr_bin_scan", pr_bin_scan, dim3(1), dim3(1024), 0, nblk, cnt, nb));
  SPM_TRY(launch(g, "pr_bin_place", pr_bin_place, dim3(nblk), dim3(kBinBlock), 0, V, roff, t, nblk,
                 (const int32_t*)cnt, rows));
  SPM_CHECK_HIP(hipMemcpyAsync(h_nb, nb, kCat * sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  PrBins& b = *out;
  b.nb[0] = h_nb[0] + h_nb[1];
  for (int i = 1; i < 4; ++i) b.nb[i] = h_nb[i + 1];
  b.zb = h_nb[0] / B;
  b.group = s.pr_group;
  b.blk[0] = 0;
  b.blk[1] = (b.nb[0] + B - 1) / B;
  b.blk[2] = b.blk[1] + (b.nb[1] + B / s.pr_group - 1) / (B / s.pr_group);
  b.blk[3] = b.blk[2] + (b.nb[2] + B / kWave - 1) / (B / kWave);
  b.blk[4] = b.blk[3] + b.nb[3];
  return SPM_OK;
}

static size_t pr_bin_scratch_ints(int32_t V) {
  return size_t(kCat) * size_t((int64_t(V) + kBinChunk - 1) / kBinChunk) + kCat;
}

// Ranks, contributions and in-degree bins over `view`. Batched jobs (multi)
// use plain pull sweeps: no push accumulators, no source segments.
static int pr_build(spm_graph* g, PrWork* w, const CsrView& view, bool multi) {
  w->view = view;
  if (view.perm) SPM_TRY(dmalloc(g, (void**)&w->out, sizeof(float) * size_t(g->V), "pr result"));
  const int32_t V = g->V;
  for (int i = 0; i < 2; ++i) {
    SPM_TRY(dmalloc(g, (void**)&w->pr[i], sizeof(float) * V, "pr ranks"));
    SPM_TRY(dmalloc(g, (void**)&w->c[i], sizeof(float) * V, "pr contributions"));
  }
  SPM_TRY(dmalloc(g, (void**)&w->degf, sizeof(float) * V, "pr outdeg"));
  if (!multi) SPM_TRY(dmalloc(g, (void**)&w->acc, sizeof(float) * V, "pr push accumulators"));
  SPM_TRY(dmalloc(g, (void**)&w->rows, sizeof(int4) * V, "pr bins"));
  SPM_TRY(dmalloc(g, (void**)&w->st, sizeof(PrState), "pr state"));
  const spm_schedule& s = g->sched;
  // Bin vertices by in-degree on the device (once per handle and schedule).
  {
    int32_t* cnt = nullptr;
    const size_t nc = pr_bin_scratch_ints(V);
    SPM_TRY(dmalloc(g, (void**)&cnt, sizeof(int32_t) * nc, "pr bin counts"));
    w->hot = !multi && s.pr_mode == 3;
    const int rc = pr_bin_device(g, view.roff, w->rows, cnt, g->h_pinned, &w->bins, w->hot ? kHotBlock : kPrBlock);
    dfree(g, cnt, sizeof(int32_t) * nc);
    SPM_TRY(rc);
  }
  w->S = 1;
  if (multi) {
    w->bins.zb = 0;  // batched jobs sweep every vertex every iteration
    return SPM_OK;
  }
  const PrBins& b = w->bins;
  w->partial_n = std::max({b.blk[4], g->num_cus * 8, 2 * kLagMaxGrid}) + 1;
  SPM_TRY(dmalloc(g, (void**)&w->partial, sizeof(float) * size_t(w->partial_n), "pr partial diffs"));
  // Source segments for cache blocking.
  if (s.pr_seg_log2 > 0) w->S = int((int64_t(V) + (int64_t(1) << s.pr_seg_log2) - 1) >> s.pr_seg_log2);
  if (w->S > 1) {
    w->sp_n = size_t(V) * size_t(w->S - 1);
    SPM_TRY(dmalloc(g, (void**)&w->sp, sizeof(int32_t) * w->sp_n, "pr segment splits"));
    const int grid = int(std::min<int64_t>((V + 255) / 256, 4096));
    SPM_TRY(launch(g, "pr_split", pr_split, dim3(grid), dim3(256), 0, V, view.roff, view.radj, w->S,
                   s.pr_seg_log2, w->sp));
  }
  return SPM_OK;
}

static int pr_prepare(spm_graph* g, const CsrView& view) {
  if (g->pr) return SPM_OK;
  if (!g->roff) return set_error(SPM_ENOGRAPH, "PageRank needs the reverse CSR (rev_indexofNodes, srcList)");
  g->pr = new PrWork();
  return pr_build(g, g->pr, view, false);
}

static void pr_free_work(spm_graph* g, PrWork* w) {
  if (w->exec) (void)hipGraphExecDestroy(w->exec);
  if (w->graph) (void)hipGraphDestroy(w->graph);
  for (int i = 0; i < 2; ++i) {
    dfree(g, w->pr[i], sizeof(float) * g->V);
    dfree(g, w->c[i], sizeof(float) * g->V);
  }
  dfree(g, w->degf, sizeof(float) * g->V);
  dfree(g, w->acc, sizeof(float) * g->V);
  dfree(g, w->partial, sizeof(float) * size_t(w->partial_n));
  dfree(g, w->sp, sizeof(int32_t) * w->sp_n);
  dfree(g, w->rows, sizeof(int4) * g->V);
  dfree(g, w->out, sizeof(float) * g->V);
  dfree(g, w->st, sizeof(PrState));
}

// Workspace of batched jobs (spm_pr_multi_device), over the original ids.
struct PrMulti {
  PrWork w;
  int32_t* job = nullptr;        // [V]
  float* dv = nullptr;           // [V]
  int32_t* bases = nullptr;      // [cap + 1]
  int32_t* live = nullptr;       // [cap]
  int32_t* iters = nullptr;      // [cap]
  float* base = nullptr;         // [cap] (1 - delta) / N_k
  float* init = nullptr;         // [cap] 1 / N_k
  unsigned int* flag = nullptr;  // [1] cross-job edge found
  int32_t cap = 0;
  int32_t n_graph = 0;           // job count the captured iterations were built for
};

static void pr_multi_free_jobs(spm_graph* g, PrMulti* m) {
  dfree(g, m->bases, sizeof(int32_t) * (size_t(m->cap) + 1));
  dfree(g, m->live, sizeof(int32_t) * size_t(m->cap));
  dfree(g, m->iters, sizeof(int32_t) * size_t(m->cap));
  dfree(g, m->base, sizeof(float) * size_t(m->cap));
  dfree(g, m->init, sizeof(float) * size_t(m->cap));
  m->bases = m->live = m->iters = nullptr;
  m->base = m->init = nullptr;
  m->cap = 0;
}

int pr_release(spm_graph* g) {
  if (PrMulti* m = g->prm) {
    pr_free_work(g, &m->w);
    pr_multi_free_jobs(g, m);
    dfree(g, m->job, sizeof(int32_t) * g->V);
    dfree(g, m->dv, sizeof(float) * g->V);
    dfree(g, m->flag, sizeof(unsigned int));
    delete m;
    g->prm = nullptr;
  }
  PrWork* w = g->pr;
  if (!w) return SPM_OK;
  pr_free_work(g, w);
  delete w;
  g->pr = nullptr;
  return SPM_OK;
}

// Schedule pr_mode 2 takes effect on pull sweeps of at most kLagMaxGrid
// workgroups (two share buffers fit in the partial array); larger sweeps run
// as pr_mode 0.
static bool pr_lagged(const spm_graph* g, const PrWork* w) {
  return g->sched.pr_mode == 2 && w->S == 1 && w->bins.blk[4] <= kLagMaxGrid && w->partial_n >= 2 * kLagMaxGrid;
}

// R iterations of w's sweep (push when `push`, else the binned pull), each
// followed by the loop test.
static int pr_enqueue_iters(spm_graph* g, PrWork* w, int R, bool push) {
  const dim3 grid(std::max(w->bins.blk[4], 1)), block(kPrBlock);
  const int pgrid = int(std::min<int64_t>((g->V + kPrBlock - 1) / kPrBlock, int64_t(g->num_cus) * 8));
  for (int r = 0; r < R; ++r) {
    const int p = r & 1;
    const float* pin = w->pr[p];
    const float* cin = w->c[p];
    if (push) {
      SPM_TRY(launch(g, "pr_push_scatter", pr_push_scatter, dim3(pgrid), block, 0, g->V, w->view.off, w->view.adj,
                     cin, w->acc, (const PrState*)w->st));
      SPM_TRY(launch(g, "pr_push_apply", pr_push_apply, dim3(pgrid), block, 0, g->V, (const float*)w->degf, w->acc,
                     pin, w->pr[1 - p], w->c[1 - p], (const PrState*)w->st, w->partial));
      SPM_TRY(launch(g, "pr_check", pr_check, dim3(1), dim3(kCheckBlock), 0, (const float*)w->partial, pgrid, w->st,
                     1));
      continue;
    }
    if (w->hot) {
      const int hgrid = std::max(1, std::min({w->bins.blk[4], kHotPerCu * g->num_cus, kLagMaxGrid}));
      const int32_t H = std::min<int32_t>(g->V, kHotMax) & ~3;
      float* mine = w->partial + (r & 1) * kLagMaxGrid;
      const float* prev = w->partial + (1 - (r & 1)) * kLagMaxGrid;
      switch (w->bins.group) {
#define SPM_PR_HOT_LAUNCH(G)                                                                                  \
  case G:                                                                                                     \
    SPM_TRY(launch(g, "pr_iter_hot", pr_iter_hot<G>, dim3(hgrid), dim3(kHotBlock), 0, w->view.roff,           \
                   w->view.radj, (const int4*)w->rows, (const float*)w->degf, w->bins, pin, cin, w->pr[1 - p], \
                   w->c[1 - p], w->st, prev, hgrid, mine, H, r));                                             \
    break;
        SPM_PR_HOT_LAUNCH(2)
        SPM_PR_HOT_LAUNCH(4)
        SPM_PR_HOT_LAUNCH(8)
        SPM_PR_HOT_LAUNCH(16)
        SPM_PR_HOT_LAUNCH(32)
#undef SPM_PR_HOT_LAUNCH
        default: return set_error(SPM_EINVAL, "pr: bad lane group");
      }
      if (r == R - 1)
        SPM_TRY(launch(g, "pr_check", pr_check, dim3(1), dim3(kCheckBlock), 0, (const float*)mine, hgrid, w->st, R));
      continue;
    }
    if (pr_lagged(g, w)) {
      // Skewed in-degrees (the largest out-degree 16x the mean or more, as
      // a proxy): sums before the verdict.
      const double avg = g->V > 0 ? double(g->E) / double(g->V) : 0.0;
      const bool overlap = double(g->max_outdeg) >= 16.0 * avg;
      float* mine = w->partial + (r & 1) * kLagMaxGrid;
      const float* prev = w->partial + (1 - (r & 1)) * kLagMaxGrid;
      switch (w->bins.group) {
#define SPM_PR_LAG_LAUNCH(G)                                                                               \
  case G:                                                                                                  \
    SPM_TRY(launch(g, "pr_iter_lag", overlap ? pr_iter_lag<G, true> : pr_iter_lag<G, false>, grid, block, 0, \
                   w->view.roff, w->view.radj,                                                             \
                   (const int4*)w->rows, (const float*)w->degf, w->bins, pin, cin, w->pr[1 - p], w->c[1 - p], \
                   w->st, prev, mine, r));                                                                 \
    break;
        SPM_PR_LAG_LAUNCH(2)
        SPM_PR_LAG_LAUNCH(4)
        SPM_PR_LAG_LAUNCH(8)
        SPM_PR_LAG_LAUNCH(16)
        SPM_PR_LAG_LAUNCH(32)
#undef SPM_PR_LAG_LAUNCH
        default: return set_error(SPM_EINVAL, "pr: bad lane group");
      }
      // The batch's last sweep is judged here, with the batch's R
      // iterations counted at once.
      if (r == R - 1)
        SPM_TRY(launch(g, "pr_check", pr_check, dim3(1), dim3(kCheckBlock), 0, (const float*)mine, int(grid.x),
                       w->st, R));
      continue;
    }
    for (int seg = 0; seg < w->S; ++seg) {
      const PrSeg sg{w->sp, w->acc, w->S, seg};
#define SPM_PR_LAUNCH(G)                                                                          \
  case G:                                                                                         \
    SPM_TRY(launch(g, "pr_iter", pr_iter<G, false>, grid, block, 0, w->view.roff,                 \
                   w->view.radj, (const int4*)w->rows, (const float*)w->degf,                     \
                   w->bins, pin, cin, w->pr[1 - p], w->c[1 - p], (const PrState*)w->st,             \
                   w->partial, sg, PrJobs{}));                                                     \
    break;
      switch (w->bins.group) {
        SPM_PR_LAUNCH(2)
        SPM_PR_LAUNCH(4)
        SPM_PR_LAUNCH(8)
        SPM_PR_LAUNCH(16)
        SPM_PR_LAUNCH(32)
        default: return set_error(SPM_EINVAL, "pr: bad lane group");
      }
#undef SPM_PR_LAUNCH
    }
    SPM_TRY(launch(g, "pr_check", pr_check, dim3(1), dim3(kCheckBlock), 0, (const float*)w->partial,
                   int(grid.x), w->st, 1));
  }
  return SPM_OK;
}

static int pr_finish_run(spm_graph* g, PrWork* w, int32_t iters, float* d_out, float* h_out, int32_t* iter_out);

int pr_run(spm_graph* g, float beta, float delta, int32_t maxIter, float* d_out, float* h_out,
           int32_t* iter_out) {
  SPM_TRY(dyn_settle(g));
  CsrView view;
  SPM_TRY(csr_view(g, &view));
  SPM_TRY(pr_prepare(g, view));
  PrWork* w = g->pr;
  const int32_t V = g->V;
  const float num_nodes = float(V);
  const float init = 1 / num_nodes;
  const float base = (1 - delta) / num_nodes;
  const int igrid = int(std::min<int64_t>((V + 255) / 256, 4096));
  SPM_TRY(launch(g, "pr_init", pr_init, dim3(igrid), dim3(256), 0, V, w->view.off, w->degf,
                 w->pr[0], w->c[0], w->acc, w->st, init, int(maxIter), beta, delta, base));
  int32_t* h = g->h_pinned;
  // Iterations per replay: even (ping-pong parity), up to 128.
  int R = std::max(2, std::min(128, (int(std::max(maxIter, 1)) + 1) & ~1));
  const bool use_graph = !g->prof.on;
  if (use_graph && (!w->exec || w->R != R)) {
    if (w->exec) { (void)hipGraphExecDestroy(w->exec); w->exec = nullptr; }
    if (w->graph) { (void)hipGraphDestroy(w->graph); w->graph = nullptr; }
    SPM_CHECK_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = pr_enqueue_iters(g, w, R, g->sched.pr_mode == 1);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g->stream, &graph);
    if (rc != SPM_OK) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return set_error(SPM_EHIP, std::string("pr capture: ") + hipGetErrorString(e));
    w->graph = graph;
    SPM_CHECK_HIP(hipGraphInstantiate(&w->exec, graph, nullptr, nullptr, 0));
    w->R = R;
  }
  for (int64_t rep = 0;; ++rep) {
    if (use_graph) SPM_CHECK_HIP(hipGraphLaunch(w->exec, g->stream));
    else SPM_TRY(pr_enqueue_iters(g, w, R, g->sched.pr_mode == 1));
    SPM_CHECK_HIP(hipMemcpyAsync(h, &w->st->done, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
    SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    if (h[0]) break;
    if (rep > int64_t(maxIter) / R + 2) return set_error(SPM_EHIP, "pr: loop did not terminate");
  }
  return pr_finish_run(g, w, h[1], d_out, h_out, iter_out);
}

// The ranks of the run's last iteration to d_out / h_out, in original ids.
static int pr_finish_run(spm_graph* g, PrWork* w, int32_t iters, float* d_out, float* h_out, int32_t* iter_out) {
  const int32_t V = g->V;
  const float* res = w->pr[iters & 1];
  if (w->view.perm) {  // ranks in original ids
    SPM_TRY(gather_back(g, w->view, res, d_out ? d_out : w->out));
    res = d_out ? d_out : w->out;
  } else if (d_out) {
    SPM_CHECK_HIP(hipMemcpyAsync(d_out, res, sizeof(float) * V, hipMemcpyDeviceToDevice, g->stream));
  }
  if (h_out) SPM_CHECK_HIP(hipMemcpyAsync(h_out, res, sizeof(float) * V, hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  if (iter_out) *iter_out = iters;
  SPM_TRY(prof_flush(g));
  return SPM_OK;
}

// R iterations of batched jobs: the sweep over live jobs, the per-job loop
// tests, then the all-stopped test.
static int pr_multi_enqueue(spm_graph* g, int R, int32_t n) {
  PrMulti* m = g->prm;
  PrWork* w = &m->w;
  const dim3 grid(std::max(w->bins.blk[4], 1)), block(kPrBlock);
  const PrJobs jb{m->job, m->live, m->base, m->dv};
  const PrSeg sg{nullptr, nullptr, 1, 0};
  for (int r = 0; r < R; ++r) {
    const int p = r & 1;
#define SPM_PR_MULTI_LAUNCH(G)                                                                          \
  case G:                                                                                               \
    SPM_TRY(launch(g, "pr_iter_multi", pr_iter<G, true>, grid, block, 0, w->view.roff, w->view.radj,    \
                   (const int4*)w->rows, (const float*)w->degf, w->bins, (const float*)w->pr[p],         \
                   (const float*)w->c[p], w->pr[1 - p], w->c[1 - p], (const PrState*)w->st,              \
                   (float*)nullptr, sg, jb));                                                            \
    break;
    switch (w->bins.group) {
      SPM_PR_MULTI_LAUNCH(2)
      SPM_PR_MULTI_LAUNCH(4)
      SPM_PR_MULTI_LAUNCH(8)
      SPM_PR_MULTI_LAUNCH(16)
      SPM_PR_MULTI_LAUNCH(32)
      default: return set_error(SPM_EINVAL, "pr: bad lane group");
    }
#undef SPM_PR_MULTI_LAUNCH
    SPM_TRY(launch(g, "pr_job_check", pr_job_check, dim3(n), block, 0, (const int32_t*)m->bases,
                   (const float*)m->dv, m->live, m->iters, (const PrState*)w->st));
    SPM_TRY(launch(g, "pr_jobs_done", pr_jobs_done, dim3(1), block, 0, (const int32_t*)m->live, n, w->st));
  }
  return SPM_OK;
}

int pr_multi_run(spm_graph* g, const int32_t* bases, int32_t n, float beta, float delta, int32_t maxIter,
                 float* d_out, int32_t* iters_out) {
  SPM_TRY(dyn_settle(g));
  if (!g->roff) return set_error(SPM_ENOGRAPH, "PageRank needs the reverse CSR (rev_indexofNodes, srcList)");
  if (!bases || n < 1 || !d_out) return set_error(SPM_EINVAL, "spm_pr_multi_device: NULL argument or no jobs");
  if (bases[0] != 0 || bases[n] != g->V)
    return set_error(SPM_EINVAL, "spm_pr_multi_device: job bases must start at 0 and end at V");
  for (int32_t k = 0; k < n; ++k)
    if (bases[k + 1] <= bases[k])
      return set_error(SPM_EINVAL, "spm_pr_multi_device: job bases must be strictly ascending (no empty job)");
  const int32_t V = g->V;
  if (!g->prm) {
    PrMulti* m = new PrMulti();
    g->prm = m;
    CsrView view;
    view.off = g->off;
    view.adj = g->adj;
    view.w = g->w;
    view.roff = g->roff;
    view.radj = g->radj;
    view.perm = nullptr;
    SPM_TRY(pr_build(g, &m->w, view, true));
    SPM_TRY(dmalloc(g, (void**)&m->job, sizeof(int32_t) * size_t(V), "pr job ids"));
    SPM_TRY(dmalloc(g, (void**)&m->dv, sizeof(float) * size_t(V), "pr vertex diffs"));
    SPM_TRY(dmalloc(g, (void**)&m->flag, sizeof(unsigned int), "pr job check"));
  }
  PrMulti* m = g->prm;
  PrWork* w = &m->w;
  if (n > m->cap) {  // the captured iterations hold the old buffers
    if (w->exec) { (void)hipGraphExecDestroy(w->exec); w->exec = nullptr; }
    if (w->graph) { (void)hipGraphDestroy(w->graph); w->graph = nullptr; }
    pr_multi_free_jobs(g, m);
    const size_t c = size_t(n);
    SPM_TRY(dmalloc(g, (void**)&m->bases, sizeof(int32_t) * (c + 1), "pr job bases"));
    SPM_TRY(dmalloc(g, (void**)&m->live, sizeof(int32_t) * c, "pr job flags"));
    SPM_TRY(dmalloc(g, (void**)&m->iters, sizeof(int32_t) * c, "pr job iterations"));
    SPM_TRY(dmalloc(g, (void**)&m->base, sizeof(float) * c, "pr job base terms"));
    SPM_TRY(dmalloc(g, (void**)&m->init, sizeof(float) * c, "pr job initial ranks"));
    m->cap = n;
  }
  // Per-job constants as the single-job run computes them (host float arithmetic).
  const size_t nj = static_cast<size_t>(n);
  std::vector<float> hbase(nj), hinit(nj);
  std::vector<int32_t> ones(nj, 1), zeros(nj, 0);
  for (int32_t k = 0; k < n; ++k) {
    const float num_nodes = float(bases[k + 1] - bases[k]);
    hinit[k] = 1 / num_nodes;
    hbase[k] = (1 - delta) / num_nodes;
  }
  SPM_CHECK_HIP(hipMemcpyAsync(m->bases, bases, sizeof(int32_t) * (size_t(n) + 1), hipMemcpyHostToDevice, g->stream));
  SPM_CHECK_HIP(hipMemcpyAsync(m->live, ones.data(), sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice, g->stream));
  SPM_CHECK_HIP(hipMemcpyAsync(m->iters, zeros.data(), sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice, g->stream));
  SPM_CHECK_HIP(hipMemcpyAsync(m->base, hbase.data(), sizeof(float) * size_t(n), hipMemcpyHostToDevice, g->stream));
  SPM_CHECK_HIP(hipMemcpyAsync(m->init, hinit.data(), sizeof(float) * size_t(n), hipMemcpyHostToDevice, g->stream));
  const int igrid = int(std::min<int64_t>((V + 255) / 256, 4096));
  SPM_TRY(launch(g, "pr_job_ids", pr_job_ids, dim3(igrid), dim3(256), 0, V, (const int32_t*)m->bases, n, m->job));
  SPM_CHECK_HIP(hipMemsetAsync(m->flag, 0, sizeof(unsigned int), g->stream));
  SPM_TRY(launch(g, "pr_job_edges", pr_job_edges, dim3(g->num_cus * 8), dim3(kPrBlock), 0, V, (const int32_t*)g->off,
                 (const int32_t*)g->adj, (const int32_t*)m->job, m->flag));
  unsigned int cross = 0;
  SPM_CHECK_HIP(hipMemcpyAsync(&cross, m->flag, sizeof cross, hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  if (cross) return set_error(SPM_EINVAL, "spm_pr_multi_device: an edge joins two jobs");
  SPM_TRY(launch(g, "pr_init_multi", pr_init_multi, dim3(igrid), dim3(256), 0, V, (const int32_t*)g->off,
                 (const int32_t*)m->job, (const float*)m->init, w->degf, w->pr[0], w->c[0], w->st, int(maxIter),
                 beta, delta));
  int R = std::max(2, std::min(128, (int(std::max(maxIter, 1)) + 1) & ~1));
  const bool use_graph = !g->prof.on;
  if (use_graph && (!w->exec || w->R != R || m->n_graph != n)) {
    if (w->exec) { (void)hipGraphExecDestroy(w->exec); w->exec = nullptr; }
    if (w->graph) { (void)hipGraphDestroy(w->graph); w->graph = nullptr; }
    SPM_CHECK_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = pr_multi_enqueue(g, R, n);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g->stream, &graph);
    if (rc != SPM_OK) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return set_error(SPM_EHIP, std::string("pr capture: ") + hipGetErrorString(e));
    w->graph = graph;
    SPM_CHECK_HIP(hipGraphInstantiate(&w->exec, graph, nullptr, nullptr, 0));
    w->R = R;
    m->n_graph = n;
  }
  int32_t* h = g->h_pinned;
  for (int64_t rep = 0;; ++rep) {
    if (use_graph) SPM_CHECK_HIP(hipGraphLaunch(w->exec, g->stream));
    else SPM_TRY(pr_multi_enqueue(g, R, n));
    SPM_CHECK_HIP(hipMemcpyAsync(h, &w->st->done, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
    SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    if (h[0]) break;
    if (rep > int64_t(maxIter) / R + 2) return set_error(SPM_EHIP, "pr: loop did not terminate");
  }
  SPM_TRY(launch(g, "pr_job_select", pr_job_select, dim3(igrid), dim3(256), 0, V, (const int32_t*)m->job,
                 (const int32_t*)m->iters, (const float*)w->pr[0], (const float*)w->pr[1], d_out));
  if (iters_out)
    SPM_CHECK_HIP(hipMemcpyAsync(iters_out, m->iters, sizeof(int32_t) * size_t(n), hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  SPM_TRY(prof_flush(g));
  return SPM_OK;
}

// ---- no-allocation entry point (spm_pr_device_noalloc) ------------------------
// Callers that cannot allocate device memory (code under graph capture, tight
// serving loops) run PageRank out of one file-scope device array instead of
// the handle's workspace. It holds the ranks and contributions (two buffers
// each), the out-degrees, the binned rows, the per-workgroup diffs, the bin
// counts and the loop state for graphs of up to kNoAllocMaxV vertices: about
// 40 MiB, which every GPU the library runs on holds until the process ends.
constexpr int32_t kNoAllocMaxV = 1 << 20;
constexpr size_t kNaAlign = 256;
constexpr size_t na_round(size_t b) { return (b + kNaAlign - 1) / kNaAlign * kNaAlign; }
constexpr size_t kNaFloats = na_round(sizeof(float) * size_t(kNoAllocMaxV));
constexpr size_t kNaRows = na_round(sizeof(int4) * size_t(kNoAllocMaxV));
// partial diffs: one per workgroup of a sweep (at most one per vertex plus the
// three partial bins) or per push workgroup, plus one
constexpr size_t kNaPartial = na_round(sizeof(float) * (size_t(kNoAllocMaxV) + 8192));
constexpr size_t kNaCnt =
    na_round(sizeof(int32_t) * (size_t(kCat) * ((size_t(kNoAllocMaxV) + kBinChunk - 1) / kBinChunk) + kCat));
constexpr size_t kNaState = na_round(sizeof(PrState));
constexpr size_t kNoAllocScratchBytes = 5 * kNaFloats + kNaRows + kNaPartial + kNaCnt + kNaState;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// The device address of `scratch` on each GPU (its own copy per device). The
// name `scratch` in host code is only a host-side stand-in, so kernels get
// these addresses, looked up once per device with that device current.
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static int noalloc_scratch(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return set_error(SPM_EINVAL, "spm_pr_device_noalloc: device id out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    SPM_CHECK_HIP(hipGetDevice(&cur));
    if (cur != dev) SPM_CHECK_HIP(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return set_error(SPM_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return SPM_OK;
}

int pr_run_noalloc(spm_graph* g, float beta, float delta, int32_t maxIter, float* d_out, int32_t* iter_out) {
  if (!g->roff) return set_error(SPM_ENOGRAPH, "PageRank needs the reverse CSR (rev_indexofNodes, srcList)");
  if (dyn_pending(g))
    return set_error(SPM_EINVAL, "spm_pr_device_noalloc: pending edge updates (fold them with spm_graph_compact first)");
  if (g->V > kNoAllocMaxV) return set_error(SPM_EINVAL, "spm_pr_device_noalloc: at most 2^20 vertices");
  unsigned char* base = nullptr;
  SPM_TRY(noalloc_scratch(g->device, &base));
  const int32_t V = g->V;
  PrWork w;
  unsigned char* p = base;
  for (int i = 0; i < 2; ++i) {
    w.pr[i] = reinterpret_cast<float*>(p), p += kNaFloats;
    w.c[i] = reinterpret_cast<float*>(p), p += kNaFloats;
  }
  w.degf = reinterpret_cast<float*>(p), p += kNaFloats;
  w.rows = reinterpret_cast<int4*>(p), p += kNaRows;
  w.partial = reinterpret_cast<float*>(p), p += kNaPartial;
  int32_t* cnt = reinterpret_cast<int32_t*>(p);
  p += kNaCnt;
  w.st = reinterpret_cast<PrState*>(p);
  w.acc = nullptr;  // pull sweeps only (push and source segments need per-handle memory)
  w.S = 1;
  w.view.off = g->off;
  w.view.adj = g->adj;
  w.view.w = g->w;
  w.view.roff = g->roff;
  w.view.radj = g->radj;
  w.view.perm = nullptr;
  SPM_TRY(pr_bin_device(g, g->roff, w.rows, cnt, g->h_pinned, &w.bins));
  w.partial_n = std::max({w.bins.blk[4], g->num_cus * 8, 2 * kLagMaxGrid}) + 1;
  if (size_t(w.partial_n) * sizeof(float) > kNaPartial) return set_error(SPM_EINVAL, "spm_pr_device_noalloc: too many workgroups");
  const float num_nodes = float(V);
  const float init = 1 / num_nodes;
  const float base_term = (1 - delta) / num_nodes;
  const int igrid = int(std::min<int64_t>((V + 255) / 256, 4096));
  SPM_TRY(launch(g, "pr_init", pr_init, dim3(igrid), dim3(256), 0, V, w.view.off, w.degf, w.pr[0], w.c[0],
                 (float*)nullptr, w.st, init, int(maxIter), beta, delta, base_term));
  int32_t* h = g->h_pinned;
  // Batches of R iterations (even, so each starts on buffer 0), launched
  // directly -- graph instantiation would allocate -- with the done flag read
  // back after each batch; iterations after convergence return at once.
  const int R = std::max(2, std::min(128, (int(std::max(maxIter, 1)) + 1) & ~1));
  for (int64_t rep = 0;; ++rep) {
    SPM_TRY(pr_enqueue_iters(g, &w, R, false));
    SPM_CHECK_HIP(hipMemcpyAsync(h, &w.st->done, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
    SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    if (h[0]) break;
    if (rep > int64_t(maxIter) / R + 2) return set_error(SPM_EHIP, "pr: loop did not terminate");
  }
  const int32_t iters = h[1];
  SPM_CHECK_HIP(hipMemcpyAsync(d_out, w.pr[iters & 1], sizeof(float) * size_t(V), hipMemcpyDeviceToDevice, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  if (iter_out) *iter_out = iters;
  SPM_TRY(prof_flush(g));
  return SPM_OK;
}

}  // namespace spm

using namespace spm;

extern "C" {

int spm_pr_device(spm_graph* g, float beta, float delta, int32_t maxIter, float* d_pageRank,
                  int32_t* iterCount) {
  if (!g || !d_pageRank) return set_error(SPM_EINVAL, "spm_pr_device: NULL argument");
  return pr_run(g, beta, delta, maxIter, d_pageRank, nullptr, iterCount);
}

int spm_pr_multi_device(spm_graph* g, const int32_t* bases, int32_t n, float beta, float delta, int32_t maxIter,
                        float* d_pageRank, int32_t* iterCounts) {
  if (!g) return set_error(SPM_EINVAL, "spm_pr_multi_device: NULL handle");
  return pr_multi_run(g, bases, n, beta, delta, maxIter, d_pageRank, iterCounts);
}

int spm_pr_device_noalloc(spm_graph* g, float beta, float delta, int32_t maxIter, float* d_pageRank,
                          int32_t* iterCount) {
  if (!g || !d_pageRank) return set_error(SPM_EINVAL, "spm_pr_device_noalloc: NULL argument");
  return pr_run_noalloc(g, beta, delta, maxIter, d_pageRank, iterCount);
}

int spm_compute_pr(spm_graph* g, float beta, float delta, int32_t maxIter, float* pageRank,
                   int32_t* iterCount) {
  if (!g || !pageRank) return set_error(SPM_EINVAL, "spm_compute_pr: NULL argument");
  return pr_run(g, beta, delta, maxIter, nullptr, pageRank, iterCount);
}

}  // extern "C"
