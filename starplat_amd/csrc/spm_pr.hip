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
// Arithmetic follows the emitted code: the per-term division pr[u]/outdeg(u)
// is the same IEEE single division, and the update is evaluated without FMA
// contraction. Only the order of the in-row summation differs (tree vs
// sequential), which keeps ranks within a few ulps of the restatement.
#include <algorithm>
#include <cstring>
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
    acc[v] = 0.f;
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

__device__ __forceinline__ void pr_store(int32_t v, float sum, const PrState& s,
                                         const float* __restrict__ pr_in, float* __restrict__ pr_out,
                                         float* __restrict__ c_out, const float* __restrict__ degf,
                                         float& mydiff) {
  const float nv = __fadd_rn(s.base, __fmul_rn(s.delta, sum));
  const float old = pr_in[v];
  pr_out[v] = nv;
  const float d = degf[v];
  c_out[v] = d > 0.f ? nv / d : 0.f;
  mydiff += nv - old >= 0.f ? nv - old : old - nv;
}

// End of an iteration inside a workgroup: the sum of |new - old| over its
// vertices goes to partial[blockIdx.x] with a plain store. No workgroup
// touches a shared counter: with ~10^5 workgroups per iteration at RMAT-22,
// same-address device-scope atomics would serialise the whole iteration.
__device__ __forceinline__ void pr_finish(float mydiff, float* __restrict__ partial, float* s_red) {
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  mydiff = wave_sum(mydiff);
  if (lane == 0) s_red[wid] = mydiff;
  __syncthreads();
  if (tid == 0) {
    float t = 0.f;
    for (int w = 0; w < kPrBlock / kWave; ++w) t += s_red[w];
    partial[blockIdx.x] = t;
  }
}

// The loop test (diff > beta && iter < maxIter), one workgroup: sums the
// per-workgroup partials in a fixed order, advances the iteration count and
// raises the done flag that makes later iterations return at once.
constexpr int kCheckBlock = 1024;
__global__ __launch_bounds__(kCheckBlock) void pr_check(const float* __restrict__ partial, int n,
                                                        PrState* __restrict__ st) {
  __shared__ float s_red[kCheckBlock / kWave];
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
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
  if (tid == 0) {
    float d = 0.f;
    for (int w = 0; w < kCheckBlock / kWave; ++w) d += s_red[w];
    const int it = st->iter + 1;
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

__device__ __forceinline__ void seg_emit(int32_t v, float sum, const PrSeg& sg, const PrState& s,
                                         const float* __restrict__ pr_in, float* __restrict__ pr_out,
                                         float* __restrict__ c_out, const float* __restrict__ degf,
                                         float& mydiff) {
  if (sg.S > 1) {
    if (sg.seg < sg.S - 1) {
      sg.acc[v] = sg.seg == 0 ? sum : sg.acc[v] + sum;
      return;
    }
    sum = sg.acc[v] + sum;
  }
  pr_store(v, sum, s, pr_in, pr_out, c_out, degf, mydiff);
}

// Sum of c[radj[e]] for e = ea+start, ea+start+stride, ... < eb, four rows
// elements at a time: four independent index loads, then four independent
// gathers, so a lane keeps several memory requests in flight instead of a
// chain of dependent ones.
__device__ __forceinline__ float gather_sum(const int32_t* __restrict__ radj, const float* __restrict__ c,
                                            int32_t ea, int32_t eb, int32_t start, int32_t stride) {
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  int32_t e = ea + start;
  for (; e + 3 * stride < eb; e += 4 * stride) {
    const int32_t a0 = radj[e], a1 = radj[e + stride], a2 = radj[e + 2 * stride], a3 = radj[e + 3 * stride];
    s0 += c[a0];
    s1 += c[a1];
    s2 += c[a2];
    s3 += c[a3];
  }
  for (; e < eb; e += stride) s0 += c[radj[e]];
  return (s0 + s1) + (s2 + s3);
}

// rows[i] = {v, first in-edge, end of in-edges, 0} for the i-th vertex in bin
// order: one 16-byte load replaces the order -> offsets chain.
__device__ __forceinline__ void row_of(const int4* __restrict__ rows, int64_t i, const int32_t* __restrict__ roff,
                                       const PrSeg& sg, int32_t& v, int32_t& ea, int32_t& eb) {
  const int4 r = rows[i];
  v = r.x;
  ea = r.y;
  eb = r.z;
  if (sg.S > 1) seg_range(v, roff, sg, ea, eb);
}

template <int G>
__global__ __launch_bounds__(kPrBlock) void pr_iter(
    const int32_t* __restrict__ roff, const int32_t* __restrict__ radj,
    const int4* __restrict__ rows, const float* __restrict__ degf, PrBins bins,
    const float* __restrict__ pr_in, const float* __restrict__ c_in, float* __restrict__ pr_out,
    float* __restrict__ c_out, const PrState* __restrict__ st, float* __restrict__ partial, PrSeg sg) {
  __shared__ float s_red[kPrBlock / kWave];
  if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const PrState s = *st;
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid / kWave;
  const int b = blockIdx.x;
  float mydiff = 0.f;
  int32_t v, ea, eb;

  if (b < bins.blk[1]) {
    // bin 0: one lane per vertex
    const int64_t i = int64_t(b) * kPrBlock + tid;
    if (i < bins.nb[0]) {
      row_of(rows, i, roff, sg, v, ea, eb);
      seg_emit(v, gather_sum(radj, c_in, ea, eb, 0, 1), sg, s, pr_in, pr_out, c_out, degf, mydiff);
    }
  } else if (b < bins.blk[2]) {
    // bin 1: G lanes per vertex
    const int64_t i = int64_t(b - bins.blk[1]) * (kPrBlock / G) + tid / G;
    const int g = tid % G;
    float sum = 0.f;
    v = -1;
    if (i < bins.nb[1]) {
      row_of(rows, bins.nb[0] + i, roff, sg, v, ea, eb);
      sum = gather_sum(radj, c_in, ea, eb, g, G);
    }
    sum = group_sum<G>(sum);
    if (v >= 0 && g == 0) seg_emit(v, sum, sg, s, pr_in, pr_out, c_out, degf, mydiff);
  } else if (b < bins.blk[3]) {
    // bin 2: one wavefront per vertex
    const int64_t i = int64_t(b - bins.blk[2]) * (kPrBlock / kWave) + wid;
    float sum = 0.f;
    v = -1;
    if (i < bins.nb[2]) {
      row_of(rows, bins.nb[0] + bins.nb[1] + i, roff, sg, v, ea, eb);
      sum = gather_sum(radj, c_in, ea, eb, lane, kWave);
    }
    sum = wave_sum(sum);
    if (v >= 0 && lane == 0) seg_emit(v, sum, sg, s, pr_in, pr_out, c_out, degf, mydiff);
  } else {
    // bin 3: one workgroup per vertex
    row_of(rows, bins.nb[0] + bins.nb[1] + bins.nb[2] + (b - bins.blk[3]), roff, sg, v, ea, eb);
    float sum = gather_sum(radj, c_in, ea, eb, tid, kPrBlock);
    sum = wave_sum(sum);
    if (lane == 0) s_red[wid] = sum;
    __syncthreads();
    if (tid == 0) {
      float t = 0.f;
      for (int w = 0; w < kPrBlock / kWave; ++w) t += s_red[w];
      seg_emit(v, t, sg, s, pr_in, pr_out, c_out, degf, mydiff);
    }
    __syncthreads();
  }
  if (sg.seg == sg.S - 1) pr_finish(mydiff, partial, s_red);
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
    pr_store(int32_t(v), sum, s, pr_in, pr_out, c_out, degf, mydiff);
  }
  pr_finish(mydiff, partial, s_red);
}

static int pr_prepare(spm_graph* g, const CsrView& view) {
  if (g->pr) return SPM_OK;
  if (!g->roff) return set_error(SPM_ENOGRAPH, "PageRank needs the reverse CSR (rev_indexofNodes, srcList)");
  PrWork* w = new PrWork();
  g->pr = w;
  w->view = view;
  if (view.perm) SPM_TRY(dmalloc(g, (void**)&w->out, sizeof(float) * size_t(g->V), "pr result"));
  const int32_t V = g->V;
  for (int i = 0; i < 2; ++i) {
    SPM_TRY(dmalloc(g, (void**)&w->pr[i], sizeof(float) * V, "pr ranks"));
    SPM_TRY(dmalloc(g, (void**)&w->c[i], sizeof(float) * V, "pr contributions"));
  }
  SPM_TRY(dmalloc(g, (void**)&w->degf, sizeof(float) * V, "pr outdeg"));
  SPM_TRY(dmalloc(g, (void**)&w->acc, sizeof(float) * V, "pr push accumulators"));
  SPM_TRY(dmalloc(g, (void**)&w->rows, sizeof(int4) * V, "pr bins"));
  SPM_TRY(dmalloc(g, (void**)&w->st, sizeof(PrState), "pr state"));
  // Bin vertices by in-degree on the host (once per handle and schedule).
  std::vector<int32_t> roff(size_t(V) + 1);
  SPM_CHECK_HIP(hipMemcpyAsync(roff.data(), view.roff, sizeof(int32_t) * (V + 1), hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  const spm_schedule& s = g->sched;
  std::vector<int32_t> order(V);
  int32_t cnt[4] = {0, 0, 0, 0};
  auto bin_of = [&](int32_t v) {
    const int32_t d = roff[v + 1] - roff[v];
    return d <= s.pr_t1 ? 0 : d <= s.pr_t2 ? 1 : d <= s.pr_t3 ? 2 : 3;
  };
  for (int32_t v = 0; v < V; ++v) ++cnt[bin_of(v)];
  int32_t start[4] = {0, cnt[0], cnt[0] + cnt[1], cnt[0] + cnt[1] + cnt[2]};
  for (int32_t v = 0; v < V; ++v) order[start[bin_of(v)]++] = v;
  std::vector<int4> rows(V);
  for (int32_t i = 0; i < V; ++i) {
    const int32_t v = order[i];
    rows[i] = make_int4(v, roff[v], roff[v + 1], 0);
  }
  SPM_CHECK_HIP(hipMemcpyAsync(w->rows, rows.data(), sizeof(int4) * V, hipMemcpyHostToDevice, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  PrBins& b = w->bins;
  for (int i = 0; i < 4; ++i) b.nb[i] = cnt[i];
  b.group = s.pr_group;
  b.blk[0] = 0;
  b.blk[1] = (cnt[0] + kPrBlock - 1) / kPrBlock;
  b.blk[2] = b.blk[1] + (cnt[1] + kPrBlock / s.pr_group - 1) / (kPrBlock / s.pr_group);
  b.blk[3] = b.blk[2] + (cnt[2] + kPrBlock / kWave - 1) / (kPrBlock / kWave);
  b.blk[4] = b.blk[3] + cnt[3];
  w->partial_n = std::max(b.blk[4], g->num_cus * 8) + 1;
  SPM_TRY(dmalloc(g, (void**)&w->partial, sizeof(float) * size_t(w->partial_n), "pr partial diffs"));
  // Source segments for cache blocking.
  w->S = 1;
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

int pr_release(spm_graph* g) {
  PrWork* w = g->pr;
  if (!w) return SPM_OK;
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
  delete w;
  g->pr = nullptr;
  return SPM_OK;
}

static int pr_enqueue_iters(spm_graph* g, int R) {
  PrWork* w = g->pr;
  const dim3 grid(std::max(w->bins.blk[4], 1)), block(kPrBlock);
  const int pgrid = int(std::min<int64_t>((g->V + kPrBlock - 1) / kPrBlock, int64_t(g->num_cus) * 8));
  for (int r = 0; r < R; ++r) {
    const int p = r & 1;
    const float* pin = w->pr[p];
    const float* cin = w->c[p];
    if (g->sched.pr_mode == 1) {
      SPM_TRY(launch(g, "pr_push_scatter", pr_push_scatter, dim3(pgrid), block, 0, g->V, w->view.off, w->view.adj,
                     cin, w->acc, (const PrState*)w->st));
      SPM_TRY(launch(g, "pr_push_apply", pr_push_apply, dim3(pgrid), block, 0, g->V, (const float*)w->degf, w->acc,
                     pin, w->pr[1 - p], w->c[1 - p], (const PrState*)w->st, w->partial));
      SPM_TRY(launch(g, "pr_check", pr_check, dim3(1), dim3(kCheckBlock), 0, (const float*)w->partial, pgrid, w->st));
      continue;
    }
    for (int seg = 0; seg < w->S; ++seg) {
      const PrSeg sg{w->sp, w->acc, w->S, seg};
#define SPM_PR_LAUNCH(G)                                                                          \
  case G:                                                                                         \
    SPM_TRY(launch(g, "pr_iter", pr_iter<G>, grid, block, 0, w->view.roff,                        \
                   w->view.radj, (const int4*)w->rows, (const float*)w->degf,                     \
                   w->bins, pin, cin, w->pr[1 - p], w->c[1 - p], (const PrState*)w->st,             \
                   w->partial, sg));                                                               \
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
                   int(grid.x), w->st));
  }
  return SPM_OK;
}

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
  // Iterations per replay: even (ping-pong parity), up to 128.
  int R = std::max(2, std::min(128, (int(std::max(maxIter, 1)) + 1) & ~1));
  const bool use_graph = !g->prof.on;
  if (use_graph && (!w->exec || w->R != R)) {
    if (w->exec) { (void)hipGraphExecDestroy(w->exec); w->exec = nullptr; }
    if (w->graph) { (void)hipGraphDestroy(w->graph); w->graph = nullptr; }
    SPM_CHECK_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = pr_enqueue_iters(g, R);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g->stream, &graph);
    if (rc != SPM_OK) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return set_error(SPM_EHIP, std::string("pr capture: ") + hipGetErrorString(e));
    w->graph = graph;
    SPM_CHECK_HIP(hipGraphInstantiate(&w->exec, graph, nullptr, nullptr, 0));
    w->R = R;
  }
  int32_t* h = g->h_pinned;
  for (int64_t rep = 0;; ++rep) {
    if (use_graph) SPM_CHECK_HIP(hipGraphLaunch(w->exec, g->stream));
    else SPM_TRY(pr_enqueue_iters(g, R));
    SPM_CHECK_HIP(hipMemcpyAsync(h, &w->st->done, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, g->stream));
    SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    if (h[0]) break;
    if (rep > int64_t(maxIter) / R + 2) return set_error(SPM_EHIP, "pr: loop did not terminate");
  }
  const int32_t iters = h[1];
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

}  // namespace spm

using namespace spm;

extern "C" {

int spm_pr_device(spm_graph* g, float beta, float delta, int32_t maxIter, float* d_pageRank,
                  int32_t* iterCount) {
  if (!g || !d_pageRank) return set_error(SPM_EINVAL, "spm_pr_device: NULL argument");
  return pr_run(g, beta, delta, maxIter, d_pageRank, nullptr, iterCount);
}

int spm_compute_pr(spm_graph* g, float beta, float delta, int32_t maxIter, float* pageRank,
                   int32_t* iterCount) {
  if (!g || !pageRank) return set_error(SPM_EINVAL, "spm_compute_pr: NULL argument");
  return pr_run(g, beta, delta, maxIter, nullptr, pageRank, iterCount);
}

}  // extern "C"
