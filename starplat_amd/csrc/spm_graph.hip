// spm_graph.hip -- device CSR mirror of graphcode/graph.hpp (:44-57) and the
// handle plumbing (errors, allocation accounting, profiling, schedules).
#include <algorithm>
#include <cstdio>
#include <cstring>

#include "spm_internal.hpp"

namespace spm {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int dmalloc(spm_graph* g, void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (bytes == 0) bytes = 4;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    char buf[256];
    std::snprintf(buf, sizeof buf, "hipMalloc(%s, %zu bytes): %s", what, bytes, hipGetErrorString(e));
    return set_error(SPM_ENOMEM, buf);
  }
  g->bytes += (int64_t)bytes;
  return SPM_OK;
}

void dfree(spm_graph* g, void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  g->bytes -= (int64_t)(bytes ? bytes : 4);
}

static hipEvent_t pool_get(spm_graph* g) {
  if (!g->prof.pool.empty()) {
    hipEvent_t e = g->prof.pool.back();
    g->prof.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void prof_begin(spm_graph* g, const char*, hipEvent_t* a, hipEvent_t* b) {
  *a = pool_get(g);
  *b = pool_get(g);
}

void prof_end(spm_graph* g, const char* name, hipEvent_t a, hipEvent_t b) {
  g->prof.pending.push_back({name, a, b});
}

int prof_flush(spm_graph* g) {
  if (g->prof.pending.empty()) return SPM_OK;
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  for (auto& p : g->prof.pending) {
    float ms = 0.f;
    SPM_CHECK_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto& t = g->prof.totals[p.name];
    t.first += ms;
    t.second += 1;
    g->prof.pool.push_back(p.a);
    g->prof.pool.push_back(p.b);
  }
  g->prof.pending.clear();
  return SPM_OK;
}

}  // namespace spm

using namespace spm;

namespace {

// The fixed heuristic schedule. Degree thresholds follow the wavefront width:
// rows shorter than a few elements are summed by one lane, rows up to a
// wavefront's worth by a lane group, rows up to a few thousand by a
// wavefront, and hubs by a whole workgroup.
void heuristic_schedule(const spm_graph* g, spm_schedule* s) {
  std::memset(s, 0, sizeof *s);
  s->pr_t1 = 4;
  s->pr_t2 = 48;
  s->pr_t3 = 2048;
  s->pr_group = 8;
  // From 2^17 vertices the pull sweep waits on its scattered contribution
  // reads, and spreading a row over more lanes keeps more of them in flight:
  // rows above 256 in-edges get a workgroup, 128..256 a wavefront, 5..128
  // four lanes. Measured (100 iterations, ms, old -> new): RMAT-18 2.90 ->
  // 2.38, RMAT-20 10.15 -> 9.06, ER 2^20 7.47 -> 6.46, RMAT-22 41.9 -> 36.0;
  // a 1024^2 grid 1.19 either way; RMAT-16 0.94 -> 1.07, so smaller graphs
  // keep the thresholds above.
  if (g->V >= (1 << 17)) {
    s->pr_t2 = 128;
    s->pr_t3 = 256;
    s->pr_group = 4;
  }
  s->sssp_mode = 0;
  s->sssp_rounds = 12;
  s->sssp_band_rounds = 8;
  s->sssp_grid = 0;
  s->tc_mode = 1;
  // Small graphs fold PageRank's loop test into the next sweep: below 2^17
  // vertices a sweep takes a few microseconds, about what the separate
  // loop-test launch costs.
  s->pr_mode = g->V < (1 << 17) ? 2 : 0;
  // One sweep: measured at RMAT-22, 2 MB source segments double the
  // iteration time (each pass repeats the per-vertex load chain).
  s->pr_seg_log2 = 0;
  // Degree-ordered relabeling pays off once the degrees are skewed enough
  // that a small set of vertices takes most accesses and the graph has from
  // 2^17 vertices: PageRank, 100 iterations, ms, original -> relabeled ids:
  // RMAT-17 1.54 -> 1.35, RMAT-18 2.40 -> 1.90, RMAT-19 4.11 -> 3.51,
  // RMAT-20 8.22 -> 6.62 (edge factor 16); SSSP equal or faster at every
  // one of these sizes (RMAT-20 0.60 -> 0.57).
  const double avg = g->V > 0 ? double(g->E) / double(g->V) : 0.0;
  s->relabel = (g->V >= (1 << 17) && double(g->max_outdeg) >= 64.0 * avg) ? 1 : 0;
  // On the relabeled copy the lowest ids are the sources most in-rows
  // gather from: from 2^19 vertices the staged sweep reads their
  // contributions from LDS, with rows up to 4096 in-edges on a wavefront
  // (100 iterations, ms: RMAT-22 33.5 -> 28.4; RMAT-20 6.62 -> 5.36;
  // RMAT-19 3.51 -> 3.46, at edge factor 32 5.68 -> 4.25; but RMAT-18
  // 1.90 -> 2.32, the staging then costing more than it saves).
  if (s->relabel && g->V >= (1 << 19)) {
    s->pr_mode = 3;
    s->pr_t3 = 4096;
  }
}

int validate_schedule(const spm_schedule* s) {
  if (s->pr_t1 < 0 || s->pr_t2 < s->pr_t1 || s->pr_t3 < s->pr_t2) return set_error(SPM_EINVAL, "schedule: PR thresholds must be non-decreasing");
  if (s->pr_group != 2 && s->pr_group != 4 && s->pr_group != 8 && s->pr_group != 16 && s->pr_group != 32)
    return set_error(SPM_EINVAL, "schedule: pr_group must be 2, 4, 8, 16 or 32");
  if (s->sssp_mode < 0 || s->sssp_mode > 3) return set_error(SPM_EINVAL, "schedule: sssp_mode must be 0, 1, 2 or 3");
  if (s->sssp_rounds < 6 || s->sssp_rounds > 96 || s->sssp_rounds % 6) return set_error(SPM_EINVAL, "schedule: sssp_rounds must be a multiple of 6 in [6, 96]");
  if (s->sssp_grid < 0 || s->sssp_grid > 65536) return set_error(SPM_EINVAL, "schedule: sssp_grid out of range");
  if (s->sssp_band_rounds < 1 || s->sssp_band_rounds > 4096)
    return set_error(SPM_EINVAL, "schedule: sssp_band_rounds must be in [1, 4096]");
  if (s->tc_mode < 0 || s->tc_mode > 1) return set_error(SPM_EINVAL, "schedule: tc_mode must be 0 (search) or 1 (hash)");
  if (s->pr_mode < 0 || s->pr_mode > 3)
    return set_error(SPM_EINVAL,
                     "schedule: pr_mode must be 0 (pull), 1 (push), 2 (pull, loop test in the next sweep) or 3 "
                     "(pull, hottest contributions staged in LDS)");
  if ((s->pr_mode == 2 || s->pr_mode == 3) && s->pr_seg_log2 != 0)
    return set_error(SPM_EINVAL, "schedule: pr_mode 2 and 3 sweep all sources at once (pr_seg_log2 must be 0)");
  if (s->pr_seg_log2 != 0 && (s->pr_seg_log2 < 10 || s->pr_seg_log2 > 30))
    return set_error(SPM_EINVAL, "schedule: pr_seg_log2 must be 0 or in [10, 30]");
  if (s->relabel < 0 || s->relabel > 1) return set_error(SPM_EINVAL, "schedule: relabel must be 0 or 1");
  return SPM_OK;
}

}  // namespace

extern "C" {

const char* spm_last_error(void) { return spm::g_last_error.c_str(); }
int spm_version(void) { return 100; }

// The CSR contract every kernel relies on: offsets non-decreasing, each row
// ascending (graph.hpp:722-734 and :810-832 sort them; edge updates and the
// triangle count binary-search rows) and its neighbour ids in [0, V). With a
// row ascending, its first and last elements bound the rest, so one pass of
// neighbour comparisons checks both. Returns 0, or 1 for offsets, 2 for an
// id out of range, 3 for a row out of order.
static int csr_check(int32_t V, const int32_t* off, const int32_t* adj) {
  int32_t bad = 0;
  for (int32_t v = 0; v < V; ++v) bad |= int32_t(off[v + 1] < off[v]);
  if (bad) return 1;
  int32_t lo = 0, hi = 0, down = 0;
  for (int32_t v = 0; v < V; ++v) {
    const int32_t a = off[v], b = off[v + 1];
    if (a == b) continue;
    lo = std::min(lo, adj[a]);
    hi = std::max(hi, adj[b - 1]);
    for (int32_t e = a + 1; e < b; ++e) down |= int32_t(adj[e] < adj[e - 1]);
  }
  if (down) return 3;
  return (lo < 0 || hi >= V) ? 2 : 0;
}

}  // extern "C"

namespace spm {

void schedule_after_fold(spm_graph* g) {
  if (g->sched_default) heuristic_schedule(g, &g->sched);
}

__device__ __forceinline__ unsigned long long pair_hash(uint32_t u, uint32_t v) {
  // splitmix64 finaliser of the packed pair (Steele, Lea, Flood 2014)
  unsigned long long z = (static_cast<unsigned long long>(u) << 32 | v) + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// sums[dir] += hash(u, v) over every edge u -> v of one direction: the
// forward CSR (rows are sources) or the reverse CSR (rows are targets).
// Equal sums mean equal edge multisets (up to a 2^-64 collision chance), i.e.
// the reverse CSR is the forward one's transpose. One wavefront per row, the
// rows visited in the spread_vertex order (RMAT's heavy rows would otherwise
// pile onto a few wavefronts), one 64-bit atomic per workgroup. Rows are
// already checked (offsets and ids), so every access stays in bounds.
constexpr int kCheckWg = 256;
__global__ __launch_bounds__(kCheckWg) void csr_pair_hash(int32_t V, const int32_t* __restrict__ off,
                                                          const int32_t* __restrict__ adj, int reverse,
                                                          unsigned long long* __restrict__ sum) {
  __shared__ unsigned long long s_part[kCheckWg / kWave];
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  const int64_t nw = int64_t(gridDim.x) * (kCheckWg / kWave);
  unsigned long long acc = 0;
  for (int64_t x = int64_t(blockIdx.x) * (kCheckWg / kWave) + wid; x < V; x += nw) {
    const int64_t r = spread_vertex(x, V);
    for (int32_t e = off[r] + lane; e < off[r + 1]; e += kWave) {
      const uint32_t a = uint32_t(r), b = uint32_t(adj[e]);
      acc += reverse ? pair_hash(b, a) : pair_hash(a, b);
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) s_part[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kCheckWg / kWave; ++w) t += s_part[w];
    atomicAdd(sum, t);
  }
}

// The reverse CSR given at creation is the transpose of the forward one.
static int check_transpose(spm_graph* g) {
  unsigned long long* d = nullptr;
  SPM_TRY(dmalloc(g, (void**)&d, 2 * sizeof(unsigned long long), "transpose check"));
  int rc = SPM_OK;
  unsigned long long h[2] = {0, 0};
  hipError_t e = hipMemsetAsync(d, 0, sizeof h, g->stream);
  const int grid = int(std::min<int64_t>((int64_t(g->V) + 3) / 4, int64_t(g->num_cus) * 8));
  if (e == hipSuccess) rc = launch(g, "csr_pair_hash", csr_pair_hash, dim3(grid), dim3(kCheckWg), 0, g->V,
                                   (const int32_t*)g->off, (const int32_t*)g->adj, 0, d);
  if (e == hipSuccess && rc == SPM_OK)
    rc = launch(g, "csr_pair_hash", csr_pair_hash, dim3(grid), dim3(kCheckWg), 0, g->V, (const int32_t*)g->roff,
                (const int32_t*)g->radj, 1, d + 1);
  if (e == hipSuccess && rc == SPM_OK) e = hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, g->stream);
  if (e == hipSuccess && rc == SPM_OK) e = hipStreamSynchronize(g->stream);
  dfree(g, d, 2 * sizeof(unsigned long long));
  if (e != hipSuccess) return set_error(SPM_EHIP, std::string("transpose check: ") + hipGetErrorString(e));
  SPM_TRY(rc);
  if (h[0] != h[1])
    return set_error(SPM_EINVAL, "spm_graph_create: rev_indexofNodes/srcList is not the transpose of indexofNodes/edgeList");
  return SPM_OK;
}

}  // namespace spm

extern "C" {

int spm_graph_create(int32_t V, int64_t E, const int32_t* off, const int32_t* adj,
                     const int32_t* w, const int32_t* roff, const int32_t* radj,
                     spm_graph** out) {
  if (!out) return set_error(SPM_EINVAL, "spm_graph_create: out is NULL");
  *out = nullptr;
  if (V < 1 || E < 0 || E > 0x7fffffffLL) return set_error(SPM_EINVAL, "spm_graph_create: need V >= 1 and 0 <= E < 2^31");
  if (!off || (E > 0 && !adj)) return set_error(SPM_EINVAL, "spm_graph_create: indexofNodes/edgeList missing");
  if ((roff == nullptr) != (radj == nullptr)) return set_error(SPM_EINVAL, "spm_graph_create: rev_indexofNodes and srcList go together");
  if (off[0] != 0 || off[V] != E) return set_error(SPM_EINVAL, "spm_graph_create: indexofNodes[0] must be 0 and indexofNodes[V] == E");
  if (roff && (roff[0] != 0 || roff[V] != E)) return set_error(SPM_EINVAL, "spm_graph_create: rev_indexofNodes[0] must be 0 and rev_indexofNodes[V] == E");
  // Every kernel indexes rows and vertex arrays by these values without
  // bounds checks, so a malformed CSR is refused here rather than faulting
  // the device later.
  static const char* const what[] = {"", "offsets must be non-decreasing", "ids must lie in [0, V)",
                                      "rows must be ascending"};
  if (int c = csr_check(V, off, adj))
    return set_error(SPM_EINVAL, std::string("spm_graph_create: indexofNodes/edgeList: ") + what[c]);
  if (int c = roff ? csr_check(V, roff, radj) : 0)
    return set_error(SPM_EINVAL, std::string("spm_graph_create: rev_indexofNodes/srcList: ") + what[c]);
  int dev = 0;
  SPM_CHECK_HIP(hipGetDevice(&dev));
  spm_graph* g = new spm_graph();
  g->device = dev;
  g->V = V;
  g->E = E;
  for (int32_t v = 0; v < V; ++v) g->max_outdeg = std::max(g->max_outdeg, off[v + 1] - off[v]);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) g->num_cus = prop.multiProcessorCount;
  hipError_t he = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (he != hipSuccess) { delete g; return set_error(SPM_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(he)); }
  he = hipHostMalloc((void**)&g->h_pinned, 64, hipHostMallocDefault);
  if (he != hipSuccess) { (void)hipStreamDestroy(g->stream); delete g; return set_error(SPM_EHIP, "hipHostMalloc failed"); }
  int rc = SPM_OK;
  const size_t vb = sizeof(int32_t) * (size_t(V) + 1), eb = sizeof(int32_t) * size_t(E);
  if (rc == SPM_OK) rc = dmalloc(g, (void**)&g->off, vb, "indexofNodes");
  if (rc == SPM_OK) rc = dmalloc(g, (void**)&g->adj, eb, "edgeList");
  if (rc == SPM_OK && w) rc = dmalloc(g, (void**)&g->w, eb, "edgeLen");
  if (rc == SPM_OK && roff) rc = dmalloc(g, (void**)&g->roff, vb, "rev_indexofNodes");
  if (rc == SPM_OK && roff) rc = dmalloc(g, (void**)&g->radj, eb, "srcList");
  if (rc == SPM_OK) {
    hipError_t e = hipMemcpyAsync(g->off, off, vb, hipMemcpyHostToDevice, g->stream);
    if (e == hipSuccess && E) e = hipMemcpyAsync(g->adj, adj, eb, hipMemcpyHostToDevice, g->stream);
    if (e == hipSuccess && w && E) e = hipMemcpyAsync(g->w, w, eb, hipMemcpyHostToDevice, g->stream);
    if (e == hipSuccess && roff) e = hipMemcpyAsync(g->roff, roff, vb, hipMemcpyHostToDevice, g->stream);
    if (e == hipSuccess && roff && E) e = hipMemcpyAsync(g->radj, radj, eb, hipMemcpyHostToDevice, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = set_error(SPM_EHIP, std::string("spm_graph_create upload: ") + hipGetErrorString(e));
  }
  if (rc == SPM_OK && roff) rc = check_transpose(g);
  if (rc != SPM_OK) {
    spm_graph_destroy(g);
    return rc;
  }
  heuristic_schedule(g, &g->sched);
  *out = g;
  return SPM_OK;
}

int spm_graph_destroy(spm_graph* g) {
  if (!g) return SPM_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  pr_release(g);
  sssp_release(g);
  tc_release(g);
  bc_release(g);
  dyn_release(g);
  relabel_release(g);
  const size_t vb = sizeof(int32_t) * (size_t(g->V) + 1), eb = sizeof(int32_t) * size_t(g->E);
  dfree(g, g->off, vb);
  dfree(g, g->adj, eb);
  dfree(g, g->w, eb);
  dfree(g, g->roff, vb);
  dfree(g, g->radj, eb);
  for (auto& p : g->prof.pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto e : g->prof.pool) (void)hipEventDestroy(e);
  if (g->h_pinned) (void)hipHostFree(g->h_pinned);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return SPM_OK;
}

int spm_graph_set_weights(spm_graph* g, const int32_t* weight) {
  if (!g || !weight) return set_error(SPM_EINVAL, "spm_graph_set_weights: NULL argument");
  SPM_TRY(dyn_settle(g));
  dyn_weights_changed(g);
  const size_t eb = sizeof(int32_t) * size_t(g->E);
  if (!g->w) SPM_TRY(dmalloc(g, (void**)&g->w, eb, "edgeLen"));
  if (g->E) SPM_CHECK_HIP(hipMemcpyAsync(g->w, weight, eb, hipMemcpyHostToDevice, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  SPM_TRY(relabel_weights(g));
  return sssp_weights_changed(g);
}

void* spm_graph_stream(spm_graph* g) { return g ? (void*)g->stream : nullptr; }

int spm_graph_sync(spm_graph* g) {
  if (!g) return set_error(SPM_EINVAL, "spm_graph_sync: NULL handle");
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  return SPM_OK;
}

int64_t spm_graph_device_bytes(spm_graph* g) { return g ? g->bytes : -1; }

int spm_get_schedule(spm_graph* g, spm_schedule* s) {
  if (!g || !s) return set_error(SPM_EINVAL, "spm_get_schedule: NULL argument");
  *s = g->sched;
  return SPM_OK;
}

int spm_set_schedule(spm_graph* g, const spm_schedule* s) {
  if (!g || !s) return set_error(SPM_EINVAL, "spm_set_schedule: NULL argument");
  SPM_TRY(validate_schedule(s));
  const bool pr_changed = s->pr_t1 != g->sched.pr_t1 || s->pr_t2 != g->sched.pr_t2 ||
                          s->pr_t3 != g->sched.pr_t3 || s->pr_group != g->sched.pr_group ||
                          s->pr_mode != g->sched.pr_mode || s->pr_seg_log2 != g->sched.pr_seg_log2;
  const bool sssp_changed = s->sssp_rounds != g->sched.sssp_rounds || s->sssp_grid != g->sched.sssp_grid ||
                            s->sssp_mode != g->sched.sssp_mode ||
                            s->sssp_band_rounds != g->sched.sssp_band_rounds;
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  if (pr_changed) pr_release(g);
  if (sssp_changed) sssp_release(g);
  g->sched = *s;
  g->sched_default = false;
  return SPM_OK;
}

int spm_default_schedule(spm_graph* g, spm_schedule* s) {
  if (!g || !s) return set_error(SPM_EINVAL, "spm_default_schedule: NULL argument");
  heuristic_schedule(g, s);
  return SPM_OK;
}

int spm_profile_enable(spm_graph* g, int enable) {
  if (!g) return set_error(SPM_EINVAL, "spm_profile_enable: NULL handle");
  SPM_TRY(prof_flush(g));
  g->prof.on = enable != 0;
  return SPM_OK;
}

int spm_profile_reset(spm_graph* g) {
  if (!g) return set_error(SPM_EINVAL, "spm_profile_reset: NULL handle");
  SPM_TRY(prof_flush(g));
  g->prof.totals.clear();
  return SPM_OK;
}

int spm_profile_query(spm_graph* g, const char* kernel, double* total_ms, int64_t* launches) {
  if (!g || !kernel) return set_error(SPM_EINVAL, "spm_profile_query: NULL argument");
  SPM_TRY(prof_flush(g));
  auto it = g->prof.totals.find(kernel);
  if (total_ms) *total_ms = it == g->prof.totals.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == g->prof.totals.end() ? 0 : it->second.second;
  return SPM_OK;
}

}  // extern "C"

// ---- gather-rate probe --------------------------------------------------------
namespace spm {

__device__ __forceinline__ uint32_t probe_mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

constexpr int kProbeGrid = 8192, kProbeIters = 16;

template <bool LOCAL>
__global__ __launch_bounds__(256) void gather_probe(const float* __restrict__ a, uint32_t n,
                                                    float* __restrict__ out) {
  const uint32_t tid = blockIdx.x * 256u + threadIdx.x;
  const uint32_t span = LOCAL ? n / 8u : n, lo = LOCAL ? (blockIdx.x % 8u) * span : 0u;
  float s = 0.f;
  for (int it = 0; it < kProbeIters; ++it) {
    uint32_t idx[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) idx[k] = lo + probe_mix(tid * 977u + uint32_t(it * 8 + k) * 0x9E3779B9u) % span;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[idx[k]];
  }
  out[tid] = s;
}

}  // namespace spm

extern "C" int spm_gather_rate(int32_t n, int32_t xcd_local, double* rate) {
  if (n < 8 || !rate) return set_error(SPM_EINVAL, "spm_gather_rate: need n >= 8 and a result pointer");
  float *a = nullptr, *out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&a, sizeof(float) * size_t(n));
  if (e == hipSuccess) e = hipMalloc((void**)&out, sizeof(float) * size_t(kProbeGrid) * 256);
  if (e == hipSuccess) e = hipMemset(a, 0, sizeof(float) * size_t(n));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  float ms = 0.f;
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {  // the first run warms the caches and the code object
    e = hipEventRecord(e0, st);
    if (xcd_local)
      hipLaunchKernelGGL(gather_probe<true>, dim3(kProbeGrid), dim3(256), 0, st, (const float*)a, uint32_t(n), out);
    else
      hipLaunchKernelGGL(gather_probe<false>, dim3(kProbeGrid), dim3(256), 0, st, (const float*)a, uint32_t(n), out);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (out) (void)hipFree(out);
  if (e != hipSuccess) return set_error(SPM_EHIP, std::string("spm_gather_rate: ") + hipGetErrorString(e));
  *rate = double(kProbeGrid) * 256.0 * kProbeIters * 8.0 / (double(ms) * 1e-3);
  return SPM_OK;
}
