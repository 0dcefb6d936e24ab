// spm_internal.hpp -- shared definitions of the gfx950 kernel library:
// the device graph handle, error plumbing, launch/profiling helper and
// wavefront primitives. Wavefronts are 64 lanes on CDNA4.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/starplat_amd.h"

namespace spm {

constexpr int kWave = 64;

// ---- errors ---------------------------------------------------------------
int set_error(int code, const std::string& msg);
#define SPM_CHECK_HIP(expr)                                                       \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::spm::set_error(SPM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define SPM_TRY(expr)              \
  do {                             \
    int r_ = (expr);               \
    if (r_ != SPM_OK) return r_;   \
  } while (0)

// ---- device buffers -------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct SsspWork;
struct PrWork;
struct PrMulti;
struct TcWork;
struct BcWork;
struct DynWork;
struct Relabel;

struct Profiler {
  bool on = false;
  struct Pending { std::string name; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, int64_t>> totals;
};

}  // namespace spm

struct spm_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  int32_t V = 0;
  int64_t E = 0;
  int32_t* off = nullptr;   // indexofNodes [V+1]
  int32_t* adj = nullptr;   // edgeList     [E]
  int32_t* w = nullptr;     // weights      [E]
  int32_t* roff = nullptr;  // rev_indexofNodes [V+1]
  int32_t* radj = nullptr;  // srcList      [E]
  int num_cus = 256;
  int64_t bytes = 0;
  spm_schedule sched{};
  spm::SsspWork* sssp = nullptr;
  spm::PrWork* pr = nullptr;
  spm::PrMulti* prm = nullptr;  // batched PageRank jobs (original ids)
  spm::TcWork* tc = nullptr;
  spm::BcWork* bc = nullptr;
  spm::DynWork* dyn = nullptr;  // edge updates since the CSR was last clean
  spm::Relabel* rl = nullptr;   // degree-ordered copy of the CSR (PR, SSSP)
  int32_t max_outdeg = 0;       // of the current clean CSR (creation, then each fold)
  bool sched_default = true;    // the schedule is still the heuristic one (no spm_set_schedule)
  spm::Profiler prof;
  int32_t* h_pinned = nullptr;  // small pinned scratch for flag readback (16 ints)
};

namespace spm {

// Allocation accounted on the handle.
int dmalloc(spm_graph* g, void** p, size_t bytes, const char* what);
void dfree(spm_graph* g, void* p, size_t bytes);

// Profiling: events for one launch on g->stream, and their collection.
void prof_begin(spm_graph* g, const char* name, hipEvent_t* a, hipEvent_t* b);
void prof_end(spm_graph* g, const char* name, hipEvent_t a, hipEvent_t b);
int prof_flush(spm_graph* g);

// With profiling on, the launch carries a start and a stop event that the
// runtime stamps with the dispatch's own begin and end times (the timestamps
// a kernel trace reads), so a microsecond kernel is timed without event
// packets of its own around it.
template <typename Kernel, typename... Args>
inline int launch(spm_graph* g, const char* name, Kernel k, dim3 grid, dim3 block, size_t shmem,
                  Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (g->prof.on) {
    prof_begin(g, name, &a, &b);
    hipExtLaunchKernelGGL(k, grid, block, uint32_t(shmem), g->stream, a, b, 0, args...);
  } else {
    hipLaunchKernelGGL(k, grid, block, shmem, g->stream, args...);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(SPM_EHIP, std::string("launch ") + name + ": " + hipGetErrorString(e));
  if (g->prof.on) prof_end(g, name, a, b);
  return SPM_OK;
}

int pr_release(spm_graph* g);
// After the CSR changed (a fold): a handle still on the heuristic schedule
// gets the heuristic of the new CSR (its size and max out-degree).
void schedule_after_fold(spm_graph* g);
// Static triangle count (Compute_TC) of the handle's graph, updates folded first.
int tc_run(spm_graph* g, uint64_t* out);
int sssp_release(spm_graph* g);
int tc_release(spm_graph* g);
int bc_release(spm_graph* g);
int dyn_release(spm_graph* g);
// Folds pending edge updates into a clean sorted CSR; every static algorithm
// calls it first (a no-op when there are none).
int dyn_settle(spm_graph* g);
// True while edge updates are waiting to be folded (dyn_settle has work).
bool dyn_pending(const spm_graph* g);
// The handle's weights changed: weights derived from them are stale.
void dyn_weights_changed(spm_graph* g);
int sssp_run(spm_graph* g, const int32_t* srcs, int32_t n, int32_t* d_out, int32_t* rounds_out);

// The CSR PageRank and SSSP run on: the handle's own arrays, or (schedule
// option relabel) a copy with vertices renumbered by descending out-degree,
// built on first use. perm maps original ids to new ones (null: identity).
struct CsrView {
  const int32_t* off;
  const int32_t* adj;
  const int32_t* w;
  const int32_t* roff;
  const int32_t* radj;
  const int32_t* perm;    // device [V]
};
int csr_view(spm_graph* g, CsrView* v);
// Drops the relabeled copy (and the PR / SSSP workspaces that refer to it).
int relabel_release(spm_graph* g);
// Re-derives the relabeled copy's weights after the handle's weights changed.
int relabel_weights(spm_graph* g);
// out[v] = in[perm[v]] for V 4-byte elements (a copy without relabeling);
// in and out must differ.
int gather_back(spm_graph* g, const CsrView& v, const void* in, void* out);
// Sorted CSR (rows ascending by column) and, when roff is given, reverse CSR
// from n (row << 32 | column, value) pairs in k0/v0; k1/v1 are scratch of the
// same size. w / rw receive the values (may be null).
int csr_from_pairs(spm_graph* g, int64_t n, unsigned long long* k0, unsigned long long* k1, int32_t* v0, int32_t* v1,
                   int32_t* off, int32_t* adj, int32_t* w, int32_t* roff, int32_t* radj, int32_t* rw);
// Distances of the last SSSP run on g (device memory, V entries).
int32_t* sssp_dist(spm_graph* g);
// The weights SSSP runs on changed: refreshes SSSP's derived edge layout.
int sssp_weights_changed(spm_graph* g);

// ---- wavefront primitives ---------------------------------------------------

// Kernels with one wavefront per vertex visit x = wave, wave + nw, ... and
// handle vertex spread_vertex(x, V): a fixed bijection of [0, V) (two
// xorshift-multiply rounds on the bit width of V, which are bijections of
// [0, 2^k), cycle-walked back into [0, V)). With the identity order,
// wavefront k gets every vertex congruent to k mod nw, and the heavy
// vertices of RMAT graphs -- ids with many zero bits -- all land on the same
// few wavefronts (15-70x the mean load at RMAT-21/22).
__host__ __device__ __forceinline__ int64_t spread_vertex(int64_t x, int32_t V) {
  const uint32_t k = V > 1 ? 32u - uint32_t(__builtin_clz(uint32_t(V - 1))) : 0u;
  const uint32_t mask = k >= 32 ? ~0u : (1u << k) - 1u;
  const uint32_t h = (k + 1) / 2;
  uint32_t y = uint32_t(x);
  do {
    y ^= y >> h;
    y = (y * 0x9E3779B1u) & mask;
    y ^= y >> h;
    y = (y * 0x85EBCA6Bu) & mask;
  } while (y >= uint32_t(V));
  return int64_t(y);
}
__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Inclusive scan across the 64 lanes of a wavefront.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = __lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T t = __shfl_up(v, o, kWave);
    if (l >= o) v += t;
  }
  return v;
}

// Sum over groups of G consecutive lanes (G a power of two <= 64); every
// lane of the group receives the total.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

}  // namespace spm
