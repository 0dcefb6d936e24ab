// spm_sssp.hip -- Bellman-Ford SSSP for gfx950, the replacement of the code
// StarPlat generates for sssp_dslV2 (graphcode/staticDSLCodes/sssp_dslV2:1-15):
// fixedPoint until(finished: !modified) { forall v (modified) forall nbr:
//   <nbr.dist, nbr.modified> = <Min(nbr.dist, v.dist + e.weight), True> }
// (OMP emission: backend_omp/dsl_cpp_generator.cpp:297-527, :2637-2708;
//  CUDA host loop with two copies per round: backend_cuda/...cpp:3405-3530).
//
// Design. The frontier ("modified" set) is a queue of vertices, each stored
// with the exclusive prefix of out-degrees of the queue entries before it.
// One 64-bit atomic per workgroup reserves (entries, edges) together, so
// entry positions and edge prefixes grow in the same order. The next round
// splits the frontier's edges into chunks of kChunk edges; the pusher of a
// vertex records, for every chunk that starts inside its edge range, the
// queue position that chunk starts at. A workgroup then stages its chunk's
// vertices (at most kChunk+1) in LDS, every entry marks the slot of its
// first edge and a max-scan over the chunk's slots gives each edge its entry
// (lanes take consecutive edges, so loads stay coalesced). Hubs of 10^5
// edges are spread over ~100 workgroups, and every relaxation lane does
// useful work. Frontier distances are read with
// device-scope loads (a stale one could lose an improvement); the per-edge
// test against the neighbour's distance is a cached pre-filter in front of
// the authoritative atomicMin.
//
// Large rounds (frontier over kSparseEdges edges) do not touch a shared
// counter per pushed vertex, nor write anything per improved vertex beyond
// the atomicMin: two tiled kernels queue, in ascending vertex order, every
// vertex whose distance is below the one it was last queued with (a second
// distance array, written by the compaction itself; each tile sums the
// counts of the tiles before it).
//
// Three queues rotate (read k, write k+1, clear k+2), so a round is one
// kernel and R rounds are captured into one hipGraph; the host reads one
// 8-byte counter per R rounds instead of two copies per round. A round whose
// input queue is empty exits at once. Per-vertex round stamps deduplicate
// pushes without clearing flags between rounds.
//
// With non-negative weights the seeds' round runs before the captured ones
// and queues its frontier in distance bands (nearest first), so the big next
// round's chunks run roughly in distance order and its pre-filter discards
// more. Each handle remembers how many rounds its last run needed and which
// of them were big: the first replay of a run is fitted to that, without
// trailing empty rounds and without compaction kernels in small rounds.
//
// For graphs of high diameter with local ids (grids, meshes) a band-local
// mode (sssp_band, schedule sssp_mode 3) lets each workgroup run several
// rounds inside its own range of vertices out of LDS per launch.
//
// Distances are exact shortest paths, so the result is bit-identical to any
// correct schedule of the reference's fixed point.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "spm_internal.hpp"

namespace spm {

constexpr int kSsspBlock = 256;
// 1024-edge chunks keep the LDS stage at 12 KB, so 8 wavefronts fit per SIMD
// (2048-edge chunks: 24.6 KB, 6 per SIMD; measured at RMAT-22 1.66 vs 1.73 ms).
constexpr int kChunk = 1024;
constexpr int kPer = kChunk / kSsspBlock;
constexpr int kSsspWaves = kSsspBlock / kWave;
// A round whose frontier has fewer edges pushes through a counter (at most
// kSparseEdges / kChunk reservations); larger rounds are compacted by tiles.
constexpr uint32_t kSparseEdges = 1u << 19;
constexpr int kTileV = kSsspBlock * 16;  // vertices per compaction tile
// Distance bands of the first frontier (see sssp_count_b).
constexpr int kBands = 32;

struct SsspQueue {
  int32_t* vtx;     // [V] frontier vertices (each pushed at most once a round)
  uint32_t* eoff;   // [V] exclusive out-degree prefix of the entry
  int32_t* cfirst;  // [E/kChunk + 2] first entry of every edge chunk
};

struct SsspWork {
  SsspQueue q[3];
  unsigned long long* ctr = nullptr;  // [4] (entries<<32 | edges) x3, rounds
  uint32_t* stamp = nullptr;          // [V] round of the last push (rounds below kSparseEdges)
  int32_t* dprev = nullptr;           // [V] distance each vertex was last queued with (>= dist)
  uint16_t* qmask = nullptr;          // [tiles * kSsspBlock] queued vertices of a compaction lane
  uint32_t* round_base = nullptr;     // [1]
  int32_t* dist = nullptr;            // [V] working distances
  uint8_t* active = nullptr;          // [2][V] dense-mode frontier flags
  uint2* tile_sum = nullptr;          // [tiles] (vertices, edges) per compaction tile
  int32_t* srcbuf = nullptr;          // multi-source seeds
  int32_t* dist_out = nullptr;        // [V] distances in original ids (relabeled runs)
  uint32_t* pk = nullptr;             // [E] (neighbour << wb | weight) when both fit 32 bits
  int wb = 0;                         // weight bits of pk (0: not packed)
  unsigned long long* wstat = nullptr;  // [2] weight range of the view (any negative, max)
  uint2* tile_b = nullptr;            // [kBands][tiles] (vertices, edges) before each tile within its band
  unsigned long long* band_tot = nullptr;  // [kBands] (vertices << 32 | edges) per band
  int64_t maxw = -1;                  // largest weight when none is negative, else -1
  bool dist_out_ok = false;
  CsrView view{};                     // the CSR the kernels run on
  size_t qcap = 0, ccap = 0, src_cap = 0;
  // Captured rounds: [0] the schedule's R rounds, [1] the first replay's
  // length fitted to the previous run (a multiple of 3 below R).
  struct Captured {
    int rounds = 0;
    int mode = -1;              // effective sssp mode the rounds were captured for
    std::vector<uint8_t> comp;  // rounds with compaction kernels (empty: all)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } cap[2];
  int last_rounds = 0;  // non-empty captured rounds of the previous queue-mode run (0: none)
  unsigned long long big_hist[2] = {~0ull, ~0ull};  // big rounds of the previous run (bit r)
  uint32_t* run_start = nullptr;         // [1] round stamp of the first captured round
  unsigned long long* big = nullptr;     // [2] big rounds of this run (bit r)
  uint32_t host_base = 1;
};

__global__ void sssp_init(int32_t V, int32_t* __restrict__ dist, int32_t* __restrict__ dprev,
                          uint8_t* __restrict__ active) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < V; i += stride) {
    dist[i] = INT_MAX;
    if (dprev) dprev[i] = INT_MAX;
    if (active) {
      active[i] = 0;
      active[V + i] = 0;
    }
  }
}

__global__ void sssp_advance(uint32_t* round_base, uint32_t by) { *round_base += by; }

// Weight range: st[0] = 1 if any weight is negative, st[1] = max weight.
__global__ void sssp_wrange(int64_t E, const int32_t* __restrict__ w, unsigned long long* __restrict__ st) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  int32_t mx = 0;
  bool neg = false;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < E; i += stride) {
    mx = max(mx, w[i]);
    neg |= w[i] < 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, kWave));
  const bool any_neg = __any(neg);
  if (__lane_id() == 0) {
    atomicMax(&st[1], static_cast<unsigned long long>(mx));
    if (any_neg) atomicOr(&st[0], 1ull);
  }
}

__global__ void sssp_pack(int64_t E, const int32_t* __restrict__ adj, const int32_t* __restrict__ w, int wb,
                          uint32_t* __restrict__ pk) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < E; i += stride)
    pk[i] = (uint32_t(adj[i]) << wb) | uint32_t(w[i]);
}

// Block-wide exclusive scan of two counters (4 wavefronts of 64 lanes).
__device__ __forceinline__ void block_scan2(uint32_t a, uint32_t b, uint32_t& ea, uint32_t& eb,
                                            uint32_t& ta, uint32_t& tb, uint32_t* s_tot) {
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  const uint32_t ia = wave_incl_scan(a), ib = wave_incl_scan(b);
  if (lane == kWave - 1) {
    s_tot[wid] = ia;
    s_tot[kSsspWaves + wid] = ib;
  }
  __syncthreads();
  uint32_t pa = 0, pb = 0;
  ta = 0;
  tb = 0;
#pragma unroll
  for (int w = 0; w < kSsspWaves; ++w) {
    const uint32_t xa = s_tot[w], xb = s_tot[kSsspWaves + w];
    if (w < wid) { pa += xa; pb += xb; }
    ta += xa;
    tb += xb;
  }
  ea = pa + ia - a;
  eb = pb + ib - b;
}

// Seeds the first frontier with n distinct sources (one workgroup): dist 0,
// queue entries with their degree prefix, and the chunk index of every chunk
// starting inside a source's edges, in queue `slot` (counter ctr[slot]; the
// other counters, rounds included, start at zero; so do the band totals and
// the big-round log, and run_start takes the first captured round's stamp).
// srcs == nullptr means the single source `one`.
__global__ __launch_bounds__(kSsspBlock) void sssp_seed(const int32_t* __restrict__ srcs, int32_t n, int32_t one,
                                                        const int32_t* __restrict__ perm,
                                                        const int32_t* __restrict__ off, int32_t* __restrict__ dist,
                                                        int32_t* __restrict__ dprev, SsspQueue q0, unsigned long long* __restrict__ ctr, int slot,
                                                        unsigned long long* __restrict__ band_tot,
                                                        uint32_t* __restrict__ run_start, uint32_t start_val,
                                                        unsigned long long* __restrict__ big,
                                                        uint8_t* __restrict__ active) {
  if (band_tot && threadIdx.x < kBands) band_tot[threadIdx.x] = 0ull;
  if (threadIdx.x < 2) big[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) *run_start = start_val;
  __shared__ uint32_t s_tot[2 * kSsspWaves];
  uint32_t cnt = 0, edges = 0;  // running totals (uniform across the block)
  for (int32_t base = 0; base < n; base += kSsspBlock) {
    const int32_t j = base + int32_t(threadIdx.x);
    int32_t s = -1, deg = 0;
    if (j < n) {
      s = srcs ? srcs[j] : one;
      if (perm) s = perm[s];  // original id -> the relabeled copy's id
      deg = off[s + 1] - off[s];
      dist[s] = 0;
      if (dprev) dprev[s] = 0;
      if (active) active[s] = 1;
    }
    const uint32_t push = deg > 0 ? 1u : 0u;
    uint32_t ep, ed, tp, td;
    block_scan2(push, uint32_t(deg), ep, ed, tp, td, s_tot);
    if (push) {
      const uint32_t pos = cnt + ep, eo = edges + ed;
      q0.vtx[pos] = s;
      q0.eoff[pos] = eo;
      for (uint32_t cc = (eo + kChunk - 1) / kChunk; cc <= (eo + uint32_t(deg) - 1) / kChunk; ++cc)
        q0.cfirst[cc] = int32_t(pos);
    }
    cnt += tp;
    edges += td;
    __syncthreads();  // s_tot reuse
  }
  if (threadIdx.x < 4)
    ctr[threadIdx.x] = int(threadIdx.x) == slot ? (static_cast<unsigned long long>(cnt) << 32) | edges : 0ull;
}

// PACKED: the frontier's edges are read as one 32-bit word (neighbour << wb |
// weight) instead of two arrays -- half the streamed bytes per edge.
template <bool PACKED>
__global__ __launch_bounds__(kSsspBlock) void sssp_relax(
    const int32_t* __restrict__ off, const int32_t* __restrict__ adj,
    const int32_t* __restrict__ wgt, const uint32_t* __restrict__ pk, int wb,
    int32_t* __restrict__ dist, uint32_t* __restrict__ stamp, int32_t* __restrict__ dprev,
    SsspQueue qin, SsspQueue qout, unsigned long long* __restrict__ ctr_in,
    unsigned long long* __restrict__ ctr_out, unsigned long long* __restrict__ ctr_clear,
    unsigned long long* __restrict__ rounds, const uint32_t* __restrict__ round_base, uint32_t rloc,
    uint32_t sparse_max, const uint32_t* __restrict__ run_start, unsigned long long* __restrict__ big) {
  __shared__ uint32_t s_off[kChunk + 1];
  __shared__ int32_t s_dv[kChunk + 1];
  __shared__ int32_t s_base[kChunk + 1];
  __shared__ int16_t s_own[kChunk];  // entry owning each edge slot of the chunk
  __shared__ int32_t s_wmax[kSsspWaves];
  __shared__ uint32_t s_tot[2 * kSsspWaves];
  __shared__ unsigned long long s_res;

  const unsigned long long c = *ctr_in;
  const uint32_t n = uint32_t(c >> 32), T = uint32_t(c);
  const uint32_t sv = *round_base + rloc;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *ctr_clear = 0ull;
    if (T) *rounds += 1;
    // Rounds of this run (counted from the captured ones) whose frontier was
    // large enough to want compaction kernels next time.
    const uint32_t ar = sv - *run_start;
    if (big && ar < 128u && T >= kSparseEdges / 4) atomicOr(&big[ar >> 6], 1ull << (ar & 63u));
  }
  if (T == 0) return;
  const uint32_t nchunks = (T + kChunk - 1) / kChunk;
  const int tid = threadIdx.x;
  // Small rounds (below sparse_max edges) push through the block-aggregated
  // counter; large rounds only lower distances and leave the queue to
  // sssp_count/sssp_emit (or sssp_count_b/sssp_emit_b).
  const bool sparse = T < sparse_max;

  for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const int32_t p0 = qin.cfirst[chunk];
    const int32_t p1 = chunk + 1 < nchunks ? qin.cfirst[chunk + 1] : int32_t(n) - 1;
    const int cnt = p1 - p0 + 1;
    const uint32_t x0 = chunk * kChunk;
#pragma unroll
    for (int q = 0; q < kPer; ++q) s_own[tid * kPer + q] = -1;
    __syncthreads();
    for (int i = tid; i < cnt; i += kSsspBlock) {
      const int32_t v = qin.vtx[p0 + i];
      const uint32_t o = qin.eoff[p0 + i];
      s_off[i] = o;
      s_dv[i] = __hip_atomic_load(&dist[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_base[i] = off[v];
      // Each entry marks the chunk slot of its first edge (the chunk's first
      // entry may start before it: slot 0; the last may start at the next
      // chunk: no slot); entries have at least one edge, so marks never
      // collide.
      if (o < x0 + kChunk) s_own[o > x0 ? o - x0 : 0] = int16_t(i);
    }
    __syncthreads();
    // Owner of every slot = the last mark at or before it: a max-scan over
    // the chunk, four consecutive slots per lane (entries are ascending).
    {
      int m[kPer];
      int run = -1;
#pragma unroll
      for (int q = 0; q < kPer; ++q) {
        run = max(run, int(s_own[tid * kPer + q]));
        m[q] = run;
      }
      int incl = run;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int y = __shfl_up(incl, d, kWave);
        if (__lane_id() >= d) incl = max(incl, y);
      }
      if (__lane_id() == kWave - 1) s_wmax[tid / kWave] = incl;
      int excl = __shfl_up(incl, 1, kWave);
      if (__lane_id() == 0) excl = -1;
      __syncthreads();
      for (int w = 0; w < tid / kWave; ++w) excl = max(excl, s_wmax[w]);
#pragma unroll
      for (int q = 0; q < kPer; ++q) s_own[tid * kPer + q] = int16_t(max(excl, m[q]));
    }
    __syncthreads();

    // Pushes live in statically indexed registers (one slot per edge a lane
    // relaxes), so no private-memory array is needed.
    int32_t pv[kPer];
    int32_t pd[kPer];
    uint32_t pc = 0, sd = 0;
    // Lanes take consecutive edges (edge x0 + k*256 + tid in step k), so each
    // load instruction of a wavefront reads one contiguous run of the
    // frontier's edges. The entry owning edge x is s_own[x - x0].
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const uint32_t x = x0 + k * kSsspBlock + tid;
      pv[k] = -1;
      pd[k] = 0;
      if (x < T) {
        const int lo = s_own[k * kSsspBlock + tid];
        const int32_t e = s_base[lo] + int32_t(x - s_off[lo]);
        int32_t nb, wt;
        if (PACKED) {
          const uint32_t word = pk[e];
          nb = int32_t(word >> wb);
          wt = int32_t(word & ((1u << wb) - 1u));
        } else {
          nb = adj[e];
          wt = wgt[e];
        }
        const int32_t nd = s_dv[lo] + wt;
        // A cached (possibly stale) read is enough to skip hopeless updates:
        // distances only decrease, so a stale value is never below the true
        // one and atomicMin below stays the authority. In a large round that
        // is all: the compaction finds every vertex lowered in this round
        // (its distance is then below the one it was last queued with). In a
        // small round every lane that passes the test pushes the neighbour
        // unless this round's stamp is already there, without waiting for
        // the atomic's old value: the value read is at most the distance at
        // the start of the round, so if this lane's atomicMin does not lower
        // the distance, another lane lowered it during this round and pushes
        // it (once, by the stamp).
        if (nd < __hip_atomic_load(&dist[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
          atomicMin(&dist[nb], nd);  // no return value: the wave does not wait
          if (sparse && atomicExch(&stamp[nb], sv) != sv) {
            // nd bounds the distance it is queued with from above; a lower
            // one from later in this round queues it once more in the next
            // compacted round at worst.
            dprev[nb] = nd;
            const int32_t deg = off[nb + 1] - off[nb];
            if (deg > 0) {
              pv[k] = nb;
              pd[k] = deg;
              ++pc;
              sd += uint32_t(deg);
            }
          }
        }
      }
    }
    if (!sparse) {
      __syncthreads();  // LDS reuse by the next chunk
      continue;
    }

    uint32_t epc, esd, tpc, tsd;
    block_scan2(pc, sd, epc, esd, tpc, tsd, s_tot);
    if (tid == 0 && tpc) s_res = atomicAdd(ctr_out, (static_cast<unsigned long long>(tpc) << 32) | tsd);
    __syncthreads();
    if (pc) {
      const unsigned long long r = s_res;
      uint32_t pos = uint32_t(r >> 32) + epc;
      uint32_t eo = uint32_t(r) + esd;
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        if (pv[k] < 0) continue;
        qout.vtx[pos] = pv[k];
        qout.eoff[pos] = eo;
        const uint32_t d = uint32_t(pd[k]);
        for (uint32_t cc = (eo + kChunk - 1) / kChunk; cc <= (eo + d - 1) / kChunk; ++cc) qout.cfirst[cc] = int32_t(pos);
        ++pos;
        eo += d;
      }
    }
    __syncthreads();  // LDS reuse by the next chunk
  }
}

// Out-degrees of the lane's 16 vertices v0 .. v0+15 (v0 a multiple of 16,
// so off + v0 is 64-byte aligned): off[v0 .. v0+16] as four 16-byte loads and
// one more, all issued at once (loading only the degrees of marked vertices,
// each under its own condition, put every load's wait before the next).
__device__ __forceinline__ void tile_degrees(int32_t V, const int32_t* __restrict__ off, int64_t v0,
                                             uint32_t (&od)[16]) {
  if (v0 + 16 <= V) {
    const int4* p = reinterpret_cast<const int4*>(off + v0);
    int32_t o[17];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int4 x = p[q];
      o[4 * q] = x.x;
      o[4 * q + 1] = x.y;
      o[4 * q + 2] = x.z;
      o[4 * q + 3] = x.w;
    }
    o[16] = off[v0 + 16];
#pragma unroll
    for (int k = 0; k < 16; ++k) od[k] = uint32_t(o[k + 1] - o[k]);
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) od[k] = v0 + k < V ? uint32_t(off[v0 + k + 1] - off[v0 + k]) : 0u;
  }
}

// Frontier compaction for large rounds, step 1: each tile of kTileV vertices
// (16 consecutive per lane) counts the vertices lowered since they were last
// queued (dist < dprev) that have out-edges, and their out-degrees; dv
// receives the distances. Every vertex lowered in the round just relaxed is
// among them; so, at worst, is a vertex pushed by a small round whose
// distance fell again later in that round (relaxing it once more is harmless).
__device__ __forceinline__ void tile_flags(int32_t V, const int32_t* __restrict__ off,
                                           const int32_t* __restrict__ dist, const int32_t* __restrict__ dprev,
                                           int64_t v0, uint32_t (&deg)[16], int32_t (&dv)[16],
                                           bool (&lowered)[4]) {
  if (v0 + 16 <= V) {
    const int4* pd = reinterpret_cast<const int4*>(dist + v0);
    const int4* pp = reinterpret_cast<const int4*>(dprev + v0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int4 d = pd[q], b = pp[q];
      dv[4 * q + 0] = d.x;
      dv[4 * q + 1] = d.y;
      dv[4 * q + 2] = d.z;
      dv[4 * q + 3] = d.w;
      deg[4 * q + 0] = d.x < b.x;
      deg[4 * q + 1] = d.y < b.y;
      deg[4 * q + 2] = d.z < b.z;
      deg[4 * q + 3] = d.w < b.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      dv[k] = v0 + k < V ? dist[v0 + k] : INT_MAX;
      deg[k] = (v0 + k < V) && dv[k] < dprev[v0 + k];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) lowered[q] = deg[4 * q] | deg[4 * q + 1] | deg[4 * q + 2] | deg[4 * q + 3];
  uint32_t od[16];
  tile_degrees(V, off, v0, od);
#pragma unroll
  for (int k = 0; k < 16; ++k) deg[k] = deg[k] ? od[k] : 0u;
}

// dprev := dist over the lane's 16 vertices: each is now queued with its
// present distance or was not lowered (dist == dprev, as dprev >= dist
// always). A lane stores 16-byte words, and only those holding a lowered
// vertex (lowered[q]), so a wavefront writes whole runs of lines where the
// round lowered many distances and little where it lowered few.
__device__ __forceinline__ void tile_queued(int32_t V, int32_t* __restrict__ dprev, int64_t v0,
                                            const int32_t (&dv)[16], const bool (&lowered)[4]) {
  if (v0 + 16 <= V) {
    int4* p = reinterpret_cast<int4*>(dprev + v0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (lowered[q]) p[q] = make_int4(dv[4 * q], dv[4 * q + 1], dv[4 * q + 2], dv[4 * q + 3]);
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (v0 + k < V) dprev[v0 + k] = dv[k];
  }
}

// The lane's queued vertices also go to qmask (bit k: vertex v0 + k), which
// is all the emit step reads, and their distances to dprev.
__global__ __launch_bounds__(kSsspBlock) void sssp_count(int32_t V, const int32_t* __restrict__ off,
                                                         const int32_t* __restrict__ dist,
                                                         int32_t* __restrict__ dprev,
                                                         const unsigned long long* __restrict__ ctr_in,
                                                         uint2* __restrict__ tile_sum, uint16_t* __restrict__ qmask) {
  __shared__ uint32_t s_tot[2 * kSsspWaves];
  const uint32_t T = uint32_t(*ctr_in);
  if (T < kSparseEdges) return;  // this round pushed through the counter
  const int64_t v0 = int64_t(blockIdx.x) * kTileV + int64_t(threadIdx.x) * 16;
  uint32_t deg[16];
  int32_t dv[16];
  bool lowered[4];
  tile_flags(V, off, dist, dprev, v0, deg, dv, lowered);
  tile_queued(V, dprev, v0, dv, lowered);
  uint32_t c = 0, e = 0, m = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    c += deg[k] > 0;
    e += deg[k];
    m |= uint32_t(deg[k] > 0) << k;
  }
  qmask[int64_t(blockIdx.x) * kSsspBlock + threadIdx.x] = uint16_t(m);
  uint32_t ec, ee, tc, te;
  block_scan2(c, e, ec, ee, tc, te, s_tot);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = make_uint2(tc, te);
}

// Step 2: every tile sums the counts of the tiles before it (no atomics, no
// waiting on other workgroups), then writes the vertices of its lane masks in
// ascending order with their degree prefix and the chunk index; the last tile
// publishes the frontier size.
__global__ __launch_bounds__(kSsspBlock) void sssp_emit(int32_t V, const int32_t* __restrict__ off,
                                                        const uint16_t* __restrict__ qmask,
                                                        const unsigned long long* __restrict__ ctr_in,
                                                        const uint2* __restrict__ tile_sum, SsspQueue qout,
                                                        unsigned long long* __restrict__ ctr_out) {
  __shared__ uint32_t s_tot[2 * kSsspWaves];
  __shared__ int32_t s_vtx[kTileV];
  __shared__ uint32_t s_eoff[kTileV];
  const uint32_t T = uint32_t(*ctr_in);
  if (T < kSparseEdges) return;
  uint32_t pc = 0, pe = 0;
  for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kSsspBlock) {
    const uint2 x = tile_sum[t];
    pc += x.x;
    pe += x.y;
  }
  uint32_t d0, d1, bc, be;
  block_scan2(pc, pe, d0, d1, bc, be, s_tot);  // bc, be: sums over the earlier tiles
  __syncthreads();
  const int64_t v0 = int64_t(blockIdx.x) * kTileV + int64_t(threadIdx.x) * 16;
  const uint32_t m = qmask[int64_t(blockIdx.x) * kSsspBlock + threadIdx.x];
  uint32_t deg[16];
  tile_degrees(V, off, v0, deg);
  uint32_t c = 0, e = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    deg[k] = (m >> k) & 1u ? deg[k] : 0u;
    c += deg[k] > 0;
    e += deg[k];
  }
  uint32_t ec, ee, tc, te;
  block_scan2(c, e, ec, ee, tc, te, s_tot);
  // The tile's entries are staged in LDS in queue order and then written out
  // as two contiguous runs (a lane's own 16 vertices would make every store
  // instruction touch 64 scattered lines).
  uint32_t pos = ec, eo = be + ee;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint32_t d = deg[k];
    if (!d) continue;
    s_vtx[pos] = int32_t(v0 + k);
    s_eoff[pos] = eo;
    for (uint32_t cc = (eo + kChunk - 1) / kChunk; cc <= (eo + d - 1) / kChunk; ++cc)
      qout.cfirst[cc] = int32_t(bc + pos);
    ++pos;
    eo += d;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < tc; i += kSsspBlock) {
    qout.vtx[bc + i] = s_vtx[i];
    qout.eoff[bc + i] = s_eoff[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    *ctr_out = (static_cast<unsigned long long>(bc + tc) << 32) | (be + te);
}

// First frontier in distance bands. The seeds' round lowers the distances
// of its vertices (all of them to seed + one edge, so within [0, maxw]);
// instead of ascending vertex order they are queued by band
// b = min(kBands - 1, dist / delta), tile after tile within a band. The
// next round's chunks are dispatched roughly nearest first, so the cached
// pre-filter of its later, farther relaxations sees more final distances
// (RMAT-22, simulated with 1536 resident 1024-edge chunks: 12% fewer edges
// scanned over the whole run). Two kernels: per tile and band counts, each
// tile reserving its place within every band it occupies with one atomic on
// the band's total (tiles of a band may sit in any order), and the emit.
//
// With a single source the lowered vertices are among its out-neighbours, so
// the tiles then cover the source's row instead of all V vertices (a row is
// non-decreasing: a slot holding the same vertex as the slot before it is
// skipped, so each vertex is queued once).
struct BandTiles {
  const int32_t* row;  // the single source's out-row (null: tiles of vertex ids)
  int32_t n;           // its length
};
// A row tile is 1024 slots, four per lane: a hub's row of ~10^5 slots then
// spreads over ~100 workgroups (vertex tiles are kTileV = 4096, 16 per lane).
constexpr int kRowPer = 4;
constexpr int kRowTile = kSsspBlock * kRowPer;

__device__ __forceinline__ int64_t band_tile_start(const BandTiles& bt) {
  return bt.row ? int64_t(blockIdx.x) * kRowTile + int64_t(threadIdx.x) * kRowPer
                : int64_t(blockIdx.x) * kTileV + int64_t(threadIdx.x) * 16;
}

__device__ __forceinline__ BandTiles band_tiles(const int32_t* __restrict__ adj, const int32_t* __restrict__ off,
                                                const int32_t* __restrict__ perm, int32_t src) {
  if (src < 0) return BandTiles{nullptr, 0};
  const int32_t s = perm ? perm[src] : src;
  return BandTiles{adj + off[s], off[s + 1] - off[s]};
}

// The lane's candidates: vertex ids v0 + k (k < 16), or row slots v0 + k
// (k < kRowPer; the rest stay empty).
__device__ __forceinline__ void tile_bands(int32_t V, const int32_t* __restrict__ off,
                                           const int32_t* __restrict__ dist, const int32_t* __restrict__ dprev,
                                           const BandTiles& bt, int64_t v0, int32_t delta, uint32_t (&deg)[16],
                                           int32_t (&vid)[16], int32_t (&dv)[16], bool (&lowered)[4],
                                           uint32_t (&band)[16]) {
  if (!bt.row) {
    tile_flags(V, off, dist, dprev, v0, deg, dv, lowered);
#pragma unroll
    for (int k = 0; k < 16; ++k) vid[k] = int32_t(v0 + k);
  } else {
    // Three phases of loads, each unconditional (clamped to slot 0 or
    // vertex 0), so a lane's slots wait for three round trips in all.
    int32_t prv[kRowPer], dp[kRowPer], o0[kRowPer], o1[kRowPer];
#pragma unroll
    for (int k = 0; k < kRowPer; ++k) {
      const int64_t i = v0 + k;
      const bool in = i < bt.n;
      vid[k] = bt.row[in ? i : 0];
      prv[k] = bt.row[in && i > 0 ? i - 1 : 0];
      if (!in) vid[k] = -1;
    }
#pragma unroll
    for (int k = 0; k < kRowPer; ++k) {
      const int32_t v = max(vid[k], 0);
      dv[k] = dist[v];
      dp[k] = dprev[v];
    }
#pragma unroll
    for (int k = 0; k < kRowPer; ++k) {
      const int32_t v = max(vid[k], 0);
      o0[k] = off[v];
      o1[k] = off[v + 1];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k < kRowPer) {
        const bool first = vid[k] >= 0 && (v0 + k == 0 || prv[k] != vid[k]);
        deg[k] = first && dv[k] < dp[k] ? uint32_t(o1[k] - o0[k]) : 0u;
      } else {
        vid[k] = -1;
        dv[k] = 0;
        deg[k] = 0;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) band[k] = deg[k] ? min(uint32_t(kBands - 1), uint32_t(dv[k] / delta)) : 0u;
}

// src >= 0: tiles over the out-row of that single source (original id,
// mapped through perm); src < 0: tiles over all vertices.
__global__ __launch_bounds__(kSsspBlock) void sssp_count_b(int32_t V, const int32_t* __restrict__ off,
                                                           const int32_t* __restrict__ adj,
                                                           const int32_t* __restrict__ perm, int32_t src,
                                                           const int32_t* __restrict__ dist,
                                                           const int32_t* __restrict__ dprev, int32_t delta,
                                                           int32_t tiles, uint2* __restrict__ tile_b,
                                                           unsigned long long* __restrict__ band_tot) {
  __shared__ uint32_t s_c[kBands], s_e[kBands];
  if (threadIdx.x < kBands) {
    s_c[threadIdx.x] = 0;
    s_e[threadIdx.x] = 0;
  }
  __syncthreads();
  uint32_t deg[16], band[16];
  int32_t vid[16], dv[16];
  bool lowered[4];
  const BandTiles bt = band_tiles(adj, off, perm, src);
  tile_bands(V, off, dist, dprev, bt, band_tile_start(bt), delta, deg, vid, dv, lowered, band);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (!deg[k]) continue;
    atomicAdd(&s_c[band[k]], 1u);
    atomicAdd(&s_e[band[k]], deg[k]);
  }
  __syncthreads();
  // The tile's place within each band (tiles of a band may sit in any order).
  if (threadIdx.x < kBands && s_c[threadIdx.x]) {
    const unsigned long long at =
        atomicAdd(&band_tot[threadIdx.x], (static_cast<unsigned long long>(s_c[threadIdx.x]) << 32) | s_e[threadIdx.x]);
    tile_b[int64_t(threadIdx.x) * tiles + blockIdx.x] = make_uint2(uint32_t(at >> 32), uint32_t(at));
  }
}

// Positions within a (band, tile) come from one 64-bit LDS counter per band,
// (entries << 32 | edges), which starts at the band's start (the totals of the
// bands before it: packed sums add both halves at once, edges staying below
// 2^32) plus the tile's place within the band. Entry positions and edge
// prefixes advance together, so the queue stays consistent whatever order
// the lanes take. The first workgroup publishes the frontier size.
__global__ __launch_bounds__(kSsspBlock) void sssp_emit_b(int32_t V, const int32_t* __restrict__ off,
                                                          const int32_t* __restrict__ adj,
                                                          const int32_t* __restrict__ perm, int32_t src,
                                                          const int32_t* __restrict__ dist,
                                                          int32_t* __restrict__ dprev, int32_t delta,
                                                          int32_t tiles, const uint2* __restrict__ tile_b,
                                                          const unsigned long long* __restrict__ band_tot,
                                                          SsspQueue qout, unsigned long long* __restrict__ ctr_out) {
  __shared__ unsigned long long s_cur[kBands];
  if (threadIdx.x < kBands) {
    unsigned long long start = 0;
    for (int b = 0; b < int(threadIdx.x); ++b) start += band_tot[b];
    const uint2 p = tile_b[int64_t(threadIdx.x) * tiles + blockIdx.x];  // read only if the band occurs here
    s_cur[threadIdx.x] = start + ((static_cast<unsigned long long>(p.x) << 32) | p.y);
    if (blockIdx.x == 0 && threadIdx.x == kBands - 1) *ctr_out = start + band_tot[kBands - 1];
  }
  uint32_t deg[16], band[16];
  int32_t vid[16], dv[16];
  bool lowered[4];
  const BandTiles bt = band_tiles(adj, off, perm, src);
  const int64_t v0 = band_tile_start(bt);
  tile_bands(V, off, dist, dprev, bt, v0, delta, deg, vid, dv, lowered, band);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint32_t d = deg[k];
    if (!d) continue;
    const unsigned long long at = atomicAdd(&s_cur[band[k]], (1ull << 32) | d);
    const uint32_t pos = uint32_t(at >> 32), eo = uint32_t(at);
    qout.vtx[pos] = vid[k];
    qout.eoff[pos] = eo;
    for (uint32_t cc = (eo + kChunk - 1) / kChunk; cc <= (eo + d - 1) / kChunk; ++cc) qout.cfirst[cc] = int32_t(pos);
    if (bt.row) dprev[vid[k]] = dv[k];
  }
  if (!bt.row) tile_queued(V, dprev, v0, dv, lowered);
}

// Dense mode: the reference's full-V sweep of the modified flags, one lane per
// vertex with the out-edges of a wavefront's 64 vertices flattened across the
// lanes (wave prefix sum + LDS binary search), two flag arrays swapped per round.
__global__ __launch_bounds__(kSsspBlock) void sssp_dense(
    int32_t V, const int32_t* __restrict__ off, const int32_t* __restrict__ adj,
    const int32_t* __restrict__ wgt, int32_t* __restrict__ dist, uint8_t* __restrict__ cur,
    uint8_t* __restrict__ nxt, unsigned long long* __restrict__ ctr_in,
    unsigned long long* __restrict__ ctr_out, unsigned long long* __restrict__ ctr_clear,
    unsigned long long* __restrict__ rounds) {
  __shared__ uint32_t s_pre[kSsspWaves][kWave];
  __shared__ int32_t s_base[kSsspWaves][kWave];
  __shared__ int32_t s_dv[kSsspWaves][kWave];
  const unsigned long long c = *ctr_in;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *ctr_clear = 0ull;
    if (c) *rounds += 1;
  }
  if (c == 0) return;
  const int lane = __lane_id(), wid = threadIdx.x / kWave;
  unsigned long long found = 0;
  const int64_t nwaves = int64_t(gridDim.x) * kSsspWaves;
  for (int64_t base = (int64_t(blockIdx.x) * kSsspWaves + wid) * kWave; base < V; base += nwaves * kWave) {
    const int64_t v = base + lane;
    uint32_t deg = 0;
    int32_t b = 0, dv = 0;
    if (v < V && cur[v]) {
      cur[v] = 0;
      b = off[v];
      deg = uint32_t(off[v + 1] - b);
      dv = __hip_atomic_load(&dist[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint32_t incl = wave_incl_scan(deg);
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    s_pre[wid][lane] = incl - deg;
    s_base[wid][lane] = b;
    s_dv[wid][lane] = dv;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t x = lane; x < total; x += kWave) {
      int lo = 0, hi = kWave - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[wid][mid] <= x) lo = mid;
        else hi = mid - 1;
      }
      const int32_t e = s_base[wid][lo] + int32_t(x - s_pre[wid][lo]);
      const int32_t nb = adj[e];
      const int32_t nd = s_dv[wid][lo] + wgt[e];
      // cached pre-filter; atomicMin decides (see sssp_relax)
      if (nd < __hip_atomic_load(&dist[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
        atomicMin(&dist[nb], nd);  // marks without the old value (see sssp_relax)
        nxt[nb] = 1;
        found = 1;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  // Read before the atomic: once one wavefront has raised the flag the others
  // skip the same-address atomic.
  if (__any(found) && lane == 0 && __hip_atomic_load(ctr_out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull)
    atomicOr(ctr_out, 1ull);
}

// Band-local mode (schedule sssp_mode 3), for graphs of high diameter whose
// edges join nearby ids (meshes, grids, road networks in a locality-keeping
// order): there the fixed point takes thousands of rounds of small frontiers,
// and a round costs a launch and a chain of dependent memory trips, not
// bandwidth. Workgroup b owns the vertices [b*kBandV, (b+1)*kBandV) and runs
// up to K rounds (schedule sssp_band_rounds) inside its band in one launch:
// distances, offsets and the frontier lists live in LDS, rounds are separated
// by a workgroup barrier, and an in-band improvement queues the vertex for
// the band's next round. An edge leaving the band lowers the global distance
// with atomicMin and flags its target for the next launch. At the end the
// band's distances are written back (atomicMin: another band may have
// lowered one meanwhile) and vertices still queued are flagged. Launches
// repeat until one flags nothing. Every improvement is propagated before
// the fixed point is declared, and any relaxation order reaches the same
// shortest distances, so the result is bit-identical to the other modes.
constexpr int kBandBlock = 1024;
constexpr int kBandV = 4096;

template <bool PACKED>
__global__ __launch_bounds__(kBandBlock) void sssp_band(
    int32_t V, const int32_t* __restrict__ off, const int32_t* __restrict__ adj, const int32_t* __restrict__ wgt,
    const uint32_t* __restrict__ pk, int wb, int32_t* __restrict__ dist, uint8_t* __restrict__ cur,
    uint8_t* __restrict__ nxt, int K, unsigned long long* __restrict__ ctr_in,
    unsigned long long* __restrict__ ctr_out, unsigned long long* __restrict__ ctr_clear,
    unsigned long long* __restrict__ rounds) {
  __shared__ int32_t s_dist[kBandV];
  __shared__ int32_t s_off[kBandV + 1];
  __shared__ uint32_t s_mark[kBandV];    // 1 + the local round whose input list holds the vertex
  __shared__ int16_t s_list[2][kBandV];  // band-local indices
  __shared__ int32_t s_cnt[3];
  __shared__ int32_t s_out;

  const unsigned long long c = *ctr_in;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *ctr_clear = 0ull;
    if (c) *rounds += 1;
  }
  if (c == 0) return;
  const int tid = threadIdx.x, lane = __lane_id();
  const int32_t base = int32_t(blockIdx.x) * kBandV;
  const int32_t nv = min(kBandV, V - base);
  if (tid < 3) s_cnt[tid] = 0;
  if (tid == 0) s_out = 0;
  __syncthreads();
  // The band's flagged vertices form the first list (their flags cleared:
  // this buffer is the next launch's output).
  for (int i = tid; i < nv; i += kBandBlock) {
    uint32_t m = 0;
    if (cur[base + i]) {
      cur[base + i] = 0;
      m = 1;
      s_list[0][atomicAdd(&s_cnt[0], 1)] = int16_t(i);
    }
    s_mark[i] = m;
  }
  __syncthreads();
  int n = s_cnt[0];
  if (n == 0) return;  // nothing queued in this band (uniform)
  for (int i = tid; i < nv; i += kBandBlock) {
    s_dist[i] = dist[base + i];
    s_off[i] = off[base + i];
  }
  if (tid == 0) s_off[nv] = off[base + nv];
  __syncthreads();

  int r = 0;  // round r reads s_list[r & 1], counted in s_cnt[r % 3]
  for (; r < K && n > 0; ++r) {
    const int16_t* lin = s_list[r & 1];
    int16_t* lout = s_list[(r + 1) & 1];
    int32_t* cout = &s_cnt[(r + 1) % 3];
    if (tid == 0) s_cnt[(r + 2) % 3] = 0;  // the next round's output counter
    bool out_flag = false;
    // One lane per queued vertex, its out-edges four at a time with the four
    // edge loads in flight together (the graphs this mode is for have short
    // rows, so a round is one global round trip plus LDS work).
    for (int i = tid; i < n; i += kBandBlock) {
      const int li = lin[i];
      const int32_t dv = s_dist[li];
      const int32_t e1 = s_off[li + 1];
      for (int32_t e = s_off[li]; e < e1; e += 4) {
        // Unconditional loads (clamped to the row's last edge), so the four
        // are issued together rather than each waiting behind its branch.
        int32_t nb[4], wt[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int32_t ek = min(e + k, e1 - 1);
          if (PACKED) {
            const uint32_t word = pk[ek];
            nb[k] = int32_t(word >> wb);
            wt[k] = int32_t(word & ((1u << wb) - 1u));
          } else {
            nb[k] = adj[ek];
            wt[k] = wgt[ek];
          }
          if (e + k >= e1) nb[k] = -1;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (nb[k] < 0) continue;
          const int32_t nd = dv + wt[k];
          const int32_t ln = nb[k] - base;
          if (ln >= 0 && ln < nv) {
            // In the band: the LDS minimum decides; a lowered vertex joins
            // the next round's list once (its mark names that list).
            if (nd < s_dist[ln] && nd < atomicMin(&s_dist[ln], nd) &&
                atomicExch(&s_mark[ln], uint32_t(r + 2)) != uint32_t(r + 2))
              lout[atomicAdd(cout, 1)] = int16_t(ln);
          } else if (nd < __hip_atomic_load(&dist[nb[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
            // Out of the band: cached pre-filter, atomicMin decides, and the
            // target's band takes it up in the next launch.
            atomicMin(&dist[nb[k]], nd);
            nxt[nb[k]] = 1;
            out_flag = true;
          }
        }
      }
    }
    if (__any(out_flag) && lane == 0) s_out = 1;
    __syncthreads();
    n = s_cnt[(r + 1) % 3];
  }
  // The round bound was reached with vertices queued: the next launch goes on.
  if (n > 0) {
    const int16_t* lin = s_list[r & 1];
    for (int i = tid; i < n; i += kBandBlock) nxt[base + lin[i]] = 1;
    if (tid == 0) s_out = 1;
  }
  for (int i = tid; i < nv; i += kBandBlock) {
    const int32_t d = s_dist[i];
    if (d < __hip_atomic_load(&dist[base + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) atomicMin(&dist[base + i], d);
  }
  __syncthreads();
  if (tid == 0 && s_out && __hip_atomic_load(ctr_out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull)
    atomicOr(ctr_out, 1ull);
}

// Packs the view's edges as (neighbour << wb | weight) when the weights are
// non-negative and wb + the bits of a vertex id fit 32 (RMAT-22 with weights
// up to 100: 22 + 7); otherwise the relaxation reads the two arrays.
static int sssp_pack_edges(spm_graph* g) {
  SsspWork* w = g->sssp;
  const int64_t E = g->E;
  w->wb = 0;
  w->maxw = -1;
  if (!w->view.w || E == 0) return SPM_OK;
  SPM_CHECK_HIP(hipMemsetAsync(w->wstat, 0, sizeof(unsigned long long) * 2, g->stream));
  const int grid = int(std::min<int64_t>((E + 255) / 256, int64_t(g->num_cus) * 16));
  SPM_TRY(launch(g, "sssp_wrange", sssp_wrange, dim3(grid), dim3(256), 0, E, w->view.w, w->wstat));
  unsigned long long* h = reinterpret_cast<unsigned long long*>(g->h_pinned);
  SPM_CHECK_HIP(hipMemcpyAsync(h, w->wstat, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  if (h[0]) return SPM_OK;  // negative weights: unpacked
  w->maxw = int64_t(h[1]);
  int wb = 1, vb = 1;
  while ((1ull << wb) <= h[1]) ++wb;
  while ((int64_t(1) << vb) < int64_t(g->V)) ++vb;
  if (wb + vb > 32) return SPM_OK;
  if (!w->pk) SPM_TRY(dmalloc(g, (void**)&w->pk, sizeof(uint32_t) * size_t(E), "sssp packed edges"));
  SPM_TRY(launch(g, "sssp_pack", sssp_pack, dim3(grid), dim3(256), 0, E, w->view.adj, w->view.w, wb, w->pk));
  w->wb = wb;
  return SPM_OK;
}

// The view's weights changed (set_weights): repack, or drop the workspace
// (and its captured rounds) when the packing changes.
int sssp_weights_changed(spm_graph* g) {
  SsspWork* w = g->sssp;
  if (!w) return SPM_OK;
  const int old = w->wb;
  SPM_TRY(sssp_pack_edges(g));
  if (w->wb != old) return sssp_release(g);
  return SPM_OK;
}

// The band-local mode runs as the dense mode when a weight is negative (its
// launches are bounded by the rounds Bellman-Ford needs only when every
// improvement lowers a distance towards a finite minimum) and when a row is
// longer than kBandMaxDeg (one lane walks a vertex's row, so a hub would
// hold up its band's every round).
constexpr int32_t kBandMaxDeg = 64;
static int sssp_mode_eff(const spm_graph* g) {
  const int m = g->sched.sssp_mode;
  return (m == 3 && ((g->sssp && g->sssp->maxw < 0) || g->max_outdeg > kBandMaxDeg)) ? 1 : m;
}

static int sssp_bands(const spm_graph* g) { return int((int64_t(g->V) + kBandV - 1) / kBandV); }

// Tiles of the banded first frontier: vertex tiles, or row tiles of one
// source's out-row (at most V slots).
static size_t sssp_band_tiles(int32_t V) {
  return std::max((size_t(V) + kTileV - 1) / kTileV, (size_t(V) + kRowTile - 1) / kRowTile);
}

static int sssp_alloc(spm_graph* g, const CsrView& view) {
  if (g->sssp) return SPM_OK;
  SsspWork* w = new SsspWork();
  g->sssp = w;
  w->view = view;
  if (view.perm) SPM_TRY(dmalloc(g, (void**)&w->dist_out, sizeof(int32_t) * size_t(g->V), "sssp result"));
  const int32_t V = g->V;
  w->qcap = size_t(V) + 1;
  w->ccap = size_t(g->E) / kChunk + 3;
  for (int i = 0; i < 3; ++i) {
    SPM_TRY(dmalloc(g, (void**)&w->q[i].vtx, sizeof(int32_t) * w->qcap, "sssp queue"));
    SPM_TRY(dmalloc(g, (void**)&w->q[i].eoff, sizeof(uint32_t) * w->qcap, "sssp queue offsets"));
    SPM_TRY(dmalloc(g, (void**)&w->q[i].cfirst, sizeof(int32_t) * w->ccap, "sssp chunk index"));
  }
  SPM_TRY(dmalloc(g, (void**)&w->ctr, sizeof(unsigned long long) * 4, "sssp counters"));
  SPM_TRY(dmalloc(g, (void**)&w->stamp, sizeof(uint32_t) * V, "sssp stamps"));
  SPM_TRY(dmalloc(g, (void**)&w->round_base, sizeof(uint32_t), "sssp round base"));
  SPM_TRY(dmalloc(g, (void**)&w->dist, sizeof(int32_t) * V, "sssp dist"));
  SPM_TRY(dmalloc(g, (void**)&w->dprev, sizeof(int32_t) * V, "sssp queued distances"));
  SPM_TRY(dmalloc(g, (void**)&w->qmask, sizeof(uint16_t) * kSsspBlock * ((size_t(V) + kTileV - 1) / kTileV),
                  "sssp compaction masks"));
  SPM_TRY(dmalloc(g, (void**)&w->active, 2 * size_t(V), "sssp flags"));
  SPM_TRY(dmalloc(g, (void**)&w->tile_sum, sizeof(uint2) * ((size_t(V) + kTileV - 1) / kTileV), "sssp tiles"));
  SPM_TRY(dmalloc(g, (void**)&w->wstat, sizeof(unsigned long long) * 2, "sssp weight range"));
  SPM_TRY(dmalloc(g, (void**)&w->tile_b, sizeof(uint2) * kBands * sssp_band_tiles(V), "sssp band tiles"));
  SPM_TRY(dmalloc(g, (void**)&w->band_tot, sizeof(unsigned long long) * kBands, "sssp band totals"));
  SPM_TRY(dmalloc(g, (void**)&w->run_start, sizeof(uint32_t), "sssp run start"));
  SPM_TRY(dmalloc(g, (void**)&w->big, sizeof(unsigned long long) * 2, "sssp big rounds"));
  SPM_TRY(sssp_pack_edges(g));
  SPM_CHECK_HIP(hipMemsetAsync(w->stamp, 0, sizeof(uint32_t) * V, g->stream));
  SPM_CHECK_HIP(hipMemcpyAsync(w->round_base, &w->host_base, sizeof(uint32_t), hipMemcpyHostToDevice, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  return SPM_OK;
}

int sssp_release(spm_graph* g) {
  SsspWork* w = g->sssp;
  if (!w) return SPM_OK;
  for (auto& c : w->cap) {
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
  }
  for (int i = 0; i < 3; ++i) {
    dfree(g, w->q[i].vtx, sizeof(int32_t) * w->qcap);
    dfree(g, w->q[i].eoff, sizeof(uint32_t) * w->qcap);
    dfree(g, w->q[i].cfirst, sizeof(int32_t) * w->ccap);
  }
  dfree(g, w->ctr, sizeof(unsigned long long) * 4);
  dfree(g, w->stamp, sizeof(uint32_t) * g->V);
  dfree(g, w->round_base, sizeof(uint32_t));
  dfree(g, w->dist, sizeof(int32_t) * g->V);
  dfree(g, w->dprev, sizeof(int32_t) * g->V);
  dfree(g, w->qmask, sizeof(uint16_t) * kSsspBlock * ((size_t(g->V) + kTileV - 1) / kTileV));
  dfree(g, w->active, 2 * size_t(g->V));
  dfree(g, w->tile_sum, sizeof(uint2) * ((size_t(g->V) + kTileV - 1) / kTileV));
  dfree(g, w->srcbuf, sizeof(int32_t) * w->src_cap);
  dfree(g, w->dist_out, sizeof(int32_t) * g->V);
  dfree(g, w->pk, sizeof(uint32_t) * size_t(g->E));
  dfree(g, w->wstat, sizeof(unsigned long long) * 2);
  dfree(g, w->tile_b, sizeof(uint2) * kBands * sssp_band_tiles(g->V));
  dfree(g, w->band_tot, sizeof(unsigned long long) * kBands);
  dfree(g, w->run_start, sizeof(uint32_t));
  dfree(g, w->big, sizeof(unsigned long long) * 2);
  delete w;
  g->sssp = nullptr;
  return SPM_OK;
}

// Default: 6 workgroups per CU (measured at RMAT-22, ms per run: 1536
// workgroups 1.155, 1024 1.205, 1280 1.182, 1792 1.475, 2048 1.414, 2560
// 1.261 -- more resident relaxations only add contention).
static int sssp_grid(const spm_graph* g) {
  return g->sched.sssp_grid > 0 ? g->sched.sssp_grid : g->num_cus * 6;
}

// Enqueues R rounds followed by the round-base advance. R is a multiple of 3
// (queues rotate by 3), and of 6 in the dense and band-local modes (flags
// alternate by 2); a band-local "round" is one launch of sssp_band (up to
// kBandRounds rounds inside every band).
// comp[r] == 0 leaves round r without the compaction kernels: it pushes
// through the counter whatever its size (correct for any frontier, slower for
// a big one); comp == nullptr compacts every round.
static int sssp_enqueue_rounds(spm_graph* g, int R, const uint8_t* comp) {
  SsspWork* w = g->sssp;
  const int grid = sssp_grid(g);
  for (int r = 0; r < R; ++r) {
    const int i = r % 3, o = (r + 1) % 3, z = (r + 2) % 3;
    const int mode = sssp_mode_eff(g);
    if (mode == 1) {
      uint8_t* cur = w->active + size_t(r & 1) * g->V;
      uint8_t* nxt = w->active + size_t((r + 1) & 1) * g->V;
      SPM_TRY(launch(g, "sssp_dense", sssp_dense, dim3(grid), dim3(kSsspBlock), 0, g->V,
                     w->view.off, w->view.adj, w->view.w, w->dist,
                     cur, nxt, w->ctr + i, w->ctr + o, w->ctr + z, w->ctr + 3));
    } else if (mode == 3) {
      uint8_t* cur = w->active + size_t(r & 1) * g->V;
      uint8_t* nxt = w->active + size_t((r + 1) & 1) * g->V;
      SPM_TRY(launch(g, "sssp_band", w->wb ? sssp_band<true> : sssp_band<false>, dim3(sssp_bands(g)),
                     dim3(kBandBlock), 0, g->V, w->view.off, w->view.adj, w->view.w, (const uint32_t*)w->pk, w->wb,
                     w->dist, cur, nxt, int(g->sched.sssp_band_rounds), w->ctr + i, w->ctr + o, w->ctr + z,
                     w->ctr + 3));
    } else {
      SPM_TRY(launch(g, "sssp_relax", w->wb ? sssp_relax<true> : sssp_relax<false>, dim3(grid), dim3(kSsspBlock), 0,
                     w->view.off, w->view.adj, w->view.w, (const uint32_t*)w->pk, w->wb, w->dist,
                     w->stamp, w->dprev, w->q[i], w->q[o], w->ctr + i, w->ctr + o, w->ctr + z, w->ctr + 3,
                     (const uint32_t*)w->round_base, uint32_t(r),
                     comp && !comp[r] ? 0xFFFFFFFFu : kSparseEdges, (const uint32_t*)w->run_start, w->big));
      if (comp && !comp[r]) continue;
      const int tiles = int((int64_t(g->V) + kTileV - 1) / kTileV);
      SPM_TRY(launch(g, "sssp_count", sssp_count, dim3(tiles), dim3(kSsspBlock), 0, g->V, w->view.off,
                     (const int32_t*)w->dist, w->dprev, (const unsigned long long*)(w->ctr + i), w->tile_sum,
                     w->qmask));
      SPM_TRY(launch(g, "sssp_emit", sssp_emit, dim3(tiles), dim3(kSsspBlock), 0, g->V, w->view.off,
                     (const uint16_t*)w->qmask, (const unsigned long long*)(w->ctr + i), (const uint2*)w->tile_sum,
                     w->q[o], w->ctr + o));
    }
  }
  SPM_TRY(launch(g, "sssp_advance", sssp_advance, dim3(1), dim3(1), 0, w->round_base, uint32_t(R)));
  return SPM_OK;
}

// The executable graph of R rounds in slot k (compaction as comp says),
// captured on first use.
static int sssp_capture(spm_graph* g, int k, int R, const std::vector<uint8_t>& comp, hipGraphExec_t* out) {
  SsspWork::Captured& c = g->sssp->cap[k];
  if (!c.exec || c.rounds != R || c.comp != comp || c.mode != sssp_mode_eff(g)) {
    if (c.exec) { (void)hipGraphExecDestroy(c.exec); c.exec = nullptr; }
    if (c.graph) { (void)hipGraphDestroy(c.graph); c.graph = nullptr; }
    c.rounds = 0;
    SPM_CHECK_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = sssp_enqueue_rounds(g, R, comp.empty() ? nullptr : comp.data());
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g->stream, &graph);
    if (rc != SPM_OK) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return set_error(SPM_EHIP, std::string("sssp capture: ") + hipGetErrorString(e));
    c.graph = graph;
    SPM_CHECK_HIP(hipGraphInstantiate(&c.exec, graph, nullptr, nullptr, 0));
    c.rounds = R;
    c.mode = sssp_mode_eff(g);
    c.comp = comp;
  }
  *out = c.exec;
  return SPM_OK;
}

// srcs: n distinct sources in host memory (n == 1 passes the value directly).
// Distances of the last run in original ids (gathered back on first request
// after a relabeled run; the gather is ordered on g's stream).
int32_t* sssp_dist(spm_graph* g) {
  SsspWork* w = g->sssp;
  if (!w) return nullptr;
  if (!w->view.perm) return w->dist;
  if (!w->dist_out_ok) {
    if (gather_back(g, w->view, w->dist, w->dist_out) != SPM_OK) return nullptr;
    w->dist_out_ok = true;
  }
  return w->dist_out;
}

int sssp_run(spm_graph* g, const int32_t* srcs, int32_t n, int32_t* d_out, int32_t* rounds_out) {
  SPM_TRY(dyn_settle(g));
  if (!g->w) return set_error(SPM_ENOGRAPH, "SSSP needs edge weights (edgeLen or spm_graph_set_weights)");
  if (n < 1 || !srcs) return set_error(SPM_EINVAL, "SSSP needs at least one source");
  for (int32_t i = 0; i < n; ++i)
    if (srcs[i] < 0 || srcs[i] >= g->V) return set_error(SPM_EINVAL, "SSSP source out of range");
  CsrView view;
  SPM_TRY(csr_view(g, &view));
  SPM_TRY(sssp_alloc(g, view));
  SsspWork* w = g->sssp;
  w->dist_out_ok = false;
  const int R = g->sched.sssp_rounds;
  const int mode = sssp_mode_eff(g);
  const bool flags = mode == 1 || mode == 3;  // frontier as per-vertex flags, no queue history
  if (w->host_base > 0xF0000000u) {  // stamp wrap-around: start over
    w->host_base = 1;
    SPM_CHECK_HIP(hipMemsetAsync(w->stamp, 0, sizeof(uint32_t) * g->V, g->stream));
    SPM_CHECK_HIP(hipMemcpyAsync(w->round_base, &w->host_base, sizeof(uint32_t), hipMemcpyHostToDevice, g->stream));
  }
  if (n > 1) {
    if (size_t(n) > w->src_cap) {
      dfree(g, w->srcbuf, sizeof(int32_t) * w->src_cap);
      w->srcbuf = nullptr;
      w->src_cap = 0;
      SPM_TRY(dmalloc(g, (void**)&w->srcbuf, sizeof(int32_t) * size_t(n), "sssp sources"));
      w->src_cap = size_t(n);
    }
    SPM_CHECK_HIP(hipMemcpyAsync(w->srcbuf, srcs, sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice, g->stream));
  }
  const int igrid = int(std::min<int64_t>((g->V + 255) / 256, 4096));
  uint8_t* act = flags ? w->active : (uint8_t*)nullptr;
  SPM_TRY(launch(g, "sssp_init", sssp_init, dim3(igrid), dim3(256), 0, g->V, w->dist,
                 flags ? (int32_t*)nullptr : w->dprev, act));
  // Queue mode with non-negative weights: the seeds' round runs here, outside
  // the captured rounds, and queues its frontier in distance bands.
  const bool banded = mode == 0 && w->maxw >= 0;
  const int slot = banded ? 2 : 0;
  // sources are original ids; the seed kernel maps them through the
  // relabeling (view.perm) when there is one
  SPM_TRY(launch(g, "sssp_seed", sssp_seed, dim3(1), dim3(kSsspBlock), 0, n > 1 ? (const int32_t*)w->srcbuf : nullptr,
                 n, srcs[0], view.perm, w->view.off, w->dist, flags ? (int32_t*)nullptr : w->dprev, w->q[slot],
                 w->ctr, slot, banded ? w->band_tot : nullptr,
                 w->run_start, w->host_base, w->big, act));
  if (banded) {
    // One source: tiles over its out-row (at most max_outdeg <= V slots, so
    // at most sssp_band_tiles(V) tiles, which tile_b is sized for).
    const bool by_row = n == 1 && g->max_outdeg <= int64_t(g->V);
    const int tiles = by_row ? std::max(1, int((g->max_outdeg + kRowTile - 1) / kRowTile))
                             : int((int64_t(g->V) + kTileV - 1) / kTileV);
    const int32_t src1 = by_row ? srcs[0] : -1;
    const int32_t delta = int32_t(w->maxw / kBands + 1);
    SPM_TRY(launch(g, "sssp_relax", w->wb ? sssp_relax<true> : sssp_relax<false>, dim3(sssp_grid(g)),
                   dim3(kSsspBlock), 0, w->view.off, w->view.adj, w->view.w, (const uint32_t*)w->pk, w->wb, w->dist,
                   w->stamp, w->dprev, w->q[2], w->q[0], w->ctr + 2, w->ctr + 0, w->ctr + 1, w->ctr + 3,
                   (const uint32_t*)w->round_base, 0u, 0u, (const uint32_t*)w->run_start,
                   (unsigned long long*)nullptr));
    SPM_TRY(launch(g, "sssp_count_b", sssp_count_b, dim3(tiles), dim3(kSsspBlock), 0, g->V, w->view.off,
                   w->view.adj, view.perm, src1, (const int32_t*)w->dist, (const int32_t*)w->dprev, delta, tiles,
                   w->tile_b, w->band_tot));
    SPM_TRY(launch(g, "sssp_emit_b", sssp_emit_b, dim3(tiles), dim3(kSsspBlock), 0, g->V, w->view.off,
                   w->view.adj, view.perm, src1, (const int32_t*)w->dist, w->dprev, delta, tiles,
                   (const uint2*)w->tile_b, (const unsigned long long*)w->band_tot, w->q[0], w->ctr + 0));
    // The seeds' round takes no round stamp (it neither pushes through the
    // stamps nor enters the big-round log), so the captured rounds start at
    // the present base.
  }
  // Rounds after the frontier empties still cost their launches (~13 us a
  // round at RMAT-22), so in queue mode the first replay is as long as the
  // previous run on this handle needed plus one round of slack (the count
  // varies by a round between runs of the same source), rounded up to a
  // multiple of 3, the queue rotation. A longer run simply continues with
  // replays of R rounds.
  // The same history drops the compaction kernels from the first replay's
  // rounds whose frontier stayed below a quarter of the compaction threshold
  // (a small round pushes through the counter and skips them anyway; their
  // launches cost ~9 us a round).
  int R1 = R;
  std::vector<uint8_t> comp1;
  if (!flags && w->last_rounds > 0) {
    R1 = std::min(R, 3 * ((w->last_rounds + 3) / 3));
    comp1.resize(size_t(R1));
    for (int r = 0; r < R1; ++r) comp1[r] = r >= 128 || ((w->big_hist[r >> 6] >> (r & 63)) & 1ull);
  }
  const std::vector<uint8_t> all;
  const bool use_graph = !g->prof.on;
  unsigned long long* h = reinterpret_cast<unsigned long long*>(g->h_pinned);
  const int64_t max_replays = int64_t(g->V) / R + 3;
  int64_t replays = 0;
  for (;;) {
    const bool first = replays == 0 && !comp1.empty();
    const int Rk = first ? R1 : R;
    const std::vector<uint8_t>& comp = first ? comp1 : all;
    if (use_graph) {
      hipGraphExec_t exec = nullptr;
      SPM_TRY(sssp_capture(g, first ? 1 : 0, Rk, comp, &exec));
      SPM_CHECK_HIP(hipGraphLaunch(exec, g->stream));
    } else {
      SPM_TRY(sssp_enqueue_rounds(g, Rk, comp.empty() ? nullptr : comp.data()));
    }
    w->host_base += uint32_t(Rk);
    SPM_CHECK_HIP(hipMemcpyAsync(h, w->ctr, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, g->stream));
    SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
    if (h[0] == 0ull) break;
    if (++replays > max_replays) return set_error(SPM_EDIVERGE, "SSSP did not converge within V rounds (negative weights?)");
  }
  SPM_CHECK_HIP(hipMemcpyAsync(h, w->ctr + 3, sizeof(unsigned long long), hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipMemcpyAsync(h + 1, w->big, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, g->stream));
  if (d_out) SPM_TRY(gather_back(g, w->view, w->dist, d_out));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  if (rounds_out) *rounds_out = int32_t(h[0]);
  if (!flags) {
    w->last_rounds = int(std::max<int64_t>(1, int64_t(h[0]) - (banded ? 1 : 0)));
    w->big_hist[0] = h[1];
    w->big_hist[1] = h[2];
  }
  SPM_TRY(prof_flush(g));
  return SPM_OK;
}

}  // namespace spm

using namespace spm;

extern "C" {

int spm_sssp_device(spm_graph* g, int32_t src, int32_t* d_dist, int32_t* rounds) {
  if (!g || !d_dist) return set_error(SPM_EINVAL, "spm_sssp_device: NULL argument");
  return sssp_run(g, &src, 1, d_dist, rounds);
}

int spm_sssp_multi_device(spm_graph* g, const int32_t* srcs, int32_t n, int32_t* d_dist, int32_t* rounds) {
  if (!g || !d_dist || !srcs || n < 1) return set_error(SPM_EINVAL, "spm_sssp_multi_device: bad argument");
  std::vector<int32_t> s(srcs, srcs + n);
  std::sort(s.begin(), s.end());
  s.erase(std::unique(s.begin(), s.end()), s.end());
  return sssp_run(g, s.data(), int32_t(s.size()), d_dist, rounds);
}

int spm_compute_sssp(spm_graph* g, int32_t* dist, const int32_t* weight, int32_t src) {
  if (!g || !dist) return set_error(SPM_EINVAL, "spm_compute_sssp: NULL argument");
  SPM_TRY(dyn_settle(g));
  if (weight) SPM_TRY(spm_graph_set_weights(g, weight));
  SPM_TRY(sssp_run(g, &src, 1, nullptr, nullptr));
  const int32_t* res = sssp_dist(g);
  if (!res) return set_error(SPM_EHIP, "spm_compute_sssp: result gather failed");
  SPM_CHECK_HIP(hipMemcpyAsync(dist, res, sizeof(int32_t) * g->V, hipMemcpyDeviceToHost, g->stream));
  SPM_CHECK_HIP(hipStreamSynchronize(g->stream));
  return SPM_OK;
}

}  // extern "C"
