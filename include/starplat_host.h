/*
 * starplat_host.h -- host-side graph data layer of the MI355X StarPlat backend.
 *
 * Builds the CSR that StarPlat's `class graph` holds after parseGraph()
 * (graphcode/graph.hpp:44-57, filled by parseEdges :464-515 and
 * parseEdgesContent :720-859) without graph.hpp's std::map edge table:
 *
 *   indexofNodes      int32[V+1]  out-edge offsets            (graph.hpp:44)
 *   edgeList          int32[E]    destinations, ascending per row (:46,:722-734)
 *   edgeLen           int32[E]    weights by CSR position      (:37, getEdgeLen :134-137)
 *   rev_indexofNodes  int32[V+1]  in-edge offsets              (:45,:786-807)
 *   srcList           int32[E]    sources, ascending per row   (:48,:810-832)
 *
 * V follows graph::num_nodes() = max vertex id + 1 (graph.hpp:139-142), and
 * duplicate (src,dst) pairs are merged by summing their weights (graph.hpp:497).
 *
 * Seeded generators (RMAT, Erdos-Renyi, 2-D grid) are counter-based, so a
 * graph depends only on its parameters and seed, never on the thread count.
 * All functions are plain C; no GPU is touched.
 */
#ifndef STARPLAT_HOST_H
#define STARPLAT_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spmh_graph {
  int32_t V;               /* num_nodes(): max id + 1                      */
  int64_t E;               /* num_edges(): directed CSR entries            */
  int32_t* indexofNodes;   /* [V+1]                                        */
  int32_t* edgeList;       /* [E]                                          */
  int32_t* edgeLen;        /* [E] weights (0 when the input had none)      */
  int32_t* rev_indexofNodes; /* [V+1]                                      */
  int32_t* srcList;        /* [E]                                          */
} spmh_graph;

/* RMAT: edge_factor * 2^scale samples, quadrant probabilities (a,b,c,1-a-b-c),
 * no noise, no relabelling (vertex 0 is the hub). Self-loops dropped,
 * duplicates removed; undirected != 0 symmetrises before deduplication.
 * wmax > 0 assigns int32 weights uniform in [wmin,wmax] per surviving edge
 * (undirected graphs get the same weight in both directions). */
spmh_graph* spmh_rmat(int scale, int edge_factor, double a, double b, double c,
                      uint64_t seed, int undirected, int wmin, int wmax);

/* Erdos-Renyi G(n, m): m uniform endpoint pairs over [0,n), self-loops
 * dropped, deduplicated, optionally symmetrised. */
spmh_graph* spmh_er(int64_t n, int64_t m, uint64_t seed, int undirected,
                    int wmin, int wmax);

/* rows x cols 4-neighbour grid, both directions stored. */
spmh_graph* spmh_grid(int rows, int cols, uint64_t seed, int wmin, int wmax);

/* CSR from an explicit edge list with graph.hpp semantics: self-loops kept,
 * duplicates merged by summing weights, V = max id + 1. w may be NULL (0). */
spmh_graph* spmh_from_edges(int64_t n, const int32_t* src, const int32_t* dst,
                            const int32_t* w);

/* graph::parseGraph() on a text file (graph.hpp:464-515): lines whose first
 * character is not a digit are skipped, "src dst [w]" per line, a missing
 * weight reads as 0. Returns NULL when the file cannot be opened. */
spmh_graph* spmh_parse_text(const char* path);

/* Disjoint union of n graphs: graph i's vertex v becomes base_i + v with
 * base_i = sum of the V of graphs before it (written to bases[n+1] when not
 * NULL). Rows stay ascending, weights are kept. */
spmh_graph* spmh_union(const spmh_graph* const* gs, int32_t n, int64_t* bases);

/* Writes "src dst w" lines (or "src dst" when weights == 0). 0 on success. */
int spmh_write_text(const spmh_graph* g, const char* path, int weights);

/* Binary CSR file: a 64-byte header (magic "SPMCSR01", V, E, flags,
 * checksum) followed by indexofNodes, edgeList, edgeLen, rev_indexofNodes,
 * srcList as little-endian int32. Replaces re-parsing graph.hpp text
 * (graph.hpp:464-515, a std::map of every pair) for large graphs. Returns 0 on
 * success; the reader returns NULL on a missing, short or corrupt file. */
int spmh_write_binary(const spmh_graph* g, const char* path);
spmh_graph* spmh_read_binary(const char* path);

void spmh_free(spmh_graph* g);

/* n distinct vertices drawn uniformly from [0,V) with the given seed, written
 * ascending (std::set<int> iteration order, as Compute_BC's sourceSet). */
int spmh_sources(int32_t V, int32_t n, uint64_t seed, int32_t* out);

/* graph.hpp's edgeMap (:57, built at :835-846): for every reverse-CSR slot p,
 * the CSR position of the same edge, written to edgeMap[E]. Rows must be
 * ascending; the k-th copy of an edge in the reverse row maps to its k-th
 * copy in the CSR row. Returns 0, or -1 if the two CSRs disagree. */
int spmh_edge_map(int32_t V, const int32_t* indexofNodes, const int32_t* edgeList,
                  const int32_t* rev_indexofNodes, const int32_t* srcList, int32_t* edgeMap);

/* The 16 graph features of the schedule selector, written to out[16]:
 *  0 log2 V            1 log2(E+1)          2 mean out-degree
 *  3 log2(1+var deg)   4 log2(1+max outdeg) 5 log2(1+max indeg)
 *  6 signed log(1+|skew of outdeg|)         7 fraction of outdeg-0 vertices
 *  8 fraction of indeg<=4 vertices          9 share of in-edges on indeg>64 vertices
 * 10 log2(1+diameter estimate; double-sweep BFS)  11 fraction reached by the sweep
 * 12 log2(1+BFS levels per 1000 reached)          13 widest BFS level / V
 * 14 job (0 SSSP, 1 PageRank; the caller's value) 15 reciprocity (fraction of
 *    edges whose reverse exists) */
int spmh_features(const spmh_graph* g, float job, float* out);

/* OpenMP threads used by this library's parallel regions started from the
 * calling thread (n > 0; the other threads keep their own setting). A
 * thread that builds graphs in the background while another thread times
 * GPU work leaves cores to it this way. */
void spmh_set_threads(int n);

/* Algorithmic-byte models used for roofline accounting. */
/* PR: bytes of one pull iteration = 8E + 20V + 4. */
double spmh_model_pr_bytes(const spmh_graph* g);
/* SSSP: replays the synchronous-frontier Bellman-Ford trace from src and
 * returns sum_k (12|F_k| + 12 M_k + 4|F_k+1|); also the rounds, the
 * reachable out-edge count m_reach and the relaxation count. */
double spmh_model_sssp(const spmh_graph* g, const int32_t* weight, int32_t src,
                       int64_t* rounds, int64_t* m_reach, int64_t* relaxations);
/* TC: 8(V+1) + 4 m_und + 4 sum_{oriented (u,v)} (d+(u) + d+(v)) where the
 * orientation is by (degree, id), the one the GPU kernel counts with. */
double spmh_model_tc_bytes(const spmh_graph* g, int64_t* m_und, int64_t* wedge_work);
/* BC: per source 32 V_r + 16 E_r + 20 D summed over the sources. */
double spmh_model_bc_bytes(const spmh_graph* g, const int32_t* sources, int32_t n,
                           int64_t* edges_traversed);

/* The access pattern of one 64-source BC batch (sources[0..n), n <= 64),
 * from the sources' BFS levels: out[0] DAG edges summed over the sources;
 * out[1] row gathers -- (in-edge (v, w), level d) pairs with a non-empty set
 * S of sources for which w is at level d and v at d - 1, i.e. the per-source
 * rows the sigma sweep fetches (the reverse sweep fetches the same set);
 * out[2] / out[3] 128-byte lines those gathers touch with 8-byte / 4-byte
 * per-source entries (64-entry rows); out[4] (vertex, level) listings;
 * out[5] in-edges read by the sigma sweep (each listing reads the vertex's
 * whole in-row); out[6] levels. Mean sources per gather = out[0] / out[1].
 * Returns 0, or -1 on bad arguments. */
int spmh_bc_gather_profile(const spmh_graph* g, const int32_t* sources, int32_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif
