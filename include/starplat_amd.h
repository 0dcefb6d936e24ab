/*
 * starplat_amd.h -- C ABI of the MI355X (gfx950) StarPlat kernel library.
 *
 * This is the boundary under StarPlat's generated entry points. The OMP
 * backend emits (src/backends/backend_omp/dsl_cpp_generator.cpp:3056-3151,
 * type mapping :2875-3054):
 *
 *   void Compute_PR  (graph& g, float beta, float delta, int maxIter, float* pageRank);
 *   void Compute_SSSP(graph& g, int* dist, int* weight, int src);
 *   auto Compute_TC  (graph& g);                  // returns long
 *   void Compute_BC  (graph& g, float* BC, std::set<int>& sourceSet);
 *
 * over graphcode/graph.hpp's CSR (graph.hpp:44-57). Code generated for this
 * backend keeps those C++ signatures and calls the functions below; nothing
 * here takes C++ or torch types. All pointers are plain host pointers unless
 * the name says d_ (device memory). Every function returns SPM_OK (0) or a
 * negative status; spm_last_error() gives the message of the last failure on
 * the calling thread.
 *
 * Threading: a handle owns one HIP stream; calls on the same handle must not
 * overlap. Different handles are independent.
 */
#ifndef STARPLAT_AMD_H
#define STARPLAT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPM_OK 0
#define SPM_EINVAL (-1)      /* bad argument                                   */
#define SPM_EHIP (-2)        /* HIP runtime error                              */
#define SPM_ENOMEM (-3)      /* device allocation failed                       */
#define SPM_ENOGRAPH (-4)    /* algorithm needs an array the handle lacks       */
#define SPM_EDIVERGE (-5)    /* fixed point did not converge within V rounds    */

typedef struct spm_graph spm_graph;

/* Message of the last failed call on this thread ("" if none). */
const char* spm_last_error(void);
/* Library ABI version (major*100 + minor). */
int spm_version(void);

/* ---- device CSR mirror ------------------------------------------------
 * Replaces: graph::parseGraph()'s arrays as consumed by generated code,
 * graph.hpp:44-57 (indexofNodes, edgeList, rev_indexofNodes, srcList) and
 * getEdgeLen() :134-137; the CUDA backend's per-call cudaMalloc/cudaMemcpy
 * of the same arrays (backend_cuda/dsl_cpp_generator.cpp:4440-4500).
 * Arrays are copied to HBM once; rev_* may be NULL (PR and BC then fail with
 * SPM_ENOGRAPH); edgeLen may be NULL (SSSP then needs spm_graph_set_weights
 * or a weight argument). Offsets must be non-decreasing, every row ascending
 * (as graph.hpp's parseGraph leaves them, :722-734, :810-832) and every
 * neighbour id in [0, num_nodes), else SPM_EINVAL (checked on the host before
 * any device work); a reverse CSR that is not the transpose of the forward
 * one is refused the same way (an edge-multiset hash of both, compared on the
 * device after the upload). */
int spm_graph_create(int32_t num_nodes, int64_t num_edges,
                     const int32_t* indexofNodes, const int32_t* edgeList,
                     const int32_t* edgeLen,
                     const int32_t* rev_indexofNodes, const int32_t* srcList,
                     spm_graph** out);
int spm_graph_destroy(spm_graph* g);
/* Replaces: the propEdge<int> weight argument of Compute_SSSP
 * (graphcode/main_functions/sssp_dslv2main.cpp:33-41 builds it per CSR
 * position). Uploads E weights. */
int spm_graph_set_weights(spm_graph* g, const int32_t* weight);
/* The hipStream_t (as void*) all work on g is ordered on. */
void* spm_graph_stream(spm_graph* g);
int spm_graph_sync(spm_graph* g);
/* Device bytes currently held by the handle (CSR plus workspaces). */
int64_t spm_graph_device_bytes(spm_graph* g);

/* ---- PageRank ----------------------------------------------------------
 * Replaces: Compute_PR(graph& g, float beta, float delta, int maxIter,
 * float* pageRank) generated from graphcode/staticDSLCodes/PageRankDSLV2:1-19
 * with PageRankDSLV3:15-19 diff semantics. pageRank: host float[V].
 * iterCount (may be NULL) receives the number of iterations run. */
int spm_compute_pr(spm_graph* g, float beta, float delta, int32_t maxIter,
                   float* pageRank, int32_t* iterCount);
/* Same, result left in device memory d_pageRank (float[V]). */
int spm_pr_device(spm_graph* g, float beta, float delta, int32_t maxIter,
                  float* d_pageRank, int32_t* iterCount);
/* Same as spm_pr_device, but allocates nothing: no hipMalloc, hipFree or
 * allocating library call (graph instantiation included) on any call. No
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid it). The workspace is one file-scope device array of ~40 MiB that
 * each GPU the library runs on holds until the process ends, so graphs may
 * have at most 2^20 vertices (else SPM_EINVAL); the handle must have its
 * reverse CSR and no pending edge updates (SPM_EINVAL: fold them with
 * spm_graph_compact first). Runs the pull sweep on the original ids with the
 * handle's lane-mapping thresholds, in the same summation order as
 * spm_pr_device on graphs it does not relabel (with the default schedule:
 * below 2^17 vertices, or degrees without a dominant hub; bit-identical
 * ranks and iteration count); push and
 * source segmentation are not used. Every call on one GPU shares that
 * array: calls that overlap on one GPU (from any handles or threads) race. */
int spm_pr_device_noalloc(spm_graph* g, float beta, float delta, int32_t maxIter,
                          float* d_pageRank, int32_t* iterCount);
/* Many PageRank jobs in one launch sequence: g is the disjoint union of the
 * jobs' graphs, job k owning vertices [bases[k], bases[k+1]) (bases[0] = 0,
 * bases[n] = V, no empty job, no edge between jobs: checked, SPM_EINVAL).
 * Each job runs Compute_PR on its own graph -- its own N_k in the 1/N_k
 * start and (1-delta)/N_k term, its own diff and stopping test -- and stops
 * independently; d_pageRank (device float[V]) receives every job's ranks and
 * iterCounts (host int32[n], may be NULL) each job's iteration count. Runs on
 * the original ids (no relabeled copy). No counterpart in the reference,
 * which runs one Compute_PR per call. */
int spm_pr_multi_device(spm_graph* g, const int32_t* bases, int32_t n, float beta, float delta,
                        int32_t maxIter, float* d_pageRank, int32_t* iterCounts);

/* ---- SSSP --------------------------------------------------------------
 * Replaces: Compute_SSSP(graph& g, int* dist, int* weight, int src) generated
 * from graphcode/staticDSLCodes/sssp_dslV2:1-15 (fixedPoint Bellman-Ford).
 * dist: host int32[V] (INT_MAX = unreachable); weight: host int32[E] by CSR
 * position, or NULL to use the handle's weights. */
int spm_compute_sssp(spm_graph* g, int32_t* dist, const int32_t* weight, int32_t src);
/* Same, handle weights, result in device memory d_dist (int32[V]). rounds
 * (may be NULL) receives the number of non-empty relaxation rounds. */
int spm_sssp_device(spm_graph* g, int32_t src, int32_t* d_dist, int32_t* rounds);
/* Same fixed point seeded from n sources at once (dist 0 at each; duplicates
 * ignored). On a disjoint union of graphs with one source per component this
 * runs many SSSP jobs in one launch sequence with per-job results unchanged.
 * No counterpart in the reference, which runs one Compute_SSSP per call. */
int spm_sssp_multi_device(spm_graph* g, const int32_t* srcs, int32_t n, int32_t* d_dist, int32_t* rounds);

/* ---- Triangle counting --------------------------------------------------
 * Replaces: auto Compute_TC(graph& g) (returns long) generated from
 * graphcode/staticDSLCodes/triangle_counting_dsl:1-13; is_an_edge ->
 * graph::check_if_nbr (graph.hpp:233-287). Requires a symmetric graph with
 * ascending rows (graph.hpp:722-734 sorts them). */
int spm_compute_tc(spm_graph* g, uint64_t* triangle_count);

/* ---- Betweenness centrality ---------------------------------------------
 * Replaces: Compute_BC(graph& g, float* BC, std::set<int>& sourceSet)
 * generated from graphcode/staticDSLCodes/bc_dslV2:1-21. sources: the set's
 * elements (any order, duplicates rejected); BC: host float[V]. Path counts
 * are FP64, dependencies and BC FP32. */
int spm_compute_bc(spm_graph* g, float* BC, const int32_t* sources, int32_t nsrc);
/* Same, result in device memory d_BC (float[V]). */
int spm_bc_device(spm_graph* g, const int32_t* sources, int32_t nsrc, float* d_BC);
/* Forward results of one source, as bc_dslV2's iterateInBFS leaves them:
 * bfsDist (int32[V], -1 unreached) and sigma (FP64[V] shortest-path counts).
 * Replaces reading the generated code's bfsDist/sigma locals
 * (backend_omp/dsl_cpp_generator.cpp:66-130); host buffers. */
int spm_bc_paths(spm_graph* g, int32_t src, int32_t* bfsDist, double* sigma);

/* ---- edge updates (dynamic graphs) -----------------------------------------
 * Replace graph::updateCSRDel and graph::updateCSRAdd (graphcode/graph.hpp:
 * 869-975 over graph_ompv2.hpp:13-454), which generated Dynamic functions call
 * once per batch (backend_omp/dsl_dyn_cpp_generator.cpp:301-316). Edge
 * (src[i], dst[i]); vertices must lie in 0..num_nodes()-1 (no new vertices).
 * A deletion removes every copy of the edge and ignores absent ones; an
 * addition appends an edge even if one exists (weight[i], or 1 when weight is
 * NULL, as update.hpp:56-60). Updates are kept beside the CSR (removed-slot
 * bits, per-vertex overflow rows, as getNeighbors' diff_edgeList,
 * graph.hpp:982-1062); the next static algorithm, spm_graph_compact or
 * spm_graph_export folds them into a clean sorted CSR first. */
int spm_graph_delete_edges(spm_graph* g, const int32_t* src, const int32_t* dst, int32_t n);
int spm_graph_add_edges(spm_graph* g, const int32_t* src, const int32_t* dst, const int32_t* weight, int32_t n);
int spm_graph_compact(spm_graph* g);
/* Edges of the current graph (num_edges() after the updates, graph.hpp:145-153). */
int spm_graph_num_edges(spm_graph* g, int64_t* num_edges);
/* Copies the current (folded) arrays to host buffers sized num_nodes()+1 and
 * num_edges; any pointer may be NULL. Rows ascending, copies of an edge kept. */
int spm_graph_export(spm_graph* g, int32_t* indexofNodes, int32_t* edgeList, int32_t* edgeLen,
                     int32_t* rev_indexofNodes, int32_t* srcList);

/* ---- dynamic SSSP ------------------------------------------------------------
 * Replaces the Dynamic function DynSSSP(Graph g, propNode<int> dist,
 * propNode<int> parent, propEdge<int> weight, updates<g> updateBatch,
 * int batchSize, int src) of graphcode/dynamicDSLCodes/dynamicBatchSSSP:67-94
 * with its staticSSSP (:1-18), Incremental (:20-35) and Decremental (:37-65).
 * Per batch: OnDelete (a deleted parent edge cuts its head off), updateCSRDel,
 * Decremental (cut-off subtrees, then a pull fixed point from in-neighbours),
 * OnAdd (an added edge that shortens its head marks both ends), updateCSRAdd
 * (added edges weigh 1), Incremental (push fixed point from the marks).
 * Needs weights >= 1 and the reverse CSR. dist: INT_MAX never reached,
 * INT_MAX/2 cut off and not reached again (the reference's markers); parent:
 * -1 for the source and unreached vertices, else the smallest-id predecessor
 * on a shortest path among those the run relaxed through. */
int spm_dynsssp_init(spm_graph* g, int32_t src);
/* One batch of n updates: type[i] is 'a' (add) or 'd' (delete). */
int spm_dynsssp_batch(spm_graph* g, const char* type, const int32_t* src, const int32_t* dst, int32_t n);
/* Current dist / parent (host buffers of num_nodes(); either may be NULL). */
int spm_dynsssp_result(spm_graph* g, int32_t* dist, int32_t* parent);
int spm_dynsssp_device(spm_graph* g, int32_t* d_dist, int32_t* d_parent);
/* The whole Dynamic function: init from src, then the updates in batches of
 * batchSize (the generated loop, dsl_dyn_cpp_generator.cpp:132-150). */
int spm_dyn_sssp(spm_graph* g, int32_t src, const char* type, const int32_t* usrc, const int32_t* udst, int32_t n,
                 int32_t batchSize, int32_t* dist, int32_t* parent);

/* ---- dynamic triangle counting -------------------------------------------
 * Replaces: the code generated for graphcode/dynamicDSLCodes/dynamicBatchTCV2
 * -- DynTC (:137-187): triangleCount = staticTC(g) (:1-21); per batch
 * OnDelete marks the deleted edges, Decremental (:81-135) counts on the graph
 * before the deletions, updateCSRDel, updateCSRAdd, OnAdd marks the added
 * edges, Incremental (:23-78) counts on the updated graph; each phase moves
 * the count by class1/2 + class2/4 + class3/6 as the DSL writes it. The
 * count is kept in 64 bits (the DSL's int would wrap past 2^31); DynTC
 * returns nothing in the DSL, spm_dyntc_count reads it. Updates change the
 * handle's graph as spm_graph_delete_edges / spm_graph_add_edges do. */
int spm_dyntc_init(spm_graph* g);
/* One batch of n updates: type[i] is 'a' (add) or 'd' (delete). */
int spm_dyntc_batch(spm_graph* g, const char* type, const int32_t* src, const int32_t* dst, int32_t n);
int spm_dyntc_count(spm_graph* g, int64_t* triangleCount);
/* The whole Dynamic function, updates in batches of batchSize. */
int spm_dyn_tc(spm_graph* g, const char* type, const int32_t* usrc, const int32_t* udst, int32_t n,
               int32_t batchSize, int64_t* triangleCount);

/* ---- dynamic PageRank ------------------------------------------------------
 * Replaces: the code generated for graphcode/dynamicDSLCodes/dynamicBatchPR
 * -- DynPR (:97-135): staticPR (:1-31), then per batch OnDelete flags the
 * deletions' destinations, propagateNodeFlags (graph.hpp:444-462) flags every
 * vertex with a live in-edge, updateCSRDel, Decremental (:65-95) sweeps the
 * flagged vertices; the same with the additions (graph after the deletions),
 * updateCSRAdd, Incremental (:32-63). As the dynamic OMP generator emits it:
 * count_outNbrs is the out-degree in the CSR the run starts from
 * (dsl_dyn_cpp_generator.cpp:279-288; the reference's updates never move
 * indexofNodes), nodes_to the live in-edges, the diff signed (val - pr).
 * Needs the reverse CSR. Ranks float[V]. */
int spm_dynpr_init(spm_graph* g, float beta, float delta, int32_t maxIter);
int spm_dynpr_batch(spm_graph* g, const char* type, const int32_t* src, const int32_t* dst, int32_t n);
int spm_dynpr_result(spm_graph* g, float* pageRank);
/* Sweeps run so far: staticPR's, then Decremental's and Incremental's per
 * batch; *count receives how many there are, at most cap are copied. */
int spm_dynpr_iterations(spm_graph* g, int32_t* iterCounts, int32_t cap, int32_t* count);
/* The whole Dynamic function, updates in batches of batchSize. */
int spm_dyn_pr(spm_graph* g, float beta, float delta, int32_t maxIter, float* pageRank, const char* type,
               const int32_t* usrc, const int32_t* udst, int32_t n, int32_t batchSize);

/* ---- schedules -----------------------------------------------------------
 * The schedule changes how work maps to lanes, never the arithmetic. No
 * counterpart in the reference (its CUDA backend hard-codes 512 threads per
 * vertex grid, backend_cuda/dsl_cpp_generator.cpp:56-77). */
typedef struct spm_schedule {
  int32_t pr_t1;          /* in-degree <= t1: one lane per vertex            */
  int32_t pr_t2;          /* <= t2: pr_group lanes per vertex                */
  int32_t pr_t3;          /* <= t3: one wavefront per vertex; above: block   */
  int32_t pr_group;       /* lanes per vertex in the second bin (2..32)      */
  int32_t sssp_mode;      /* 0 edge-balanced frontier queue with the first
                             frontier in distance bands, 1 dense sweep,
                             2 queue in vertex order throughout, 3 band-
                             local rounds in LDS (high-diameter graphs with
                             local ids; as 1 when a weight is negative or
                             a row is longer than 64)                      */
  int32_t sssp_rounds;    /* rounds per captured graph (multiple of 6)       */
  int32_t sssp_grid;      /* workgroups per relaxation launch (0 = auto)     */
  int32_t tc_mode;        /* 1 per-vertex LDS hash (default), 0 edge search */
  int32_t pr_mode;        /* 0 pull over in-rows, 1 push (scatter + apply),
                             2 pull whose sweeps evaluate the loop test of
                             the sweep before (one launch per iteration;
                             sweeps of <= 2048 workgroups, else as 0),
                             3 pull by persistent 1024-lane workgroups that
                             read the contributions of the lowest ids (the
                             highest out-degrees when relabeled) from LDS */
  int32_t pr_seg_log2;    /* pull: sweep sources in segments of 2^k vertices
                             so gathers stay in L2 (0 = one sweep; 10..30)   */
  int32_t relabel;        /* PR and SSSP on a copy of the CSR with vertices
                             renumbered by descending out-degree (built once
                             per handle; results in original ids); 0 or 1   */
  int32_t sssp_band_rounds; /* sssp_mode 3: rounds a 4096-vertex band runs in
                             LDS per launch before handing on (1..4096)      */
  int32_t reserved[4];
} spm_schedule;
int spm_get_schedule(spm_graph* g, spm_schedule* s);
int spm_set_schedule(spm_graph* g, const spm_schedule* s);
/* The fixed heuristic schedule for g (fallback and judge of the learned one). */
int spm_default_schedule(spm_graph* g, spm_schedule* s);

/* ---- learned schedule selector --------------------------------------------
 * No counterpart in the reference: a 16 -> 64 -> 64 -> 8 perceptron mapping
 * graph features to a schedule class, trained and evaluated on the GPU by
 * this library's own kernels. FP32 master parameters, BF16 weight copies
 * read by forward and backward, FP32 accumulation. Parameter order:
 * W1[64][16], b1[64], W2[64][64], b2[64], W3[8][64], b3[8]. */
typedef struct spm_mlp spm_mlp;
int spm_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint64_t seed, spm_mlp** out);
int spm_mlp_destroy(spm_mlp* m);
int spm_mlp_num_params(void);
int spm_mlp_set_params(spm_mlp* m, const float* params);
int spm_mlp_get_params(spm_mlp* m, float* params);
/* logits: float[n][8] for x: float[n][16] (host buffers). */
int spm_mlp_forward(spm_mlp* m, const float* x, int32_t n, float* logits);
int spm_mlp_predict(spm_mlp* m, const float* x, int32_t n, int32_t* cls);
/* Mean softmax cross-entropy over the n samples and its gradient w.r.t. the
 * FP32 masters (evaluated with the BF16 weights). */
int spm_mlp_grad(spm_mlp* m, const float* x, const int32_t* y, int32_t n, float* loss, float* grads);
/* Adam, seeded shuffles, `epochs` passes of mini-batches of `batch`. */
int spm_mlp_train(spm_mlp* m, const float* x, const int32_t* y, int32_t n, int32_t epochs, int32_t batch,
                  float lr, uint64_t seed, float* final_loss);

/* ---- profiling -----------------------------------------------------------
 * With profiling on, each kernel launch carries start/stop events that the
 * dispatch itself stamps (hipExtLaunchKernel) on the handle's stream and
 * graphs are not used, so per-kernel device time can be read back. Off by
 * default. */
int spm_profile_enable(spm_graph* g, int enable);
int spm_profile_reset(spm_graph* g);
/* total_ms and launches of the named kernel since the last reset. */
int spm_profile_query(spm_graph* g, const char* kernel, double* total_ms, int64_t* launches);
/* No counterpart in the reference: the rate of random 4-byte loads the GPU
 * sustains (loads per second into *rate), the ceiling of gather-bound
 * kernels. 2^28 loads at uniformly random indices into an array of n floats
 * (n >= 8), eight independent loads in flight per lane; with xcd_local the
 * indices of workgroup b fall in the eighth (b mod 8) of the array, which
 * keeps each eighth in one XCD's L2 (workgroups are dispatched to the 8 XCDs
 * round-robin). */
int spm_gather_rate(int32_t n, int32_t xcd_local, double* rate);

#ifdef __cplusplus
}
#endif
#endif
