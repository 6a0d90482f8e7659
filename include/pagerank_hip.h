/*
 * pagerank_hip.h -- C ABI of the MI355X (gfx950) PageRank engine: libpagerank_hip.so.
 *
 * This is the drop-in boundary for the GPU PageRank path of puzzlef/pagerank-cuda-dynamic
 * (inc/pagerankCuda.hxx). The reference's entry points are C++ templates that marshal a host DiGraph into
 * uint32 CSR arrays, allocate device memory, copy, run, copy back and free, all in one call
 * (pagerankInvokeCuda, pagerankCuda.hxx:713-849). Here the same work is split at the host->device boundary:
 *   - prh_graph_create uploads the CSR arrays the reference builds at pagerankCuda.hxx:756-775 and copies at
 *     :798-805 (plain pointers and sizes, no C++ types);
 *   - prh_solve runs one of the five solvers with the reference's options and returns its result fields.
 * The C++ header pagerankHip.hxx (package host/ directory) rebuilds the reference's template signatures on top.
 *
 * Layout (the reference's, pagerankCuda.hxx:715-761): vertices are dense ids 0..N-1; offsets are uint32
 * (so M < 2^32); the graph handed to a solver is the updated graph y, given as its transpose's CSR
 * (in_offsets[N+1], in_sources[M], in-neighbours ascending), out_degrees[N] (the transpose's vertex value,
 * transpose.hxx:109-134) and, for the dynamic solvers, y's own CSR (out_offsets[N+1], out_targets[M]).
 * Vertices with out-degree 0 (dead ends) are accepted by the Static, Naive-dynamic, Dynamic Traversal and Dynamic
 * Frontier solvers and the batched solver, as in the reference (a dead end is never an in-neighbour, so nothing divides
 * by its degree); DF-P (whose closed-form update divides by the out-degree) and the weighted solves reject a graph
 * with dead ends (PRH_ERROR_INVALID_GRAPH). The reference's experiment adds a self-loop per vertex (main.cu:168).
 *
 * Errors: every call returns a prh_status; nothing aborts the process (the reference exits on any CUDA
 * error, TRY_CUDA, inc/_cuda.hxx:136-150). prh_context_last_error() describes the last failure.
 * Threading: a context owns one HIP stream; calls on one context are serialised by the caller, separate
 * contexts may be used from separate threads.
 */
#ifndef PAGERANK_HIP_H
#define PAGERANK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum prh_status {
  PRH_SUCCESS = 0,
  PRH_ERROR_INVALID_ARGUMENT = 1, /* null pointer, bad method, bad option, id out of range, NaN rank */
  PRH_ERROR_INVALID_GRAPH = 2,    /* offsets not 0..M or not monotone, id >= N; dead ends for DF-P / weighted */
  PRH_ERROR_OUT_OF_MEMORY = 3,    /* device allocation failed or exceeded the context's memory limit */
  PRH_ERROR_DEVICE = 4,           /* HIP runtime error */
  PRH_ERROR_UNSUPPORTED = 5       /* e.g. a dynamic solver on a graph uploaded without its out-CSR */
} prh_status;

/* The five solvers of pagerankCuda.hxx. */
typedef enum prh_method {
  PRH_STATIC = 0,                 /* pagerankStaticCuda, pagerankCuda.hxx:864-875 */
  PRH_NAIVE_DYNAMIC = 1,          /* pagerankNaiveDynamicCuda, pagerankCuda.hxx:891-901 */
  PRH_DYNAMIC_TRAVERSAL = 2,      /* pagerankDynamicTraversalCuda, pagerankCuda.hxx:921-930 */
  PRH_DYNAMIC_FRONTIER = 3,       /* pagerankDynamicFrontierCuda, pagerankCuda.hxx:950-959 */
  PRH_PRUNE_DYNAMIC_FRONTIER = 4  /* pagerankPruneDynamicFrontierCuda, pagerankCuda.hxx:979-988 */
} prh_method;

/* PagerankOptions<double> (pagerank.hxx:24-55); prh_default_options gives the reference's defaults
 * (repeat 1, tolerance 1e-10, frontier and prune tolerance 1e-6, damping 0.85, 500 iterations). */
typedef struct prh_options {
  int repeat;
  double tolerance;
  double frontier_tolerance;
  double prune_tolerance;
  double damping;
  int max_iterations;
} prh_options;

/* PagerankResult<double> (pagerank.hxx:64-113) minus the rank vector: iterations of the last repeat and
 * phase times in milliseconds averaged over repeat (pagerankCuda.hxx:807-824, :848). The first solve that runs on
 * a graph's degree-ordered layout or on a launch plan not used before also reports their one-time build
 * (preparation_time, wall clock) and adds it to initialization_time and time, as the reference times its per-solve
 * partition in its initialization time (pagerankCuda.hxx:811-814). */
typedef struct prh_stats {
  int iterations;
  float time;
  float initialization_time;
  float marking_time;
  float computation_time;
  float preparation_time;
} prh_stats;

/* Launch tuning (no counterpart in the reference, whose values are fixed: switch degree 64, 256-thread blocks).
 * Results do not depend on these beyond floating-point summation order. */
typedef struct prh_tuning {
  uint32_t switch_degree; /* in-degree <= this: several vertices per wave; above: one wave per vertex [64] */
  uint32_t huge_degree;   /* in-degree > this: split into chunks of this many edges, one workgroup per chunk,
                             summed by a second launch [16384] */
  int low_width;          /* lanes per vertex of in-degree <= switch_degree: 1, 2, 4, 8, 16 or 32; 0 = by degree
                             band, W lanes for in-degree in (W, 2W] (the weighted solves always use bands) [0] */
  int check_interval;     /* iterations launched between host reads of the device convergence flag [8] */
  int source_slices;      /* Static / Naive-dynamic: 0 = row kernel; 1, 2, 4 or 8 = sum each row slice by slice,
                             source u in slice (u / 16) mod source_slices, slice x on the workgroups b with
                             b mod source_slices = x (8: one slice per XCD and its L2) [0] */
  uint32_t block_edges;   /* edges per workgroup of the rank kernels (>= 256) [2048] */
  int extension_direction; /* DF / DF-P extension: 0 push along out-edges of frontier vertices; 1 push or pull
                              (unaffected rows look for a frontier in-neighbour), whichever the last rank launch
                              found cheaper; 2 pull. The flags, hence the results, are the same for all [1] */
  int blocking;            /* every method, the weighted solves too: 0 = pull kernels above; 1024, 2048, 4096,
                              8192 or 16384 = propagation blocking with source blocks and destination bins of this
                              many vertices (each iteration expands contributions through LDS into a buffer with
                              one entry per (source, destination bin) pair, then adds them per bin in LDS; takes
                              precedence over source_slices; sums repeat to rounding, not bit for bit) [0] */
  int layout;              /* vertex order of the rows every plan runs over: 0 = by in-degree, descending (each
                              degree class a contiguous row range); 1 = the caller's order, which keeps the
                              neighbourhoods of a graph with locality (a road network, a grid) close together; the
                              row plan then runs one lane width for all rows and the sliced plan is not used. A graph
                              changes layout at its next solve (counted as preparation); one updated in place keeps
                              its layout [0] */
  int window;              /* Static / Naive-dynamic / DT / DF / DF-P with layout = 1 and blocking = 0: 1 = the banded
                              plan, in which each workgroup sweeps a contiguous range of rows chunk by chunk and gathers
                              from a window of the contributions around the chunk held in LDS (for graphs whose edges
                              join vertices close together in the caller's order: road networks, meshes, grids; sources
                              outside the window are gathered directly, so every graph gets the same results); 0 = the
                              row plan. Needs layout = 1; the weighted solves and the streaming session ignore it [0] */
} prh_tuning;

typedef struct prh_context prh_context;
typedef struct prh_graph prh_graph;

/* Replaces the PagerankOptions constructor defaults (pagerank.hxx:52). */
void prh_default_options(prh_options *options);
/* Replaces the fixed launch constants of pagerankCuda.hxx / _cuda.hxx:62-120. */
void prh_default_tuning(prh_tuning *tuning);

/* Context: device, stream and reusable device workspace. Replaces the per-call cudaSetDeviceFlags and
 * cudaMalloc/cudaFree of pagerankInvokeCuda (pagerankCuda.hxx:777-796, :829-847). The environment variable
 * PRH_GRAPH_CAPTURE=0 at creation makes the context issue the launches it would replay from captured HIP graphs
 * (batched PPR, the streaming session) one by one. */
int prh_context_create(int device, prh_context **context);
void prh_context_destroy(prh_context *context);
const char *prh_context_last_error(const prh_context *context);
/* Upper bound on device bytes this context may allocate (graphs + workspace); 0 = unlimited. */
int prh_context_set_memory_limit(prh_context *context, size_t bytes);
size_t prh_context_memory_used(const prh_context *context);
int prh_context_set_tuning(prh_context *context, const prh_tuning *tuning);
int prh_context_get_tuning(const prh_context *context, prh_tuning *tuning);
/* The hipStream_t the context's work runs on. */
void *prh_context_stream(prh_context *context);

/* Upload y's CSR arrays (host pointers; out_offsets/out_targets may be NULL for static-only use).
 * Replaces the host->device copies of pagerankInvokeCuda (pagerankCuda.hxx:798-802). The arrays are checked
 * on the device: offsets 0..M monotone, ids < N, out-degree equal to the out-CSR degree. */
int prh_graph_create(prh_context *context, uint32_t n, uint64_t m, const uint32_t *in_offsets,
                     const uint32_t *in_sources, const uint32_t *out_degrees, const uint32_t *out_offsets,
                     const uint32_t *out_targets, prh_graph **graph);
void prh_graph_destroy(prh_graph *graph);
uint32_t prh_graph_order(const prh_graph *graph);
uint64_t prh_graph_size(const prh_graph *graph);

/* Apply an edge batch to an uploaded graph on the device (no counterpart in the reference, whose experiment builds
 * the updated graph y on the host with tidyBatchUpdateU and applyBatchUpdateOmpU, batch.hxx:201-247, main.cu:80-86,
 * and uploads it). The batch (host arrays, the caller's vertex ids) is tidied as tidyBatchUpdateU tidies it:
 * deletions of existing edges only, insertions of absent edges only, each edge once. The graph must have been
 * uploaded with its out-CSR and afterwards is y for every solver (prh_solve, prh_mark_affected, prh_stream_create);
 * its propagation-blocked plan (blocks of at most 8192 vertices) is patched in place rather than rebuilt, and the
 * update's time (also returned in *update_ms, may be NULL) is added to the initialization and preparation times of
 * the next solve. While that plan is patched the in-CSR is built only when something reads it (a solve with another
 * plan, prh_graph_csr, prh_stream_create), whose first such call then also reports the build as preparation. The vertex set is fixed: an id >= N fails and leaves the graph unchanged. After an update the
 * batched and weighted solvers refuse the graph (PRH_ERROR_UNSUPPORTED): upload y for them. */
int prh_graph_update(prh_context *context, prh_graph *graph, const uint32_t *deletion_sources,
                     const uint32_t *deletion_targets, size_t n_deletions, const uint32_t *insertion_sources,
                     const uint32_t *insertion_targets, size_t n_insertions, float *update_ms);

/* The graph's current CSR in the caller's vertex ids (host arrays): which = 0 the in-CSR (in_offsets[N+1], in-neighbour
 * ids ascending), 1 the out-CSR (out_offsets[N+1], targets ascending). ids receives M values. No counterpart in the
 * reference; it reads back what prh_graph_update built. */
int prh_graph_csr(prh_context *context, const prh_graph *graph, int which, uint32_t *offsets, uint32_t *ids);

/* Run a solver on y. Replaces pagerankInvokeCuda (pagerankCuda.hxx:713-849) as called by the five entry
 * points (:864-988). Host pointers. The batch is (deletion_sources[i], deletion_targets[i]) and
 * insertion_sources[i] in dense ids (the reference's delu/delv/insu, :765-775); it is used by the dynamic
 * traversal / frontier methods only. initial_ranks (N values, may be NULL for PRH_STATIC) is the reference's q.
 * ranks receives N values. partition is the reference's PARTITION switch; results do not depend on it.
 * With an empty batch the dynamic methods mark nothing (the reference leaves its flags uninitialised). */
int prh_solve(prh_context *context, const prh_graph *graph, prh_method method, int partition,
              const prh_options *options, const uint32_t *deletion_sources, const uint32_t *deletion_targets,
              size_t n_deletions, const uint32_t *insertion_sources, size_t n_insertions,
              const double *initial_ranks, double *ranks, prh_stats *stats);

/* Same as prh_solve with every array (batch, initial_ranks, ranks) in device memory of the context's device;
 * no host copies happen and inputs stay resident. */
int prh_solve_device(prh_context *context, const prh_graph *graph, prh_method method, int partition,
                     const prh_options *options, const uint32_t *deletion_sources,
                     const uint32_t *deletion_targets, size_t n_deletions, const uint32_t *insertion_sources,
                     size_t n_insertions, const double *initial_ranks, double *ranks, prh_stats *stats);

/* The degree partition of a CSR given by its offsets (host pointers): order[N] receives the vertices with
 * degree <= switch_degree in ascending id order followed by the rest in ascending id order, and
 * *partition_point the size of the first part. Replaces pagerankPartitionVerticesCuW (pagerankCuda.hxx:625-638),
 * which the reference calls with switch degree 64. */
int prh_partition(prh_context *context, uint32_t n, const uint32_t *offsets, uint32_t switch_degree,
                  uint32_t *order, uint32_t *partition_point);

/* prh_partition with every array in device memory of the context's device and no device allocation: the partition's
 * scratch (per-tile counts and offsets) is a fixed array of 1.5 MiB that the library keeps on each device until the
 * process ends. No counterpart in the reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). order receives N vertex ids and
 * *partition_point the size of the first part (one uint32 in device memory). The work is enqueued on the context's
 * stream (prh_context_stream) and the call returns without waiting for it. At most 2^28 vertices
 * (PRH_ERROR_INVALID_ARGUMENT above that). Calls on one device share the scratch: calls that overlap in time on one
 * device (other streams, contexts or threads) race and must be ordered by the caller. */
int prh_partition_noalloc(prh_context *context, uint32_t n, const uint32_t *offsets, uint32_t switch_degree,
                          uint32_t *order, uint32_t *partition_point);

/* Affected flags right after the marking phase of a dynamic method (host pointers; affected receives N bytes,
 * 0 or 1). PRH_DYNAMIC_TRAVERSAL gives the traversal closure (pagerankAffectedTraversalCuW,
 * pagerankCuda.hxx:438-456); PRH_DYNAMIC_FRONTIER and PRH_PRUNE_DYNAMIC_FRONTIER give the initial frontier
 * (pagerankAffectedFrontierCuW, pagerankCuda.hxx:405-412). */
int prh_mark_affected(prh_context *context, const prh_graph *graph, prh_method method,
                      const uint32_t *deletion_sources, const uint32_t *deletion_targets, size_t n_deletions,
                      const uint32_t *insertion_sources, size_t n_insertions, uint8_t *affected);

/* Batched personalized PageRank (no counterpart in the reference): k <= 64 teleport columns solved in one
 * multi-column pull, each column j with teleport s_j uniform on its seed set
 * seeds[seed_offsets[j] .. seed_offsets[j+1]) (host arrays; a repeated seed adds its share again). Column j
 * iterates r_j <- (1-damping) s_j + damping * P^T r_j from r_j = s_j and stops at the first iteration whose L-inf
 * change is below options->tolerance, as a single solve would (pagerankLoopCuU's rule, pagerankCuda.hxx:680-690).
 * ranks receives N*k values, row-major by vertex (ranks[v*k + j]); iterations receives k counts. Iterations run as
 * a HIP graph of check_interval iterations replayed until every column has converged (with PRH_GRAPH_CAPTURE=0 in the
 * environment when the context was created: the same launches issued one by one, for tools that trace kernels; the
 * streaming session below honours it too). */
int prh_ppr_batched(prh_context *context, const prh_graph *graph, int k, const uint32_t *seed_offsets,
                    const uint32_t *seeds, const prh_options *options, double *ranks, int *iterations,
                    prh_stats *stats);

/* Weighted PageRank and its adjoint (no counterpart in the reference). Edge weights are FP32, one per edge of the
 * graph's out-CSR (weights[e] belongs to out_targets[e]), and are row-normalised to FP64 transition probabilities
 * P[u,v] = w_uv / W_u, W_u = sum of u's out-weights. The graph needs its out-CSR with ascending rows.
 *
 * prh_weighted_pagerank: solves r = (1-damping) s + damping P^T r by the pull iteration over the in-CSR from
 * r = s, stopping at the first iteration whose L-inf change is below options->tolerance (max_iterations caps it).
 *
 * prh_weighted_adjoint: given the forward solution r and g = dL/dr, solves lambda = g + damping P lambda by the
 * same pull over the out-CSR from lambda = g, then writes dL/dw (grad_weights, FP64, out-CSR edge order) and
 * dL/ds = (1-damping) lambda (grad_teleport); lambda may be NULL.
 *
 * device_pointers = 0: every array is host memory; 1: every array is device memory of the context's device (the
 * call synchronises the context stream before returning, callers order their own streams before the call). */
int prh_weighted_pagerank(prh_context *context, const prh_graph *graph, const float *weights, const double *teleport,
                          const prh_options *options, double *ranks, prh_stats *stats, int device_pointers);
int prh_weighted_adjoint(prh_context *context, const prh_graph *graph, const float *weights, const double *ranks,
                         const double *grad_ranks, const prh_options *options, double *lambda,
                         double *grad_weights, double *grad_teleport, prh_stats *stats, int device_pointers);

/* Autotuner (no counterpart in the reference, whose launch constants are fixed: switch degree 64, 256-thread
 * blocks, pagerankCuda.hxx:93, _cuda.hxx:62-120). Times candidate launch plans of the static update on this graph
 * (row plan: switch degree 64/128/256 x chunk threshold 4096/16384 x 2048/4096 edges per workgroup; sliced plan:
 * 1 or 8 source slices x 2048/4096 edges per workgroup), `iterations` iterations each after a warm-up, and sets
 * the fastest as the context's tuning (other fields keep their values). With cache_path (may be NULL) the choice
 * is read from / written to a JSON file keyed by the device architecture, its compute-unit count, N, M and a hash
 * of the graph's degree sequence; a hit skips the timing (*from_cache = 1). Results of every plan agree within
 * rounding, so tuning changes run time only. */
int prh_autotune(prh_context *context, const prh_graph *graph, const char *cache_path, int iterations,
                 prh_tuning *chosen, double *ms_per_iteration, int *from_cache);

/* Streaming dynamic session (no counterpart in the reference, whose experiment rebuilds y on the host and uploads
 * it for every batch, main.cu:80-143): a device-resident copy of `graph` (uploaded with its out-CSR) that update
 * batches modify on the device, with the ranks kept on the device between batches. Each update tidies the batch as
 * tidyBatchUpdateU does (batch.hxx:201-208: deletions of existing edges only, insertions of absent edges only,
 * duplicates dropped), applies it (applyBatchUpdateU, batch.hxx:219-247), and solves with `method` from the previous
 * ranks: PRH_STATIC (from 1/N), PRH_NAIVE_DYNAMIC, PRH_DYNAMIC_TRAVERSAL (the closure's levels run on the device,
 * each only while the one before it marked a vertex), PRH_DYNAMIC_FRONTIER or PRH_PRUNE_DYNAMIC_FRONTIER. The
 * batch application, the update plan's patch and the marking run as one captured HIP graph, the traversal closure as
 * a captured graph of 8 levels and the iterations as a captured graph of 8 iterations, each replayed until the
 * device reports it done. When the context's tuning asks for propagation blocking with blocks of at most 8192
 * vertices, the session builds that plan of the base graph and patches it per batch (deleted slots flagged dead,
 * other inserted edges kept in an overlay), as prh_graph_update does; otherwise it runs the row plan. Batch ids are
 * the caller's vertex ids (host arrays); each side holds at most max_batch edges; max_edges bounds the graph's growth
 * (0: |E| + 16 max_batch). initial_ranks (host, caller order) may be NULL: a static solve of `graph` then provides
 * them. A batch that names a vertex outside the graph or leaves a vertex without out-edges fails and leaves the
 * stream unusable. stats->initialization_time covers the batch upload, application, patch and initial marking,
 * marking_time the traversal closure, computation_time the iterations. The context's tuning at the first update of
 * each buffer parity is captured with the graphs. */
typedef struct prh_stream prh_stream;
int prh_stream_create(prh_context *context, const prh_graph *graph, prh_method method, const prh_options *options,
                      uint32_t max_batch, uint64_t max_edges, const double *initial_ranks, prh_stream **stream);
int prh_stream_update(prh_stream *stream, const uint32_t *deletion_sources, const uint32_t *deletion_targets,
                      size_t n_deletions, const uint32_t *insertion_sources, const uint32_t *insertion_targets,
                      size_t n_insertions, double *ranks, prh_stats *stats);
/* The current ranks in the caller's vertex order (host). */
int prh_stream_ranks(prh_stream *stream, double *ranks);
/* The current number of edges. */
uint64_t prh_stream_size(const prh_stream *stream);
/* The session's current CSR in the caller's vertex ids, as prh_graph_csr gives a graph's (which = 0 in-CSR, 1 out-CSR;
 * offsets N + 1 values, ids prh_stream_size values). */
int prh_stream_csr(prh_stream *stream, int which, uint32_t *offsets, uint32_t *ids);
void prh_stream_destroy(prh_stream *stream);

/* Per-launch timing of the rank-update kernel with HIP events on the context's stream (no counterpart in the
 * reference). While enabled, every solve accumulates the summed duration and count of the rank-update
 * launches that did work; reading resets the accumulators. */
int prh_kernel_timing_enable(prh_context *context, int enable);
int prh_kernel_timing_read(prh_context *context, uint64_t *launches, double *total_ms);

#ifdef __cplusplus
}
#endif

#endif /* PAGERANK_HIP_H */
