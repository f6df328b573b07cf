/*
 * dynlp_c.h -- C-ABI of the MI355X DynLP batch-update library (libdynlp.so).
 *
 * Plain C: pointers, sizes and int status codes; no GPU-runtime or C++ types.
 * `void* stream` is a hipStream_t (NULL = the default stream).  Every entry
 * point returns DLP_OK (0) or a DLP_ERR_* code; dlp_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Groups of entry points:
 *   1. Step API -- one entry per reference step launcher, same argument order
 *      and meaning; raw device pointers; host scalars that the reference
 *      returns through `int&` come back through `int*` (these calls
 *      synchronise `stream`, as the reference's do).
 *   2. runDynLP -- the reference's batch loop over host inputs.
 *   3. Context API -- the same batch loop with the graph and the whole batch
 *      stream resident in HBM, for replay, tracing and measurement.
 *   4. Input generation -- the reference's synthetic graph and batch
 *      generators (host code), so that callers get identical inputs.
 *   5. Many independent graphs -- one propagation launch per batch index
 *      over many resident graphs, and an arrival-driven serving loop.
 *   6. Device-memory plumbing (allocation and copies) for callers without a
 *      GPU runtime of their own.
 *   7. IrLP -- the reference's full-recompute validation baseline.
 *
 * Precision contract: FP32 with IEEE division and no FMA contraction; the
 * propagation sum follows the reference kernel's 256-slot block reduction
 * tree position for position, so results are bit-identical to the CPU
 * restatement that follows the same order.
 */
#ifndef DYNLP_C_H
#define DYNLP_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define DLP_OK 0
#define DLP_ERR_INVALID 1   /* bad argument / precondition violated */
#define DLP_ERR_HIP 2       /* HIP runtime error                    */
#define DLP_ERR_NOMEM 3     /* device or host allocation failed     */
#define DLP_ERR_STATE 4     /* call out of order / bad handle       */
#define DLP_ERR_NODEVICE 5  /* no GPU visible                        */
#define DLP_ERR_INTERNAL 6  /* a device-side bounded wait expired    */

/* ---- vertex status values (reference graphTypes.h:191-195) ----------- */
#define DLP_VERTEX_ACTIVE 0
#define DLP_VERTEX_DELETED 1
#define DLP_VERTEX_GROUND_TRUTH 2

/* Message of the last failing call on this thread ("" if none). */
const char *dlp_last_error(void);

/* Library version string. */
const char *dlp_version(void);

/* Number of visible GPUs (0 when none; never an error). */
int dlp_device_count(int *count);

/* ====================================================================== */
/* 1. Step API                                                            */
/* ====================================================================== */

/* Replaces launchMarkDeletions (sparsifySubgraph.h:41-43, .cu:185-203).
 * For each deleted u: isAffected[v] = 1 for every valid v in row(u), row(u)
 * entries set to -1, vertexStatus[u] = DELETED; then every entry of a
 * non-deleted row that points at a DELETED vertex is set to -1. */
int dlp_mark_deletions(int *rowPtr, int *colInd, int *vertexStatus,
                       int *isAffected, const int *deletedNodes,
                       int numDeleted, int numNodes, void *stream);

/* Replaces launchBuildSparsifiedSubgraph (sparsifySubgraph.h:78-84,
 * .cu:205-287).  Edges with w > tau and both ends new go to G' in both
 * directions, in input order: (u,v,w) then (v,u,w).  subColInd / subValues
 * receive the destinations / weights like the reference; subEdgeSrc
 * (nullable, not in the reference signature) receives the sources, which the
 * reference frees before its CC step.  Endpoints u of every w > tau edge
 * are marked affected iff 0 <= u < numInserted (the reference bounds marks
 * by numSubNodes == numInserted, .cu:177-178, :259-263).  subRowPtr is not
 * written (the reference never writes it either).  Output arrays need room
 * for 2 * numNewEdges entries.  Synchronises `stream`. */
int dlp_build_sparsified_subgraph(
    int *rowPtr, int *colInd, float *values, int *isAffected,
    const int *vertexStatus, const int *newEdgeSrc, const int *newEdgeDst,
    const float *newEdgeWeight, int numNewEdges, const int *insertedNodes,
    int numInserted, const int *isNewVertex, float tau, int *subRowPtr,
    int *subColInd, float *subValues, int *subEdgeSrc, int *numSubNodes,
    int *numSubEdges, void *stream);

/* Replaces launchConnectedComponents (connectedComponents.h:50-54,
 * .cu:162-290).  Components of the inserted vertices under the directed
 * edge list (edgeSrc[i], edgeDst[i]) -- G', per the header contract
 * connectedComponents.h:33-36.  componentId[v] = rank of the component's
 * minimum vertex id among all component minima (ascending), -1 for vertices
 * not inserted.  numEdges == 0 follows the reference (.cu:191-205): all
 * componentIds -1 and numComponents = numInserted.  Edge endpoints must be
 * inserted vertices.  Synchronises `stream`. */
int dlp_connected_components(const int *edgeSrc, const int *edgeDst,
                             int numEdges, const int *insertedNodes,
                             int numInserted, int *componentId,
                             int *numComponents, int numNodes, void *stream);

/* Replaces launchLabelInitialization (labelInitialization.h:54-60,
 * .cu:148-187).  W0[c] / W1[c] = weight from component c to ground-truth
 * class 0 / 1 vertices; F_u = W1/(W0+W1) if W0+W1 > 1e-10f else 0.5;
 * ground truth -> 0 / 1; componentId < 0 -> 0.5.  Deterministic order: each
 * member's row is summed left to right, then members' sums are combined by a
 * bottom-up pairwise tree in ascending vertex id.  insertedNodes must be
 * distinct. */
int dlp_label_initialization(const int *rowPtr, const int *colInd,
                             const float *values, float *labels,
                             const int *vertexStatus, const int *componentId,
                             int numComponents, const int *insertedNodes,
                             int numInserted, const int *gtClass,
                             int numNodes, void *stream);

/* Replaces launchIterativePropagation (iterativePropagation.h:73-80,
 * .cu:241-346).  Frontier = {v : isAffected[v] == 1, status ACTIVE, not
 * ground truth}; each iteration F_new[u] = clamp(sum w F_old / sum w) over
 * non-deleted neighbours (F_old[u] if the weight sum is <= 1e-10f); a vertex
 * whose label moved by more than delta re-enters the next frontier with its
 * eligible neighbours.  Double-buffered exactly like the reference: labels
 * and labelsBuffer alternate, non-frontier entries of the write buffer are
 * not refreshed, and the last written buffer ends in `labels`, with ground
 * truth enforced.  isAffected holds the reference's exit state.
 * Synchronises `stream`. */
int dlp_iterative_propagation(const int *rowPtr, const int *colInd,
                              const float *values, float *labels,
                              float *labelsBuffer, const int *vertexStatus,
                              const int *gtClass, int *isAffected,
                              int numNodes, float delta, int maxIterations,
                              int *iterationsUsed, void *stream);

/* dlp_iterative_propagation without any device allocation.  No counterpart
 * in the reference: for callers that cannot allocate device memory (serving
 * loops and harnesses that forbid hipMalloc).  Same arguments, semantics and
 * results as dlp_iterative_propagation; its scratch (counters, stamps, the
 * two frontier queues and the LDS-kernel bookkeeping: 28 bytes per vertex)
 * lives in a static device array of 14.1 MiB per GPU, held until the
 * process ends, so numNodes may be at most 524,288 (DLP_ERR_INVALID above).
 * Every call on a GPU rewrites that GPU's one copy of the array (it keeps
 * the last call's contents; each call clears what must start at zero), so
 * two calls that overlapped on one GPU would race: the library serialises
 * the calls of a process, and each returns with `stream` synchronised. */
int dlp_iterative_propagation_noalloc(const int *rowPtr, const int *colInd,
                                      const float *values, float *labels,
                                      float *labelsBuffer, const int *vertexStatus,
                                      const int *gtClass, int *isAffected,
                                      int numNodes, float delta, int maxIterations,
                                      int *iterationsUsed, void *stream);

/* ====================================================================== */
/* 2. runDynLP over host inputs                                           */
/* ====================================================================== */

/* DynLPConfig minus its output-directory string (graphTypes.h:124-154). */
typedef struct dlp_config {
  int totalNodes;
  int avgDegree;
  int numBatches;
  unsigned int seed;
  float tau;  /* < 0: auto (mean full-graph edge weight, dynLP.cu:186-190) */
  float delta;
  int maxIterations;
  float groundTruthFraction;
  float unlabeledFraction;
  float deletedFraction;
  int validate;
  int verbose;
} dlp_config;

/* Reference defaults (graphTypes.h:147-153). */
void dlp_config_default(dlp_config *cfg);

/* Host CSR (CSRGraph). */
typedef struct dlp_graph_view {
  int numNodes;
  int numEdges;
  const int *rowPtr;   /* numNodes + 1 */
  const int *colInd;   /* numEdges     */
  const float *values; /* numEdges     */
} dlp_graph_view;

/* std::vector<BatchData> flattened: batch b owns [off[b], off[b+1]) of each
 * list. */
typedef struct dlp_batch_view {
  int numBatches;
  const int64_t *insertedOffsets; /* numBatches + 1 */
  const int *insertedNodeIds;
  const int64_t *groundTruthOffsets;
  const int *groundTruthNodeIds;
  const float *groundTruthLabels;
  const int64_t *deletedOffsets;
  const int *deletedNodeIds;
  const int64_t *edgeOffsets;
  const int *newEdgeSrc;
  const int *newEdgeDst;
  const float *newEdgeWeight;
} dlp_batch_view;

/* Replaces runDynLP (dynLP.h:69-74, dynLP.cu:175-587).  finalLabels: N
 * floats; batchTimes / batchAccuracy: numBatches floats each (nullable).
 * Batch times are device time per batch (HIP events). */
int dlp_run_dynlp(const dlp_config *cfg, const dlp_graph_view *graph,
                  const float *trueLabels, const dlp_batch_view *batches,
                  float *finalLabels, float *batchTimes, float *batchAccuracy);

/* ====================================================================== */
/* 3. Context API: device-resident graph + batch stream                   */
/* ====================================================================== */

typedef struct dlp_context dlp_context;

/* Per-batch counters of the last run of that batch. */
typedef struct dlp_batch_stats {
  int batch;
  int numInserted;
  int numDeleted;
  int numNewEdges;
  int numSubEdges;      /* directed G' entries                         */
  int numComponents;
  int numMarked;        /* |affected set| before the eligibility filter */
  int frontier0;        /* first frontier size                           */
  int iterations;       /* iterationsUsed                                */
  int numTouched;       /* distinct vertices written by propagation      */
  int64_t updates;      /* sum of frontier sizes over iterations         */
  float timeMs;         /* device time of the batch                      */
  float accuracy;
} dlp_batch_stats;

/* Tuning knobs of the device path (0 = built-in default). */
typedef struct dlp_tuning {
  int propGroupLanes;   /* lanes per frontier vertex: 4, 8, 16, 32 or 64 */
  int propBlocks;       /* most blocks of a propagation launch (a launch
                           sizes its grid to the frontier); 0 = by graph
                           size: N / 4096 within [512, 8192]            */
  int chunkIterations;  /* most multi-block iterations enqueued between
                           host checks (chunks double up to it; 1024)    */
  int localThreshold;   /* frontier size at or below which iterations run
                           in the LDS-resident single-work-group kernel;
                           < 0 disables it                               */
  int localLdsBytes;    /* LDS the local kernel may use, at most the
                           device's per-work-group limit (the default)   */
  int localShape;       /* work-group shape of the local kernel, 1-based
                           (threads / records per padded row): 1 = 1024/8,
                           2 = 1024/16, 3 = 512/8, 4 = 256/8, 5 = 1024/4 */
  int localOwnerAfter;  /* iterations one launch of the local kernel runs
                           on its frontier queue before the propagation
                           goes on with the owner-mode local kernel (one
                           thread per core vertex, frontier as flags;
                           for long-lived frontiers such as limit
                           cycles): 1024 by default, -1 never            */
  int waveRowMin;       /* multi-block kernel: rows longer than this (at
                           most 64, the default) are summed by a whole
                           wave instead of the vertex's lane group       */
  int blockRowMin;      /* multi-block kernel: rows longer than this
                           (>= 64; default 1024) are summed by a whole
                           work-group; -1 never                          */
  int maskOwnerMax;     /* owner mode on a frontier of at most this many
                           vertices (<= 512, the default) runs in the
                           register-mask LDS kernels (one core vertex per
                           thread, rows in registers, frontier as 64-bit
                           masks, one or no barrier per iteration); -1
                           never                                         */
} dlp_tuning;

/* Uploads the full graph, true labels and every batch, allocates all device
 * state once (no allocation while batches run). */
int dlp_context_create(dlp_context **out, const dlp_config *cfg,
                       const dlp_graph_view *graph, const float *trueLabels,
                       const dlp_batch_view *batches, void *stream);
int dlp_context_destroy(dlp_context *ctx);
/* fields left 0 keep their current value */
int dlp_context_set_tuning(dlp_context *ctx, const dlp_tuning *tuning);
int dlp_context_get_tuning(const dlp_context *ctx, dlp_tuning *tuning);
/* The tau actually used (auto tau resolved). */
int dlp_context_tau(const dlp_context *ctx, float *tau);
/* Device bytes held by the context. */
int dlp_context_device_bytes(const dlp_context *ctx, int64_t *bytes);

/* Back to the state before batch 0 (device fills only). */
int dlp_context_reset(dlp_context *ctx);
/* Runs batch b (must be the next batch).  Returns when the batch's work is
 * enqueued; its propagation loop synchronises with the host between chunks of
 * iterations. */
int dlp_context_run_batch(dlp_context *ctx, int b);
/* reset + every batch. */
int dlp_context_run_all(dlp_context *ctx);
int dlp_context_synchronize(dlp_context *ctx);
/* Copies the current labels (N floats) to host. */
int dlp_context_read_labels(dlp_context *ctx, float *labels);
/* Stats of batch b from the most recent run (synchronises). */
int dlp_context_batch_stats(dlp_context *ctx, int b, dlp_batch_stats *out);

/* Snapshot of the last batch's intermediate state, for parity checks.
 * All pointers are host buffers and nullable.  Sizes: *_N arrays hold N
 * entries; csrRowPtr N+1; csrColInd/csrValues room for the full graph's
 * numEdges; sub* room for 2 * numNewEdges of that batch; frontierSizes
 * room for maxFrontierSizes entries. */
typedef struct dlp_trace {
  int *csrRowPtr;        /* device CSR during steps 1-3, standard layout   */
  int *csrColInd;
  float *csrValues;
  int *affected_N;       /* 1 where marked by step 1 (before filtering)   */
  int *subSrc;           /* G' in build order                             */
  int *subDst;
  float *subW;
  int *componentId_N;
  float *initLabels_N;   /* labels entering propagation                   */
  int *frontierSizes;
  int maxFrontierSizes;
  /* outputs */
  int numSubEdges;
  int numComponents;
  int numFrontierSizes;
} dlp_trace;

/* Enables (t != NULL) or disables tracing.  When enabled, each run_batch
 * synchronises and fills t; the struct must outlive the batches traced. */
int dlp_context_set_trace(dlp_context *ctx, dlp_trace *t);

/* Propagation-kernel measurement: with profiling on, every propagation
 * launch is bracketed by HIP events on the context's stream and the kernel
 * counts the work it does. */
typedef struct dlp_prop_profile {
  /* multi-block kernel (one launch per iteration) */
  int64_t launches;     /* launches measured                                */
  double kernelMs;      /* sum of their event-measured durations            */
  int64_t edges;        /* row entries read by frontier vertices            */
  int64_t changedEdges; /* sum over changed vertices of (row length + 1)    */
  int64_t appended;     /* next-frontier queue appends                      */
  /* LDS-resident kernel (many iterations per launch) */
  int64_t localLaunches;
  double localKernelMs;
  int64_t localUpdates; /* frontier vertices it processed                   */
  int64_t localEdges;
  int64_t localChangedEdges;
  int64_t localAppended;
  int64_t localOwnerUpdates; /* of localUpdates: processed in owner mode   */
  int64_t localBuilds;       /* LDS working sets built (passes)           */
  int64_t localWaveUpdates;  /* of localOwnerUpdates: processed by the
                                register-mask owner kernels (cores of at
                                most 512 vertices)                        */
} dlp_prop_profile;

/* Sets the profiling level and clears the accumulated profile:
 * 0 off; 1 HIP-event timing of every LDS-resident launch and of every chunk
 * of back-to-back multi-block launches (no work counters; the host does not
 * record events between launches); 2 events around every launch plus the
 * work counters (edges, changed edges, appends). */
int dlp_context_set_profiling(dlp_context *ctx, int level);
/* Profile accumulated since profiling was last turned on (synchronises). */
int dlp_context_prop_profile(dlp_context *ctx, dlp_prop_profile *out);

/* Wrap-around guards of the two device counters that only grow (no
 * counterpart in the reference, whose loop restarts them per call).
 * Propagation stamps iteration k of a batch stampBase + k + 1 and then
 * advances stampBase by iterationsUsed + 1; the LDS-resident kernels add one
 * to a generation counter per working set they build.  Before every batch
 * the library checks upper bounds of both and, when the batch could reach
 * the limit, rebases the stamps (eligible vertices' stamps to 0, stampBase
 * to 1) or restarts the generations; results do not change.
 * dlp_counter_guard_plan is that decision for given bounds (host only). */
int dlp_counter_guard_plan(int64_t stampBound, int64_t genBound, int maxIterations,
                           int *rebaseStamps, int *restartGens, int64_t *stampBoundAfter,
                           int64_t *genBoundAfter);
/* Moves a context's stampBase and generation counter up to the given
 * values at a batch boundary (tests of the guards start near the limits);
 * values may only grow. */
int dlp_context_set_counter_bases(dlp_context *ctx, int stampBase, int localGen);

/* How a batch's CSR update chooses rows (the reference rebuilds the whole
 * host CSR every batch, dynLP.cu:111-167): mode 0 = only the rows the
 * batch can change -- those naming a vertex of its lists or of the
 * previous batch's deletions -- when the full graph is symmetric and they
 * are at most a quarter of the rows, else every row (the default); 1 =
 * every row; 2 = only the rows the batch can change whenever the full
 * graph is symmetric.  mode < 0 keeps the mode.  *symmetric (nullable)
 * receives 1 if the full graph names v in u's row exactly as often as u in
 * v's row.  The CSR, and everything after it, is the same in every mode. */
int dlp_context_row_update(dlp_context *ctx, int mode, int *symmetric);
/* The same symmetry test on a host CSR, without a GPU: *symmetric = 1 if
 * the entries with column >= 0, as (row, column) pairs, form the same
 * multiset as their reverses (no counterpart in the reference). */
int dlp_graph_symmetric(const dlp_graph_view *graph, int *symmetric);
/* Current device values of both counters (synchronises). */
int dlp_context_counter_bases(dlp_context *ctx, int *stampBase, int *localGen);

/* ====================================================================== */
/* 4. Synthetic inputs (reference generateSparseGraph.cu / generateBatches.cu) */
/* ====================================================================== */

typedef struct dlp_inputs dlp_inputs;

/* generateSparseGraph (generateSparseGraph.h:42-44) + generateBatches
 * (generateBatches.h:56-59), same RNG call sequence. */
int dlp_inputs_generate(dlp_inputs **out, int totalNodes, int avgDegree,
                        int numBatches, unsigned int seed,
                        float groundTruthFraction, float deletedFraction);
/* Arrival-driven variant (no counterpart in the reference): batch 0 is
 * the GT seed batch of dlp_inputs_generate; batch b >= 1 inserts the next
 * batchSizes[b] vertices of one shuffled order with the same GT fraction,
 * deletions drawn uniformly from the active non-GT vertices, and edges in
 * batch max(b_u, b_v). */
int dlp_inputs_generate_stream(dlp_inputs **out, int totalNodes, int avgDegree,
                               const int *batchSizes, int numBatches,
                               unsigned int seed, float groundTruthFraction,
                               float deletedFraction);
/* Views into the handle's memory (valid until dlp_inputs_free). */
int dlp_inputs_graph(const dlp_inputs *in, dlp_graph_view *graph,
                     const float **trueLabels);
int dlp_inputs_batches(const dlp_inputs *in, dlp_batch_view *batches);
void dlp_inputs_free(dlp_inputs *in);

/* The generator's CSR text dump (generateSparseGraph.cu:217-274):
 * <prefix>RowPtr.txt, ColInd.txt, Values.txt (6 decimals) and
 * TrueLabels.txt ("id label"); directories of the prefix are created. */
int dlp_write_graph_files(const dlp_graph_view *graph, const float *trueLabels,
                          const char *prefix);
/* A graph read back from such files (weights as written) with the
 * reference's batch generator run over it (generateBatches.cu:44-223,
 * seed + 12345), so a stored graph can be replayed. */
int dlp_inputs_from_graph_files(dlp_inputs **out, const char *prefix, int numBatches,
                                unsigned int seed, float groundTruthFraction,
                                float deletedFraction);

/* ====================================================================== */
/* 5. Many independent graphs: one propagation launch per batch index     */
/*    (no counterpart in the reference, which updates one graph at a     */
/*    time; its per-graph semantics are runDynLP's, dynLP.cu:175-587)     */
/* ====================================================================== */
typedef struct dlp_multi dlp_multi;

typedef struct dlp_multi_tuning {
  int localShape;       /* work-group shape of the propagation launch,
                           1-based as dlp_tuning.localShape (default 3)  */
  int thresholdPercent; /* scales each graph's LDS-pass frontier
                           threshold (default 100); -1 = no LDS passes   */
  int backoff;          /* HBM iterations after an LDS overflow (8)      */
  int useGraphs;        /* 1 = capture each batch in a HIP graph and
                           replay it (default), 2 = enqueue every time   */
  int heavyFrontier;    /* a graph whose frontier grows beyond this many
                           vertices leaves the shared launch and finishes
                           its propagation with the single-graph engines
                           on a stream of its own; -1 = never.  Default:
                           never for dlp_multi_run_batch / run_next,
                           8192 for dlp_multi_serve                      */
} dlp_multi_tuning;

/* One context per graph (arrays of numGraphs configs, graphs, true-label
 * pointers and batch views), all on one internal stream. */
int dlp_multi_create(dlp_multi **out, int numGraphs, const dlp_config *cfgs,
                     const dlp_graph_view *graphs, const float *const *trueLabels,
                     const dlp_batch_view *batches);
int dlp_multi_destroy(dlp_multi *m);
/* fields left 0 keep their value; resets every graph's state */
int dlp_multi_set_tuning(dlp_multi *m, const dlp_multi_tuning *t);
int dlp_multi_get_tuning(const dlp_multi *m, dlp_multi_tuning *t);
int dlp_multi_reset(dlp_multi *m);
/* batch b of every graph that has one; batches run in order from a reset
 * (DLP_ERR_STATE after dlp_multi_run_next or dlp_multi_serve advanced
 * graphs one by one, until dlp_multi_reset) */
int dlp_multi_run_batch(dlp_multi *m, int b);
/* reset, then every batch */
int dlp_multi_run_all(dlp_multi *m);
/* waits for the stream and checks every graph's device error flags */
int dlp_multi_synchronize(dlp_multi *m);
int dlp_multi_read_labels(dlp_multi *m, int graph, float *labels);
/* per-graph record of batch b; timeMs is the all-graph batch time */
int dlp_multi_batch_stats(dlp_multi *m, int graph, int b, dlp_batch_stats *out);
int dlp_multi_batch_time(dlp_multi *m, int b, float *ms);
int dlp_multi_device_bytes(const dlp_multi *m, int64_t *bytes);

/* Scratch pool (no counterpart in the reference, which runs one graph).
 * poolBytes > 0: the graphs give up their own per-batch scratch (frontier
 * queues, candidate and touched lists, the sparsified subgraph, sort and
 * reduction buffers) and every batch draws it, while it runs, from one
 * device allocation of poolBytes shared by all graphs: dlp_multi_serve
 * admits a batch only when its scratch fits the pool's free space (else it
 * is deferred), and dlp_multi_run_batch / run_next fail with DLP_ERR_NOMEM
 * when the step's batches do not fit at once.  0 = every graph its own
 * scratch again (the default).  Graph state is kept either way. */
int dlp_multi_set_scratch_pool(dlp_multi *m, int64_t poolBytes);
typedef struct dlp_pool_stats {
  int64_t capacity;  /* bytes of the pool; 0 = per-graph scratch          */
  int64_t inUse;     /* bytes lent to running batches now                 */
  int64_t peak;      /* most bytes lent at once since the pool was set    */
  int64_t leases;    /* batches that drew their scratch from the pool     */
  int64_t deferrals; /* admissions dlp_multi_serve deferred for lack of
                        pool space                                        */
} dlp_pool_stats;
int dlp_multi_pool_stats(dlp_multi *m, dlp_pool_stats *out);
/* bytes of scratch batch b of graph g draws from a pool */
int dlp_multi_batch_scratch_bytes(const dlp_multi *m, int graph, int b, int64_t *bytes);

/* dlp_context_set_counter_bases / dlp_context_counter_bases for graph g. */
int dlp_multi_set_counter_bases(dlp_multi *m, int graph, int stampBase, int localGen);
int dlp_multi_counter_bases(dlp_multi *m, int graph, int *stampBase, int *localGen);

/* Serving.  Runs the next batch of each listed graph (distinct indices)
 * together: their pipelines, one propagation launch, their batch ends. */
int dlp_multi_run_next(dlp_multi *m, const int *graphs, int count);

typedef struct dlp_serve_request {
  int graph;       /* the request is this graph's next batch          */
  double arrival;  /* seconds after the start of dlp_multi_serve      */
} dlp_serve_request;

typedef struct dlp_serve_config {
  int maxSlots;           /* batches in the shared launch (>= 1)       */
  int64_t hbmBudgetBytes; /* > 0: size of the scratch pool the running
                             batches draw from (dlp_multi_set_scratch_
                             pool); 0 = the runner's current scratch   */
  int plain;              /* 1: one batch at a time, single-graph path */
  int handBackLanes;      /* host threads (each with a stream) that
                             finish handed-back graphs (multi tuning
                             heavyFrontier > 0) beside later steps;
                             0 = 16                                    */
  int sliceIterations;    /* a step stops each graph after this many
                             iterations (0 = 2048, -1 = no limit) ...  */
  double sliceMs;         /* ... or after this many milliseconds of its
                             work-group HBM iterations (0 = 2.0, < 0 =
                             no limit); the next step resumes it       */
  int handBackAfter;      /* with hand-back on, a graph stopped by its
                             slice at or past this iteration of its
                             batch goes to a hand-back lane if one is
                             free: the lane's owner-mode LDS kernels
                             run a long-lived frontier (a limit cycle)
                             faster than the shared launch (0 = 1024,
                             -1 = never)                               */
} dlp_serve_config;

/* Arrival-driven continuous batching with iteration-level admission:
 * requests (sorted by arrival) are released at their arrival times.  A
 * step is one shared propagation launch in which every running batch
 * advances by at most one slice (sliceIterations / sliceMs) and stops at
 * an iteration boundary; before each step, arrived requests are admitted
 * in arrival order -- at most one batch per graph, up to maxSlots running
 * batches, and only when the batch's scratch fits the pool (else it is
 * deferred; a batch larger than the whole pool is an error, DLP_ERR_NOMEM)
 * -- and join the running ones.  completion[i] receives the time (seconds
 * from the start) at which request i's labels were ready.  With hand-back
 * on, a graph whose frontier outgrows heavyFrontier, or that is still
 * running at iteration handBackAfter of its batch, leaves the shared
 * launches and is finished by a hand-back lane while the steps go on; its
 * tenant's next request waits for it.  Stopping and resuming a graph does
 * not change its iterations or labels. */
int dlp_multi_serve(dlp_multi *m, const dlp_serve_request *reqs, int numReqs,
                    const dlp_serve_config *cfg, double *completion, int *steps);

/* ====================================================================== */
/* 6. Device-memory plumbing for callers without a GPU runtime of their   */
/*    own (the reference calls cudaMalloc/cudaMemcpy directly)            */
/* ====================================================================== */
int dlp_device_malloc(void **ptr, size_t bytes);
int dlp_device_free(void *ptr);
int dlp_copy_to_device(void *dst, const void *src, size_t bytes);
int dlp_copy_to_host(void *dst, const void *src, size_t bytes);
/* device -> device with the library's own copy kernel (16-byte vectors when
 * both pointers are 16-byte aligned); synchronous */
int dlp_copy_device(void *dst, const void *src, size_t bytes);
/* Random 4-byte accesses per second over an array of arrayInts ints
 * (`accesses` per measurement): rates[0] returning atomicMax, rates[1]
 * loads, rates[2] stores -- the ceilings of the gather- and atomic-bound
 * propagation kernels (no counterpart in the reference). */
int dlp_random_access_rates(int64_t arrayInts, int64_t accesses, double *rates);
/* Cost of a device-wide sync between two dependent rounds of work on
 * `blocks` work-groups of 512 threads (each round: read the word another
 * work-group wrote in the round before, write one's own): us[0] per round
 * with one launch per round, us[1] per round inside one cooperative launch
 * with a bounded counter barrier between rounds, us[2] the same with a
 * barrier of per-work-group flags; ok[i] = 1 when every round saw all of
 * the previous round's writes (no counterpart in the reference). */
int dlp_grid_sync_cost(int blocks, int rounds, double *us, int *ok);

/* ====================================================================== */
/* 7. IrLP: the full-recompute baseline (validation.h, main.cu:120-217)  */
/* ====================================================================== */

/* Replaces launchIrLP (validation.h, validation.cu:144-205).  Device
 * pointers; `labels` in / out.  Each iteration sets every vertex that is
 * ACTIVE and not ground truth (gtClass < 0) to the clamped weighted mean of
 * its non-deleted neighbours' labels (its old label when the weight sum is
 * <= 1e-10f), keeps every other label, then forces gtClass 0 / 1 vertices
 * to 0 / 1.  Stops after the first iteration in which no such label moved
 * by more than delta, or after maxIterations; iterationsUsed counts the
 * iterations run.  Same floats as the reference.  The second label buffer
 * and the changed flags are library workspace (the reference allocates and
 * frees them per call).  Synchronises `stream`. */
int dlp_irlp(const int *rowPtr, const int *colInd, const float *values, float *labels,
             const int *vertexStatus, const int *gtClass, int numNodes, float delta,
             int maxIterations, int *iterationsUsed, void *stream);

/* Launch shape of the IrLP sweep (no counterpart in the reference). */
typedef struct dlp_irlp_tuning {
  int groupLanes; /* lanes per vertex: 1 (default: one lane per row of up
                     to 16 entries, a wave per longer row), 4, 8, 16, 32
                     or 64                                                 */
  int blocks;     /* 256-thread blocks per sweep; <= 0: one vertex group
                     per lane group, at most 8192 blocks                   */
  int firstChunk; /* iterations queued before the first convergence check;
                     each later chunk doubles up to maxChunk (<= 256)      */
  int maxChunk;
} dlp_irlp_tuning;
void dlp_irlp_tuning_default(dlp_irlp_tuning *t);

/* dlp_irlp with a launch shape (nullable) and the stream time of the
 * iteration loop in *kernelMs (nullable). */
int dlp_irlp_ex(const int *rowPtr, const int *colInd, const float *values, float *labels,
                const int *vertexStatus, const int *gtClass, int numNodes, float delta,
                int maxIterations, int *iterationsUsed, void *stream,
                const dlp_irlp_tuning *tuning, float *kernelMs);

/* Replaces runIrLPValidation (validation.h, validation.cu:210-276): host
 * CSR of the active graph, host status / gtClass, host labels in / out.
 * *skipped = 1 (labels untouched, iterationsUsed 0) when numNodes <= 0 or
 * the graph has no entries, where the reference prints a warning and
 * returns true. */
int dlp_irlp_validation(const dlp_graph_view *activeGraph, const int *vertexStatus,
                        const int *gtClass, float *labels, int numNodes, float delta,
                        int maxIterations, int *iterationsUsed, int *skipped);
/* dlp_irlp_validation with a launch shape (nullable) and the stream time of
 * the iteration loop alone in *loopMs (nullable). */
int dlp_irlp_validation_ex(const dlp_graph_view *activeGraph, const int *vertexStatus,
                           const int *gtClass, float *labels, int numNodes, float delta,
                           int maxIterations, int *iterationsUsed, int *skipped,
                           const dlp_irlp_tuning *tuning, float *loopMs);

typedef struct dlp_irlp_result {
  int iterations;      /* iterationsUsed                                   */
  int skipped;         /* empty final active graph                         */
  int activeEdges;     /* entries of the final active CSR                  */
  float timeMs;        /* wall time of the validation call (upload, run,
                          download), the span main.cu:177-197 times        */
  float loopMs;        /* stream time of the iteration loop alone         */
  float accuracy;      /* over non-deleted, non-ground-truth vertices       */
} dlp_irlp_result;

/* The reference's --validate step (main.cu:124-217): replays every batch
 * to the final vertex status, ground-truth classes and initial labels (0.5,
 * ground truth its label), builds the active CSR (entries of the full graph
 * between inserted, non-deleted vertices), runs dlp_irlp_validation with
 * cfg->delta / cfg->maxIterations and scores the result against trueLabels.
 * irlpLabels: numNodes floats; tuning nullable. */
int dlp_irlp_baseline(const dlp_config *cfg, const dlp_graph_view *fullGraph,
                      const float *trueLabels, const dlp_batch_view *batches,
                      const dlp_irlp_tuning *tuning, float *irlpLabels,
                      dlp_irlp_result *out);

/* ====================================================================== */
/* 8. Autotuner (no counterpart in the reference, which launches a        */
/*    256-thread block per frontier vertex whatever the degrees,          */
/*    iterativePropagation.cu:292-304)                                    */
/* ====================================================================== */

/* The shape class a tuning is remembered under. */
typedef struct dlp_shape_key {
  int logNodes;     /* ceil(log2 numNodes)                              */
  int meanDegree;   /* stored entries / numNodes, rounded               */
  int logMaxDegree; /* ceil(log2 (longest row + 1))                     */
} dlp_shape_key;

/* The degree-histogram part (host arithmetic, no GPU): the graph's shape
 * key and the tuning a context of it starts from -- lanes per frontier
 * vertex of the multi-block kernel (a lane group covers the
 * 90th-percentile row in two passes, 4 to 32 lanes), the wave and
 * work-group row tiers (the work-group tier only when a row is longer than
 * 1,024 entries), the grid cap from the vertex count.  The LDS kernel's
 * fields that need the device are left 0. */
int dlp_autotune_heuristic(const dlp_graph_view *graph, dlp_shape_key *key, dlp_tuning *out);

/* Times candidate launch shapes (lanes per vertex half, equal and double
 * the current ones x LDS work-group shapes 1 and 3) on the context's first
 * trialBatches batches (all when <= 0), keeps the fastest, stores it in the
 * process's tuning table under the context's shape key and leaves the
 * context reset.  Every candidate gives the same results. */
int dlp_context_autotune(dlp_context *ctx, int trialBatches, dlp_tuning *chosen, float *bestMs);
/* The context's shape key and where its tuning came from: 0 the
 * heuristic, 1 the tuning table, 2 dlp_context_autotune. */
int dlp_context_shape_key(const dlp_context *ctx, dlp_shape_key *key, int *source);

/* Times candidate work-group shapes (3, 1) and lockstep hand-back
 * frontiers (never, 65,536, 16,384) of the batched runner on its first
 * trialBatches batch indices (all when <= 0), keeps the fastest, stores it
 * in the tuning table under (log2 graph count, log2 total vertices) and
 * leaves the runner reset. */
int dlp_multi_autotune(dlp_multi *m, int trialBatches, dlp_multi_tuning *chosen, float *bestMs);
/* Where the runner's tuning came from: 0 defaults, 1 table, 2 autotune. */
int dlp_multi_tuning_source(const dlp_multi *m, int *source);

/* The process's tuning table.  Contexts and runners created while it holds
 * an entry for their shape start with that entry's tuning.  The library
 * loads <its directory>/tuning/<gfx arch>.tune (if present) when the first
 * context or runner is created.  Text, one entry per line. */
int dlp_tuning_table_load(const char *path, int *entries);
int dlp_tuning_table_save(const char *path, int *entries);
int dlp_tuning_table_clear(void);

#ifdef __cplusplus
}
#endif

#endif /* DYNLP_C_H */
