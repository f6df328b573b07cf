/*
 * dlp_kernels.h -- host launchers of the batch-update kernels (internal).
 */
#ifndef DLP_KERNELS_H
#define DLP_KERNELS_H

#include <vector>

#include "dlp_internal.h"

/* ------------------------------------------------------------------------ */
/* propagation (dlp_propagate.hip)                                          */
/* ------------------------------------------------------------------------ */
struct PropArgs {
  const int2 *rowInfo; /* nullable: {row start, row length} per vertex; the
                          multi-block and work-group HBM paths read it in
                          place of rowBeg / rowLen                           */
  const int *rowBeg;  /* row start per vertex                                  */
  const int *rowEnd;  /* standard CSR: rowPtr + 1 (used when rowLen == null)   */
  const int *rowLen;  /* slack layout: active row length                       */
  const int *colInd;
  const float *values;
  const int *status;  /* non-null: skip DELETED neighbours (general CSR input) */
  float *lab0;        /* buffer read by even iterations                        */
  float *lab1;
  int *iterStamp;     /* INT_MAX = may not enter a frontier                    */
  int *q0, *q1;       /* frontier queues                                       */
  int queueCap;       /* entries each queue holds (0: unknown); lets the
                         multi-block kernel load its first queue entry
                         before it knows the frontier size                    */
  int *touched;       /* nullable: first write of a vertex in this batch       */
  DevCounters *ctr;
  int *iterSizes;     /* nullable: frontier size of every iteration            */
  int iterSizesCap;
  int profile;        /* accumulate the prof* work counters                    */
  float delta;
  /* row tiers of the multi-block kernel: rows longer than waveRowMin (at
   * most 64) are summed by a whole wave instead of a lane group, rows longer
   * than blockRowMin (0 = never) by a whole work-group */
  int waveRowMin;
  int blockRowMin;
  /* 1: a row names v exactly when v's row names it (the pipeline's active
   * CSR over a symmetric full graph); the multi-wave register-mask kernels
   * read the next frontier off the rows and need it */
  int symmetric;
};
constexpr int kDefaultWaveRowMin = 64;
constexpr int kDefaultBlockRowMin = 1024;

/* one propagation iteration k (reads qcount[k%3]); groupLanes in
 * {4,8,16,32,64}; blocks of kPropagateThreads threads */
constexpr int kPropagateThreads = 512;
constexpr int kMinPropagateBlocks = 512;
/* Default grid cap of the multi-block kernel for a graph of N vertices:
 * while the per-vertex arrays stay in the caches (up to a few million
 * vertices) launches are short and more blocks only add dispatch and
 * per-block flush work (C2, 500,000 vertices: 566 ms per step at 512 blocks,
 * 603 ms at 8,192); once the label and stamp gathers go to HBM, more waves
 * in flight hide their latency (C3, 50 M vertices: 37.0 s at 512, 23.0 s at
 * 8,192). */
inline int dlp_auto_prop_blocks(long long numNodes) {
  const long long b = numNodes / 4096;
  return (int)(b < kMinPropagateBlocks ? kMinPropagateBlocks : (b > 8192 ? 8192 : b));
}
int dlp_launch_propagate(const PropArgs &a, int k, int groupLanes, int blocks, hipStream_t s);

/* global buffers of the LDS-resident propagation kernel */
struct LocalArgs {
  int numNodes;          /* length of coreGen / coreIdx                        */
  int *coreGen;          /* N: generation in which the vertex is in the core;
                            zero-filled at allocation, with ctr->localGen = 0  */
  int *coreIdx;          /* N: its core index in that generation               */
  int *haloGen;          /* N: generation in which the vertex is a neighbour of
                            the core outside it (zero-filled like coreGen)     */
  int *haloIdx;          /* N: its halo index in that generation (bit 30:
                            ineligible)                                        */
  int *coreList;         /* capacity coreListCap: core of the last build       */
  int coreListCap;
  int *ovf;              /* capacity ovfCap: marked vertices outside the core  */
  int ovfCap;            /* >= ldsWords / 4                                    */
  int ldsWords;          /* dynamic LDS of the launch, in 4-byte words          */
  int shape;             /* work-group shape, [0, dlp_local_shape_count())      */
  int ownerAfter;        /* iterations of a launch of the queue kernel before the
                            host goes on with the owner kernel; < 0: never     */
};
/* LocalArgs::shape of the register-mask owner kernels (rows padded to 16
 * records, one core vertex per thread): kLocalMaskShape + level, level 0 =
 * one wave (cores of at most 64 vertices), 1 = 256 threads, 2 = 512;
 * chosen by the host for owner mode on small frontiers, not tuning shapes */
constexpr int kLocalMaskShape = 100;
constexpr int kMaskLevels = 3;
constexpr int kMaskLevelThreads[kMaskLevels] = {64, 256, 512};
constexpr int kMaskOwnerMax = 512;
/* default LocalArgs::ownerAfter: only long-lived frontiers (C2's limit
 * cycles run 100,000 iterations in one launch) go on in owner mode */
constexpr int kOwnerAfterDefault = 1024;
/* allocation sizes of the LocalArgs buffers for ldsWords words of LDS */
inline int dlp_local_list_cap(int ldsWords) { return ldsWords / 4 + 64; }
/* default frontier threshold of the local kernel: the frontier whose
 * working set (about 26 words per core vertex with rows padded to 8
 * records, plus labels of its neighbours outside the core) fills the LDS */
inline int dlp_local_default_threshold(int ldsWords, long long meanDeg) {
  const long long t = ldsWords / (26 + 2 * meanDeg);
  return (int)(t < 1 ? 1 : t);
}

/* Runs propagation iterations from k in one work-group with the frontier's
 * rows and neighbour labels held in LDS, until the frontier empties, k
 * reaches maxIter, or the working set outgrows LDS.  Leaves the same global
 * state the multi-block kernel would (labels in both buffers, next queue and
 * its count, stamps, touched list, iteration counters) and records the stop
 * iteration and reason in ctr->nextIter / ctr->localReason. */
int dlp_launch_propagate_local(const PropArgs &a, const LocalArgs &l, int k, int maxIter,
                               bool owner,
                               hipStream_t s);
/* dynamic LDS bytes the local kernel can use on the current device */
int dlp_local_lds_bytes(int *bytes);
/* number of work-group shapes of the local kernel (0 = 1024 threads, rows
 * padded to 8 records) */
int dlp_local_shape_count();

/* ------------------------------------------------------------------------ */
/* IrLP full recompute (dlp_irlp.hip)                                       */
/* ------------------------------------------------------------------------ */
struct IrlpArgs {
  const int *rowPtr;  /* standard CSR, numNodes + 1                          */
  const int *colInd;
  const float *values;
  const int *status;
  const int *gtClass;
  float *lab0;        /* the caller's labels: read by even iterations        */
  float *lab1;        /* second buffer, numNodes floats                       */
  int *flags;         /* kIrlpMaxChunk + 1 words                              */
  int *colMasked;     /* nullable, numEntries ints: filled by dlp_run_irlp
                         with colInd, every entry that points at a DELETED
                         vertex set to -1, so sweeps read no neighbour status */
  long long numEntries; /* rowPtr[numNodes] (with colMasked)                  */
  int numNodes;
  float delta;
};
constexpr int kIrlpMaxChunk = 256;
struct IrlpLoopConfig {
  int groupLanes = 1;  /* lanes per vertex: 1 (one lane per row up to 16
                          entries, a wave per longer row) or 4 .. 64         */
  int blocks = 0;      /* <= 0: enough 256-thread blocks for one vertex group
                          each, at most 8192                                  */
  int firstChunk = 8;  /* iterations queued before the first host check; each
                          later chunk doubles, up to maxChunk                 */
  int maxChunk = 64;
};
int dlp_irlp_blocks_default(int numNodes, int groupLanes);
/* Runs IrLP on device buffers until an iteration changes nothing or
 * maxIterations; the final labels end in a.lab0.  hostFlags: pinned, at
 * least cfg.maxChunk + 1 words.  kernelMs (nullable): stream time of the
 * whole loop.  Synchronises s. */
int dlp_run_irlp(const IrlpArgs &a, int maxIterations, const IrlpLoopConfig &cfg, int *hostFlags,
                 hipStream_t s, int *iterationsUsed, float *kernelMs);

/* one graph of a multi-graph propagation launch */
struct GraphProp {
  PropArgs a;
  LocalArgs l;      /* ldsWords == 0: HBM iterations only             */
  int maxIter;
  int threshold;    /* frontier size at or below which LDS passes run */
  int backoff;      /* HBM iterations after an LDS overflow           */
  int heavy;        /* > 0: a frontier larger than this hands the graph
                       back to the host at that iteration              */
  int startIter;    /* iteration the launch starts the graph at (0 for a
                       new batch; the queue and counters of startIter
                       are on the device)                               */
  int stopIter;     /* the launch stops the graph at this iteration
                       (<= maxIter) ...                                 */
  long long sliceTicks; /* ... or, once it ran an iteration, at the first
                       iteration boundary after this many wall-clock
                       ticks of the launch (0 = no time limit)          */
  int *exitIter;    /* device int: -1 = propagation finished (frontier
                       empty or maxIter reached); k >= 0 = handed back
                       at iteration k (resume with the full-GPU loop);
                       -2 - k = stopped at iteration k by stopIter or
                       sliceTicks (resume in a later launch)            */
};
/* exitIter value of a graph stopped at iteration k by its slice */
__host__ __device__ inline int dlp_slice_exit(int k) { return -2 - k; }

/* Runs the propagation of every graph in `graphs` (device array, count
 * numGraphs) to completion in one launch, one work-group per graph, with
 * work-group shape `shape` (as LocalArgs::shape) and dynamic LDS of
 * ldsWords words (every graph's l.ldsWords must be 0 or at most that).
 * Each graph's queue, counters and labels must be set up as for
 * dlp_run_propagation. */
int dlp_launch_propagate_multi(const GraphProp *graphs, int numGraphs, int shape, int ldsWords,
                               hipStream_t s);

/* HIP-event timing of propagation launches (optional) */
struct PropTiming {
  std::vector<hipEvent_t> ev; /* pairs, reused                          */
  int used = 0;
  int64_t launches = 0, localLaunches = 0;
  double ms = 0.0, localMs = 0.0;
  std::vector<char> isLocal;  /* per pair: local kernel launch          */
  std::vector<int> count;     /* per pair: launches it brackets         */
  bool perChunk = false;      /* one pair per chunk of multi-block
                                 launches instead of one per launch     */
};

struct PropLoopConfig {
  int groupLanes;     /* multi-block kernel lanes per vertex                 */
  int blocks;         /* multi-block kernel grid, at most: a chunk launches
                         enough blocks for one lane group per vertex of twice
                         the frontier it starts from, at least
                         kMinPropagateBlocks                                 */
  int maxChunk;       /* multi-block iterations between host checks          */
  int localThreshold; /* frontier size at or below which the LDS kernel runs */
  int maskOwnerMax;   /* owner mode starts in a register-mask kernel on
                         frontiers of at most this many vertices (<= 0 never) */
};

/* Runs propagation iterations 0 .. until the frontier empties or maxIter,
 * starting from the queue and count already on the device (qcount[0]).
 * `l` may be null (multi-block only).  `pin` is pinned host memory for the
 * counter read-backs; `timing` may be null. */
int dlp_run_propagation(const PropArgs &a, const LocalArgs *l, int maxIter,
                        const PropLoopConfig &cfg, DevCounters *pin, hipStream_t s,
                        PropTiming *timing, int startK = 0);

/* ------------------------------------------------------------------------ */
/* batch pipeline (dlp_batch.hip)                                           */
/* ------------------------------------------------------------------------ */
struct DState {
  int N;
  /* full graph */
  const int *fRow;
  const int *fCol;
  const float *fVal;
  /* active CSR, slack layout: row u occupies [fRow[u], fRow[u] + aLen[u]) */
  int *aLen;
  int2 *aRow;     /* {fRow[u], aLen[u]}: one load per frontier vertex      */
  int *aCol;
  float *aVal;
  /* vertex state */
  int *hStatus;   /* reference h_vertexStatus                               */
  int *dStatus;   /* reference d_vertexStatus during the batch              */
  int *gtClass;
  const int *batchMap;
  uint8_t *vflags;
  float *lab0, *lab1;
  const float *trueLab;
  int *iterStamp, *markStamp, *compId, *par;
  int *rowStamp;  /* batch tag of the last CSR rebuild that listed the row  */
  int *accState;  /* accuracy contribution: bit 0 counted (host status
                     ACTIVE), bit 1 predicted class == true class          */
  uint32_t *newBits;
  int numWords;
  /* per-batch scratch */
  int *S;          /* inserted vertices, ascending                          */
  int *cand;       /* marked vertices                                       */
  int *q0, *q1, *touched;
  int *subSrc, *subDst;
  float *subW;
  float *p0, *p1;  /* per-member weight to GT class 0 / 1                    */
  int *keyA, *valA, *keyB, *valB, *keyC, *valC;
  int *segStart, *segEnd;
  float *red0, *red1; /* per-component tree scratch                          */
  float *W0, *W1;
  int *scratch;    /* scan / sort scratch                                   */
  DevCounters *ctr;
  DevBatchRecord *records;
};

/* a batch's inputs on the device */
struct DBatch {
  int b;
  const int *ins;
  int nIns;
  const int *gtIds;
  const float *gtLab;
  int nGt;
  const int *del;
  int nDel;
  const int *eSrc, *eDst;
  const float *eW;
  int nE;
  /* previous batch's lists (flags to clear, device status to restore) */
  const int *prevIns;
  int nPrevIns;
  const int *prevDel;
  int nPrevDel;
  const int *prevGt;
  int nPrevGt;
};

/* ------------------------------------------------------------------------ */
/* counters that grow across batches                                        */
/* ------------------------------------------------------------------------ */
/* Iteration stamps: batch b's propagation uses stamps stampBase + k + 1 for
 * k < maxIterations and then advances stampBase by iterationsUsed + 1, and
 * INT_MAX marks an ineligible vertex; the LDS kernels add 1 to localGen per
 * working-set build.  Both only grow, so a long-lived context (a resident
 * serving tenant) would wrap them.  The host keeps an upper bound of each
 * and, before a batch that could cross the limit, enqueues a rebase:
 * eligible stamps <- 0 with stampBase <- 1, or every core / halo generation
 * <- 0 with localGen <- 0.  Both keep every invariant the kernels rely on
 * (eligible stamps below stampBase between batches, generations below
 * localGen). */
constexpr int kMaxIterationsLimit = 1 << 28; /* larger maxIterations are rejected */
constexpr long long kGenLimit = (1LL << 31) - 2;
/* working-set builds of one batch, at most: a build either runs at least one
 * iteration, ends the launch, or is retried once without the saved core */
inline long long dlp_gens_per_batch(int maxIter) { return 4LL * ((long long)maxIter + 1) + 4; }
struct CounterGuardPlan {
  bool rebaseStamps, restartGens;
  long long stampBound, genBound; /* bounds after the batch */
};
/* the guard decision before one batch, from the bounds before it */
inline CounterGuardPlan dlp_guard_plan(long long stampBound, long long genBound,
                                               int maxIter) {
  CounterGuardPlan p{};
  const long long perStamp = (long long)maxIter + 1;
  p.rebaseStamps = stampBound + perStamp >= (long long)DLP_INT_MAX - 1;
  p.stampBound = (p.rebaseStamps ? 1 : stampBound) + perStamp;
  const long long perGen = dlp_gens_per_batch(maxIter);
  p.restartGens = genBound + perGen >= kGenLimit;
  p.genBound = (p.restartGens ? 0 : genBound) + perGen;
  return p;
}
/* eligible stamps <- 0, ctr->stampBase <- 1 */
int dlp_rebase_stamps(const DState &d, hipStream_t s);
/* core and halo generations <- 0, ctr->localGen <- 0 */
int dlp_restart_local_gens(const LocalArgs &l, DevCounters *ctr, hipStream_t s);

int dlp_reset_state(const DState &d, hipStream_t s);
int dlp_batch_begin(const DState &d, const DBatch &bt, hipStream_t s);
/* The active CSR of batch bt.  rowBound < 0: every row from the full graph
 * (one pass over it); else only the rows naming a vertex of the batch's
 * lists or of the previous batch's deletion list (at most rowBound rows):
 * valid when the full graph is symmetric, so that those are the listed
 * vertices' own rows and their neighbours' rows. */
int dlp_csr_update(const DState &d, const DBatch &bt, int markTag, long long rowBound,
                   hipStream_t s);
int dlp_sorted_inserted(const DState &d, hipStream_t s);
int dlp_sparsify(const DState &d, const DBatch &bt, float tau, int markTag, hipStream_t s);
int dlp_components(const DState &d, const DBatch &bt, hipStream_t s);
int dlp_clear_component_ids(const DState &d, const int *ins, int n, hipStream_t s);
int dlp_label_init(const DState &d, const DBatch &bt, hipStream_t s);
int dlp_labels_no_edges(const DState &d, const DBatch &bt, hipStream_t s);
int dlp_frontier0(const DState &d, int bound, hipStream_t s);
int dlp_batch_end(const DState &d, const DBatch &bt, hipStream_t s);
/* active CSR -> standard CSR (trace export); rowPtr N+1 */
int dlp_export_csr(const DState &d, int *rowPtr, int *colInd, float *values, hipStream_t s);
/* affected flags (markStamp == tag) -> int[N] */
int dlp_export_marks(const DState &d, int tag, int *out, hipStream_t s);

/* ------------------------------------------------------------------------ */
/* step API kernels on a standard CSR (dlp_steps.hip)                       */
/* ------------------------------------------------------------------------ */
int dlp_k_mark_deletions(int *rowPtr, int *colInd, int *status, int *isAffected,
                         const int *deleted, int numDeleted, int numNodes, hipStream_t s);
/* scratch: dlp_sparsify_scratch_ints(numNewEdges) ints; errFlag gets bit 16
 * if the one-pass compaction gave up waiting on a tile */
int dlp_k_sparsify(const int *src, const int *dst, const float *w, int numNewEdges,
                   int numInserted, const int *isNew, float tau, int *isAffected, int *subSrc,
                   int *subDst, float *subW, int *countDev, int *scratch, int *errFlag,
                   hipStream_t s);
int dlp_k_fill_int(int *p, int value, long long n, hipStream_t s);
int dlp_k_copy_float(float *dst, const float *src, long long n, hipStream_t s);
/* ascending list of distinct inserted ids; its length goes to ctr->numNew */
int dlp_k_sorted_list(const int *inserted, int numInserted, int numNodes, uint32_t *bits,
                      int *sorted, DevCounters *ctr, int *scratch, hipStream_t s);
/* `sorted` must hold numInserted distinct ids (checked by the caller) */
int dlp_k_cc(const int *src, const int *dst, int numEdges, const int *inserted, int numInserted,
             int *componentId, int numNodes, int *par, const int *sorted, DevCounters *ctr,
             int *scratch, hipStream_t s);
int dlp_k_label_init(const int *rowPtr, const int *colInd, const float *values, float *labels,
                     const int *componentId, int numComponents, const int *inserted,
                     int numInserted, const int *gtClass, const int *sorted, float *p0,
                     float *p1, int *keyA, int *valA, int *keyB, int *valB, int *keyC, int *valC,
                     int *segStart, int *segEnd, float *red0, float *red1, float *W0, float *W1,
                     int *scratch, hipStream_t s);
int dlp_k_prop_prepare(const int *isAffected, const int *status, const int *gtClass,
                       int numNodes, int *iterStamp, int *q0, DevCounters *ctr, hipStream_t s);
int dlp_k_prop_finish(float *labels, const int *gtClass, int numNodes, int *isAffected,
                      const int *pending, DevCounters *ctr, int pendingIndex, int capped,
                      hipStream_t s);

#endif /* DLP_KERNELS_H */
