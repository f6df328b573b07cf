/*
 * dlp_context.h -- the device context of one evolving graph (internal).
 */
#ifndef DLP_CONTEXT_H
#define DLP_CONTEXT_H

#include <algorithm>
#include <string>
#include <vector>

#include "dlp_kernels.h"
#include "dynlp/dynlp_c.h"

/* ------------------------------------------------------------------------ */
/* context                                                                  */
/* ------------------------------------------------------------------------ */
struct dlp_context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool ownStream = false;
  dlp_config cfg{};
  float tau = 0.5f;
  int N = 0, B = 0;
  long long nnzFull = 0;
  /* the full graph names v in u's row exactly as often as u in v's row */
  bool symmetric = false;
  /* per batch: rows the batch's lists (and the previous batch's deletions)
   * can change, with their full-graph neighbours, counted with repeats */
  std::vector<int64_t> rowsBound;
  int rowUpdate = 0; /* dlp_context_row_update mode */
  std::vector<int64_t> insOff, gtOff, delOff, eOff;
  std::vector<void *> allocs;
  int64_t deviceBytes = 0;
  DState d{};
  int *dIns = nullptr, *dGtIds = nullptr, *dDel = nullptr, *dESrc = nullptr, *dEDst = nullptr;
  float *dGtLab = nullptr, *dEW = nullptr;
  int *iterSizes = nullptr;
  int iterSizesCap = 0;
  DevCounters *pinCtr = nullptr;
  /* propBlocks 0 = by graph size; chunks of up to 1024 multi-block
   * launches (C2: 8 -> 471.9, 64 -> 498.9, 256 -> 523.5, 1024 -> 539.8 M
   * updates/s: fewer host round trips, and after a working set outgrew LDS
   * the LDS kernel is retried only after a chunk); local threshold and LDS
   * set at creation */
  dlp_tuning tuning{8, 0, 1024, 0, 0, 1, kOwnerAfterDefault, kDefaultWaveRowMin,
                    kDefaultBlockRowMin, kMaskOwnerMax};
  LocalArgs local{};
  int localMaxBytes = 0;
  /* per-batch scratch (queues, candidate / touched lists, G', sort and
   * reduction buffers): the context's own allocation, sized for its largest
   * batch, or none while a runner's scratch pool lends it per batch */
  void *scratchAlloc = nullptr;
  int64_t scratchAllocBytes = 0;
  int maxIns = 0, maxE = 0;
  int nextBatch = 0;
  /* upper bounds of the device's stampBase and localGen (dlp_guard_plan) */
  long long stampBound = 1, genBound = 0;
  int markSeq = 0;
  int lastCCBatch = -1;
  std::vector<hipEvent_t> evStart, evEnd;
  std::vector<DevBatchRecord> records;
  std::vector<char> ran;
  dlp_trace *trace = nullptr;
  /* shape class and where the tuning came from (0 heuristic, 1 table, 2
   * autotune) */
  dlp_shape_key shape{};
  int tuningSource = 0;
  /* profiling: event-timed propagation launches */
  int profiling = 0;
  PropTiming timing;

  template <class T>
  int alloc(T **p, long long count) {
    size_t bytes = (size_t)std::max<long long>(count, 1) * sizeof(T);
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      dlp_set_errorf_("hipMalloc of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
      return DLP_ERR_NOMEM;
    }
    allocs.push_back(q);
    deviceBytes += (int64_t)bytes;
    *p = (T *)q;
    return DLP_OK;
  }
  template <class T>
  int upload(T **p, const T *host, long long count) {
    int rc = alloc(p, count);
    if (rc != DLP_OK) return rc;
    if (count > 0) DLP_HIP(hipMemcpy(*p, host, (size_t)count * sizeof(T), hipMemcpyHostToDevice));
    return DLP_OK;
  }
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void *p : allocs) (void)hipFree(p);
    allocs.clear();
    if (scratchAlloc) (void)hipFree(scratchAlloc);
    scratchAlloc = nullptr;
    for (hipEvent_t e : timing.ev) (void)hipEventDestroy(e);
    timing.ev.clear();
    for (hipEvent_t e : evStart) (void)hipEventDestroy(e);
    for (hipEvent_t e : evEnd) (void)hipEventDestroy(e);
    evStart.clear();
    evEnd.clear();
    if (pinCtr) (void)hipHostFree(pinCtr);
    pinCtr = nullptr;
    if (ownStream && stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};


/* Per-batch scratch of a batch with `ins` inserted vertices and `e` new
 * edges: its size in bytes and, when base is non-null, the per-batch
 * pointers of c->d set into [base, base + size) (each buffer 256-byte
 * aligned).  base == nullptr with bind leaves the pointers null. */
int64_t dlp_ctx_scratch_layout(dlp_context *c, long long ins, long long e, char *base, bool bind);
/* scratch bytes of batch b */
int64_t dlp_ctx_batch_scratch_bytes(const dlp_context *c, int b);
/* own = true: allocate the context's own scratch (its largest batch) and
 * bind it; false: free it and leave the per-batch pointers null (stream
 * idle) */
int dlp_ctx_own_scratch(dlp_context *c, bool own);

/* before a batch of the context: rebases its iteration stamps and restarts
 * its LDS generations on stream s when the batch could wrap them (enqueued
 * outside any captured graph: the decision is the host's) */
int dlp_ctx_counter_guard(dlp_context *c, hipStream_t s);

/* the batch pipeline of a context, in three phases (same stream) */
DBatch dlp_ctx_make_batch(const dlp_context *c, int b);
/* status update .. first frontier; `markTag` receives the batch's mark tag */
int dlp_ctx_batch_pre(dlp_context *c, int b, const DBatch &bt, int *markTag);
/* the arguments of the context's propagation kernels */
PropArgs dlp_ctx_prop_args(const dlp_context *c);
/* the context's propagation loop (its tuning, engines and timing) from
 * iteration startK, the queue and counters of startK being on the device */
int dlp_ctx_run_propagation(dlp_context *c, int startK);
/* GT enforcement, accuracy, batch record; marks the batch as run */
int dlp_ctx_batch_post(dlp_context *c, int b, const DBatch &bt);
/* the host bookkeeping batch b's pre and post phases perform (mark tag
 * sequence, last batch with a CC step, run flags), for a batch replayed
 * from a captured graph */
void dlp_ctx_advance_host(dlp_context *c, int b, const DBatch &bt);
/* the batch record of batch b (no timing); synchronises the stream */
int dlp_ctx_fill_stats(dlp_context *c, int b, dlp_batch_stats *out);

#endif /* DLP_CONTEXT_H */
