/*
 * dlp_runtime.hip -- device context, batch sequencing and the C-ABI of
 * runDynLP / the context API.
 *
 * The reference's batch loop (dynLP.cu:284-558) rebuilds the active CSR on
 * the host, uploads it and every per-vertex array each batch, and its steps
 * malloc/free scratch and read flags back every iteration.  Here the full
 * graph, the true labels and every batch's lists are uploaded once at
 * context creation, all state and scratch is allocated once, and a batch is
 * a sequence of kernels on one stream; the host only reads the frontier
 * count between chunks of propagation iterations to know when to stop.
 */
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "dlp_kernels.h"
#include "dlp_tune.h"

/* ------------------------------------------------------------------------ */
/* errors                                                                   */
/* ------------------------------------------------------------------------ */
static thread_local std::string g_lastError;

extern "C" void dlp_set_error_(const char *msg) { g_lastError = msg ? msg : ""; }

void dlp_set_errorf_(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_lastError = buf;
}

int dlp_hip_fail_(hipError_t e, const char *what) {
  dlp_set_errorf_("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return e == hipErrorOutOfMemory ? DLP_ERR_NOMEM : DLP_ERR_HIP;
}

extern "C" const char *dlp_last_error(void) { return g_lastError.c_str(); }
extern "C" const char *dlp_version(void) { return "libdynlp 0.1 (gfx950)"; }

extern "C" int dlp_device_count(int *count) {
  if (!count) {
    dlp_set_error_("dlp_device_count: NULL argument");
    return DLP_ERR_INVALID;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  (void)hipGetLastError();
  *count = n;
  return DLP_OK;
}

static int need_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    dlp_set_error_("no GPU visible to HIP");
    return DLP_ERR_NODEVICE;
  }
  return DLP_OK;
}

#include "dlp_context.h"

namespace {

/* splitmix64 finaliser: a well-mixed 64-bit hash of an ordered vertex pair */
uint64_t pair_hash(uint32_t u, uint32_t v) {
  uint64_t x = ((uint64_t)u << 32 | v) + 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

/* Whether the full graph's entries, (row, column) pairs with column >= 0,
 * form the same multiset as their reverses: then a row names v exactly
 * when v's row names it.  Compared through sums (mod 2^64) of a hash of
 * each pair and of its reverse -- equal for every symmetric graph, and
 * equal for an asymmetric one only by a hash collision (probability about
 * 2^-64).  Threads over row ranges. */
bool csr_symmetric(const dlp_graph_view *g) {
  const int N = g->numNodes;
  if (N <= 0 || g->numEdges <= 0) return true;
  const int T = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<uint64_t> fwd(T, 0), rev(T, 0);
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      const int lo = (int)((long long)N * t / T), hi = (int)((long long)N * (t + 1) / T);
      uint64_t f = 0, r = 0;
      for (int u = lo; u < hi; ++u)
        for (int e = g->rowPtr[u]; e < g->rowPtr[u + 1]; ++e) {
          const int v = g->colInd[e];
          if (v < 0) continue;
          f += pair_hash((uint32_t)u, (uint32_t)v);
          r += pair_hash((uint32_t)v, (uint32_t)u);
        }
      fwd[t] = f;
      rev[t] = r;
    });
  for (std::thread &x : th) x.join();
  uint64_t f = 0, r = 0;
  for (int t = 0; t < T; ++t) {
    f += fwd[t];
    r += rev[t];
  }
  return f == r;
}

/* host-side validation of the inputs (the reference trusts them; bad ids
 * would make device kernels index out of bounds) */
int validate_inputs(const dlp_config *cfg, const dlp_graph_view *g, const float *trueLabels,
                    const dlp_batch_view *bv) {
  if (!cfg || !g || !bv || !trueLabels) {
    dlp_set_error_("NULL config, graph, batches or trueLabels");
    return DLP_ERR_INVALID;
  }
  const int N = g->numNodes;
  if (N <= 0 || !g->rowPtr || (g->numEdges > 0 && (!g->colInd || !g->values))) {
    dlp_set_error_("graph must have numNodes > 0 and non-NULL arrays");
    return DLP_ERR_INVALID;
  }
  if (g->rowPtr[0] != 0 || g->rowPtr[N] != g->numEdges) {
    dlp_set_error_("graph rowPtr must start at 0 and end at numEdges");
    return DLP_ERR_INVALID;
  }
  for (int u = 0; u < N; ++u)
    if (g->rowPtr[u + 1] < g->rowPtr[u]) {
      dlp_set_errorf_("graph rowPtr decreases at row %d", u);
      return DLP_ERR_INVALID;
    }
  for (int e = 0; e < g->numEdges; ++e)
    if (g->colInd[e] < -1 || g->colInd[e] >= N) {
      dlp_set_errorf_("graph colInd[%d] = %d out of range", e, g->colInd[e]);
      return DLP_ERR_INVALID;
    }
  if (cfg->numBatches <= 0 || cfg->numBatches > bv->numBatches) {
    dlp_set_errorf_("config.numBatches = %d but %d batches given", cfg->numBatches,
                    bv->numBatches);
    return DLP_ERR_INVALID;
  }
  const int B = cfg->numBatches;
  std::vector<int> seen(N, -1), seenGt(N, -1);
  auto checkIds = [&](const int64_t *off, const int *ids, const char *what, std::vector<int> *uniq,
                      int b) -> int {
    if (off[b + 1] < off[b] || (off[b + 1] > off[b] && !ids)) {
      dlp_set_errorf_("batch %d: bad %s offsets", b, what);
      return DLP_ERR_INVALID;
    }
    for (int64_t i = off[b]; i < off[b + 1]; ++i) {
      int v = ids[i];
      if (v < 0 || v >= N) {
        dlp_set_errorf_("batch %d: %s id %d out of range", b, what, v);
        return DLP_ERR_INVALID;
      }
      if (uniq) {
        if ((*uniq)[v] == b) {
          dlp_set_errorf_("batch %d: %s id %d appears twice", b, what, v);
          return DLP_ERR_INVALID;
        }
        (*uniq)[v] = b;
      }
    }
    return DLP_OK;
  };
  for (int b = 0; b < B; ++b) {
    int rc;
    if ((rc = checkIds(bv->insertedOffsets, bv->insertedNodeIds, "inserted", &seen, b)) != DLP_OK)
      return rc;
    if ((rc = checkIds(bv->groundTruthOffsets, bv->groundTruthNodeIds, "ground-truth", &seenGt,
                       b)) != DLP_OK)
      return rc;
    if ((rc = checkIds(bv->deletedOffsets, bv->deletedNodeIds, "deleted", nullptr, b)) != DLP_OK)
      return rc;
    if ((rc = checkIds(bv->edgeOffsets, bv->newEdgeSrc, "edge source", nullptr, b)) != DLP_OK)
      return rc;
    if ((rc = checkIds(bv->edgeOffsets, bv->newEdgeDst, "edge destination", nullptr, b)) !=
        DLP_OK)
      return rc;
    for (int64_t i = bv->edgeOffsets[b]; i < bv->edgeOffsets[b + 1]; ++i)
      if (!std::isfinite(bv->newEdgeWeight[i])) {
        dlp_set_errorf_("batch %d: edge weight %lld is not finite", b, (long long)i);
        return DLP_ERR_INVALID;
      }
    for (int64_t i = bv->groundTruthOffsets[b]; i < bv->groundTruthOffsets[b + 1]; ++i)
      if (std::isnan(bv->groundTruthLabels[i])) {
        dlp_set_errorf_("batch %d: ground-truth label is NaN", b);
        return DLP_ERR_INVALID;
      }
  }
  if (cfg->maxIterations < 0 || cfg->maxIterations > kMaxIterationsLimit) {
    dlp_set_errorf_("maxIterations must be in [0, %d]", kMaxIterationsLimit);
    return DLP_ERR_INVALID;
  }
  return DLP_OK;
}

}  // namespace

DBatch dlp_ctx_make_batch(const dlp_context *c, int b) {
  DBatch bt{};
  bt.b = b;
  auto span = [](const std::vector<int64_t> &off, int b2, long long &n) {
    n = off[b2 + 1] - off[b2];
    return off[b2];
  };
  long long n;
  long long o = span(c->insOff, b, n);
  bt.ins = c->dIns + o;
  bt.nIns = (int)n;
  o = span(c->gtOff, b, n);
  bt.gtIds = c->dGtIds + o;
  bt.gtLab = c->dGtLab + o;
  bt.nGt = (int)n;
  o = span(c->delOff, b, n);
  bt.del = c->dDel + o;
  bt.nDel = (int)n;
  o = span(c->eOff, b, n);
  bt.eSrc = c->dESrc + o;
  bt.eDst = c->dEDst + o;
  bt.eW = c->dEW + o;
  bt.nE = (int)n;
  if (b > 0) {
    o = span(c->insOff, b - 1, n);
    bt.prevIns = c->dIns + o;
    bt.nPrevIns = (int)n;
    o = span(c->delOff, b - 1, n);
    bt.prevDel = c->dDel + o;
    bt.nPrevDel = (int)n;
    o = span(c->gtOff, b - 1, n);
    bt.prevGt = c->dGtIds + o;
    bt.nPrevGt = (int)n;
  }
  return bt;
}

namespace {

int run_propagation(dlp_context *c, const PropArgs &a, int maxIter, int startK = 0) {
  PropLoopConfig cfg;
  cfg.groupLanes = c->tuning.propGroupLanes;
  cfg.blocks = c->tuning.propBlocks > 0 ? c->tuning.propBlocks : dlp_auto_prop_blocks(c->N);
  cfg.maxChunk = c->tuning.chunkIterations;
  cfg.localThreshold = c->tuning.localThreshold;
  cfg.maskOwnerMax = c->tuning.maskOwnerMax;
  /* the local kernel encodes outside vertices as -2 - (2 * id + 1) */
  const bool local = c->tuning.localThreshold > 0 && c->N < (1 << 30);
  return dlp_run_propagation(a, local ? &c->local : nullptr, maxIter, cfg, c->pinCtr, c->stream,
                             c->profiling ? &c->timing : nullptr, startK);
}

int trace_pre_propagation(dlp_context *c, const DBatch &bt, int markTag) {
  dlp_trace *t = c->trace;
  const int N = c->N;
  DLP_HIP(hipStreamSynchronize(c->stream));
  std::vector<void *> tmp;
  auto devbuf = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess) return nullptr;
    tmp.push_back(p);
    return p;
  };
  auto cleanup = [&]() {
    for (void *p : tmp) (void)hipFree(p);
  };
  int rc = DLP_OK;
  if (t->csrRowPtr && t->csrColInd && t->csrValues) {
    int *rp = (int *)devbuf(sizeof(int) * (N + 1));
    int *ci = (int *)devbuf(sizeof(int) * (size_t)std::max<long long>(c->nnzFull, 1));
    float *vv = (float *)devbuf(sizeof(float) * (size_t)std::max<long long>(c->nnzFull, 1));
    if (!rp || !ci || !vv) {
      cleanup();
      dlp_set_error_("trace: hipMalloc failed");
      return DLP_ERR_NOMEM;
    }
    rc = dlp_export_csr(c->d, rp, ci, vv, c->stream);
    if (rc == DLP_OK && hipStreamSynchronize(c->stream) == hipSuccess) {
      (void)hipMemcpy(t->csrRowPtr, rp, sizeof(int) * (N + 1), hipMemcpyDeviceToHost);
      int nnz = t->csrRowPtr[N];
      (void)hipMemcpy(t->csrColInd, ci, sizeof(int) * nnz, hipMemcpyDeviceToHost);
      (void)hipMemcpy(t->csrValues, vv, sizeof(float) * nnz, hipMemcpyDeviceToHost);
    }
  }
  if (rc == DLP_OK && t->affected_N) {
    int *m = (int *)devbuf(sizeof(int) * N);
    rc = m ? dlp_export_marks(c->d, markTag, m, c->stream) : DLP_ERR_NOMEM;
    if (rc == DLP_OK && hipStreamSynchronize(c->stream) == hipSuccess)
      (void)hipMemcpy(t->affected_N, m, sizeof(int) * N, hipMemcpyDeviceToHost);
  }
  cleanup();
  if (rc != DLP_OK) return rc;
  DevCounters ctr;
  DLP_HIP(hipMemcpy(&ctr, c->d.ctr, sizeof(DevCounters), hipMemcpyDeviceToHost));
  const bool ccBranch = bt.nIns > 0 && bt.nE > 0;
  t->numSubEdges = ccBranch ? 2 * ctr.subPairs : 0;
  t->numComponents = ccBranch ? ctr.numComponents : bt.nIns;
  if (t->numSubEdges > 0) {
    if (t->subSrc)
      DLP_HIP(hipMemcpy(t->subSrc, c->d.subSrc, sizeof(int) * t->numSubEdges,
                        hipMemcpyDeviceToHost));
    if (t->subDst)
      DLP_HIP(hipMemcpy(t->subDst, c->d.subDst, sizeof(int) * t->numSubEdges,
                        hipMemcpyDeviceToHost));
    if (t->subW)
      DLP_HIP(hipMemcpy(t->subW, c->d.subW, sizeof(float) * t->numSubEdges,
                        hipMemcpyDeviceToHost));
  }
  if (t->componentId_N)
    DLP_HIP(hipMemcpy(t->componentId_N, c->d.compId, sizeof(int) * N, hipMemcpyDeviceToHost));
  if (t->initLabels_N)
    DLP_HIP(hipMemcpy(t->initLabels_N, c->d.lab0, sizeof(float) * N, hipMemcpyDeviceToHost));
  return DLP_OK;
}

}  // namespace

/* ------------------------------------------------------------------------ */
/* per-batch scratch                                                        */
/* ------------------------------------------------------------------------ */
int64_t dlp_ctx_scratch_layout(dlp_context *c, long long ins, long long e, char *base,
                               bool bind) {
  DState &d = c->d;
  const long long N = std::max(c->N, 1);
  const long long mi = std::max(ins, 1LL), me = std::max(e, 1LL);
  const long long scanInts = (long long)std::max(
      {dlp_scan_scratch_ints(d.numWords), dlp_scan_scratch_ints(me), dlp_scan_scratch_ints(mi),
       dlp_scan_scratch_ints(N), dlp_radix_scratch_ints(mi), dlp_sparsify_scratch_ints(me)});
  int64_t off = 0;
  auto carve = [&](auto **p, long long count) {
    using Ptr = std::remove_reference_t<decltype(*p)>;
    if (bind) *p = base ? reinterpret_cast<Ptr>(base + off) : nullptr;
    off += ((int64_t)count * 4 + 255) / 256 * 256;  // every buffer holds 4-byte items
  };
  carve(&d.cand, N);
  carve(&d.q0, N);
  carve(&d.q1, N);
  carve(&d.touched, N);
  carve(&d.S, mi);
  carve(&d.subSrc, 2 * me);
  carve(&d.subDst, 2 * me);
  carve(&d.subW, 2 * me);
  for (float **p : {&d.p0, &d.p1, &d.red0, &d.red1, &d.W0, &d.W1}) carve(p, mi);
  for (int **p : {&d.keyA, &d.valA, &d.keyB, &d.valB, &d.keyC, &d.valC, &d.segStart, &d.segEnd})
    carve(p, mi);
  carve(&d.scratch, scanInts);
  return off;
}

int64_t dlp_ctx_batch_scratch_bytes(const dlp_context *c, int b) {
  return dlp_ctx_scratch_layout(const_cast<dlp_context *>(c), c->insOff[b + 1] - c->insOff[b],
                                c->eOff[b + 1] - c->eOff[b], nullptr, false);
}

int dlp_ctx_own_scratch(dlp_context *c, bool own) {
  if (!own) {
    if (c->scratchAlloc) {
      DLP_HIP(hipStreamSynchronize(c->stream));
      DLP_HIP(hipFree(c->scratchAlloc));
      c->deviceBytes -= c->scratchAllocBytes;
    }
    c->scratchAlloc = nullptr;
    c->scratchAllocBytes = 0;
    dlp_ctx_scratch_layout(c, c->maxIns, c->maxE, nullptr, true);
    return DLP_OK;
  }
  if (c->scratchAlloc) return DLP_OK;
  const int64_t bytes = dlp_ctx_scratch_layout(c, c->maxIns, c->maxE, nullptr, false);
  void *p = nullptr;
  const hipError_t e = hipMalloc(&p, (size_t)bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    dlp_set_errorf_("hipMalloc of %lld bytes of batch scratch failed (%s)", (long long)bytes,
                    hipGetErrorString(e));
    return DLP_ERR_NOMEM;
  }
  c->scratchAlloc = p;
  c->scratchAllocBytes = bytes;
  c->deviceBytes += bytes;
  dlp_ctx_scratch_layout(c, c->maxIns, c->maxE, (char *)p, true);
  return DLP_OK;
}

extern "C" int dlp_context_create(dlp_context **out, const dlp_config *cfg,
                                  const dlp_graph_view *g, const float *trueLabels,
                                  const dlp_batch_view *bv, void *stream) {
  if (!out) {
    dlp_set_error_("dlp_context_create: out is NULL");
    return DLP_ERR_INVALID;
  }
  *out = nullptr;
  int rc = validate_inputs(cfg, g, trueLabels, bv);
  if (rc != DLP_OK) return rc;
  if ((rc = need_device()) != DLP_OK) return rc;

  dlp_context *c = new dlp_context();
  c->cfg = *cfg;
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    dlp_set_error_("hipGetDevice failed");
    return DLP_ERR_HIP;
  }
  c->N = g->numNodes;
  c->B = cfg->numBatches;
  c->nnzFull = g->numEdges;
  const int N = c->N, B = c->B;

  /* tau: mean full-graph weight when not given (dynLP.cu:90-102, :186-190) */
  float tau = cfg->tau;
  if (tau < 0.0f) {
    double sum = 0.0;
    long long count = 0;
    for (long long i = 0; i < g->numEdges; ++i)
      if (g->colInd[i] >= 0) {
        sum += g->values[i];
        ++count;
      }
    tau = (g->numEdges == 0 || count == 0) ? 0.5f : (float)(sum / (double)count);
  }
  c->tau = tau;
  c->symmetric = csr_symmetric(g);

  /* batch offsets (host) and concatenated device lists */
  auto offs = [&](const int64_t *o, std::vector<int64_t> &dst) {
    dst.assign(o, o + B + 1);
    for (int b = B; b >= 0; --b) dst[b] -= o[0];
  };
  offs(bv->insertedOffsets, c->insOff);
  offs(bv->groundTruthOffsets, c->gtOff);
  offs(bv->deletedOffsets, c->delOff);
  offs(bv->edgeOffsets, c->eOff);
  int maxIns = 0, maxE = 0;
  for (int b = 0; b < B; ++b) {
    maxIns = std::max<int>(maxIns, (int)(c->insOff[b + 1] - c->insOff[b]));
    maxE = std::max<int>(maxE, (int)(c->eOff[b + 1] - c->eOff[b]));
  }
  /* rows each batch can change (the incremental CSR update's bound) */
  c->rowsBound.assign(B, 0);
  for (int b = 0; b < B; ++b) {
    auto rows = [&](const int64_t *o, const int *ids, int bb) {
      int64_t r = 0;
      for (int64_t i = o[bb]; i < o[bb + 1]; ++i) {
        const int x = ids[i];
        r += 1 + (x >= 0 && x < N ? g->rowPtr[x + 1] - g->rowPtr[x] : 0);
      }
      return r;
    };
    c->rowsBound[b] = rows(bv->insertedOffsets, bv->insertedNodeIds, b) +
                      rows(bv->deletedOffsets, bv->deletedNodeIds, b) +
                      rows(bv->groundTruthOffsets, bv->groundTruthNodeIds, b) +
                      (b > 0 ? rows(bv->deletedOffsets, bv->deletedNodeIds, b - 1) : 0);
  }
  std::vector<int> batchMap(N, -1);
  for (int b = 0; b < B; ++b)
    for (int64_t i = bv->insertedOffsets[b]; i < bv->insertedOffsets[b + 1]; ++i)
      batchMap[bv->insertedNodeIds[i]] = b;

  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      dlp_set_error_("hipStreamCreate failed");
      return DLP_ERR_HIP;
    }
    c->ownStream = true;
  }

  DState &d = c->d;
  d.N = N;
  d.numWords = (N + 31) / 32;
  const long long E = std::max<long long>(g->numEdges, 1);
  int *fRow, *fCol, *bm;
  float *fVal, *tl;
#define TRY(x)            \
  do {                    \
    rc = (x);             \
    if (rc != DLP_OK) {   \
      c->release();       \
      delete c;           \
      return rc;          \
    }                     \
  } while (0)
  TRY(c->upload(&fRow, g->rowPtr, N + 1));
  TRY(c->upload(&fCol, g->colInd, g->numEdges));
  TRY(c->upload(&fVal, g->values, g->numEdges));
  TRY(c->upload(&tl, trueLabels, N));
  TRY(c->upload(&bm, batchMap.data(), N));
  d.fRow = fRow;
  d.fCol = fCol;
  d.fVal = fVal;
  d.trueLab = tl;
  d.batchMap = bm;
  TRY(c->upload(&c->dIns, bv->insertedNodeIds + bv->insertedOffsets[0], c->insOff[B]));
  TRY(c->upload(&c->dGtIds, bv->groundTruthNodeIds + bv->groundTruthOffsets[0], c->gtOff[B]));
  TRY(c->upload(&c->dGtLab, bv->groundTruthLabels + bv->groundTruthOffsets[0], c->gtOff[B]));
  TRY(c->upload(&c->dDel, bv->deletedNodeIds + bv->deletedOffsets[0], c->delOff[B]));
  TRY(c->upload(&c->dESrc, bv->newEdgeSrc + bv->edgeOffsets[0], c->eOff[B]));
  TRY(c->upload(&c->dEDst, bv->newEdgeDst + bv->edgeOffsets[0], c->eOff[B]));
  TRY(c->upload(&c->dEW, bv->newEdgeWeight + bv->edgeOffsets[0], c->eOff[B]));

  TRY(c->alloc(&d.aLen, N));
  TRY(c->alloc(&d.aRow, N));
  TRY(c->alloc(&d.aCol, E));
  TRY(c->alloc(&d.aVal, E));
  TRY(c->alloc(&d.hStatus, N));
  TRY(c->alloc(&d.dStatus, N));
  TRY(c->alloc(&d.gtClass, N));
  TRY(c->alloc(&d.vflags, N));
  TRY(c->alloc(&d.lab0, N));
  TRY(c->alloc(&d.lab1, N));
  TRY(c->alloc(&d.iterStamp, N));
  TRY(c->alloc(&d.markStamp, N));
  TRY(c->alloc(&d.compId, N));
  TRY(c->alloc(&d.par, N));
  TRY(c->alloc(&d.rowStamp, N));
  TRY(c->alloc(&d.accState, N));
  TRY(c->alloc(&d.newBits, d.numWords));
  c->maxIns = maxIns;
  c->maxE = maxE;
  TRY(dlp_ctx_own_scratch(c, true));
  TRY(c->alloc(&d.ctr, 1));
  TRY(c->alloc(&d.records, B));
  c->iterSizesCap = std::max(1, std::min(cfg->maxIterations, 1 << 20));
  TRY(c->alloc(&c->iterSizes, c->iterSizesCap));
  /* LDS-resident propagation: LDS budget, core bookkeeping, default threshold
   * = the frontier whose rows (mean degree) fill the LDS once */
  {
    int ldsBytes = 0;
    TRY(dlp_local_lds_bytes(&ldsBytes));
    LocalArgs &l = c->local;
    l.ldsWords = ldsBytes / 4;
    l.numNodes = N;
    l.ownerAfter = c->tuning.localOwnerAfter;
    l.coreListCap = l.ovfCap = dlp_local_list_cap(l.ldsWords);
    TRY(c->alloc(&l.coreGen, N));
    TRY(c->alloc(&l.haloGen, N));
    TRY(c->alloc(&l.haloIdx, N));
    TRY(c->alloc(&l.coreIdx, N));
    TRY(c->alloc(&l.coreList, l.coreListCap));
    TRY(c->alloc(&l.ovf, l.ovfCap));
    DLP_HIP(hipMemsetAsync(l.coreGen, 0, (size_t)N * sizeof(int), c->stream));
    DLP_HIP(hipMemsetAsync(l.haloGen, 0, (size_t)N * sizeof(int), c->stream));
    DLP_HIP(hipMemsetAsync(d.ctr, 0, sizeof(DevCounters), c->stream));
    const long long meanDeg = (g->numEdges + N - 1) / N;
    c->tuning.localThreshold = dlp_local_default_threshold(l.ldsWords, meanDeg);
    c->tuning.localLdsBytes = c->localMaxBytes = ldsBytes;
  }
#undef TRY
  /* launch shapes: the degree-histogram heuristic, then the tuning table's
   * entry for this graph shape if there is one */
  {
    dlp_tuning h{};
    if ((rc = dlp_autotune_heuristic(g, &c->shape, &h)) != DLP_OK ||
        (rc = dlp_context_set_tuning(c, &h)) != DLP_OK) {
      c->release();
      delete c;
      return rc;
    }
    dlp_tuning_table_autoload();
    dlp_tuning t{};
    if (dlp_tuning_lookup_ctx(c->shape, &t) && dlp_context_set_tuning(c, &t) == DLP_OK)
      c->tuningSource = 1;
  }
  if (hipHostMalloc((void **)&c->pinCtr, sizeof(DevCounters), hipHostMallocDefault) !=
      hipSuccess) {
    c->release();
    delete c;
    dlp_set_error_("hipHostMalloc failed");
    return DLP_ERR_NOMEM;
  }
  c->evStart.resize(B);
  c->evEnd.resize(B);
  for (int b = 0; b < B; ++b) {
    if (hipEventCreate(&c->evStart[b]) != hipSuccess || hipEventCreate(&c->evEnd[b]) != hipSuccess) {
      c->release();
      delete c;
      dlp_set_error_("hipEventCreate failed");
      return DLP_ERR_HIP;
    }
  }
  c->records.assign(B, DevBatchRecord{});
  c->ran.assign(B, 0);
  rc = dlp_context_reset(c);
  if (rc != DLP_OK) {
    c->release();
    delete c;
    return rc;
  }
  DLP_HIP(hipStreamSynchronize(c->stream));
  *out = c;
  return DLP_OK;
}

extern "C" int dlp_context_destroy(dlp_context *c) {
  if (!c) return DLP_OK;
  c->release();
  delete c;
  return DLP_OK;
}

extern "C" int dlp_context_set_tuning(dlp_context *c, const dlp_tuning *t) {
  if (!c || !t) {
    dlp_set_error_("dlp_context_set_tuning: NULL argument");
    return DLP_ERR_INVALID;
  }
  dlp_tuning n = c->tuning;
  if (t->propGroupLanes) {
    int g = t->propGroupLanes;
    if (g != 4 && g != 8 && g != 16 && g != 32 && g != 64) {
      dlp_set_errorf_("propGroupLanes %d not in {4,8,16,32,64}", g);
      return DLP_ERR_INVALID;
    }
    n.propGroupLanes = g;
  }
  if (t->propBlocks) {
    if (t->propBlocks < 1 || t->propBlocks > 65536) {
      dlp_set_error_("propBlocks must be in [1, 65536]");
      return DLP_ERR_INVALID;
    }
    n.propBlocks = t->propBlocks;
  }
  if (t->chunkIterations) {
    if (t->chunkIterations < 1 || t->chunkIterations > 4096) {
      dlp_set_error_("chunkIterations must be in [1, 4096]");
      return DLP_ERR_INVALID;
    }
    n.chunkIterations = t->chunkIterations;
  }
  if (t->localThreshold) n.localThreshold = t->localThreshold < 0 ? -1 : t->localThreshold;
  if (t->localLdsBytes) {
    if (t->localLdsBytes < 4096 || t->localLdsBytes > c->localMaxBytes) {
      dlp_set_errorf_("localLdsBytes must be in [4096, %d]", c->localMaxBytes);
      return DLP_ERR_INVALID;
    }
    n.localLdsBytes = t->localLdsBytes & ~255;
  }
  if (t->localShape) {
    if (t->localShape < 1 || t->localShape > dlp_local_shape_count()) {
      dlp_set_errorf_("localShape must be in [1, %d]", dlp_local_shape_count());
      return DLP_ERR_INVALID;
    }
    n.localShape = t->localShape;
  }
  if (t->localOwnerAfter) n.localOwnerAfter = t->localOwnerAfter < 0 ? -1 : t->localOwnerAfter;
  if (t->waveRowMin) {
    if (t->waveRowMin < 1 || t->waveRowMin > 64) {
      dlp_set_error_("waveRowMin must be in [1, 64]");
      return DLP_ERR_INVALID;
    }
    n.waveRowMin = t->waveRowMin;
  }
  if (t->maskOwnerMax) {
    if (t->maskOwnerMax != -1 && (t->maskOwnerMax < 1 || t->maskOwnerMax > kMaskOwnerMax)) {
      dlp_set_errorf_("maskOwnerMax must be -1 or in [1, %d]", kMaskOwnerMax);
      return DLP_ERR_INVALID;
    }
    n.maskOwnerMax = t->maskOwnerMax;
  }
  if (t->blockRowMin) {
    if (t->blockRowMin != -1 && t->blockRowMin < 64) {
      dlp_set_error_("blockRowMin must be -1 or at least 64");
      return DLP_ERR_INVALID;
    }
    n.blockRowMin = t->blockRowMin;
  }
  c->tuning = n;
  c->local.ldsWords = n.localLdsBytes / 4;
  c->local.shape = n.localShape - 1;
  c->local.ownerAfter = n.localOwnerAfter;
  return DLP_OK;
}

extern "C" int dlp_context_get_tuning(const dlp_context *c, dlp_tuning *t) {
  if (!c || !t) {
    dlp_set_error_("dlp_context_get_tuning: NULL argument");
    return DLP_ERR_INVALID;
  }
  *t = c->tuning;
  return DLP_OK;
}

extern "C" int dlp_context_tau(const dlp_context *c, float *tau) {
  if (!c || !tau) {
    dlp_set_error_("dlp_context_tau: NULL argument");
    return DLP_ERR_INVALID;
  }
  *tau = c->tau;
  return DLP_OK;
}

extern "C" int dlp_context_device_bytes(const dlp_context *c, int64_t *bytes) {
  if (!c || !bytes) {
    dlp_set_error_("dlp_context_device_bytes: NULL argument");
    return DLP_ERR_INVALID;
  }
  *bytes = c->deviceBytes;
  return DLP_OK;
}

extern "C" int dlp_context_reset(dlp_context *c) {
  if (!c) {
    dlp_set_error_("dlp_context_reset: NULL context");
    return DLP_ERR_STATE;
  }
  int rc = dlp_reset_state(c->d, c->stream);
  if (rc != DLP_OK) return rc;
  c->stampBound = 1;  // k_reset sets stampBase to 1 and keeps localGen
  c->nextBatch = 0;
  c->lastCCBatch = -1;
  std::fill(c->ran.begin(), c->ran.end(), 0);
  return DLP_OK;
}

extern "C" int dlp_context_set_trace(dlp_context *c, dlp_trace *t) {
  if (!c) {
    dlp_set_error_("dlp_context_set_trace: NULL context");
    return DLP_ERR_STATE;
  }
  c->trace = t;
  return DLP_OK;
}

extern "C" int dlp_context_run_batch(dlp_context *c, int b) {
  if (!c) {
    dlp_set_error_("dlp_context_run_batch: NULL context");
    return DLP_ERR_STATE;
  }
  if (b != c->nextBatch || b < 0 || b >= c->B) {
    dlp_set_errorf_("batch %d requested, next batch is %d (reset to start over)", b,
                    c->nextBatch);
    return DLP_ERR_STATE;
  }
  hipStream_t s = c->stream;
  DBatch bt = dlp_ctx_make_batch(c, b);
  int rc, markTag = 0;
  if ((rc = dlp_ctx_counter_guard(c, s)) != DLP_OK) return rc;
  DLP_HIP(hipEventRecord(c->evStart[b], s));
  if ((rc = dlp_ctx_batch_pre(c, b, bt, &markTag)) != DLP_OK) return rc;
  if (c->trace && (rc = trace_pre_propagation(c, bt, markTag)) != DLP_OK) return rc;
  if ((rc = run_propagation(c, dlp_ctx_prop_args(c), c->cfg.maxIterations)) != DLP_OK) return rc;
  if ((rc = dlp_ctx_batch_post(c, b, bt)) != DLP_OK) return rc;
  DLP_HIP(hipEventRecord(c->evEnd[b], s));
  if (c->trace) {
    DLP_HIP(hipStreamSynchronize(s));
    DevBatchRecord r;
    DLP_HIP(hipMemcpy(&r, c->d.records + b, sizeof(r), hipMemcpyDeviceToHost));
    c->trace->numFrontierSizes = r.iterations;
    int nc = std::min(std::min(r.iterations, c->trace->maxFrontierSizes), c->iterSizesCap);
    if (c->trace->frontierSizes && nc > 0)
      DLP_HIP(hipMemcpy(c->trace->frontierSizes, c->iterSizes, sizeof(int) * nc,
                        hipMemcpyDeviceToHost));
  }
  return DLP_OK;
}

int dlp_ctx_counter_guard(dlp_context *c, hipStream_t s) {
  const CounterGuardPlan p =
      dlp_guard_plan(c->stampBound, c->genBound, c->cfg.maxIterations);
  int rc;
  if (p.rebaseStamps && (rc = dlp_rebase_stamps(c->d, s)) != DLP_OK) return rc;
  if (p.restartGens && (rc = dlp_restart_local_gens(c->local, c->d.ctr, s)) != DLP_OK) return rc;
  c->stampBound = p.stampBound;
  c->genBound = p.genBound;
  return DLP_OK;
}

extern "C" int dlp_counter_guard_plan(int64_t stampBound, int64_t genBound, int maxIterations,
                                        int *rebaseStamps, int *restartGens,
                                        int64_t *stampBoundAfter, int64_t *genBoundAfter) {
  if (!rebaseStamps || !restartGens || !stampBoundAfter || !genBoundAfter ||
      maxIterations < 0 || maxIterations > kMaxIterationsLimit) {
    dlp_set_error_("dlp_counter_guard_plan: bad argument");
    return DLP_ERR_INVALID;
  }
  const CounterGuardPlan p = dlp_guard_plan(stampBound, genBound, maxIterations);
  *rebaseStamps = p.rebaseStamps;
  *restartGens = p.restartGens;
  *stampBoundAfter = p.stampBound;
  *genBoundAfter = p.genBound;
  return DLP_OK;
}

int dlp_ctx_set_counter_bases(dlp_context *c, int stampBase, int localGen) {
  if (stampBase < 1 || localGen < 0 || (long long)localGen >= kGenLimit) {
    dlp_set_error_("counter bases: stampBase must be >= 1 and localGen in [0, 2^31 - 2)");
    return DLP_ERR_INVALID;
  }
  DLP_HIP(hipStreamSynchronize(c->stream));
  DevCounters ctr;
  DLP_HIP(hipMemcpy(&ctr, c->d.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
  if (stampBase < ctr.stampBase || localGen < ctr.localGen) {
    dlp_set_errorf_("counter bases may only grow (stampBase %d, localGen %d now)", ctr.stampBase,
                    ctr.localGen);
    return DLP_ERR_INVALID;
  }
  ctr.stampBase = stampBase;
  ctr.localGen = localGen;
  DLP_HIP(hipMemcpy(c->d.ctr, &ctr, sizeof(ctr), hipMemcpyHostToDevice));
  c->stampBound = stampBase;
  c->genBound = localGen;
  return DLP_OK;
}

int dlp_ctx_counter_bases(dlp_context *c, int *stampBase, int *localGen) {
  DLP_HIP(hipStreamSynchronize(c->stream));
  DevCounters ctr;
  DLP_HIP(hipMemcpy(&ctr, c->d.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
  if (stampBase) *stampBase = ctr.stampBase;
  if (localGen) *localGen = ctr.localGen;
  return DLP_OK;
}

extern "C" int dlp_graph_symmetric(const dlp_graph_view *g, int *symmetric) {
  if (!g || !symmetric || g->numNodes < 0 || g->numEdges < 0 ||
      (g->numNodes > 0 && !g->rowPtr) || (g->numEdges > 0 && !g->colInd)) {
    dlp_set_error_("dlp_graph_symmetric: bad argument");
    return DLP_ERR_INVALID;
  }
  for (int u = 0; u < g->numNodes; ++u)
    if (g->rowPtr[u] > g->rowPtr[u + 1] || g->rowPtr[u + 1] > g->numEdges) {
      dlp_set_errorf_("dlp_graph_symmetric: rowPtr not monotone within numEdges at row %d", u);
      return DLP_ERR_INVALID;
    }
  *symmetric = csr_symmetric(g) ? 1 : 0;
  return DLP_OK;
}

extern "C" int dlp_context_row_update(dlp_context *c, int mode, int *symmetric) {
  if (!c || mode > 2) {
    dlp_set_error_("dlp_context_row_update: NULL context or mode not in {0, 1, 2}");
    return DLP_ERR_INVALID;
  }
  if (mode >= 0) c->rowUpdate = mode;
  if (symmetric) *symmetric = c->symmetric ? 1 : 0;
  return DLP_OK;
}

extern "C" int dlp_context_set_counter_bases(dlp_context *c, int stampBase, int localGen) {
  if (!c) {
    dlp_set_error_("dlp_context_set_counter_bases: NULL context");
    return DLP_ERR_STATE;
  }
  return dlp_ctx_set_counter_bases(c, stampBase, localGen);
}

extern "C" int dlp_context_counter_bases(dlp_context *c, int *stampBase, int *localGen) {
  if (!c) {
    dlp_set_error_("dlp_context_counter_bases: NULL context");
    return DLP_ERR_STATE;
  }
  return dlp_ctx_counter_bases(c, stampBase, localGen);
}

int dlp_ctx_batch_pre(dlp_context *c, int b, const DBatch &bt, int *markTag) {
  hipStream_t s = c->stream;
  const DState &d = c->d;
  int rc;
  if ((rc = dlp_batch_begin(d, bt, s)) != DLP_OK) return rc;
  *markTag = ++c->markSeq;
  /* only the rows the batch can change, when they are few and the full
   * graph is symmetric (rows naming a vertex = rows of its neighbours) */
  const bool listed = c->symmetric && c->rowUpdate != 1 &&
                      (c->rowUpdate == 2 || c->rowsBound[b] <= (int64_t)c->N / 4);
  const long long rowBound = listed ? c->rowsBound[b] : -1;
  if ((rc = dlp_csr_update(d, bt, *markTag, rowBound, s)) != DLP_OK) return rc;
  if (bt.nIns > 0 && bt.nE > 0) {
    if (c->lastCCBatch >= 0) {
      DBatch prev = dlp_ctx_make_batch(c, c->lastCCBatch);
      if ((rc = dlp_clear_component_ids(d, prev.ins, prev.nIns, s)) != DLP_OK) return rc;
    }
    if ((rc = dlp_sorted_inserted(d, s)) != DLP_OK) return rc;
    if ((rc = dlp_sparsify(d, bt, c->tau, *markTag, s)) != DLP_OK) return rc;
    if ((rc = dlp_components(d, bt, s)) != DLP_OK) return rc;
    if ((rc = dlp_label_init(d, bt, s)) != DLP_OK) return rc;
    c->lastCCBatch = b;
  } else if (bt.nIns > 0) {
    if ((rc = dlp_labels_no_edges(d, bt, s)) != DLP_OK) return rc;
  }
  return dlp_frontier0(d, c->N, s);
}

int dlp_ctx_run_propagation(dlp_context *c, int startK) {
  return run_propagation(c, dlp_ctx_prop_args(c), c->cfg.maxIterations, startK);
}

PropArgs dlp_ctx_prop_args(const dlp_context *c) {
  const DState &d = c->d;
  PropArgs a{};
  a.queueCap = c->N;
  a.rowInfo = d.aRow;
  a.rowBeg = d.fRow;
  a.rowEnd = nullptr;
  a.rowLen = d.aLen;
  a.colInd = d.aCol;
  a.values = d.aVal;
  a.status = nullptr;  // every remaining entry points at a non-deleted vertex
  a.lab0 = d.lab0;
  a.lab1 = d.lab1;
  a.iterStamp = d.iterStamp;
  a.q0 = d.q0;
  a.q1 = d.q1;
  a.touched = d.touched;
  a.ctr = d.ctr;
  /* per-iteration frontier sizes only for a trace: a global store per
   * iteration would also hold up the LDS kernels' work-group barriers */
  a.iterSizes = c->trace ? c->iterSizes : nullptr;
  a.iterSizesCap = c->trace ? c->iterSizesCap : 0;
  a.profile = c->profiling >= 2;
  a.delta = c->cfg.delta;
  a.waveRowMin = c->tuning.waveRowMin;
  a.blockRowMin = c->tuning.blockRowMin > 0 ? c->tuning.blockRowMin : 0;
  /* the active CSR keeps exactly the entries whose two ends are present
   * (holes for this batch's deletions on both sides), so it is symmetric
   * when the full graph is */
  a.symmetric = c->symmetric ? 1 : 0;
  return a;
}

void dlp_ctx_advance_host(dlp_context *c, int b, const DBatch &bt) {
  ++c->markSeq;
  if (bt.nIns > 0 && bt.nE > 0) c->lastCCBatch = b;
  c->ran[b] = 1;
  c->nextBatch = b + 1;
}

int dlp_ctx_batch_post(dlp_context *c, int b, const DBatch &bt) {
  int rc = dlp_batch_end(c->d, bt, c->stream);
  if (rc != DLP_OK) return rc;
  c->ran[b] = 1;
  c->nextBatch = b + 1;
  return DLP_OK;
}

extern "C" int dlp_context_set_profiling(dlp_context *c, int enable) {
  if (!c) {
    dlp_set_error_("dlp_context_set_profiling: NULL context");
    return DLP_ERR_STATE;
  }
  DLP_HIP(hipStreamSynchronize(c->stream));
  DevCounters ctr;
  DLP_HIP(hipMemcpy(&ctr, c->d.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
  ctr.profEdges = ctr.profChangedEdges = ctr.profAppended = 0;
  ctr.profLocalUpdates = ctr.profLocalEdges = ctr.profLocalChangedEdges = 0;
  ctr.profLocalAppended = ctr.profLocalOwnerUpdates = ctr.profLocalBuilds = 0;
  ctr.profLocalWaveUpdates = 0;
  DLP_HIP(hipMemcpy(c->d.ctr, &ctr, sizeof(ctr), hipMemcpyHostToDevice));
  c->profiling = enable;
  c->timing.perChunk = enable == 1;
  c->timing.used = 0;
  c->timing.launches = c->timing.localLaunches = 0;
  c->timing.ms = c->timing.localMs = 0.0;
  return DLP_OK;
}

extern "C" int dlp_context_prop_profile(dlp_context *c, dlp_prop_profile *out) {
  if (!c || !out) {
    dlp_set_error_("dlp_context_prop_profile: NULL argument");
    return DLP_ERR_INVALID;
  }
  DLP_HIP(hipStreamSynchronize(c->stream));
  DevCounters ctr;
  DLP_HIP(hipMemcpy(&ctr, c->d.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
  out->launches = c->timing.launches;
  out->kernelMs = c->timing.ms;
  out->edges = (int64_t)ctr.profEdges;
  out->changedEdges = (int64_t)ctr.profChangedEdges;
  out->appended = (int64_t)ctr.profAppended;
  out->localLaunches = c->timing.localLaunches;
  out->localKernelMs = c->timing.localMs;
  out->localUpdates = (int64_t)ctr.profLocalUpdates;
  out->localEdges = (int64_t)ctr.profLocalEdges;
  out->localChangedEdges = (int64_t)ctr.profLocalChangedEdges;
  out->localAppended = (int64_t)ctr.profLocalAppended;
  out->localOwnerUpdates = (int64_t)ctr.profLocalOwnerUpdates;
  out->localBuilds = (int64_t)ctr.profLocalBuilds;
  out->localWaveUpdates = (int64_t)ctr.profLocalWaveUpdates;
  return DLP_OK;
}

extern "C" int dlp_context_run_all(dlp_context *c) {
  int rc = dlp_context_reset(c);
  if (rc != DLP_OK) return rc;
  for (int b = 0; b < c->B; ++b)
    if ((rc = dlp_context_run_batch(c, b)) != DLP_OK) return rc;
  return DLP_OK;
}

extern "C" int dlp_context_synchronize(dlp_context *c) {
  if (!c) {
    dlp_set_error_("dlp_context_synchronize: NULL context");
    return DLP_ERR_STATE;
  }
  DLP_HIP(hipStreamSynchronize(c->stream));
  return DLP_OK;
}

extern "C" int dlp_context_read_labels(dlp_context *c, float *labels) {
  if (!c || !labels) {
    dlp_set_error_("dlp_context_read_labels: NULL argument");
    return DLP_ERR_INVALID;
  }
  DLP_HIP(hipMemcpyAsync(labels, c->d.lab0, sizeof(float) * c->N, hipMemcpyDeviceToHost,
                         c->stream));
  DLP_HIP(hipStreamSynchronize(c->stream));
  return DLP_OK;
}

extern "C" int dlp_context_batch_stats(dlp_context *c, int b, dlp_batch_stats *out) {
  if (!c || !out || b < 0 || b >= c->B) {
    dlp_set_error_("dlp_context_batch_stats: bad argument");
    return DLP_ERR_INVALID;
  }
  if (!c->ran[b]) {
    dlp_set_errorf_("batch %d has not run since the last reset", b);
    return DLP_ERR_STATE;
  }
  int rc = dlp_ctx_fill_stats(c, b, out);
  if (rc != DLP_OK) return rc;
  float ms = 0.0f;
  DLP_HIP(hipEventElapsedTime(&ms, c->evStart[b], c->evEnd[b]));
  out->timeMs = ms;
  return DLP_OK;
}

int dlp_ctx_fill_stats(dlp_context *c, int b, dlp_batch_stats *out) {
  DLP_HIP(hipStreamSynchronize(c->stream));
  DevBatchRecord r;
  DLP_HIP(hipMemcpy(&r, c->d.records + b, sizeof(r), hipMemcpyDeviceToHost));
  DBatch bt = dlp_ctx_make_batch(c, b);
  out->batch = b;
  out->numInserted = bt.nIns;
  out->numDeleted = bt.nDel;
  out->numNewEdges = bt.nE;
  const bool cc = bt.nIns > 0 && bt.nE > 0;
  out->numSubEdges = cc ? r.numSub : 0;
  out->numComponents = cc ? r.numComponents : bt.nIns;
  out->numMarked = r.numCand;
  out->frontier0 = r.frontier0;
  out->iterations = r.iterations;
  out->numTouched = r.touched;
  out->updates = (int64_t)r.updates;
  out->timeMs = 0.0f;
  out->accuracy = r.total > 0 ? (float)r.correct / (float)r.total : 1.0f;
  return DLP_OK;
}

/* ------------------------------------------------------------------------ */
/* runDynLP                                                                 */
/* ------------------------------------------------------------------------ */
extern "C" int dlp_run_dynlp(const dlp_config *cfg, const dlp_graph_view *g,
                             const float *trueLabels, const dlp_batch_view *bv,
                             float *finalLabels, float *batchTimes, float *batchAccuracy) {
  if (!finalLabels) {
    dlp_set_error_("dlp_run_dynlp: finalLabels is NULL");
    return DLP_ERR_INVALID;
  }
  dlp_context *c = nullptr;
  int rc = dlp_context_create(&c, cfg, g, trueLabels, bv, nullptr);
  if (rc != DLP_OK) return rc;
  if (cfg->tau < 0.0f) std::printf("Auto tau (average edge weight): %g\n", c->tau);
  for (int b = 0; b < c->B && rc == DLP_OK; ++b) {
    rc = dlp_context_run_batch(c, b);
    if (rc != DLP_OK) break;
    dlp_batch_stats st;
    rc = dlp_context_batch_stats(c, b, &st);
    if (rc != DLP_OK) break;
    if (batchTimes) batchTimes[b] = st.timeMs;
    if (batchAccuracy) batchAccuracy[b] = st.accuracy;
    if (cfg->verbose) {
      DBatch bt = dlp_ctx_make_batch(c, b);
      std::printf("--- Batch %d ---\n  Inserting %d vertices (%d GT), deleting %d\n", b, bt.nIns,
                  bt.nGt, bt.nDel);
      if (bt.nIns > 0 && bt.nE > 0) {
        std::printf("  Step 1: %d connected components, %d subgraph edges (tau=%g)\n",
                    st.numComponents, st.numSubEdges, c->tau);
        std::printf("  Step 2: Labels initialized for %d new vertices\n", bt.nIns);
      }
      std::printf("  Step 3: Converged in %d iterations\n", st.iterations);
      std::printf("  Accuracy: %g%%, Time: %g ms\n\n", st.accuracy * 100.0f, st.timeMs);
      std::fflush(stdout);
    }
  }
  if (rc == DLP_OK) rc = dlp_context_read_labels(c, finalLabels);
  std::string err = g_lastError;
  dlp_context_destroy(c);
  g_lastError = err;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* device-memory plumbing                                                   */
/* ------------------------------------------------------------------------ */
extern "C" int dlp_device_malloc(void **ptr, size_t bytes) {
  if (!ptr) {
    dlp_set_error_("dlp_device_malloc: NULL argument");
    return DLP_ERR_INVALID;
  }
  int rc = need_device();
  if (rc != DLP_OK) return rc;
  DLP_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return DLP_OK;
}

extern "C" int dlp_device_free(void *ptr) {
  if (ptr) DLP_HIP(hipFree(ptr));
  return DLP_OK;
}

extern "C" int dlp_copy_to_device(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return DLP_OK;
  if (!dst || !src) {
    dlp_set_error_("dlp_copy_to_device: NULL argument");
    return DLP_ERR_INVALID;
  }
  DLP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return DLP_OK;
}

extern "C" int dlp_copy_to_host(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return DLP_OK;
  if (!dst || !src) {
    dlp_set_error_("dlp_copy_to_host: NULL argument");
    return DLP_ERR_INVALID;
  }
  DLP_HIP(hipDeviceSynchronize());
  DLP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return DLP_OK;
}

namespace {
__global__ void k_copy_vec16(uint4 *dst, const uint4 *src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
__global__ void k_copy_bytes(unsigned char *dst, const unsigned char *src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

extern "C" int dlp_copy_device(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return DLP_OK;
  if (!dst || !src) {
    dlp_set_error_("dlp_copy_device: NULL argument");
    return DLP_ERR_INVALID;
  }
  size_t done = 0;
  if ((((uintptr_t)dst | (uintptr_t)src) & 15) == 0 && bytes >= 16) {
    const size_t n = bytes / 16;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy_vec16, dim3(grid), dim3(256), 0, 0, (uint4 *)dst,
                       (const uint4 *)src, n);
    DLP_LAUNCHED("k_copy_vec16");
    done = n * 16;
  }
  if (done < bytes) {
    const size_t n = bytes - done;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy_bytes, dim3(grid), dim3(256), 0, 0, (unsigned char *)dst + done,
                       (const unsigned char *)src + done, n);
    DLP_LAUNCHED("k_copy_bytes");
  }
  DLP_HIP(hipDeviceSynchronize());
  return DLP_OK;
}
