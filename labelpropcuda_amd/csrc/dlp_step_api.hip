/*
 * dlp_step_api.hip -- C-ABI of the five step launchers (standard CSR input).
 *
 * Each entry point keeps the reference launcher's arguments and semantics
 * (see dynlp_c.h).  Scratch comes from one workspace per device that grows
 * on demand and is reused across calls, instead of the reference's
 * malloc/free per call; calls that use it are serialised by a mutex and end
 * with a stream synchronisation, so the workspace is free when they return.
 */
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "dlp_kernels.h"

namespace {

struct Workspace {
  std::vector<void *> bufs;
  std::vector<size_t> sizes;
  DevCounters *ctr = nullptr;
  DevCounters *pin = nullptr;
  /* LDS-resident propagation: core generations must outlive counter resets */
  int localGen = 0;
  void *coreGenBuf = nullptr; /* the slot buffer localGen refers to */
  size_t coreGenBytes = 0;
  int ldsBytes = 0;
  int *pinFlags = nullptr; /* pinned changed flags of the IrLP loop */
};

std::mutex g_wsMutex;
std::map<int, Workspace> g_ws;

/* buffer `slot` of at least `bytes`, grown when too small */
int ws_get(Workspace &w, int slot, size_t bytes, void **out) {
  if ((int)w.bufs.size() <= slot) {
    w.bufs.resize(slot + 1, nullptr);
    w.sizes.resize(slot + 1, 0);
  }
  bytes = std::max<size_t>(bytes, 256);
  if (w.sizes[slot] < bytes) {
    if (w.bufs[slot]) (void)hipFree(w.bufs[slot]);
    w.bufs[slot] = nullptr;
    w.sizes[slot] = 0;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      dlp_set_errorf_("workspace hipMalloc of %zu bytes failed", bytes);
      return DLP_ERR_NOMEM;
    }
    w.bufs[slot] = p;
    w.sizes[slot] = bytes;
  }
  *out = w.bufs[slot];
  return DLP_OK;
}

int ws_open(Workspace **out) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    dlp_set_error_("no GPU visible to HIP");
    return DLP_ERR_NODEVICE;
  }
  int dev = 0;
  DLP_HIP(hipGetDevice(&dev));
  Workspace &w = g_ws[dev];
  if (!w.ctr) {
    void *p = nullptr;
    DLP_HIP(hipMalloc(&p, sizeof(DevCounters)));
    w.ctr = (DevCounters *)p;
    DLP_HIP(hipHostMalloc((void **)&w.pin, sizeof(DevCounters), hipHostMallocDefault));
  }
  *out = &w;
  return DLP_OK;
}

/* counters <- zero, stampBase = 1 */
int ws_reset_counters(Workspace &w, hipStream_t s) {
  DevCounters z = {};
  z.stampBase = 1;
  z.localGen = w.localGen;
  *w.pin = z;
  DLP_HIP(hipMemcpyAsync(w.ctr, w.pin, sizeof(DevCounters), hipMemcpyHostToDevice, s));
  DLP_HIP(hipStreamSynchronize(s));
  return DLP_OK;
}

int ws_read_counters(Workspace &w, hipStream_t s) {
  DLP_HIP(hipMemcpyAsync(w.pin, w.ctr, sizeof(DevCounters), hipMemcpyDeviceToHost, s));
  DLP_HIP(hipStreamSynchronize(s));
  return DLP_OK;
}

/* ascending list of the inserted ids; fails unless they are distinct (the
 * list is only complete then) */
int sorted_distinct(Workspace &w, const int *inserted, int numInserted, int numNodes,
                    uint32_t *bits, int *sorted, int *scratch, hipStream_t s, const char *who) {
  int rc = dlp_k_sorted_list(inserted, numInserted, numNodes, bits, sorted, w.ctr, scratch, s);
  if (rc != DLP_OK) return rc;
  if ((rc = ws_read_counters(w, s)) != DLP_OK) return rc;
  if (w.pin->numNew != numInserted) {
    dlp_set_errorf_("%s: insertedNodes must be distinct", who);
    return DLP_ERR_INVALID;
  }
  return DLP_OK;
}

#define WS_BUF(T, name, slot, count)                                    \
  T *name = nullptr;                                                    \
  do {                                                                  \
    void *p_ = nullptr;                                                 \
    int rc_ = ws_get(*w, slot, sizeof(T) * (size_t)(count), &p_);       \
    if (rc_ != DLP_OK) return rc_;                                      \
    name = (T *)p_;                                                     \
  } while (0)

}  // namespace

extern "C" int dlp_mark_deletions(int *rowPtr, int *colInd, int *vertexStatus, int *isAffected,
                                  const int *deletedNodes, int numDeleted, int numNodes,
                                  void *stream) {
  if (numDeleted == 0) return DLP_OK; /* sparsifySubgraph.cu:188 */
  if (numDeleted < 0 || numNodes < 0 || !rowPtr || !colInd || !vertexStatus || !isAffected ||
      !deletedNodes) {
    dlp_set_error_("dlp_mark_deletions: bad argument");
    return DLP_ERR_INVALID;
  }
  return dlp_k_mark_deletions(rowPtr, colInd, vertexStatus, isAffected, deletedNodes, numDeleted,
                              numNodes, (hipStream_t)stream);
}

extern "C" int dlp_build_sparsified_subgraph(
    int *rowPtr, int *colInd, float *values, int *isAffected, const int *vertexStatus,
    const int *newEdgeSrc, const int *newEdgeDst, const float *newEdgeWeight, int numNewEdges,
    const int *insertedNodes, int numInserted, const int *isNewVertex, float tau, int *subRowPtr,
    int *subColInd, float *subValues, int *subEdgeSrc, int *numSubNodes, int *numSubEdges,
    void *stream) {
  (void)rowPtr;
  (void)colInd;
  (void)values;
  (void)vertexStatus;
  (void)insertedNodes;
  (void)subRowPtr;
  if (!numSubNodes || !numSubEdges) {
    dlp_set_error_("dlp_build_sparsified_subgraph: NULL output count");
    return DLP_ERR_INVALID;
  }
  if (numNewEdges == 0) { /* sparsifySubgraph.cu:213-217 */
    *numSubNodes = numInserted;
    *numSubEdges = 0;
    return DLP_OK;
  }
  if (numNewEdges < 0 || !isAffected || !newEdgeSrc || !newEdgeDst || !newEdgeWeight ||
      !isNewVertex || !subColInd || !subValues) {
    dlp_set_error_("dlp_build_sparsified_subgraph: bad argument");
    return DLP_ERR_INVALID;
  }
  *numSubNodes = numInserted;
  std::lock_guard<std::mutex> lock(g_wsMutex);
  Workspace *w;
  int rc = ws_open(&w);
  if (rc != DLP_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = ws_reset_counters(*w, s)) != DLP_OK) return rc;
  WS_BUF(int, scratch, 0, dlp_sparsify_scratch_ints(numNewEdges));
  rc = dlp_k_sparsify(newEdgeSrc, newEdgeDst, newEdgeWeight, numNewEdges, numInserted,
                      isNewVertex, tau, isAffected, subEdgeSrc, subColInd, subValues,
                      &w->ctr->scanTotal, scratch, &w->ctr->errorFlag, s);
  if (rc != DLP_OK) return rc;
  if ((rc = ws_read_counters(*w, s)) != DLP_OK) return rc;
  if (w->pin->errorFlag) {
    dlp_set_errorf_("dlp_build_sparsified_subgraph: device error flags 0x%x",
                    w->pin->errorFlag);
    return DLP_ERR_INTERNAL;
  }
  *numSubEdges = 2 * w->pin->scanTotal;
  return DLP_OK;
}

extern "C" int dlp_connected_components(const int *edgeSrc, const int *edgeDst, int numEdges,
                                        const int *insertedNodes, int numInserted,
                                        int *componentId, int *numComponents, int numNodes,
                                        void *stream) {
  if (!numComponents) {
    dlp_set_error_("dlp_connected_components: NULL numComponents");
    return DLP_ERR_INVALID;
  }
  if (numInserted == 0) { /* connectedComponents.cu:167-170 */
    *numComponents = 0;
    return DLP_OK;
  }
  if (numInserted < 0 || numEdges < 0 || numNodes <= 0 || !insertedNodes || !componentId ||
      (numEdges > 0 && (!edgeSrc || !edgeDst))) {
    dlp_set_error_("dlp_connected_components: bad argument");
    return DLP_ERR_INVALID;
  }
  hipStream_t s = (hipStream_t)stream;
  if (numEdges == 0) { /* connectedComponents.cu:191-205 */
    int rc = dlp_k_fill_int(componentId, -1, numNodes, s);
    if (rc != DLP_OK) return rc;
    *numComponents = numInserted;
    DLP_HIP(hipStreamSynchronize(s));
    return DLP_OK;
  }
  std::lock_guard<std::mutex> lock(g_wsMutex);
  Workspace *w;
  int rc = ws_open(&w);
  if (rc != DLP_OK) return rc;
  if ((rc = ws_reset_counters(*w, s)) != DLP_OK) return rc;
  const int words = (numNodes + 31) / 32;
  WS_BUF(int, par, 1, numNodes);
  WS_BUF(uint32_t, bits, 2, words);
  WS_BUF(int, sorted, 3, numInserted);
  WS_BUF(int, scratch, 0,
         std::max(dlp_scan_scratch_ints(words), dlp_scan_scratch_ints(numInserted)));
  if ((rc = sorted_distinct(*w, insertedNodes, numInserted, numNodes, bits, sorted, scratch, s,
                            "dlp_connected_components")) != DLP_OK)
    return rc;
  rc = dlp_k_cc(edgeSrc, edgeDst, numEdges, insertedNodes, numInserted, componentId, numNodes, par,
                sorted, w->ctr, scratch, s);
  if (rc != DLP_OK) return rc;
  if ((rc = ws_read_counters(*w, s)) != DLP_OK) return rc;
  if (w->pin->errorFlag) {
    dlp_set_errorf_("dlp_connected_components: edge endpoint is not an inserted vertex "
                    "(flags 0x%x)",
                    w->pin->errorFlag);
    return DLP_ERR_INVALID;
  }
  *numComponents = w->pin->numComponents;
  return DLP_OK;
}

extern "C" int dlp_label_initialization(const int *rowPtr, const int *colInd,
                                        const float *values, float *labels,
                                        const int *vertexStatus, const int *componentId,
                                        int numComponents, const int *insertedNodes,
                                        int numInserted, const int *gtClass, int numNodes,
                                        void *stream) {
  (void)vertexStatus;
  if (numInserted == 0) return DLP_OK; /* labelInitialization.cu:155 */
  if (numInserted < 0 || numNodes <= 0 || numComponents < 0 || !rowPtr || !labels ||
      !componentId || !insertedNodes || !gtClass) {
    dlp_set_error_("dlp_label_initialization: bad argument");
    return DLP_ERR_INVALID;
  }
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> lock(g_wsMutex);
  Workspace *w;
  int rc = ws_open(&w);
  if (rc != DLP_OK) return rc;
  if ((rc = ws_reset_counters(*w, s)) != DLP_OK) return rc;
  const int words = (numNodes + 31) / 32;
  const long long ni = numInserted, nc = std::max(numComponents, 1);
  WS_BUF(uint32_t, bits, 2, words);
  WS_BUF(int, sorted, 3, ni);
  WS_BUF(float, p0, 4, ni);
  WS_BUF(float, p1, 5, ni);
  WS_BUF(int, keyA, 6, ni);
  WS_BUF(int, valA, 7, ni);
  WS_BUF(int, keyB, 8, ni);
  WS_BUF(int, valB, 9, ni);
  WS_BUF(int, keyC, 10, ni);
  WS_BUF(int, valC, 11, ni);
  WS_BUF(int, segStart, 12, nc);
  WS_BUF(int, segEnd, 13, nc);
  WS_BUF(float, red0, 14, ni);
  WS_BUF(float, red1, 15, ni);
  WS_BUF(float, W0, 16, nc);
  WS_BUF(float, W1, 17, nc);
  WS_BUF(int, scratch, 0,
         std::max({dlp_scan_scratch_ints(words), dlp_scan_scratch_ints(ni),
                   dlp_radix_scratch_ints(ni)}));
  if ((rc = sorted_distinct(*w, insertedNodes, numInserted, numNodes, bits, sorted, scratch, s,
                            "dlp_label_initialization")) != DLP_OK)
    return rc;
  rc = dlp_k_label_init(rowPtr, colInd, values, labels, componentId, numComponents, insertedNodes,
                        numInserted, gtClass, sorted, p0, p1, keyA, valA, keyB, valB, keyC, valC,
                        segStart, segEnd, red0, red1, W0, W1, scratch, s);
  if (rc != DLP_OK) return rc;
  DLP_HIP(hipStreamSynchronize(s));
  return DLP_OK;
}

namespace {

/* scratch of one propagation call: counters (device, and a host copy for the
 * read-backs), stamps, the two queues and the LDS-kernel bookkeeping */
struct PropScratch {
  DevCounters *ctr;
  DevCounters *host;
  int *stamp, *q0, *q1;
  int *coreGen; /* 2 * numNodes: core generations, then halo generations */
  int *coreIdx, *haloIdx, *coreList, *ovf;
  int listCap;
  int ldsBytes;
};

/* launchIterativePropagation (iterativePropagation.cu:241-346) on the given
 * scratch; the counters must hold zeros with stampBase 1 and the localGen
 * the generations in coreGen were written under */
int propagate_on(const PropScratch &ps, const int *rowPtr, const int *colInd, const float *values,
                 float *labels, float *labelsBuffer, const int *vertexStatus, const int *gtClass,
                 int *isAffected, int numNodes, float delta, int maxIterations,
                 int *iterationsUsed, hipStream_t s) {
  int rc;
  /* both buffers equal on entry (iterativePropagation.cu:268-270) */
  if ((rc = dlp_k_copy_float(labelsBuffer, labels, numNodes, s)) != DLP_OK) return rc;
  if ((rc = dlp_k_prop_prepare(isAffected, vertexStatus, gtClass, numNodes, ps.stamp, ps.q0,
                               ps.ctr, s)) != DLP_OK)
    return rc;
  PropArgs a{};
  a.queueCap = numNodes;
  a.rowBeg = rowPtr;
  a.rowEnd = rowPtr + 1;
  a.rowLen = nullptr;
  a.colInd = colInd;
  a.values = values;
  a.status = vertexStatus;
  a.lab0 = labels;
  a.lab1 = labelsBuffer;
  a.iterStamp = ps.stamp;
  a.q0 = ps.q0;
  a.q1 = ps.q1;
  a.touched = nullptr;
  a.ctr = ps.ctr;
  a.iterSizes = nullptr;
  a.iterSizesCap = 0;
  a.profile = 0;
  a.delta = delta;
  a.waveRowMin = kDefaultWaveRowMin;
  a.blockRowMin = kDefaultBlockRowMin;

  LocalArgs l{};
  l.ldsWords = ps.ldsBytes / 4;
  l.numNodes = numNodes;
  l.shape = 0;
  l.ownerAfter = kOwnerAfterDefault;
  l.coreListCap = l.ovfCap = ps.listCap;
  l.coreGen = ps.coreGen;
  l.haloGen = ps.coreGen + numNodes;
  l.coreIdx = ps.coreIdx;
  l.haloIdx = ps.haloIdx;
  l.coreList = ps.coreList;
  l.ovf = ps.ovf;
  PropLoopConfig cfg;
  cfg.groupLanes = 8;
  cfg.blocks = dlp_auto_prop_blocks(numNodes);
  cfg.maxChunk = 1024;
  cfg.maskOwnerMax = kMaskOwnerMax;
  {
    int nnz = 0;
    DLP_HIP(hipMemcpyAsync(&nnz, rowPtr + numNodes, sizeof(int), hipMemcpyDeviceToHost, s));
    DLP_HIP(hipStreamSynchronize(s));
    const long long meanDeg = ((long long)std::max(nnz, 0) + numNodes - 1) / numNodes;
    cfg.localThreshold = dlp_local_default_threshold(l.ldsWords, meanDeg);
  }
  rc = dlp_run_propagation(a, numNodes < (1 << 30) ? &l : nullptr, maxIterations, cfg, ps.host, s,
                           nullptr);
  if (rc != DLP_OK) return rc;
  DLP_HIP(hipMemcpyAsync(ps.host, ps.ctr, sizeof(DevCounters), hipMemcpyDeviceToHost, s));
  DLP_HIP(hipStreamSynchronize(s));
  const int m = ps.host->iterUsed;
  *iterationsUsed = m;
  /* last written buffer ends in labels (iterativePropagation.cu:330-334) */
  if (m & 1) {
    if ((rc = dlp_k_copy_float(labels, labelsBuffer, numNodes, s)) != DLP_OK) return rc;
  }
  const int capped = (m == maxIterations && m > 0 && ps.host->qcount[m % 3] > 0) ? 1 : 0;
  rc = dlp_k_prop_finish(labels, gtClass, numNodes, m > 0 ? isAffected : nullptr,
                         (m & 1) ? ps.q1 : ps.q0, ps.ctr, m % 3, capped, s);
  if (rc != DLP_OK) return rc;
  DLP_HIP(hipStreamSynchronize(s));
  return DLP_OK;
}

int check_propagation_args(const char *who, const int *rowPtr, const float *labels,
                           const float *labelsBuffer, const int *vertexStatus, const int *gtClass,
                           const int *isAffected, int numNodes, int maxIterations,
                           const int *iterationsUsed) {
  if (!iterationsUsed || !rowPtr || !labels || !labelsBuffer || !vertexStatus || !gtClass ||
      !isAffected || numNodes < 0) {
    dlp_set_errorf_("%s: bad argument", who);
    return DLP_ERR_INVALID;
  }
  if (maxIterations > kMaxIterationsLimit) {  // iteration stamps are ints
    dlp_set_errorf_("%s: maxIterations above %d", who, kMaxIterationsLimit);
    return DLP_ERR_INVALID;
  }
  return DLP_OK;
}

}  // namespace

extern "C" int dlp_iterative_propagation(const int *rowPtr, const int *colInd,
                                         const float *values, float *labels,
                                         float *labelsBuffer, const int *vertexStatus,
                                         const int *gtClass, int *isAffected, int numNodes,
                                         float delta, int maxIterations, int *iterationsUsed,
                                         void *stream) {
  int rc = check_propagation_args("dlp_iterative_propagation", rowPtr, labels, labelsBuffer,
                                  vertexStatus, gtClass, isAffected, numNodes, maxIterations,
                                  iterationsUsed);
  if (rc != DLP_OK) return rc;
  *iterationsUsed = 0;
  if (numNodes == 0) return DLP_OK;
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> lock(g_wsMutex);
  Workspace *w;
  if ((rc = ws_open(&w)) != DLP_OK) return rc;
  if ((rc = ws_reset_counters(*w, s)) != DLP_OK) return rc;
  if (!w->ldsBytes && (rc = dlp_local_lds_bytes(&w->ldsBytes)) != DLP_OK) return rc;
  const int listCap = dlp_local_list_cap(w->ldsBytes / 4);
  WS_BUF(int, stamp, 1, numNodes);
  WS_BUF(int, q0, 3, numNodes);
  WS_BUF(int, q1, 18, numNodes);
  WS_BUF(int, coreGen, 19, 2LL * numNodes);  // core generations, then halo generations
  WS_BUF(int, coreIdx, 20, numNodes);
  WS_BUF(int, haloIdx, 23, numNodes);
  WS_BUF(int, coreList, 21, listCap);
  WS_BUF(int, ovf, 22, listCap);
  if (coreGen != w->coreGenBuf || w->sizes[19] != w->coreGenBytes) {
    /* new or regrown buffer: restart the generations (the counters were
     * already uploaded with the old localGen) */
    DLP_HIP(hipMemsetAsync(coreGen, 0, w->sizes[19], s));
    DLP_HIP(hipMemsetAsync(&w->ctr->localGen, 0, sizeof(int), s));
    w->coreGenBuf = coreGen;
    w->coreGenBytes = w->sizes[19];
    w->localGen = 0;
  }
  const PropScratch ps{w->ctr, w->pin, stamp, q0, q1, coreGen, coreIdx, haloIdx,
                       coreList, ovf, listCap, w->ldsBytes};
  rc = propagate_on(ps, rowPtr, colInd, values, labels, labelsBuffer, vertexStatus, gtClass,
                    isAffected, numNodes, delta, maxIterations, iterationsUsed, s);
  w->localGen = w->pin->localGen;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* propagation without allocation: scratch in a static device array         */
/* ------------------------------------------------------------------------ */
namespace {

constexpr int kNoAllocMaxNodes = 1 << 19;   /* 524,288 vertices              */
constexpr int kNoAllocListCap = 16384;      /* LDS-kernel core / overflow lists */
constexpr long long kNoAllocCtrInts = 64;   /* 256 bytes for the counters    */
constexpr long long kNoAllocInts =
    kNoAllocCtrInts + 7LL * kNoAllocMaxNodes + 2LL * kNoAllocListCap; /* 14.1 MiB */
static_assert(sizeof(DevCounters) <= 4 * kNoAllocCtrInts, "counters outgrow their slot");

/* The scratch of dlp_iterative_propagation_noalloc, one copy per GPU, held
 * until the process ends.  Host code never uses this name as an address:
 * the per-device address comes from hipGetSymbolAddress (g_noAllocBase). */
static __device__ __align__(256) int scratch[kNoAllocInts];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noAllocMutex;
int *g_noAllocBase[kNoAllocMaxDevices];  /* device address of `scratch` per GPU */
int g_noAllocLds[kNoAllocMaxDevices];    /* LDS bytes of the local kernel per GPU */

/* counters <- zero with stampBase 1 and localGen 0; generations <- 0 */
__global__ void k_noalloc_clear(DevCounters *ctr, int *gens, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long long v = i; v < n; v += (long long)gridDim.x * blockDim.x) gens[v] = 0;
  if (i == 0) {
    DevCounters z = {};
    z.stampBase = 1;
    *ctr = z;
  }
}

}  // namespace

extern "C" int dlp_iterative_propagation_noalloc(const int *rowPtr, const int *colInd,
                                                 const float *values, float *labels,
                                                 float *labelsBuffer, const int *vertexStatus,
                                                 const int *gtClass, int *isAffected,
                                                 int numNodes, float delta, int maxIterations,
                                                 int *iterationsUsed, void *stream) {
  int rc = check_propagation_args("dlp_iterative_propagation_noalloc", rowPtr, labels,
                                  labelsBuffer, vertexStatus, gtClass, isAffected, numNodes,
                                  maxIterations, iterationsUsed);
  if (rc != DLP_OK) return rc;
  if (numNodes > kNoAllocMaxNodes) {
    dlp_set_errorf_("dlp_iterative_propagation_noalloc: numNodes %d above the static scratch's "
                    "limit of %d vertices",
                    numNodes, kNoAllocMaxNodes);
    return DLP_ERR_INVALID;
  }
  *iterationsUsed = 0;
  if (numNodes == 0) return DLP_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    dlp_set_error_("no GPU visible to HIP");
    return DLP_ERR_NODEVICE;
  }
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> lock(g_noAllocMutex);
  int dev = 0;
  DLP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    dlp_set_errorf_("dlp_iterative_propagation_noalloc: device %d above %d", dev,
                    kNoAllocMaxDevices - 1);
    return DLP_ERR_INVALID;
  }
  if (!g_noAllocBase[dev]) {  // first call on this device: its copy's address
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      dlp_set_errorf_("hipGetSymbolAddress of the static scratch failed (%s)",
                      hipGetErrorString(e));
      return DLP_ERR_HIP;
    }
    if ((rc = dlp_local_lds_bytes(&g_noAllocLds[dev])) != DLP_OK) return rc;
    g_noAllocBase[dev] = (int *)p;
  }
  const int listCap = dlp_local_list_cap(g_noAllocLds[dev] / 4);
  if (listCap > kNoAllocListCap) {
    dlp_set_errorf_("dlp_iterative_propagation_noalloc: LDS lists of %d entries above %d",
                    listCap, kNoAllocListCap);
    return DLP_ERR_INVALID;
  }
  int *base = g_noAllocBase[dev];
  const long long N = numNodes;
  PropScratch ps{};
  ps.ctr = (DevCounters *)base;
  ps.stamp = base + kNoAllocCtrInts;
  ps.q0 = ps.stamp + N;
  ps.q1 = ps.q0 + N;
  ps.coreGen = ps.q1 + N;
  ps.coreIdx = ps.coreGen + 2 * N;
  ps.haloIdx = ps.coreIdx + N;
  ps.coreList = ps.haloIdx + N;
  ps.ovf = ps.coreList + listCap;
  ps.listCap = listCap;
  ps.ldsBytes = g_noAllocLds[dev];
  if (ps.ovf + listCap > base + kNoAllocInts) {
    dlp_set_error_("dlp_iterative_propagation_noalloc: scratch layout overflow");
    return DLP_ERR_INTERNAL;
  }
  DevCounters host{};
  ps.host = &host;
  /* the array keeps the previous call's contents: counters and generations
   * start from zero on every call */
  const unsigned grid = (unsigned)std::min<long long>((2 * N + 255) / 256, 4096);
  hipLaunchKernelGGL(k_noalloc_clear, dim3(grid), dim3(256), 0, s, ps.ctr, ps.coreGen, 2 * N);
  DLP_LAUNCHED("k_noalloc_clear");
  return propagate_on(ps, rowPtr, colInd, values, labels, labelsBuffer, vertexStatus, gtClass,
                      isAffected, numNodes, delta, maxIterations, iterationsUsed, s);
}

/* ------------------------------------------------------------------------ */
/* IrLP full recompute                                                      */
/* ------------------------------------------------------------------------ */
extern "C" void dlp_irlp_tuning_default(dlp_irlp_tuning *t) {
  if (!t) return;
  const IrlpLoopConfig d;
  t->groupLanes = d.groupLanes;
  t->blocks = d.blocks;
  t->firstChunk = d.firstChunk;
  t->maxChunk = d.maxChunk;
}

extern "C" int dlp_irlp_ex(const int *rowPtr, const int *colInd, const float *values,
                           float *labels, const int *vertexStatus, const int *gtClass,
                           int numNodes, float delta, int maxIterations, int *iterationsUsed,
                           void *stream, const dlp_irlp_tuning *tuning, float *kernelMs) {
  if (!iterationsUsed) {
    dlp_set_error_("dlp_irlp: NULL iterationsUsed");
    return DLP_ERR_INVALID;
  }
  *iterationsUsed = 0;
  if (kernelMs) *kernelMs = 0.0f;
  if (numNodes < 0 || (numNodes > 0 && (!rowPtr || !labels || !vertexStatus || !gtClass))) {
    dlp_set_error_("dlp_irlp: bad argument");
    return DLP_ERR_INVALID;
  }
  if (numNodes == 0 || maxIterations <= 0) return DLP_OK; /* validation.cu:168 loop not entered */
  IrlpLoopConfig cfg;
  if (tuning) {
    cfg.groupLanes = tuning->groupLanes;
    cfg.blocks = tuning->blocks;
    cfg.firstChunk = tuning->firstChunk;
    cfg.maxChunk = tuning->maxChunk;
  }
  std::lock_guard<std::mutex> lock(g_wsMutex);
  Workspace *w;
  int rc = ws_open(&w);
  if (rc != DLP_OK) return rc;
  if (!w->pinFlags)
    DLP_HIP(hipHostMalloc((void **)&w->pinFlags, sizeof(int) * (kIrlpMaxChunk + 1),
                          hipHostMallocDefault));
  WS_BUF(float, buffer, 24, numNodes);
  WS_BUF(int, flags, 25, kIrlpMaxChunk + 1);
  hipStream_t s = (hipStream_t)stream;
  int numEntries = 0;
  DLP_HIP(hipMemcpyAsync(&numEntries, rowPtr + numNodes, sizeof(int), hipMemcpyDeviceToHost, s));
  DLP_HIP(hipStreamSynchronize(s));
  if (numEntries < 0 || (numEntries > 0 && !colInd)) {
    dlp_set_error_("dlp_irlp: rowPtr[numNodes] negative or colInd NULL");
    return DLP_ERR_INVALID;
  }
  WS_BUF(int, colMasked, 26, numEntries);
  /* the second buffer starts as a copy of the labels (validation.cu:157-159) */
  if ((rc = dlp_k_copy_float(buffer, labels, numNodes, s)) != DLP_OK) return rc;
  IrlpArgs a;
  a.rowPtr = rowPtr;
  a.colInd = colInd;
  a.values = values;
  a.status = vertexStatus;
  a.gtClass = gtClass;
  a.lab0 = labels;
  a.lab1 = buffer;
  a.flags = flags;
  a.colMasked = colMasked;
  a.numEntries = numEntries;
  a.numNodes = numNodes;
  a.delta = delta;
  return dlp_run_irlp(a, maxIterations, cfg, w->pinFlags, s, iterationsUsed, kernelMs);
}

extern "C" int dlp_irlp(const int *rowPtr, const int *colInd, const float *values, float *labels,
                        const int *vertexStatus, const int *gtClass, int numNodes, float delta,
                        int maxIterations, int *iterationsUsed, void *stream) {
  return dlp_irlp_ex(rowPtr, colInd, values, labels, vertexStatus, gtClass, numNodes, delta,
                     maxIterations, iterationsUsed, stream, nullptr, nullptr);
}

extern "C" int dlp_irlp_validation_ex(const dlp_graph_view *activeGraph,
                                      const int *vertexStatus, const int *gtClass, float *labels,
                                      int numNodes, float delta, int maxIterations,
                                      int *iterationsUsed, int *skipped,
                                      const dlp_irlp_tuning *tuning, float *loopMs) {
  if (!activeGraph || !iterationsUsed) {
    dlp_set_error_("dlp_irlp_validation: NULL argument");
    return DLP_ERR_INVALID;
  }
  *iterationsUsed = 0;
  if (skipped) *skipped = 0;
  if (loopMs) *loopMs = 0.0f;
  if (numNodes <= 0 || activeGraph->numEdges <= 0) { /* validation.cu:216-219 */
    if (skipped) *skipped = 1;
    return DLP_OK;
  }
  if (activeGraph->numNodes != numNodes || !activeGraph->rowPtr || !activeGraph->colInd ||
      !activeGraph->values || !vertexStatus || !gtClass || !labels) {
    dlp_set_error_("dlp_irlp_validation: bad argument");
    return DLP_ERR_INVALID;
  }
  const size_t N = (size_t)numNodes, E = (size_t)activeGraph->numEdges;
  void *buf[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  const size_t bytes[6] = {sizeof(int) * (N + 1), sizeof(int) * E, sizeof(float) * E,
                           sizeof(float) * N,     sizeof(int) * N, sizeof(int) * N};
  const void *src[6] = {activeGraph->rowPtr, activeGraph->colInd, activeGraph->values,
                        labels,              vertexStatus,        gtClass};
  hipStream_t s = nullptr;
  int rc = DLP_OK;
  hipError_t e = hipStreamCreate(&s);
  for (int i = 0; i < 6 && e == hipSuccess; ++i) e = hipMalloc(&buf[i], bytes[i]);
  for (int i = 0; i < 6 && e == hipSuccess; ++i)
    e = hipMemcpyAsync(buf[i], src[i], bytes[i], hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    dlp_set_errorf_("dlp_irlp_validation: HIP error %d (%s)", (int)e, hipGetErrorString(e));
    rc = DLP_ERR_HIP;
  }
  if (rc == DLP_OK)
    rc = dlp_irlp_ex(static_cast<const int *>(buf[0]), static_cast<const int *>(buf[1]),
                     static_cast<const float *>(buf[2]), static_cast<float *>(buf[3]),
                     static_cast<const int *>(buf[4]), static_cast<const int *>(buf[5]),
                     numNodes, delta, maxIterations, iterationsUsed, s, tuning, loopMs);
  if (rc == DLP_OK) {
    e = hipMemcpyAsync(labels, buf[3], bytes[3], hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      dlp_set_errorf_("dlp_irlp_validation: HIP error %d (%s)", (int)e, hipGetErrorString(e));
      rc = DLP_ERR_HIP;
    }
  }
  for (int i = 0; i < 6; ++i)
    if (buf[i]) (void)hipFree(buf[i]);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

extern "C" int dlp_irlp_validation(const dlp_graph_view *activeGraph, const int *vertexStatus,
                                   const int *gtClass, float *labels, int numNodes, float delta,
                                   int maxIterations, int *iterationsUsed, int *skipped) {
  return dlp_irlp_validation_ex(activeGraph, vertexStatus, gtClass, labels, numNodes, delta,
                                maxIterations, iterationsUsed, skipped, nullptr, nullptr);
}
