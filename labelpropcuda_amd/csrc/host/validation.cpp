/*
 * validation.cpp -- the --validate step (main.cu:120-217): final-state replay
 * of the batch stream, active CSR, IrLP run (dlp_irlp_validation) and its
 * accuracy.  Host code; the IrLP iterations run on the GPU.
 */
#include <chrono>
#include <new>
#include <vector>

#include "dynlp/dynlp_c.h"

extern "C" void dlp_set_error_(const char *msg);  // defined in the runtime

namespace {

bool batchesValid(const dlp_batch_view *b, int N) {
  if (b->numBatches < 0) return false;
  if (b->numBatches == 0) return true;
  if (!b->insertedOffsets || !b->groundTruthOffsets || !b->deletedOffsets) return false;
  const int64_t *offs[3] = {b->insertedOffsets, b->groundTruthOffsets, b->deletedOffsets};
  const int *ids[3] = {b->insertedNodeIds, b->groundTruthNodeIds, b->deletedNodeIds};
  for (int k = 0; k < 3; ++k) {
    if (offs[k][0] != 0) return false;
    for (int i = 0; i < b->numBatches; ++i)
      if (offs[k][i + 1] < offs[k][i]) return false;
    const int64_t n = offs[k][b->numBatches];
    if (n > 0 && !ids[k]) return false;
    for (int64_t i = 0; i < n; ++i)
      if (ids[k][i] < 0 || ids[k][i] >= N) return false;
  }
  return b->groundTruthOffsets[b->numBatches] == 0 || b->groundTruthLabels;
}

}  // namespace

extern "C" int dlp_irlp_baseline(const dlp_config *cfg, const dlp_graph_view *fullGraph,
                                 const float *trueLabels, const dlp_batch_view *batches,
                                 const dlp_irlp_tuning *tuning, float *irlpLabels,
                                 dlp_irlp_result *out) {
  if (!cfg || !fullGraph || !trueLabels || !batches || !irlpLabels || !out) {
    dlp_set_error_("dlp_irlp_baseline: NULL argument");
    return DLP_ERR_INVALID;
  }
  const int N = fullGraph->numNodes;
  if (N < 0 || (N > 0 && (!fullGraph->rowPtr || (fullGraph->numEdges > 0 &&
                                                  (!fullGraph->colInd || !fullGraph->values)))) ||
      !batchesValid(batches, N)) {
    dlp_set_error_("dlp_irlp_baseline: bad graph or batch view");
    return DLP_ERR_INVALID;
  }
  *out = dlp_irlp_result{};
  try {
    /* final status, ground-truth classes and starting labels (main.cu:127-146) */
    std::vector<int> status((size_t)N, DLP_VERTEX_DELETED), gtClass((size_t)N, -1);
    std::vector<int> lastInsert((size_t)N, -1); /* vertexBatchMap, main.cu:149-155 */
    for (int i = 0; i < N; ++i) irlpLabels[i] = 0.5f;
    for (int b = 0; b < batches->numBatches; ++b) {
      for (int64_t i = batches->insertedOffsets[b]; i < batches->insertedOffsets[b + 1]; ++i) {
        status[batches->insertedNodeIds[i]] = DLP_VERTEX_ACTIVE;
        lastInsert[batches->insertedNodeIds[i]] = b;
      }
      for (int64_t i = batches->deletedOffsets[b]; i < batches->deletedOffsets[b + 1]; ++i)
        status[batches->deletedNodeIds[i]] = DLP_VERTEX_DELETED;
      for (int64_t i = batches->groundTruthOffsets[b]; i < batches->groundTruthOffsets[b + 1];
           ++i) {
        const int v = batches->groundTruthNodeIds[i];
        const float lab = batches->groundTruthLabels[i];
        status[v] = DLP_VERTEX_GROUND_TRUTH;
        gtClass[v] = lab < 0.5f ? 0 : 1;
        irlpLabels[v] = lab;
      }
    }
    /* active CSR: entries of the full graph between inserted, non-deleted
     * vertices (main.cu:157-178) */
    auto inGraph = [&](int v) { return status[v] != DLP_VERTEX_DELETED && lastInsert[v] >= 0; };
    std::vector<int> rowPtr((size_t)N + 1, 0), colInd;
    std::vector<float> values;
    for (int u = 0; u < N; ++u) {
      if (inGraph(u)) {
        for (int e = fullGraph->rowPtr[u]; e < fullGraph->rowPtr[u + 1]; ++e) {
          const int v = fullGraph->colInd[e];
          if (v < 0 || v >= N || !inGraph(v)) continue;
          colInd.push_back(v);
          values.push_back(fullGraph->values[e]);
        }
      }
      rowPtr[(size_t)u + 1] = (int)colInd.size();
    }
    dlp_graph_view active;
    active.numNodes = N;
    active.numEdges = (int)colInd.size();
    active.rowPtr = rowPtr.data();
    active.colInd = colInd.data();
    active.values = values.data();
    out->activeEdges = active.numEdges;

    const auto t0 = std::chrono::steady_clock::now();
    int rc = dlp_irlp_validation_ex(&active, status.data(), gtClass.data(), irlpLabels, N,
                                    cfg->delta, cfg->maxIterations, &out->iterations,
                                    &out->skipped, tuning, &out->loopMs);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc != DLP_OK) return rc;
    out->timeMs = std::chrono::duration<float, std::milli>(t1 - t0).count();

    /* accuracy over vertices neither deleted nor ground truth (main.cu:199-209) */
    long long correct = 0, total = 0;
    for (int i = 0; i < N; ++i) {
      if (status[i] == DLP_VERTEX_DELETED || status[i] == DLP_VERTEX_GROUND_TRUTH) continue;
      const int predicted = irlpLabels[i] >= 0.5f ? 1 : 0;
      const int truth = trueLabels[i] >= 0.5f ? 1 : 0;
      correct += predicted == truth;
      ++total;
    }
    out->accuracy = total > 0 ? (float)correct / (float)total : 1.0f;
  } catch (const std::bad_alloc &) {
    dlp_set_error_("dlp_irlp_baseline: out of host memory");
    return DLP_ERR_NOMEM;
  }
  return DLP_OK;
}
