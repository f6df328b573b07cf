/*
 * dynLP.h -- C++ drop-in API of the MI355X DynLP batch-update path.
 *
 * The reference's C++ entry points, same names, argument order and meaning,
 * implemented over the C-ABI of dynlp_c.h (libdynlp.so).  Device pointers
 * are plain; `hipStream_t` is taken as `void*` so this header needs no GPU
 * runtime include.  Errors behave like the reference's CUDA_CHECK: the
 * message is printed and the process exits.
 *
 *   runDynLP                      <- dynLP.h:69-74
 *   launchMarkDeletions           <- sparsifySubgraph.h:41-43
 *   launchBuildSparsifiedSubgraph <- sparsifySubgraph.h:78-84
 *   launchConnectedComponents     <- connectedComponents.h:50-54
 *   launchLabelInitialization     <- labelInitialization.h:54-60
 *   launchIterativePropagation    <- iterativePropagation.h:73-80
 *   launchIrLP, runIrLPValidation <- validation.h (validation.cu:144-276)
 *   generateSparseGraph           <- generateSparseGraph.h:42-44
 *   generateBatches               <- generateBatches.h:56-59
 *   parseArguments .. writeResultsCSV <- utils.h:41-79
 */
#ifndef DYNLP_DYNLP_H
#define DYNLP_DYNLP_H

#include <string>
#include <vector>

#include "graphTypes.h"

/* ---- driver API --------------------------------------------------------- */
bool runDynLP(const DynLPConfig &config, const CSRGraph &graph,
              const std::vector<float> &trueLabels,
              const std::vector<BatchData> &batches,
              std::vector<float> &finalLabels, std::vector<float> &batchTimes,
              std::vector<float> &batchAccuracy);

/* ---- step API ------------------------------------------------------------ */
inline void launchMarkDeletions(int *d_rowPtr, int *d_colInd,
                                int *d_vertexStatus, int *d_isAffected,
                                const int *d_deletedNodes, int numDeleted,
                                int numNodes, void *stream) {
  DLP_CHECK(dlp_mark_deletions(d_rowPtr, d_colInd, d_vertexStatus,
                               d_isAffected, d_deletedNodes, numDeleted,
                               numNodes, stream));
}

/* d_subEdgeSrc (not in the reference signature) may be nullptr. */
inline void launchBuildSparsifiedSubgraph(
    int *d_rowPtr, int *d_colInd, float *d_values, int *d_isAffected,
    const int *d_vertexStatus, const int *d_newEdgeSrc,
    const int *d_newEdgeDst, const float *d_newEdgeWeight, int numNewEdges,
    const int *d_insertedNodes, int numInserted, const int *d_isNewVertex,
    float tau, int *d_subRowPtr, int *d_subColInd, float *d_subValues,
    int &numSubNodes, int &numSubEdges, void *stream,
    int *d_subEdgeSrc = nullptr) {
  DLP_CHECK(dlp_build_sparsified_subgraph(
      d_rowPtr, d_colInd, d_values, d_isAffected, d_vertexStatus,
      d_newEdgeSrc, d_newEdgeDst, d_newEdgeWeight, numNewEdges,
      d_insertedNodes, numInserted, d_isNewVertex, tau, d_subRowPtr,
      d_subColInd, d_subValues, d_subEdgeSrc, &numSubNodes, &numSubEdges,
      stream));
}

inline void launchConnectedComponents(const int *d_edgeSrc,
                                      const int *d_edgeDst, int numEdges,
                                      const int *d_insertedNodes,
                                      int numInserted, int *d_componentId,
                                      int &numComponents, int numNodes,
                                      void *stream) {
  DLP_CHECK(dlp_connected_components(d_edgeSrc, d_edgeDst, numEdges,
                                     d_insertedNodes, numInserted,
                                     d_componentId, &numComponents, numNodes,
                                     stream));
}

inline void launchLabelInitialization(
    const int *d_rowPtr, const int *d_colInd, const float *d_values,
    float *d_labels, const int *d_vertexStatus, const int *d_componentId,
    int numComponents, const int *d_insertedNodes, int numInserted,
    const int *d_gtClass, int numNodes, void *stream) {
  DLP_CHECK(dlp_label_initialization(
      d_rowPtr, d_colInd, d_values, d_labels, d_vertexStatus, d_componentId,
      numComponents, d_insertedNodes, numInserted, d_gtClass, numNodes,
      stream));
}

inline void launchIterativePropagation(
    const int *d_rowPtr, const int *d_colInd, const float *d_values,
    float *d_labels, float *d_labelsBuffer, const int *d_vertexStatus,
    const int *d_gtClass, int *d_isAffected, int numNodes, float delta,
    int maxIterations, int &iterationsUsed, void *stream) {
  DLP_CHECK(dlp_iterative_propagation(
      d_rowPtr, d_colInd, d_values, d_labels, d_labelsBuffer, d_vertexStatus,
      d_gtClass, d_isAffected, numNodes, delta, maxIterations,
      &iterationsUsed, stream));
}

/* ---- IrLP full-recompute baseline ---------------------------------------- */
inline void launchIrLP(const int *d_rowPtr, const int *d_colInd, const float *d_values,
                       float *d_labels, const int *d_vertexStatus, const int *d_gtClass,
                       int numNodes, float delta, int maxIterations, int &iterationsUsed,
                       void *stream) {
  DLP_CHECK(dlp_irlp(d_rowPtr, d_colInd, d_values, d_labels, d_vertexStatus, d_gtClass,
                     numNodes, delta, maxIterations, &iterationsUsed, stream));
}

/* Host arrays in, labels in / out; prints the reference's warning and
 * returns true for an empty graph. */
bool runIrLPValidation(const CSRGraph &activeGraph, const std::vector<int> &h_vertexStatus,
                       const std::vector<int> &h_gtClass, std::vector<float> &h_labels,
                       int numNodes, float delta, int maxIterations, int &iterationsUsed);

/* ---- synthetic inputs (host) --------------------------------------------- */
bool generateSparseGraph(int totalNodes, int avgDegree, unsigned int seed,
                         CSRGraph &graph, std::vector<float> &trueLabels,
                         const std::string &outputPrefix);

bool generateBatches(const CSRGraph &graph,
                     const std::vector<float> &trueLabels,
                     const DynLPConfig &config,
                     std::vector<BatchData> &batches);

/* The generator's CSR text dump (<prefix>RowPtr.txt, ColInd.txt, Values.txt,
 * TrueLabels.txt; generateSparseGraph.cu:217-274). */
bool writeCSRGraph(const std::string &prefix, const CSRGraph &graph,
                   const std::vector<float> &trueLabels);

/* ---- readers of the written files (no counterpart in the reference) ------ */
/* the CSR dump above (weights as written, 6 decimals) */
bool readCSRGraph(const std::string &prefix, CSRGraph &graph, std::vector<float> &trueLabels);
/* dynlp_labels.txt / irlp_labels.txt */
bool readLabelsFile(const std::string &path, std::vector<float> &labels);
/* results.csv */
bool readResultsCSV(const std::string &path, std::vector<float> &batchTimes,
                    std::vector<float> &batchAccuracy);

/* ---- CLI / file utilities ------------------------------------------------ */
bool parseArguments(int argc, char *argv[], DynLPConfig &config);
void printConfig(const DynLPConfig &config);
bool ensureDirectory(const std::string &dirPath);
bool writeLabelsToFile(const std::string &path, const float *labels,
                       int numNodes);
bool writeResultsCSV(const std::string &path, const float *batchTimes,
                     const float *batchAccuracy, int numBatches);

/* Flattens batches into the C-ABI view; the storage vectors own the data. */
struct FlatBatches {
  std::vector<int64_t> insOff, gtOff, delOff, edgeOff;
  std::vector<int> ins, gtIds, del, eSrc, eDst;
  std::vector<float> gtLab, eW;
  dlp_batch_view view() const;
};
FlatBatches flattenBatches(const std::vector<BatchData> &batches);
dlp_graph_view graphView(const CSRGraph &graph);
dlp_config toCConfig(const DynLPConfig &config);

#endif /* DYNLP_DYNLP_H */
