/*
 * graphTypes.h -- host-side data model of the DynLP batch-update path.
 *
 * Same types, fields and defaults as the reference's headers/graphTypes.h,
 * without any GPU runtime include: everything device-side lives behind the
 * C-ABI in dynlp_c.h.
 *
 *   CSRGraph      <- graphTypes.h:90-98   (colInd == -1 marks a soft-deleted entry)
 *   DynLPConfig   <- graphTypes.h:124-154 (defaults :147-153)
 *   BatchData     <- graphTypes.h:167-182
 *   VertexStatus  <- graphTypes.h:191-195
 *
 * The reference's CUDA_CHECK / CUDA_KERNEL_CHECK macros (graphTypes.h:54-77)
 * are replaced by DLP_CHECK, which checks a C-ABI status code and, like the
 * reference macros, prints the failing call and exits.
 */
#ifndef DYNLP_GRAPH_TYPES_H
#define DYNLP_GRAPH_TYPES_H

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "dynlp_c.h"

#define DLP_CHECK(call)                                                        \
  do {                                                                         \
    int dlp_status_ = (call);                                                  \
    if (dlp_status_ != DLP_OK) {                                               \
      std::fprintf(stderr, "DynLP error at %s:%d - %s\n", __FILE__, __LINE__,  \
                   dlp_last_error());                                          \
      std::exit(EXIT_FAILURE);                                                 \
    }                                                                          \
  } while (0)

/* Host CSR graph (reference graphTypes.h:90-98). */
struct CSRGraph {
  std::vector<int> rowPtr;   ///< size numNodes + 1
  std::vector<int> colInd;   ///< size numEdges, -1 = soft-deleted
  std::vector<float> values; ///< edge similarity weights
  int numNodes;
  int numEdges;              ///< CSR entries, soft-deleted ones included

  CSRGraph() : numNodes(0), numEdges(0) { rowPtr.push_back(0); }
};

/* Run configuration (reference graphTypes.h:124-154, same defaults). */
struct DynLPConfig {
  int totalNodes;
  int avgDegree;
  int numBatches;
  unsigned int seed;

  float tau;            ///< < 0 means "auto": mean edge weight of the full graph
  float delta;          ///< convergence threshold |F' - F| > delta
  int maxIterations;

  float groundTruthFraction;
  float unlabeledFraction;
  float deletedFraction;

  std::string outputDir;
  bool validate;
  bool verbose;

  DynLPConfig()
      : totalNodes(50000), avgDegree(5), numBatches(10), seed(42), tau(-1.0f),
        delta(0.0001f), maxIterations(100000), groundTruthFraction(0.01f),
        unlabeledFraction(0.90f), deletedFraction(0.09f), outputDir("output/"),
        validate(true), verbose(true) {}
};

/* One dynamic batch (reference graphTypes.h:167-182). */
struct BatchData {
  std::vector<int> insertedNodeIds;
  std::vector<int> groundTruthNodeIds;
  std::vector<float> groundTruthLabels;
  std::vector<int> deletedNodeIds;
  std::vector<int> newEdgeSrc;   ///< u < v, in full-graph CSR order
  std::vector<int> newEdgeDst;
  std::vector<float> newEdgeWeight;
  int batchId;

  BatchData() : batchId(0) {}
};

/* Vertex status (reference graphTypes.h:191-195). */
enum VertexStatus : int {
  VERTEX_ACTIVE = DLP_VERTEX_ACTIVE,
  VERTEX_DELETED = DLP_VERTEX_DELETED,
  VERTEX_GROUND_TRUTH = DLP_VERTEX_GROUND_TRUTH
};

#endif /* DYNLP_GRAPH_TYPES_H */
