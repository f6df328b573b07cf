/*
 * dynlp_oracle.h -- CPU oracle for the DynLP batch-update path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or as the
 * timed CPU baseline), never as the product.
 *
 * A literal, single-threaded restatement of the reference's runDynLP
 * (src/dynLP.cu:175-587) and of its step launchers, in plain C.  The same
 * source is compiled twice: orc32_* computes in FP32 (the parity target),
 * orc64_* computes labels and weight sums in FP64 (error statistics only).
 *
 * Deviations from the literal reference, each forced:
 *   - CC runs over the sparsified subgraph G' (the contract of
 *     connectedComponents.h:33-36 / sparsifySubgraph.h:45-56); the literal
 *     call at dynLP.cu:437-440 passes the unfiltered new-edge list, which is
 *     undefined behaviour (it hooks roots to -1 and later indexes par[-1]).
 *   - Floating-point atomics of the reference (labelInitialization.cu:83-85)
 *     have no fixed order; here each member's row is summed left to right and
 *     the members of a component are combined by a bottom-up pairwise tree in
 *     ascending vertex id.
 *   - Division is IEEE (the reference builds with --use_fast_math).
 */
#ifndef DYNLP_ORACLE_H
#define DYNLP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DECLARE(PFX, REAL)                                                  \
  typedef struct PFX##state PFX##state;                                         \
  /* step restatements (device pointers of the reference -> host arrays) */     \
  void PFX##mark_deletions(const int *rowPtr, int *colInd, int *status,         \
                           int *isAffected, const int *deleted, int numDeleted, \
                           int numNodes);                                       \
  int PFX##sparsify(const int *src, const int *dst, const float *w,             \
                    int numNewEdges, int numInserted, const int *isNew,         \
                    float tau, int *isAffected, int *subSrc, int *subDst,       \
                    float *subW);                                               \
  int PFX##connected_components(const int *src, const int *dst, int numEdges,   \
                                const int *inserted, int numInserted,           \
                                int *componentId, int numNodes,                 \
                                int *numComponents);                            \
  int PFX##label_initialization(const int *rowPtr, const int *colInd,           \
                                const float *values, REAL *labels,              \
                                const int *componentId, int numComponents,      \
                                const int *inserted, int numInserted,           \
                                const int *gtClass);                            \
  int PFX##propagate(const int *rowPtr, const int *colInd,                      \
                     const float *values, REAL *labels, REAL *labelsBuffer,     \
                     const int *status, const int *gtClass, int *isAffected,    \
                     int numNodes, float delta, int maxIterations,              \
                     int *frontierSizes, int maxFrontierSizes,                  \
                     int *numFrontierSizes);                                    \
  /* worker threads of propagate (1 = sequential; more: same results) for     \
   * frontiers of at least minFrontier vertices */                             \
  void PFX##set_threads(int threads, int minFrontier);                          \
  int PFX##irlp(const int *rowPtr, const int *colInd, const float *values,      \
                 REAL *labels, const int *status, const int *gtClass,           \
                 int numNodes, float delta, int maxIterations);                 \
  /* whole batch loop */                                                        \
  PFX##state *PFX##create(                                                      \
      int numNodes, const int *rowPtr, const int *colInd, const float *values,  \
      const float *trueLabels, int numBatches, const int64_t *insOff,           \
      const int *ins, const int64_t *gtOff, const int *gtIds,                   \
      const float *gtLabels, const int64_t *delOff, const int *del,             \
      const int64_t *edgeOff, const int *edgeSrc, const int *edgeDst,           \
      const float *edgeW, float tau, float delta, int maxIterations);           \
  void PFX##destroy(PFX##state *st);                                            \
  float PFX##tau(const PFX##state *st);                                         \
  int PFX##run_batch(PFX##state *st, int b);                                    \
  /* state of the last batch run */                                             \
  int PFX##nnz(const PFX##state *st);                                           \
  void PFX##get_csr(const PFX##state *st, int *rowPtr, int *colInd,             \
                    float *values);                                             \
  void PFX##get_affected(const PFX##state *st, int *isAffected);                \
  int PFX##get_sub(const PFX##state *st, int *src, int *dst, float *w);         \
  int PFX##get_component_ids(const PFX##state *st, int *componentId);           \
  void PFX##get_init_labels(const PFX##state *st, REAL *labels);                \
  void PFX##get_labels(const PFX##state *st, REAL *labels);                     \
  int PFX##get_frontier_sizes(const PFX##state *st, int *sizes, int maxSizes);  \
  /* device status during the last batch, host status, GT class */             \
  void PFX##get_status(const PFX##state *st, int *dStatus, int *hStatus,        \
                       int *gtClass);                                           \
  /* iterations, numSubEdges, numComponents, numMarked, frontier0 */            \
  void PFX##get_stats(const PFX##state *st, int *stats5, int64_t *updates,      \
                      float *accuracy);

ORC_DECLARE(orc32_, float)
ORC_DECLARE(orc64_, double)

#ifdef __cplusplus
}
#endif

#endif /* DYNLP_ORACLE_H */
