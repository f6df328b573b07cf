/* host_internal.h -- shared state of the host-side helpers (not exported). */
#ifndef DLP_HOST_INTERNAL_H
#define DLP_HOST_INTERNAL_H

#include <vector>

#include "dynlp/graphTypes.h"

namespace dlp_host {
/* Suppresses the generators' progress printing (the C-ABI input generator
 * sets it; the CLI keeps the reference's output). */
extern bool g_quiet;
}  // namespace dlp_host

/* batches of caller-chosen sizes (sizes[0] is ignored: batch 0 is the GT
 * seed batch), see graphgen.cpp */
bool generateStreamBatches(const CSRGraph &graph, const std::vector<float> &trueLabels,
                           const std::vector<int> &sizes, unsigned int seed,
                           float groundTruthFraction, float deletedFraction,
                           std::vector<BatchData> &batches);

#endif
