/*
 * cxx_api.cpp -- the reference's C++ driver entry points over the C-ABI, and
 * the C-ABI input generator.
 *
 * runDynLP (dynLP.h:69-74) keeps its signature and its error behaviour
 * (print and exit, graphTypes.h:54-62); the work happens in dlp_run_dynlp.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <new>
#include <vector>

#include "dynlp/dynLP.h"
#include "host_internal.h"

dlp_batch_view FlatBatches::view() const {
  dlp_batch_view v;
  v.numBatches = (int)insOff.size() - 1;
  v.insertedOffsets = insOff.data();
  v.insertedNodeIds = ins.data();
  v.groundTruthOffsets = gtOff.data();
  v.groundTruthNodeIds = gtIds.data();
  v.groundTruthLabels = gtLab.data();
  v.deletedOffsets = delOff.data();
  v.deletedNodeIds = del.data();
  v.edgeOffsets = edgeOff.data();
  v.newEdgeSrc = eSrc.data();
  v.newEdgeDst = eDst.data();
  v.newEdgeWeight = eW.data();
  return v;
}

FlatBatches flattenBatches(const std::vector<BatchData> &batches) {
  FlatBatches f;
  f.insOff.push_back(0);
  f.gtOff.push_back(0);
  f.delOff.push_back(0);
  f.edgeOff.push_back(0);
  for (const BatchData &b : batches) {
    f.ins.insert(f.ins.end(), b.insertedNodeIds.begin(), b.insertedNodeIds.end());
    f.gtIds.insert(f.gtIds.end(), b.groundTruthNodeIds.begin(), b.groundTruthNodeIds.end());
    f.gtLab.insert(f.gtLab.end(), b.groundTruthLabels.begin(), b.groundTruthLabels.end());
    f.del.insert(f.del.end(), b.deletedNodeIds.begin(), b.deletedNodeIds.end());
    f.eSrc.insert(f.eSrc.end(), b.newEdgeSrc.begin(), b.newEdgeSrc.end());
    f.eDst.insert(f.eDst.end(), b.newEdgeDst.begin(), b.newEdgeDst.end());
    f.eW.insert(f.eW.end(), b.newEdgeWeight.begin(), b.newEdgeWeight.end());
    f.insOff.push_back((int64_t)f.ins.size());
    f.gtOff.push_back((int64_t)f.gtIds.size());
    f.delOff.push_back((int64_t)f.del.size());
    f.edgeOff.push_back((int64_t)f.eSrc.size());
  }
  return f;
}

dlp_graph_view graphView(const CSRGraph &graph) {
  dlp_graph_view g;
  g.numNodes = graph.numNodes;
  g.numEdges = graph.numEdges;
  g.rowPtr = graph.rowPtr.data();
  g.colInd = graph.colInd.data();
  g.values = graph.values.data();
  return g;
}

dlp_config toCConfig(const DynLPConfig &config) {
  dlp_config c;
  c.totalNodes = config.totalNodes;
  c.avgDegree = config.avgDegree;
  c.numBatches = config.numBatches;
  c.seed = config.seed;
  c.tau = config.tau;
  c.delta = config.delta;
  c.maxIterations = config.maxIterations;
  c.groundTruthFraction = config.groundTruthFraction;
  c.unlabeledFraction = config.unlabeledFraction;
  c.deletedFraction = config.deletedFraction;
  c.validate = config.validate ? 1 : 0;
  c.verbose = config.verbose ? 1 : 0;
  return c;
}

bool runDynLP(const DynLPConfig &config, const CSRGraph &graph,
              const std::vector<float> &trueLabels, const std::vector<BatchData> &batches,
              std::vector<float> &finalLabels, std::vector<float> &batchTimes,
              std::vector<float> &batchAccuracy) {
  FlatBatches flat = flattenBatches(batches);
  dlp_batch_view bv = flat.view();
  dlp_graph_view gv = graphView(graph);
  dlp_config c = toCConfig(config);
  finalLabels.assign(graph.numNodes, 0.5f);
  batchTimes.assign(config.numBatches, 0.0f);
  batchAccuracy.assign(config.numBatches, 0.0f);
  DLP_CHECK(dlp_run_dynlp(&c, &gv, trueLabels.data(), &bv, finalLabels.data(),
                          batchTimes.data(), batchAccuracy.data()));
  return true;
}

bool runIrLPValidation(const CSRGraph &activeGraph, const std::vector<int> &h_vertexStatus,
                       const std::vector<int> &h_gtClass, std::vector<float> &h_labels,
                       int numNodes, float delta, int maxIterations, int &iterationsUsed) {
  dlp_graph_view gv = graphView(activeGraph);
  int skipped = 0;
  iterationsUsed = 0;
  if (numNodes > 0 && activeGraph.numEdges > 0 &&
      ((int)h_vertexStatus.size() < numNodes || (int)h_gtClass.size() < numNodes ||
       (int)h_labels.size() < numNodes)) {
    std::fprintf(stderr, "runIrLPValidation: arrays shorter than numNodes\n");
    std::exit(EXIT_FAILURE);
  }
  DLP_CHECK(dlp_irlp_validation(&gv, h_vertexStatus.data(), h_gtClass.data(), h_labels.data(),
                                numNodes, delta, maxIterations, &iterationsUsed, &skipped));
  if (skipped) std::cout << "Warning: IrLP validation skipped (empty graph).\n";
  return true;
}

/* ---- C-ABI input generator --------------------------------------------- */

struct dlp_inputs {
  CSRGraph graph;
  std::vector<float> trueLabels;
  FlatBatches flat;
};

extern "C" void dlp_set_error_(const char *msg);  // defined in the runtime
void dlp_set_errorf_(const char *fmt, ...);         // defined in the runtime

extern "C" void dlp_config_default(dlp_config *cfg) {
  if (cfg) *cfg = toCConfig(DynLPConfig());
}

extern "C" int dlp_inputs_generate(dlp_inputs **out, int totalNodes, int avgDegree,
                                   int numBatches, unsigned int seed,
                                   float groundTruthFraction, float deletedFraction) {
  if (!out) {
    dlp_set_error_("dlp_inputs_generate: out is NULL");
    return DLP_ERR_INVALID;
  }
  *out = nullptr;
  dlp_inputs *in = new (std::nothrow) dlp_inputs();
  if (!in) {
    dlp_set_error_("dlp_inputs_generate: out of host memory");
    return DLP_ERR_NOMEM;
  }
  DynLPConfig cfg;
  cfg.totalNodes = totalNodes;
  cfg.avgDegree = avgDegree;
  cfg.numBatches = numBatches;
  cfg.seed = seed;
  cfg.groundTruthFraction = groundTruthFraction;
  cfg.deletedFraction = deletedFraction;
  cfg.verbose = false;
  bool saved = dlp_host::g_quiet;
  dlp_host::g_quiet = true;
  std::vector<BatchData> batches;
  bool ok = generateSparseGraph(totalNodes, avgDegree, seed, in->graph, in->trueLabels, "") &&
            generateBatches(in->graph, in->trueLabels, cfg, batches);
  dlp_host::g_quiet = saved;
  if (!ok) {
    delete in;
    dlp_set_error_("dlp_inputs_generate: invalid generator arguments");
    return DLP_ERR_INVALID;
  }
  in->flat = flattenBatches(batches);
  *out = in;
  return DLP_OK;
}

extern "C" int dlp_write_graph_files(const dlp_graph_view *graph, const float *trueLabels,
                                     const char *prefix) {
  if (!graph || !trueLabels || !prefix || graph->numNodes < 1 || graph->numEdges < 0 ||
      !graph->rowPtr || (graph->numEdges > 0 && (!graph->colInd || !graph->values))) {
    dlp_set_error_("dlp_write_graph_files: bad argument");
    return DLP_ERR_INVALID;
  }
  try {
    CSRGraph g;
    g.numNodes = graph->numNodes;
    g.numEdges = graph->numEdges;
    g.rowPtr.assign(graph->rowPtr, graph->rowPtr + graph->numNodes + 1);
    g.colInd.assign(graph->colInd, graph->colInd + graph->numEdges);
    g.values.assign(graph->values, graph->values + graph->numEdges);
    std::vector<float> labels(trueLabels, trueLabels + graph->numNodes);
    if (!writeCSRGraph(prefix, g, labels)) {
      dlp_set_error_("dlp_write_graph_files: could not write the files");
      return DLP_ERR_INVALID;
    }
  } catch (const std::bad_alloc &) {
    dlp_set_error_("dlp_write_graph_files: out of host memory");
    return DLP_ERR_NOMEM;
  }
  return DLP_OK;
}

extern "C" int dlp_inputs_from_graph_files(dlp_inputs **out, const char *prefix,
                                           int numBatches, unsigned int seed,
                                           float groundTruthFraction, float deletedFraction) {
  if (!out || !prefix) {
    dlp_set_error_("dlp_inputs_from_graph_files: NULL argument");
    return DLP_ERR_INVALID;
  }
  *out = nullptr;
  dlp_inputs *in = new (std::nothrow) dlp_inputs();
  if (!in) {
    dlp_set_error_("dlp_inputs_from_graph_files: out of host memory");
    return DLP_ERR_NOMEM;
  }
  if (!readCSRGraph(prefix, in->graph, in->trueLabels)) {
    delete in;
    dlp_set_errorf_("dlp_inputs_from_graph_files: cannot read %s*.txt", prefix);
    return DLP_ERR_INVALID;
  }
  DynLPConfig cfg;
  cfg.totalNodes = in->graph.numNodes;
  cfg.numBatches = numBatches;
  cfg.seed = seed;
  cfg.groundTruthFraction = groundTruthFraction;
  cfg.deletedFraction = deletedFraction;
  cfg.verbose = false;
  bool saved = dlp_host::g_quiet;
  dlp_host::g_quiet = true;
  std::vector<BatchData> batches;
  const bool ok = generateBatches(in->graph, in->trueLabels, cfg, batches);
  dlp_host::g_quiet = saved;
  if (!ok) {
    delete in;
    dlp_set_error_("dlp_inputs_from_graph_files: invalid batch generator arguments");
    return DLP_ERR_INVALID;
  }
  in->flat = flattenBatches(batches);
  *out = in;
  return DLP_OK;
}

extern "C" int dlp_inputs_generate_stream(dlp_inputs **out, int totalNodes, int avgDegree,
                                          const int *batchSizes, int numBatches,
                                          unsigned int seed, float groundTruthFraction,
                                          float deletedFraction) {
  if (!out || !batchSizes || numBatches <= 0) {
    dlp_set_error_("dlp_inputs_generate_stream: bad argument");
    return DLP_ERR_INVALID;
  }
  *out = nullptr;
  dlp_inputs *in = new (std::nothrow) dlp_inputs();
  if (!in) {
    dlp_set_error_("dlp_inputs_generate_stream: out of host memory");
    return DLP_ERR_NOMEM;
  }
  bool saved = dlp_host::g_quiet;
  dlp_host::g_quiet = true;
  std::vector<BatchData> batches;
  std::vector<int> sizes(batchSizes, batchSizes + numBatches);
  bool ok = generateSparseGraph(totalNodes, avgDegree, seed, in->graph, in->trueLabels, "") &&
            generateStreamBatches(in->graph, in->trueLabels, sizes, seed, groundTruthFraction,
                                  deletedFraction, batches);
  dlp_host::g_quiet = saved;
  if (!ok) {
    delete in;
    dlp_set_error_("dlp_inputs_generate_stream: invalid generator arguments");
    return DLP_ERR_INVALID;
  }
  in->flat = flattenBatches(batches);
  *out = in;
  return DLP_OK;
}

extern "C" int dlp_inputs_graph(const dlp_inputs *in, dlp_graph_view *graph,
                                const float **trueLabels) {
  if (!in || !graph) {
    dlp_set_error_("dlp_inputs_graph: NULL argument");
    return DLP_ERR_INVALID;
  }
  *graph = graphView(in->graph);
  if (trueLabels) *trueLabels = in->trueLabels.data();
  return DLP_OK;
}

extern "C" int dlp_inputs_batches(const dlp_inputs *in, dlp_batch_view *batches) {
  if (!in || !batches) {
    dlp_set_error_("dlp_inputs_batches: NULL argument");
    return DLP_ERR_INVALID;
  }
  *batches = in->flat.view();
  return DLP_OK;
}

extern "C" void dlp_inputs_free(dlp_inputs *in) { delete in; }
