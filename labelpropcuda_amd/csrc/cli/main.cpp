/*
 * main.cpp -- dynLP command-line driver (drop-in for the reference's
 * bin/dynLP, src/main.cu:53-298): same flags, same output files
 * (dynlp_labels.txt, results.csv, graph/csr*.txt) and the same summary.
 *
 * --validate runs the IrLP full-recompute baseline on the final state of the
 * batch stream (main.cu:120-217, dlp_irlp_baseline), writes irlp_labels.txt
 * and adds the IrLP time, accuracy and speedup to the summary.
 */
#include <iomanip>
#include <iostream>
#include <numeric>
#include <vector>

#include "dynlp/dynLP.h"

int main(int argc, char *argv[]) {
  DynLPConfig config;
  if (!parseArguments(argc, argv, config)) return 1;
  printConfig(config);
  if (!ensureDirectory(config.outputDir)) return 1;
  if (!ensureDirectory(config.outputDir + "/graph")) return 1;

  std::cout << "Generating sparse graph...\n";
  CSRGraph fullGraph;
  std::vector<float> trueLabels;
  if (!generateSparseGraph(config.totalNodes, config.avgDegree, config.seed, fullGraph, trueLabels,
                           config.outputDir + "/graph/csr")) {
    std::cout << "Error: Graph generation failed.\n";
    return 1;
  }
  std::cout << "Generating batches...\n";
  std::vector<BatchData> batches;
  if (!generateBatches(fullGraph, trueLabels, config, batches)) {
    std::cout << "Error: Batch generation failed.\n";
    return 1;
  }

  std::cout << "Running DynLP...\n\n";
  std::vector<float> labels, batchTimes, batchAccuracy;
  if (!runDynLP(config, fullGraph, trueLabels, batches, labels, batchTimes, batchAccuracy)) {
    std::cout << "Error: DynLP execution failed.\n";
    return 1;
  }
  std::vector<float> irlpLabels;
  dlp_irlp_result irlp{};
  if (config.validate) {
    std::cout << "Running IrLP baseline for validation...\n";
    FlatBatches flat = flattenBatches(batches);
    dlp_batch_view bv = flat.view();
    dlp_graph_view gv = graphView(fullGraph);
    dlp_config c = toCConfig(config);
    irlpLabels.assign(config.totalNodes, 0.5f);
    DLP_CHECK(dlp_irlp_baseline(&c, &gv, trueLabels.data(), &bv, nullptr, irlpLabels.data(),
                                &irlp));
    if (irlp.skipped) std::cout << "Warning: IrLP validation skipped (empty graph).\n";
    std::cout << "IrLP baseline: " << irlp.iterations << " iterations, " << irlp.timeMs
              << " ms, accuracy " << std::fixed << std::setprecision(2)
              << (irlp.accuracy * 100.0f) << "%\n\n";
  }

  if (!writeLabelsToFile(config.outputDir + "/dynlp_labels.txt", labels.data(), config.totalNodes)) {
    std::cout << "Error: Could not write DynLP labels.\n";
    return 1;
  }
  if (!writeResultsCSV(config.outputDir + "/results.csv", batchTimes.data(), batchAccuracy.data(),
                       config.numBatches)) {
    std::cout << "Error: Could not write results CSV.\n";
    return 1;
  }
  if (config.validate &&
      !writeLabelsToFile(config.outputDir + "/irlp_labels.txt", irlpLabels.data(),
                         config.totalNodes)) {
    std::cout << "Error: Could not write IrLP labels.\n";
    return 1;
  }

  float totalTime = std::accumulate(batchTimes.begin(), batchTimes.end(), 0.0f);
  float avgAcc =
      std::accumulate(batchAccuracy.begin(), batchAccuracy.end(), 0.0f) / config.numBatches;
  std::cout << "============================================================\n"
            << " DynLP Results Summary\n"
            << "============================================================\n"
            << "  Total DynLP time:     " << std::fixed << std::setprecision(2) << totalTime
            << " ms\n"
            << "  Average accuracy:     " << std::fixed << std::setprecision(2) << (avgAcc * 100.0f)
            << "%\n"
            << "  Final batch accuracy: " << std::fixed << std::setprecision(2)
            << (batchAccuracy.back() * 100.0f) << "%\n";
  if (config.validate) {
    const float speedup = irlp.timeMs > 0.0f ? irlp.timeMs / totalTime : 0.0f;
    std::cout << "  IrLP baseline time:   " << std::fixed << std::setprecision(2) << irlp.timeMs
              << " ms\n"
              << "  IrLP accuracy:        " << std::fixed << std::setprecision(2)
              << (irlp.accuracy * 100.0f) << "%\n"
              << "  Speedup vs IrLP:      " << std::fixed << std::setprecision(1) << speedup
              << "x\n";
  }
  std::cout << "============================================================\n\n"
            << "Per-batch results:\n";
  for (int b = 0; b < config.numBatches; ++b)
    std::cout << "  Batch " << std::setw(2) << b << ": " << std::fixed << std::setprecision(2)
              << std::setw(10) << batchTimes[b] << " ms, accuracy " << std::fixed
              << std::setprecision(2) << (batchAccuracy[b] * 100.0f) << "%\n";
  std::cout << "\nOutput files written to: " << config.outputDir << "/\n"
            << "  dynlp_labels.txt  - Final predicted labels\n"
            << "  results.csv       - Per-batch timing and accuracy\n";
  if (config.validate) std::cout << "  irlp_labels.txt   - IrLP baseline labels\n";
  std::cout << "  graph/            - Generated graph in CSR format\n";
  return 0;
}
