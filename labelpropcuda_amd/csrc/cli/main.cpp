/*
 * main.cpp -- dynLP command-line driver (drop-in for the reference's
 * bin/dynLP, src/main.cu:53-298): same flags, same output files
 * (dynlp_labels.txt, results.csv, graph/csr*.txt, irlp_labels.txt with
 * --validate) and the same summary text.
 *
 * --validate runs the IrLP full-recompute baseline on the final state of the
 * batch stream through the C-ABI (dlp_irlp_baseline, main.cu:120-217).  The
 * result files are a table (name, writer, failure message, description)
 * that drives both the writing and the file list of the summary; the
 * summary's figures are rows of one label column.
 */
#include <cstdio>
#include <functional>
#include <numeric>
#include <string>
#include <vector>

#include "dynlp/dynLP.h"

namespace {

/* one file of the run's output directory */
struct OutputFile {
  std::string name;         /* relative to the output directory */
  std::string description;  /* line of the summary's file list  */
  std::string failure;      /* message when it cannot be written */
  std::function<bool(const std::string &path)> write;
};

/* "  <label, 22 wide><value>" */
void summaryRow(std::string &out, const char *label, const char *fmt, double value) {
  char buf[128];
  int n = std::snprintf(buf, sizeof(buf), "  %-22s", label);
  std::snprintf(buf + n, sizeof(buf) - n, fmt, value);
  out += buf;
  out += "\n";
}

std::string summaryText(const DynLPConfig &config, const std::vector<float> &batchTimes,
                        const std::vector<float> &batchAccuracy, const dlp_irlp_result *irlp,
                        const std::vector<OutputFile> &files) {
  const std::string rule(60, '=');
  const float total = std::accumulate(batchTimes.begin(), batchTimes.end(), 0.0f);
  const float meanAcc =
      std::accumulate(batchAccuracy.begin(), batchAccuracy.end(), 0.0f) / config.numBatches;
  std::string out = rule + "\n DynLP Results Summary\n" + rule + "\n";
  summaryRow(out, "Total DynLP time:", "%.2f ms", total);
  summaryRow(out, "Average accuracy:", "%.2f%%", meanAcc * 100.0f);
  summaryRow(out, "Final batch accuracy:", "%.2f%%", batchAccuracy.back() * 100.0f);
  if (irlp) {
    summaryRow(out, "IrLP baseline time:", "%.2f ms", irlp->timeMs);
    summaryRow(out, "IrLP accuracy:", "%.2f%%", irlp->accuracy * 100.0f);
    summaryRow(out, "Speedup vs IrLP:", "%.1fx",
               irlp->timeMs > 0.0f ? irlp->timeMs / total : 0.0f);
  }
  out += rule + "\n\nPer-batch results:\n";
  for (int b = 0; b < config.numBatches; ++b) {
    char buf[128];
    std::snprintf(buf, sizeof(buf), "  Batch %2d: %10.2f ms, accuracy %.2f%%\n", b,
                  (double)batchTimes[b], (double)(batchAccuracy[b] * 100.0f));
    out += buf;
  }
  out += "\nOutput files written to: " + config.outputDir + "/\n";
  for (const OutputFile &f : files) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "  %-18s— %s\n", f.name.c_str(), f.description.c_str());
    out += buf;
  }
  out += "  graph/            — Generated graph in CSR format\n";
  return out;
}

}  // namespace

int main(int argc, char *argv[]) {
  DynLPConfig config;
  if (!parseArguments(argc, argv, config)) return 1;
  printConfig(config);
  if (!ensureDirectory(config.outputDir) || !ensureDirectory(config.outputDir + "/graph"))
    return 1;

  std::printf("Generating sparse graph...\n");
  CSRGraph fullGraph;
  std::vector<float> trueLabels;
  if (!generateSparseGraph(config.totalNodes, config.avgDegree, config.seed, fullGraph, trueLabels,
                           config.outputDir + "/graph/csr")) {
    std::printf("Error: Graph generation failed.\n");
    return 1;
  }
  std::printf("Generating batches...\n");
  std::vector<BatchData> batches;
  if (!generateBatches(fullGraph, trueLabels, config, batches)) {
    std::printf("Error: Batch generation failed.\n");
    return 1;
  }

  std::printf("Running DynLP...\n\n");
  std::fflush(stdout);
  std::vector<float> labels, batchTimes, batchAccuracy;
  if (!runDynLP(config, fullGraph, trueLabels, batches, labels, batchTimes, batchAccuracy)) {
    std::printf("Error: DynLP execution failed.\n");
    return 1;
  }

  std::vector<float> irlpLabels;
  dlp_irlp_result irlp{};
  if (config.validate) {
    std::printf("Running IrLP baseline for validation...\n");
    std::fflush(stdout);
    FlatBatches flat = flattenBatches(batches);
    dlp_batch_view bv = flat.view();
    dlp_graph_view gv = graphView(fullGraph);
    dlp_config c = toCConfig(config);
    irlpLabels.assign(config.totalNodes, 0.5f);
    DLP_CHECK(dlp_irlp_baseline(&c, &gv, trueLabels.data(), &bv, nullptr, irlpLabels.data(),
                                &irlp));
    if (irlp.skipped) std::printf("Warning: IrLP validation skipped (empty graph).\n");
    std::printf("IrLP baseline: %d iterations, %g ms, accuracy %.2f%%\n\n", irlp.iterations,
                (double)irlp.timeMs, (double)(irlp.accuracy * 100.0f));
  }

  std::vector<OutputFile> files = {
      {"dynlp_labels.txt", "Final predicted labels", "Error: Could not write DynLP labels.",
       [&](const std::string &p) { return writeLabelsToFile(p, labels.data(), config.totalNodes); }},
      {"results.csv", "Per-batch timing and accuracy", "Error: Could not write results CSV.",
       [&](const std::string &p) {
         return writeResultsCSV(p, batchTimes.data(), batchAccuracy.data(), config.numBatches);
       }},
  };
  if (config.validate)
    files.push_back({"irlp_labels.txt", "IrLP baseline labels", "Error: Could not write IrLP labels.",
                     [&](const std::string &p) {
                       return writeLabelsToFile(p, irlpLabels.data(), config.totalNodes);
                     }});
  for (const OutputFile &f : files)
    if (!f.write(config.outputDir + "/" + f.name)) {
      std::printf("%s\n", f.failure.c_str());
      return 1;
    }

  const std::string summary = summaryText(config, batchTimes, batchAccuracy,
                                          config.validate ? &irlp : nullptr, files);
  std::fputs(summary.c_str(), stdout);
  return 0;
}
