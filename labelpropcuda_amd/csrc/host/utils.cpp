/*
 * utils.cpp -- command-line flags, the configuration banner and the two
 * result files of the DynLP driver.
 *
 * Interface kept from the reference (utils.cu:24-200) byte for byte: the
 * flag names and their meaning, the error messages, the --help text, the
 * banner layout, dynlp_labels.txt ("<id> <label>" with 6 decimals) and
 * results.csv ("batch,time_ms,accuracy", 2 and 6 decimals).  The code is
 * organised around one flag table: parsing, value checks and --help are all
 * driven by it.
 */
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <memory>
#include <sstream>
#include <string>
#include <system_error>
#include <utility>
#include <vector>

#include "dynlp/dynLP.h"

namespace fs = std::filesystem;

namespace {

/* what a flag's value must satisfy; a failed check prints
 * "Error: <flag> must be <rule text>." */
enum class Check { None, Positive, OpenUnit };

/* one command-line flag: how it is spelled, what it sets, how it is shown */
struct Flag {
  const char *name;
  const char *meta;    /* value placeholder in --help ("" = no value)    */
  const char *text;    /* --help description                             */
  Check check;
  /* applies the flag; `value` is nullptr for flags without a value */
  void (*apply)(DynLPConfig &, const char *value);
  /* the numeric value the check looks at, after apply */
  double (*checked)(const DynLPConfig &);
};

double asInt(int v) { return (double)v; }

const std::vector<Flag> &flagTable() {
  static const std::vector<Flag> table = {
      {"--totalNodes", "N", "Total vertices across all batches (default: 50000)",
       Check::Positive, [](DynLPConfig &c, const char *v) { c.totalNodes = std::atoi(v); },
       [](const DynLPConfig &c) { return asInt(c.totalNodes); }},
      {"--numBatches", "N", "Number of dynamic batches (default: 10)", Check::Positive,
       [](DynLPConfig &c, const char *v) { c.numBatches = std::atoi(v); },
       [](const DynLPConfig &c) { return asInt(c.numBatches); }},
      {"--avgDegree", "N", "Average degree for graph gen (default: 5)", Check::Positive,
       [](DynLPConfig &c, const char *v) { c.avgDegree = std::atoi(v); },
       [](const DynLPConfig &c) { return asInt(c.avgDegree); }},
      {"--tau", "F", "Sparsification threshold (default: auto)", Check::None,
       [](DynLPConfig &c, const char *v) { c.tau = (float)std::atof(v); }, nullptr},
      {"--delta", "F", "Convergence threshold (default: 0.0001)", Check::Positive,
       [](DynLPConfig &c, const char *v) { c.delta = (float)std::atof(v); },
       [](const DynLPConfig &c) { return (double)c.delta; }},
      {"--maxIter", "N", "Max propagation iterations (default: 100000)", Check::None,
       [](DynLPConfig &c, const char *v) { c.maxIterations = std::atoi(v); }, nullptr},
      {"--gtFraction", "F", "Ground-truth fraction (default: 0.01)", Check::OpenUnit,
       [](DynLPConfig &c, const char *v) { c.groundTruthFraction = (float)std::atof(v); },
       [](const DynLPConfig &c) { return (double)c.groundTruthFraction; }},
      {"--seed", "N", "RNG seed (default: 42)", Check::None,
       [](DynLPConfig &c, const char *v) { c.seed = (unsigned int)std::atoi(v); }, nullptr},
      {"--output", "DIR", "Output directory (default: output/)", Check::None,
       [](DynLPConfig &c, const char *v) { c.outputDir = v; }, nullptr},
      {"--validate", "", "Run IrLP baseline comparison", Check::None,
       [](DynLPConfig &c, const char *) { c.validate = true; }, nullptr},
      {"--noValidate", "", "Skip IrLP baseline", Check::None,
       [](DynLPConfig &c, const char *) { c.validate = false; }, nullptr},
      {"--verbose", "", "Print per-batch details", Check::None,
       [](DynLPConfig &c, const char *) { c.verbose = true; }, nullptr},
      {"--quiet", "", "Suppress per-batch details", Check::None,
       [](DynLPConfig &c, const char *) { c.verbose = false; }, nullptr},
  };
  return table;
}

/* "  <spelling, 13 wide><placeholder, 5 wide><text>" */
std::string helpLine(const std::string &spelling, const char *meta, const char *text) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), "  %-13s%-5s%s\n", spelling.c_str(), meta, text);
  return buf;
}

void printHelp() {
  std::string out = "Usage: dynLP [options]\n\nOptions:\n";
  for (const Flag &f : flagTable()) out += helpLine(f.name, f.meta, f.text);
  out += helpLine("--help, -h", "", "Show this help message");
  std::fputs(out.c_str(), stdout);
}

bool passes(Check check, double v) {
  switch (check) {
    case Check::Positive: return v > 0.0;
    case Check::OpenUnit: return v > 0.0 && v < 1.0;
    default: return true;
  }
}

const char *ruleText(Check check) {
  return check == Check::OpenUnit ? "in (0, 1)" : "positive";
}

const Flag *lookup(const std::string &arg) {
  for (const Flag &f : flagTable())
    if (arg == f.name) return &f;
  return nullptr;
}

/* value formatted the way an ostream prints it by default */
template <class T>
std::string shown(const T &v) {
  std::ostringstream os;
  os << v;
  return os.str();
}

/* "Error: <what>: <filesystem_error text>" for a failed directory creation */
std::string fsErrorText(const std::string &path, const std::error_code &ec) {
  return fs::filesystem_error("cannot create directories", fs::path(path), ec).what();
}

/* opens `path` for writing, creating its parent directory first */
std::unique_ptr<FILE, int (*)(FILE *)> openOutput(const std::string &path) {
  const fs::path parent = fs::path(path).parent_path();
  if (!parent.empty()) ensureDirectory(parent.string());
  return std::unique_ptr<FILE, int (*)(FILE *)>(std::fopen(path.c_str(), "w"), &std::fclose);
}

}  // namespace

bool parseArguments(int argc, char *argv[], DynLPConfig &config) {
  for (int i = 1; i < argc; ++i) {
    const std::string arg = argv[i];
    if (arg == "--help" || arg == "-h") {
      printHelp();
      return false;
    }
    const Flag *f = lookup(arg);
    const bool needsValue = f && f->meta[0] != '\0';
    if (!f || (needsValue && i + 1 >= argc)) {
      std::printf("Error: Unknown argument: %s\nRun with --help for usage information.\n",
                  arg.c_str());
      return false;
    }
    f->apply(config, needsValue ? argv[++i] : nullptr);
    if (f->checked && !passes(f->check, f->checked(config))) {
      std::printf("Error: %s must be %s.\n", f->name, ruleText(f->check));
      return false;
    }
  }
  if (config.groundTruthFraction + config.unlabeledFraction + config.deletedFraction > 1.001f) {
    std::printf("Error: Batch fractions (GT + unlabeled + deleted) exceed 1.0.\n");
    return false;
  }
  return true;
}

void printConfig(const DynLPConfig &config) {
  const std::string rule(60, '=');
  const std::vector<std::pair<const char *, std::string>> rows = {
      {"Total nodes:", shown(config.totalNodes)},
      {"Number of batches:", shown(config.numBatches)},
      {"Average degree:", shown(config.avgDegree)},
      {"Tau (sparsify):",
       config.tau < 0 ? std::string("auto (avg edge weight)") : std::to_string(config.tau)},
      {"Delta (converge):", shown(config.delta)},
      {"Max iterations:", shown(config.maxIterations)},
      {"GT fraction:", shown(config.groundTruthFraction)},
      {"Unlabeled fraction:", shown(config.unlabeledFraction)},
      {"Deleted fraction:", shown(config.deletedFraction)},
      {"Seed:", shown(config.seed)},
      {"Output directory:", config.outputDir},
      {"Validate (IrLP):", config.validate ? "yes" : "no"},
      {"Verbose:", config.verbose ? "yes" : "no"},
  };
  std::string out = "\n" + rule + "\n DynLP Configuration\n" + rule + "\n";
  for (const auto &row : rows) {
    char label[32];
    std::snprintf(label, sizeof(label), "  %-20s", row.first);
    out += label + row.second + "\n";
  }
  out += rule + "\n\n";
  std::fputs(out.c_str(), stdout);
}

bool ensureDirectory(const std::string &dirPath) {
  std::error_code ec;
  const bool present = fs::exists(dirPath, ec);
  if (!ec && !present) fs::create_directories(dirPath, ec);
  if (!ec) return true;
  std::printf("Error: Could not create directory %s: %s\n", dirPath.c_str(),
              fsErrorText(dirPath, ec).c_str());
  return false;
}

bool writeLabelsToFile(const std::string &path, const float *labels, int numNodes) {
  auto f = openOutput(path);
  if (!f) {
    std::printf("Error: Could not open file for writing: %s\n", path.c_str());
    return false;
  }
  for (int i = 0; i < numNodes; ++i) std::fprintf(f.get(), "%d %.6f\n", i, (double)labels[i]);
  return true;
}

bool writeResultsCSV(const std::string &path, const float *batchTimes,
                     const float *batchAccuracy, int numBatches) {
  auto f = openOutput(path);
  if (!f) {
    std::printf("Error: Could not open CSV file for writing: %s\n", path.c_str());
    return false;
  }
  std::fputs("batch,time_ms,accuracy\n", f.get());
  for (int i = 0; i < numBatches; ++i)
    std::fprintf(f.get(), "%d,%.2f,%.6f\n", i, (double)batchTimes[i], (double)batchAccuracy[i]);
  return true;
}
