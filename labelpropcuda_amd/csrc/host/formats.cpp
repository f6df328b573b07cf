/*
 * formats.cpp -- readers of the text files the driver writes.
 *
 * The reference only writes them: the CSR dump of the generated graph
 * (generateSparseGraph.cu:217-274: <prefix>RowPtr.txt, ColInd.txt,
 * Values.txt with 6 decimals, TrueLabels.txt as "id label"), the labels file
 * ("%d %.6f" per vertex, utils.cu:161-178) and results.csv
 * ("batch,time_ms,accuracy", utils.cu:180-200).  Reading them back lets a
 * stored graph (its weights rounded to the 6 written decimals) be replayed,
 * and results be compared across runs.  Every reader checks the structure
 * it relies on and returns false with a message on the first violation.
 */
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "dynlp/dynLP.h"

namespace {

bool fail(const std::string &path, const std::string &what) {
  std::cout << "Error: " << path << ": " << what << "\n";
  return false;
}

/* one value per line */
template <class T>
bool readColumn(const std::string &path, std::vector<T> &out) {
  std::ifstream f(path);
  if (!f.is_open()) return fail(path, "cannot open");
  out.clear();
  T v;
  while (f >> v) out.push_back(v);
  if (!f.eof()) return fail(path, "malformed value after line " + std::to_string(out.size()));
  return true;
}

/* "id value" per line, ids 0, 1, 2, ... in order */
bool readIndexed(const std::string &path, std::vector<float> &out) {
  std::ifstream f(path);
  if (!f.is_open()) return fail(path, "cannot open");
  out.clear();
  long long id;
  float v;
  while (f >> id >> v) {
    if (id != (long long)out.size())
      return fail(path, "vertex id " + std::to_string(id) + " on line " +
                            std::to_string(out.size() + 1));
    out.push_back(v);
  }
  if (!f.eof()) return fail(path, "malformed line " + std::to_string(out.size() + 1));
  return true;
}

}  // namespace

bool readCSRGraph(const std::string &prefix, CSRGraph &graph, std::vector<float> &trueLabels) {
  std::vector<long long> rowPtr, colInd;
  std::vector<float> values;
  if (!readColumn(prefix + "RowPtr.txt", rowPtr) || !readColumn(prefix + "ColInd.txt", colInd) ||
      !readColumn(prefix + "Values.txt", values) ||
      !readIndexed(prefix + "TrueLabels.txt", trueLabels))
    return false;
  if (rowPtr.size() < 2) return fail(prefix + "RowPtr.txt", "fewer than two row pointers");
  const long long N = (long long)rowPtr.size() - 1, E = (long long)colInd.size();
  if (N > 0x7fffffffLL || E > 0x7fffffffLL) return fail(prefix, "graph exceeds 32-bit ids");
  if (rowPtr[0] != 0 || rowPtr[N] != E)
    return fail(prefix + "RowPtr.txt", "does not span ColInd.txt (" + std::to_string(E) +
                                           " entries)");
  for (long long i = 0; i < N; ++i)
    if (rowPtr[i + 1] < rowPtr[i])
      return fail(prefix + "RowPtr.txt", "decreasing at row " + std::to_string(i));
  if ((long long)values.size() != E)
    return fail(prefix + "Values.txt", "holds " + std::to_string(values.size()) + " weights for " +
                                           std::to_string(E) + " entries");
  if ((long long)trueLabels.size() != N)
    return fail(prefix + "TrueLabels.txt", "holds " + std::to_string(trueLabels.size()) +
                                               " labels for " + std::to_string(N) + " vertices");
  for (long long e = 0; e < E; ++e)
    if (colInd[e] < 0 || colInd[e] >= N)
      return fail(prefix + "ColInd.txt", "vertex id out of range on line " +
                                             std::to_string(e + 1));
  graph.numNodes = (int)N;
  graph.numEdges = (int)E;
  graph.rowPtr.assign(rowPtr.begin(), rowPtr.end());
  graph.colInd.assign(colInd.begin(), colInd.end());
  graph.values = std::move(values);
  return true;
}

bool readLabelsFile(const std::string &path, std::vector<float> &labels) {
  return readIndexed(path, labels);
}

bool readResultsCSV(const std::string &path, std::vector<float> &batchTimes,
                    std::vector<float> &batchAccuracy) {
  std::ifstream f(path);
  if (!f.is_open()) return fail(path, "cannot open");
  std::string line;
  if (!std::getline(f, line) || line != "batch,time_ms,accuracy")
    return fail(path, "missing the batch,time_ms,accuracy header");
  batchTimes.clear();
  batchAccuracy.clear();
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    long long b;
    float t, a;
    char c1, c2;
    std::istringstream row(line);
    if (!(row >> b >> c1 >> t >> c2 >> a) || c1 != ',' || c2 != ',' ||
        b != (long long)batchTimes.size())
      return fail(path, "malformed row " + std::to_string(batchTimes.size() + 1));
    batchTimes.push_back(t);
    batchAccuracy.push_back(a);
  }
  return true;
}
