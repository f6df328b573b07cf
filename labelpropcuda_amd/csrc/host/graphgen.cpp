/*
 * graphgen.cpp -- synthetic inputs of the DynLP benchmark (host C++).
 *
 * generateSparseGraph follows generateSparseGraph.cu:118-277 and
 * generateBatches follows generateBatches.cu:44-223: the same std::mt19937
 * seeds, the same distributions and the same sequence of RNG calls, so the
 * graph and the batch stream are identical to the reference's for a given
 * seed (same libstdc++).  Only membership-only containers differ:
 *   - the edge dedup set (reference std::set<pair<int,int>>) is an
 *     open-addressing hash set of packed (min,max) keys -- it is only ever
 *     queried for membership, so the output is unchanged, and 175 M edges
 *     fit in a few GB instead of a red-black tree;
 *   - the GT / active-non-GT sets of the batch generator are byte maps; the
 *     only ordered use (the deletion candidates, copied from a std::set in
 *     ascending order, generateBatches.cu:475) is rebuilt by an ascending
 *     scan.
 */
#include <algorithm>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "dynlp/dynLP.h"
#include "host_internal.h"

namespace dlp_host {
bool g_quiet = false;
}

namespace {

/* Open-addressing set of uint64 keys (linear probing, power-of-two table). */
class EdgeSet {
 public:
  explicit EdgeSet(long long expected) {
    uint64_t cap = 16;
    while ((double)cap * 0.6 < (double)expected) cap <<= 1;
    table_.assign(cap, kEmpty);
    mask_ = cap - 1;
  }
  bool contains(uint64_t key) const {
    for (uint64_t i = hash(key) & mask_;; i = (i + 1) & mask_) {
      if (table_[i] == key) return true;
      if (table_[i] == kEmpty) return false;
    }
  }
  void insert(uint64_t key) {
    if ((double)(size_ + 1) > 0.7 * (double)table_.size()) grow();
    place(key);
  }

 private:
  static constexpr uint64_t kEmpty = ~0ull;
  static uint64_t hash(uint64_t x) {  // splitmix64 finaliser
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
  }
  void place(uint64_t key) {
    for (uint64_t i = hash(key) & mask_;; i = (i + 1) & mask_) {
      if (table_[i] == key) return;
      if (table_[i] == kEmpty) {
        table_[i] = key;
        ++size_;
        return;
      }
    }
  }
  void grow() {
    std::vector<uint64_t> old;
    old.swap(table_);
    table_.assign(old.size() * 2, kEmpty);
    mask_ = table_.size() - 1;
    size_ = 0;
    for (uint64_t k : old)
      if (k != kEmpty) place(k);
  }
  std::vector<uint64_t> table_;
  uint64_t mask_ = 0;
  uint64_t size_ = 0;
};

inline uint64_t edgeKey(int a, int b) {
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

/* Edge list -> symmetric CSR with rows sorted by column
 * (generateSparseGraph.cu:51-110).  Columns within a row are distinct, so a
 * sort on the column gives the same rows as the reference's insertion
 * sort. */
void edgeListToCSR(const std::vector<std::pair<int, int>> &edges,
                   const std::vector<float> &weights, int numNodes,
                   CSRGraph &graph) {
  std::vector<int> degree(numNodes, 0);
  for (const auto &e : edges) {
    degree[e.first]++;
    degree[e.second]++;
  }
  graph.numNodes = numNodes;
  graph.rowPtr.assign(numNodes + 1, 0);
  for (int i = 0; i < numNodes; ++i) graph.rowPtr[i + 1] = graph.rowPtr[i] + degree[i];
  int total = graph.rowPtr[numNodes];
  graph.numEdges = total;
  graph.colInd.assign(total, 0);
  graph.values.assign(total, 0.0f);
  std::vector<int> offset(numNodes, 0);
  for (size_t e = 0; e < edges.size(); ++e) {
    int u = edges[e].first, v = edges[e].second;
    float w = weights[e];
    int pu = graph.rowPtr[u] + offset[u]++;
    graph.colInd[pu] = v;
    graph.values[pu] = w;
    int pv = graph.rowPtr[v] + offset[v]++;
    graph.colInd[pv] = u;
    graph.values[pv] = w;
  }
  std::vector<std::pair<int, float>> row;
  for (int u = 0; u < numNodes; ++u) {
    int s = graph.rowPtr[u], t = graph.rowPtr[u + 1];
    row.clear();
    for (int i = s; i < t; ++i) row.emplace_back(graph.colInd[i], graph.values[i]);
    std::sort(row.begin(), row.end(),
              [](const std::pair<int, float> &a, const std::pair<int, float> &b) {
                return a.first < b.first;
              });
    for (int i = s; i < t; ++i) {
      graph.colInd[i] = row[i - s].first;
      graph.values[i] = row[i - s].second;
    }
  }
}

}  // namespace

bool generateSparseGraph(int totalNodes, int avgDegree, unsigned int seed,
                         CSRGraph &graph, std::vector<float> &trueLabels,
                         const std::string &outputPrefix) {
  if (totalNodes <= 1) {
    std::cout << "Error: totalNodes must be > 1.\n";
    return false;
  }
  if (avgDegree <= 0) {
    std::cout << "Error: avgDegree must be positive.\n";
    return false;
  }
  std::mt19937 rng(seed);

  // balanced labels, shuffled (:133-138)
  trueLabels.resize(totalNodes);
  for (int i = 0; i < totalNodes; ++i) trueLabels[i] = (i < totalNodes / 2) ? 0.0f : 1.0f;
  std::shuffle(trueLabels.begin(), trueLabels.end(), rng);

  // spanning tree over a shuffled permutation (:142-169)
  std::vector<int> perm(totalNodes);
  std::iota(perm.begin(), perm.end(), 0);
  std::shuffle(perm.begin(), perm.end(), rng);

  long long targetEdges = static_cast<long long>(totalNodes) * avgDegree / 2;
  if (targetEdges < static_cast<long long>(totalNodes) - 1) targetEdges = totalNodes - 1;

  EdgeSet edgeSet(targetEdges);
  std::vector<std::pair<int, int>> edgeList;
  std::vector<float> weightList;
  edgeList.reserve((size_t)targetEdges);
  weightList.reserve((size_t)targetEdges);

  std::uniform_real_distribution<float> sameClassDist(0.5f, 1.0f);
  std::uniform_real_distribution<float> crossClassDist(0.01f, 0.5f);

  for (int i = 0; i < totalNodes - 1; ++i) {
    int u = std::min(perm[i], perm[i + 1]);
    int v = std::max(perm[i], perm[i + 1]);
    if (u == v) continue;
    edgeSet.insert(edgeKey(u, v));
    float w = (trueLabels[u] == trueLabels[v]) ? sameClassDist(rng) : crossClassDist(rng);
    edgeList.push_back({u, v});
    weightList.push_back(w);
  }

  // uniform extra edges (:171-208)
  std::uniform_int_distribution<int> nodeDist(0, totalNodes - 1);
  int attempts = 0;
  int maxAttempts = static_cast<int>(targetEdges) * 10;
  while (static_cast<long long>(edgeList.size()) < targetEdges && attempts < maxAttempts) {
    int u = nodeDist(rng);
    int v = nodeDist(rng);
    if (u == v) {
      attempts++;
      continue;
    }
    int a = std::min(u, v);
    int b = std::max(u, v);
    uint64_t key = edgeKey(a, b);
    if (edgeSet.contains(key)) {
      attempts++;
      continue;
    }
    edgeSet.insert(key);
    float w = (trueLabels[a] == trueLabels[b]) ? sameClassDist(rng) : crossClassDist(rng);
    edgeList.push_back({a, b});
    weightList.push_back(w);
    attempts = 0;
  }

  if (!dlp_host::g_quiet)
    std::cout << "Generated graph: " << totalNodes << " nodes, " << edgeList.size()
            << " undirected edges (avg degree ~" << (2.0 * edgeList.size() / totalNodes)
            << ")\n";

  edgeListToCSR(edgeList, weightList, totalNodes, graph);

  if (!outputPrefix.empty()) {  // text dump (:217-274)
    if (!writeCSRGraph(outputPrefix, graph, trueLabels)) return false;
    if (!dlp_host::g_quiet) std::cout << "CSR graph written to: " << outputPrefix << "*.txt\n";
  }
  return true;
}

bool writeCSRGraph(const std::string &prefix, const CSRGraph &graph,
                   const std::vector<float> &trueLabels) {
  const size_t slash = prefix.find_last_of("/");
  if (slash != std::string::npos && !ensureDirectory(prefix.substr(0, slash))) return false;
  {
    std::ofstream f(prefix + "RowPtr.txt");
    if (!f.is_open()) {
      std::cout << "Error: Could not write RowPtr file.\n";
      return false;
    }
    for (int i = 0; i <= graph.numNodes; ++i) f << graph.rowPtr[i] << "\n";
  }
  {
    std::ofstream f(prefix + "ColInd.txt");
    if (!f.is_open()) {
      std::cout << "Error: Could not write ColInd file.\n";
      return false;
    }
    for (int i = 0; i < graph.numEdges; ++i) f << graph.colInd[i] << "\n";
  }
  {
    std::ofstream f(prefix + "Values.txt");
    if (!f.is_open()) {
      std::cout << "Error: Could not write Values file.\n";
      return false;
    }
    f << std::fixed;
    for (int i = 0; i < graph.numEdges; ++i) f << std::setprecision(6) << graph.values[i] << "\n";
  }
  {
    std::ofstream f(prefix + "TrueLabels.txt");
    if (!f.is_open()) {
      std::cout << "Error: Could not write TrueLabels file.\n";
      return false;
    }
    for (int i = 0; i < graph.numNodes; ++i) f << i << " " << trueLabels[i] << "\n";
  }
  return true;
}

bool generateBatches(const CSRGraph &graph, const std::vector<float> &trueLabels,
                     const DynLPConfig &config, std::vector<BatchData> &batches) {
  int totalNodes = graph.numNodes;
  int numBatches = config.numBatches;
  if (numBatches <= 0 || totalNodes <= 0) {
    std::cout << "Error: Invalid numBatches or totalNodes.\n";
    return false;
  }
  std::mt19937 rng(config.seed + 12345);

  // Phase A: class lists, batch 0 = balanced GT (:66-99)
  std::vector<int> class0, class1;
  for (int i = 0; i < totalNodes; ++i) (trueLabels[i] < 0.5f ? class0 : class1).push_back(i);
  std::shuffle(class0.begin(), class0.end(), rng);
  std::shuffle(class1.begin(), class1.end(), rng);

  int numInitialGT = std::max(2, static_cast<int>(totalNodes * config.groundTruthFraction));
  int numGTPerClass = numInitialGT / 2;

  batches.assign(numBatches, BatchData());
  batches[0].batchId = 0;
  std::vector<char> isInitialGT(totalNodes, 0);
  for (int i = 0; i < numGTPerClass && i < static_cast<int>(class0.size()); ++i) {
    int v = class0[i];
    batches[0].insertedNodeIds.push_back(v);
    batches[0].groundTruthNodeIds.push_back(v);
    batches[0].groundTruthLabels.push_back(0.0f);
    isInitialGT[v] = 1;
  }
  for (int i = 0; i < numGTPerClass && i < static_cast<int>(class1.size()); ++i) {
    int v = class1[i];
    batches[0].insertedNodeIds.push_back(v);
    batches[0].groundTruthNodeIds.push_back(v);
    batches[0].groundTruthLabels.push_back(1.0f);
    isInitialGT[v] = 1;
  }

  // remaining vertices, shuffled, split evenly (:101-136)
  std::vector<int> remaining;
  remaining.reserve(totalNodes);
  for (int i = 0; i < totalNodes; ++i)
    if (!isInitialGT[i]) remaining.push_back(i);
  std::shuffle(remaining.begin(), remaining.end(), rng);

  int numRemaining = static_cast<int>(remaining.size());
  int perBatch = numRemaining / std::max(1, numBatches - 1);
  int leftover = numRemaining - perBatch * std::max(1, numBatches - 1);
  int idx = 0;
  for (int b = 1; b < numBatches; ++b) {
    batches[b].batchId = b;
    int count = perBatch + (b <= leftover ? 1 : 0);
    for (int j = 0; j < count && idx < numRemaining; ++j) batches[b].insertedNodeIds.push_back(remaining[idx++]);
    int numBatchGT = std::max(
        1, static_cast<int>(batches[b].insertedNodeIds.size() * config.groundTruthFraction));
    std::vector<int> shuffled = batches[b].insertedNodeIds;
    std::shuffle(shuffled.begin(), shuffled.end(), rng);
    for (int j = 0; j < numBatchGT && j < static_cast<int>(shuffled.size()); ++j) {
      int v = shuffled[j];
      batches[b].groundTruthNodeIds.push_back(v);
      batches[b].groundTruthLabels.push_back(trueLabels[v]);
    }
  }

  // Phase B: each edge goes to max(batch[u], batch[v]) (:145-165)
  std::vector<int> vertexBatch(totalNodes, -1);
  for (int b = 0; b < numBatches; ++b)
    for (int v : batches[b].insertedNodeIds) vertexBatch[v] = b;
  for (int u = 0; u < totalNodes; ++u) {
    if (vertexBatch[u] < 0) continue;
    for (int e = graph.rowPtr[u]; e < graph.rowPtr[u + 1]; ++e) {
      int v = graph.colInd[e];
      if (v < 0 || vertexBatch[v] < 0) continue;
      if (u < v) {
        int ab = std::max(vertexBatch[u], vertexBatch[v]);
        batches[ab].newEdgeSrc.push_back(u);
        batches[ab].newEdgeDst.push_back(v);
        batches[ab].newEdgeWeight.push_back(graph.values[e]);
      }
    }
  }

  // Phase C: deletions from active non-GT vertices (:167-206)
  std::vector<char> inGT(isInitialGT), activeNonGT(totalNodes, 0);
  long long activeCount = 0;
  std::vector<int> candidates;
  for (int b = 1; b < numBatches; ++b) {
    for (int v : batches[b].groundTruthNodeIds) inGT[v] = 1;
    for (int v : batches[b].insertedNodeIds)
      if (!inGT[v] && !activeNonGT[v]) {
        activeNonGT[v] = 1;
        ++activeCount;
      }
    int numToDelete = static_cast<int>(batches[b].insertedNodeIds.size() * config.deletedFraction);
    numToDelete = std::min(numToDelete, static_cast<int>(activeCount));
    if (numToDelete > 0) {
      candidates.clear();
      candidates.reserve((size_t)activeCount);
      for (int v = 0; v < totalNodes; ++v)
        if (activeNonGT[v]) candidates.push_back(v);
      std::shuffle(candidates.begin(), candidates.end(), rng);
      for (int j = 0; j < numToDelete; ++j) {
        int v = candidates[j];
        batches[b].deletedNodeIds.push_back(v);
        activeNonGT[v] = 0;
        --activeCount;
      }
    }
  }

  if (!dlp_host::g_quiet) {
    std::cout << "\nBatch generation summary:\n";
    for (int b = 0; b < numBatches; ++b)
      std::cout << "  Batch " << b << ": " << batches[b].insertedNodeIds.size() << " inserted ("
                << batches[b].groundTruthNodeIds.size() << " GT), "
                << batches[b].deletedNodeIds.size() << " deleted, "
                << batches[b].newEdgeSrc.size() << " edges\n";
    std::cout << "\n";
  }
  return true;
}

/* Streaming variant for arrival-driven workloads (no counterpart in the
 * reference): batch 0 is the balanced GT seed batch of generateBatches;
 * batch b >= 1 inserts the next sizes[b] vertices of one shuffled order
 * (fewer once the graph is exhausted), with max(1, int(k * gtFraction)) GT
 * vertices drawn from the batch and min(int(k * deletedFraction), active)
 * deletions drawn uniformly from the active non-GT vertices (swap-remove
 * pool, O(k) per batch).  Edges go to batch max(b_u, b_v) as in
 * generateBatches. */
bool generateStreamBatches(const CSRGraph &graph, const std::vector<float> &trueLabels,
                           const std::vector<int> &sizes, unsigned int seed,
                           float groundTruthFraction, float deletedFraction,
                           std::vector<BatchData> &batches) {
  const int totalNodes = graph.numNodes;
  const int numBatches = (int)sizes.size();
  if (numBatches <= 0 || totalNodes <= 0) return false;
  std::mt19937 rng(seed + 12345);
  std::vector<int> class0, class1;
  for (int i = 0; i < totalNodes; ++i) (trueLabels[i] < 0.5f ? class0 : class1).push_back(i);
  std::shuffle(class0.begin(), class0.end(), rng);
  std::shuffle(class1.begin(), class1.end(), rng);
  const int numGTPerClass = std::max(2, static_cast<int>(totalNodes * groundTruthFraction)) / 2;
  batches.assign(numBatches, BatchData());
  std::vector<char> isInitialGT(totalNodes, 0);
  for (int c = 0; c < 2; ++c) {
    const std::vector<int> &cls = c ? class1 : class0;
    for (int i = 0; i < numGTPerClass && i < (int)cls.size(); ++i) {
      const int v = cls[i];
      batches[0].insertedNodeIds.push_back(v);
      batches[0].groundTruthNodeIds.push_back(v);
      batches[0].groundTruthLabels.push_back(c ? 1.0f : 0.0f);
      isInitialGT[v] = 1;
    }
  }
  std::vector<int> order;
  order.reserve(totalNodes);
  for (int i = 0; i < totalNodes; ++i)
    if (!isInitialGT[i]) order.push_back(i);
  std::shuffle(order.begin(), order.end(), rng);
  size_t next = 0;
  std::vector<int> pool;                  // active non-GT vertices
  std::vector<int> poolPos(totalNodes, -1);
  for (int b = 1; b < numBatches; ++b) {
    BatchData &bt = batches[b];
    bt.batchId = b;
    const int k = (int)std::min<size_t>(std::max(sizes[b], 0), order.size() - next);
    bt.insertedNodeIds.assign(order.begin() + next, order.begin() + next + k);
    next += k;
    if (k > 0) {
      const int numGT = std::max(1, static_cast<int>(k * groundTruthFraction));
      std::vector<int> shuffled = bt.insertedNodeIds;
      std::shuffle(shuffled.begin(), shuffled.end(), rng);
      for (int j = 0; j < numGT && j < k; ++j) {
        const int v = shuffled[j];
        bt.groundTruthNodeIds.push_back(v);
        bt.groundTruthLabels.push_back(trueLabels[v]);
      }
      std::vector<int> gtSorted = bt.groundTruthNodeIds;
      std::sort(gtSorted.begin(), gtSorted.end());
      for (int v : bt.insertedNodeIds)
        if (!std::binary_search(gtSorted.begin(), gtSorted.end(), v)) {
          poolPos[v] = (int)pool.size();
          pool.push_back(v);
        }
    }
    const int numDel = std::min(static_cast<int>(k * deletedFraction), (int)pool.size());
    for (int j = 0; j < numDel; ++j) {
      const int pick = std::uniform_int_distribution<int>(0, (int)pool.size() - 1)(rng);
      const int v = pool[pick];
      bt.deletedNodeIds.push_back(v);
      pool[pick] = pool.back();
      poolPos[pool[pick]] = pick;
      pool.pop_back();
      poolPos[v] = -1;
    }
  }
  std::vector<int> vertexBatch(totalNodes, -1);
  for (int b = 0; b < numBatches; ++b)
    for (int v : batches[b].insertedNodeIds) vertexBatch[v] = b;
  for (int u = 0; u < totalNodes; ++u) {
    if (vertexBatch[u] < 0) continue;
    for (int e = graph.rowPtr[u]; e < graph.rowPtr[u + 1]; ++e) {
      const int v = graph.colInd[e];
      if (v < 0 || vertexBatch[v] < 0 || u >= v) continue;
      const int ab = std::max(vertexBatch[u], vertexBatch[v]);
      batches[ab].newEdgeSrc.push_back(u);
      batches[ab].newEdgeDst.push_back(v);
      batches[ab].newEdgeWeight.push_back(graph.values[e]);
    }
  }
  return true;
}
