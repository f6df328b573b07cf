/*
 * dlp_autotune.hip -- the launch-shape autotuner and the tuning table.
 *
 * The reference launches every propagation iteration with one 256-thread
 * block per frontier vertex (iterativePropagation.cu:292-304), whatever the
 * degrees.  Here the shapes are chosen per graph:
 *   - from the degree histogram (host arithmetic): the lane group of the
 *     multi-block kernel from the 90th-percentile row length, the wave and
 *     work-group tiers for long and hub rows, the grid cap from the vertex
 *     count (dlp_autotune_heuristic);
 *   - by timing (dlp_context_autotune, dlp_multi_autotune): candidate lane
 *     groups and LDS work-group shapes on a graph's own first batches, or
 *     work-group shapes and hand-back frontiers of the batched runner; all
 *     candidates give identical results, so only time decides;
 *   - and remembered: the choice is stored in a process-wide table under
 *     the graph's shape (vertex-count class, mean degree, max-degree class;
 *     graph count and total vertices for the batched runner), which new
 *     contexts and runners consult at creation.  The table is a text file;
 *     the library loads <its own directory>/tuning/<gfx arch>.tune once,
 *     when the first context or runner is created.
 */
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "dlp_context.h"
#include "dlp_tune.h"

namespace {

struct CtxEntry {
  dlp_shape_key key;
  dlp_tuning t;
  float ms;
};
struct MultiEntry {
  int logGraphs, logNodes;
  dlp_multi_tuning t;
  float ms;
};

std::mutex g_mu;
std::vector<CtxEntry> g_ctx;
std::vector<MultiEntry> g_multi;
bool g_autoloaded = false;

int ceil_log2(long long x) {
  int b = 0;
  while (b < 62 && (1LL << b) < x) ++b;
  return b;
}

bool same_key(const dlp_shape_key &a, const dlp_shape_key &b) {
  return a.logNodes == b.logNodes && a.meanDegree == b.meanDegree &&
         a.logMaxDegree == b.logMaxDegree;
}

void put_ctx(const dlp_shape_key &k, const dlp_tuning &t, float ms) {
  for (CtxEntry &e : g_ctx)
    if (same_key(e.key, k)) {
      e.t = t;
      e.ms = ms;
      return;
    }
  g_ctx.push_back({k, t, ms});
}

void put_multi(int lg, int ln, const dlp_multi_tuning &t, float ms) {
  for (MultiEntry &e : g_multi)
    if (e.logGraphs == lg && e.logNodes == ln) {
      e.t = t;
      e.ms = ms;
      return;
    }
  g_multi.push_back({lg, ln, t, ms});
}

/* parses the table text; returns the number of entries or -1 */
int parse_table(FILE *f) {
  char line[512];
  int n = 0;
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    dlp_shape_key k{};
    dlp_tuning t{};
    dlp_multi_tuning mt{};
    float ms = 0.0f;
    int lg = 0, ln = 0;
    if (std::sscanf(line, "context %d %d %d : %d %d %d %d %d %d %d %d %d %d ms=%f",
                    &k.logNodes, &k.meanDegree, &k.logMaxDegree, &t.propGroupLanes,
                    &t.propBlocks, &t.chunkIterations, &t.localThreshold, &t.localLdsBytes,
                    &t.localShape, &t.localOwnerAfter, &t.waveRowMin, &t.blockRowMin,
                    &t.maskOwnerMax, &ms) == 14) {
      put_ctx(k, t, ms);
      ++n;
    } else if (std::sscanf(line, "multi %d %d : %d %d %d %d %d ms=%f", &lg, &ln, &mt.localShape,
                           &mt.thresholdPercent, &mt.backoff, &mt.useGraphs, &mt.heavyFrontier,
                           &ms) == 8) {
      put_multi(lg, ln, mt, ms);
      ++n;
    } else {
      return -1;
    }
  }
  return n;
}

/* the device architecture name without its feature suffixes ("gfx950") */
std::string device_arch() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hipGetLastError();
    return "";
  }
  std::string a = prop.gcnArchName;
  const size_t colon = a.find(':');
  return colon == std::string::npos ? a : a.substr(0, colon);
}

/* directory of this shared library */
std::string library_dir() {
  Dl_info info;
  if (!dladdr((void *)&dlp_tuning_table_load, &info) || !info.dli_fname) return "";
  std::string p = info.dli_fname;
  const size_t slash = p.find_last_of('/');
  return slash == std::string::npos ? "." : p.substr(0, slash);
}

}  // namespace

void dlp_tuning_table_autoload() {
  std::string path;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_autoloaded) return;
    g_autoloaded = true;
    const std::string arch = device_arch(), dir = library_dir();
    if (arch.empty() || dir.empty()) return;
    path = dir + "/tuning/" + arch + ".tune";
  }
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return;
  std::lock_guard<std::mutex> lock(g_mu);
  (void)parse_table(f);  // a malformed committed table is ignored
  std::fclose(f);
}

bool dlp_tuning_lookup_ctx(const dlp_shape_key &key, dlp_tuning *out) {
  std::lock_guard<std::mutex> lock(g_mu);
  for (const CtxEntry &e : g_ctx)
    if (same_key(e.key, key)) {
      *out = e.t;
      return true;
    }
  return false;
}

bool dlp_tuning_lookup_multi(int logGraphs, int logNodes, dlp_multi_tuning *out) {
  std::lock_guard<std::mutex> lock(g_mu);
  for (const MultiEntry &e : g_multi)
    if (e.logGraphs == logGraphs && e.logNodes == logNodes) {
      *out = e.t;
      return true;
    }
  return false;
}

void dlp_multi_shape(int numGraphs, long long totalNodes, int *logGraphs, int *logNodes) {
  *logGraphs = ceil_log2(numGraphs);
  *logNodes = ceil_log2(totalNodes);
}

/* ------------------------------------------------------------------------ */
/* degree-histogram heuristic                                               */
/* ------------------------------------------------------------------------ */
extern "C" int dlp_autotune_heuristic(const dlp_graph_view *g, dlp_shape_key *key,
                                      dlp_tuning *out) {
  if (!g || !key || !out || g->numNodes <= 0 || !g->rowPtr) {
    dlp_set_error_("dlp_autotune_heuristic: bad argument");
    return DLP_ERR_INVALID;
  }
  const int N = g->numNodes;
  /* histogram of row lengths in power-of-two classes: class c holds lengths
   * in [2^(c-1), 2^c), class 0 the empty rows */
  long long hist[33] = {0};
  int maxDeg = 0;
  for (int u = 0; u < N; ++u) {
    const int d = g->rowPtr[u + 1] - g->rowPtr[u];
    maxDeg = std::max(maxDeg, d);
    int c = 0;
    while (c < 32 && (1LL << c) <= d) ++c;
    ++hist[c];
  }
  /* 90th percentile: the upper end of the class that holds it */
  long long seen = 0, p90 = 0;
  for (int c = 0; c <= 32; ++c) {
    seen += hist[c];
    if (10 * seen >= 9LL * N) {
      p90 = c == 0 ? 0 : (1LL << c) - 1;
      break;
    }
  }
  key->logNodes = ceil_log2(N);
  key->meanDegree = (int)llround((double)g->numEdges / (double)N);
  key->logMaxDegree = ceil_log2((long long)maxDeg + 1);
  dlp_tuning t{};
  /* a lane group covers a typical row in two passes: fewer lanes per vertex
   * keep more vertices per wave in flight (the kernel is gather-bound) */
  long long lanes = 4;
  while (lanes < 32 && 2 * lanes < p90) lanes *= 2;
  t.propGroupLanes = (int)lanes;
  t.propBlocks = dlp_auto_prop_blocks(N);
  t.chunkIterations = 1024;
  t.localShape = 1;
  t.localOwnerAfter = kOwnerAfterDefault;
  t.waveRowMin = kDefaultWaveRowMin;
  /* the work-group tier only exists for graphs with hub rows */
  t.blockRowMin = maxDeg > kDefaultBlockRowMin ? kDefaultBlockRowMin : -1;
  t.maskOwnerMax = kMaskOwnerMax;
  *out = t;
  return DLP_OK;
}

/* ------------------------------------------------------------------------ */
/* timed trials                                                             */
/* ------------------------------------------------------------------------ */
namespace {

/* stream time of batches [0, nb) from a reset, ms */
int time_context_batches(dlp_context *c, int nb, hipEvent_t e0, hipEvent_t e1, float *ms) {
  int rc = dlp_context_reset(c);
  if (rc != DLP_OK) return rc;
  DLP_HIP(hipEventRecord(e0, c->stream));
  for (int b = 0; b < nb; ++b)
    if ((rc = dlp_context_run_batch(c, b)) != DLP_OK) return rc;
  DLP_HIP(hipEventRecord(e1, c->stream));
  DLP_HIP(hipEventSynchronize(e1));
  DLP_HIP(hipEventElapsedTime(ms, e0, e1));
  return DLP_OK;
}

}  // namespace

extern "C" int dlp_context_autotune(dlp_context *c, int trialBatches, dlp_tuning *chosen,
                                    float *bestMs) {
  if (!c) {
    dlp_set_error_("dlp_context_autotune: NULL context");
    return DLP_ERR_STATE;
  }
  if (c->trace) {
    dlp_set_error_("dlp_context_autotune: disable the trace first");
    return DLP_ERR_STATE;
  }
  const int nb = trialBatches <= 0 ? c->B : std::min(trialBatches, c->B);
  dlp_tuning base = c->tuning;
  std::vector<dlp_tuning> cand;
  for (int lanes : {base.propGroupLanes / 2, base.propGroupLanes, base.propGroupLanes * 2}) {
    if (lanes < 4 || lanes > 64) continue;
    for (int shape : {1, 3}) {
      dlp_tuning t = base;
      t.propGroupLanes = lanes;
      t.localShape = shape;
      cand.push_back(t);
    }
  }
  hipEvent_t e0, e1;
  DLP_HIP(hipEventCreate(&e0));
  DLP_HIP(hipEventCreate(&e1));
  int rc = DLP_OK, best = -1;
  float bestT = 0.0f;
  const int profiling = c->profiling;
  c->profiling = 0;
  for (size_t i = 0; i < cand.size() && rc == DLP_OK; ++i) {
    if ((rc = dlp_context_set_tuning(c, &cand[i])) != DLP_OK) break;
    float ms = 0.0f;
    if (i == 0 && (rc = time_context_batches(c, nb, e0, e1, &ms)) != DLP_OK) break;  // warm-up
    if ((rc = time_context_batches(c, nb, e0, e1, &ms)) != DLP_OK) break;
    if (best < 0 || ms < bestT) {
      best = (int)i;
      bestT = ms;
    }
  }
  c->profiling = profiling;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != DLP_OK) return rc;
  if ((rc = dlp_context_set_tuning(c, &cand[best])) != DLP_OK) return rc;
  if ((rc = dlp_context_reset(c)) != DLP_OK) return rc;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    put_ctx(c->shape, c->tuning, bestT);
  }
  c->tuningSource = 2;
  if (chosen) *chosen = c->tuning;
  if (bestMs) *bestMs = bestT;
  return DLP_OK;
}

extern "C" int dlp_context_shape_key(const dlp_context *c, dlp_shape_key *key, int *source) {
  if (!c || !key) {
    dlp_set_error_("dlp_context_shape_key: NULL argument");
    return DLP_ERR_INVALID;
  }
  *key = c->shape;
  if (source) *source = c->tuningSource;
  return DLP_OK;
}

/* ------------------------------------------------------------------------ */
/* the table                                                                */
/* ------------------------------------------------------------------------ */
extern "C" int dlp_tuning_table_load(const char *path, int *entries) {
  if (!path) {
    dlp_set_error_("dlp_tuning_table_load: NULL path");
    return DLP_ERR_INVALID;
  }
  FILE *f = std::fopen(path, "r");
  if (!f) {
    dlp_set_errorf_("dlp_tuning_table_load: cannot open %s", path);
    return DLP_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  const int n = parse_table(f);
  std::fclose(f);
  if (n < 0) {
    dlp_set_errorf_("dlp_tuning_table_load: %s is not a tuning table", path);
    return DLP_ERR_INVALID;
  }
  if (entries) *entries = n;
  return DLP_OK;
}

extern "C" int dlp_tuning_table_save(const char *path, int *entries) {
  if (!path) {
    dlp_set_error_("dlp_tuning_table_save: NULL path");
    return DLP_ERR_INVALID;
  }
  FILE *f = std::fopen(path, "w");
  if (!f) {
    dlp_set_errorf_("dlp_tuning_table_save: cannot open %s", path);
    return DLP_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  std::fprintf(f,
               "# libdynlp tuning table: timed launch shapes per graph shape\n"
               "# context <log2 N> <mean degree> <log2 (max degree + 1)> : <propGroupLanes> "
               "<propBlocks> <chunkIterations> <localThreshold> <localLdsBytes> <localShape> "
               "<localOwnerAfter> <waveRowMin> <blockRowMin> <maskOwnerMax> ms=<trial ms>\n"
               "# multi <log2 graphs> <log2 total vertices> : <localShape> <thresholdPercent> "
               "<backoff> <useGraphs> <heavyFrontier> ms=<trial ms>\n");
  for (const CtxEntry &e : g_ctx)
    std::fprintf(f, "context %d %d %d : %d %d %d %d %d %d %d %d %d %d ms=%.3f\n",
                 e.key.logNodes, e.key.meanDegree, e.key.logMaxDegree, e.t.propGroupLanes,
                 e.t.propBlocks, e.t.chunkIterations, e.t.localThreshold, e.t.localLdsBytes,
                 e.t.localShape, e.t.localOwnerAfter, e.t.waveRowMin, e.t.blockRowMin,
                 e.t.maskOwnerMax, e.ms);
  for (const MultiEntry &e : g_multi)
    std::fprintf(f, "multi %d %d : %d %d %d %d %d ms=%.3f\n", e.logGraphs, e.logNodes,
                 e.t.localShape, e.t.thresholdPercent, e.t.backoff, e.t.useGraphs,
                 e.t.heavyFrontier, e.ms);
  if (entries) *entries = (int)(g_ctx.size() + g_multi.size());
  std::fclose(f);
  return DLP_OK;
}

extern "C" int dlp_tuning_table_clear(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_ctx.clear();
  g_multi.clear();
  g_autoloaded = true;  // a cleared table stays empty
  return DLP_OK;
}

/* stores a batched-runner choice (dlp_multi.hip times the candidates) */
void dlp_tuning_put_multi(int logGraphs, int logNodes, const dlp_multi_tuning &t, float ms) {
  std::lock_guard<std::mutex> lock(g_mu);
  put_multi(logGraphs, logNodes, t, ms);
}
