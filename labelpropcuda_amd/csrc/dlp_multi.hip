/*
 * dlp_multi.hip -- many independent evolving graphs updated together.
 *
 * One context per graph (its own device state, all on one stream).  A batch
 * index b runs, for every graph that has a batch b, the pipeline up to the
 * first frontier, then ONE propagation launch for all of them (one
 * work-group per graph, k_propagate_multi), then the per-graph batch end.
 * The pipelines and the propagation launch need no host decision, so they
 * are captured once into a HIP graph per batch index and replayed on later
 * steps, and so are the batch ends; the state reset of all graphs is a graph
 * of its own.
 *
 * One work-group is slow on a large frontier, so a graph whose frontier
 * grows beyond `heavy` vertices is handed back: its work-group stops at the
 * start of that iteration, and after the launch the host continues that
 * graph's propagation with the single-graph loop (multi-block launches over
 * the whole GPU, the LDS-resident kernel once the frontier is small again)
 * before the batch ends run.  The result is the same either way.
 *
 * The reference updates one graph at a time with a host round trip per
 * propagation iteration (dynLP.cu:175-587, iterativePropagation.cu:274-327).
 */
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dlp_context.h"
#include "dlp_pool.h"
#include "dlp_tune.h"

struct dlp_multi {
  int G = 0, maxB = 0;
  hipStream_t stream = nullptr;
  std::vector<dlp_context *> ctx;
  int ldsWords = 0;
  int shape = 2;        /* LocalArgs::shape of the propagation launch */
  int backoff = 8;
  int thresholdPct = 100;
  int heavy = 0;        /* hand-back frontier size; -1 = never, 0 = the
                           default of the mode (lockstep batches: never;
                           serving: kServeHandBack)                    */
  int lanes = 8;        /* handed-back graphs finished side by side      */
  bool useGraphs = true;
  std::vector<GraphProp *> dProps; /* per batch: graphs that have it */
  std::vector<int> counts;
  std::vector<hipGraphExec_t> exec;     /* per batch: pipelines + launch   */
  std::vector<hipGraphExec_t> execPost; /* per batch: batch ends          */
  int *dExit = nullptr;                 /* per graph: hand-back iteration */
  int *hExit = nullptr;                 /* pinned copy                    */
  hipGraphExec_t resetExec = nullptr;
  std::vector<hipEvent_t> evStart, evEnd;
  std::vector<char> ran;
  int nextBatch = 0;
  int tuningSource = 0; /* 0 defaults, 1 tuning table, 2 dlp_multi_autotune */
  int logGraphs = 0, logNodes = 0;
  /* run_next: descriptors of the graphs of one step */
  std::vector<GraphProp> props;     /* per graph, batch-independent      */
  GraphProp *dStep = nullptr;       /* device, G entries                 */
  GraphProp *hStep = nullptr;       /* pinned staging, G entries         */
  /* scratch pool: one device allocation that lends each running batch its
   * per-batch scratch (poolBase null: every graph has its own) */
  void *poolBase = nullptr;
  RangePool pool;
  int64_t deferrals = 0;      /* serving admissions deferred for pool space */
  std::vector<int64_t> lease; /* per graph: offset of its lease, -1 = none  */
  int wallRateKHz = 0;        /* device wall clock (slice deadlines)        */

  void drop_graphs() {
    for (std::vector<hipGraphExec_t> *v : {&exec, &execPost})
      for (hipGraphExec_t &e : *v)
        if (e) {
          (void)hipGraphExecDestroy(e);
          e = nullptr;
        }
    if (resetExec) (void)hipGraphExecDestroy(resetExec);
    resetExec = nullptr;
  }
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    for (GraphProp *p : dProps)
      if (p) (void)hipFree(p);
    dProps.clear();
    if (dStep) (void)hipFree(dStep);
    dStep = nullptr;
    if (hStep) (void)hipHostFree(hStep);
    hStep = nullptr;
    if (dExit) (void)hipFree(dExit);
    dExit = nullptr;
    if (hExit) (void)hipHostFree(hExit);
    hExit = nullptr;
    for (dlp_context *c : ctx) dlp_context_destroy(c);
    ctx.clear();
    if (poolBase) (void)hipFree(poolBase);
    poolBase = nullptr;
    for (hipEvent_t e : evStart) (void)hipEventDestroy(e);
    for (hipEvent_t e : evEnd) (void)hipEventDestroy(e);
    evStart.clear();
    evEnd.clear();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};

namespace {

/* hand-back frontier of serving steps when the tuning leaves it at 0: a
 * graph whose frontier outgrows this leaves the step for a hand-back lane.
 * Saturated 64-tenant serving (52.9 M vertices, 128 batches), median of
 * three runs: 326 M updates/s at 2,048, 344 M at 4,096, 376 M at 8,192,
 * 212 M at 16,384 (a single work-group per graph is slow on frontiers of
 * ten thousand vertices). */
constexpr int kServeHandBack = 8192;

/* effective hand-back frontier (0 = none) of lockstep batches / serving */
int heavy_lockstep(const dlp_multi *m) { return m->heavy > 0 ? m->heavy : 0; }
int heavy_serving(const dlp_multi *m) {
  return m->heavy > 0 ? m->heavy : (m->heavy == 0 ? kServeHandBack : 0);
}

/* per-graph descriptors of batch b's propagation launch */
GraphProp graph_prop(const dlp_multi *m, int g);

int upload_props(dlp_multi *m) {
  m->props.clear();
  for (int g = 0; g < m->G; ++g) m->props.push_back(graph_prop(m, g));
  for (int b = 0; b < m->maxB; ++b) {
    std::vector<GraphProp> h;
    for (int gi = 0; gi < m->G; ++gi) {
      if (b >= m->ctx[gi]->B) continue;
      h.push_back(m->props[gi]);
    }
    m->counts[b] = (int)h.size();
    if (!m->dProps[b]) {
      void *p = nullptr;
      DLP_HIP(hipMalloc(&p, sizeof(GraphProp) * std::max<size_t>(h.size(), 1)));
      m->dProps[b] = (GraphProp *)p;
    }
    if (!h.empty())
      DLP_HIP(hipMemcpy(m->dProps[b], h.data(), sizeof(GraphProp) * h.size(),
                        hipMemcpyHostToDevice));
  }
  return DLP_OK;
}

/* the propagation descriptor of one graph (the same for every batch) */
GraphProp graph_prop(const dlp_multi *m, int gi) {
  const dlp_context *c = m->ctx[gi];
  GraphProp g{};
  g.a = dlp_ctx_prop_args(c);
  g.l = c->local;
  g.l.ldsWords = m->ldsWords;
  g.maxIter = c->cfg.maxIterations;
  const long long th = (long long)c->tuning.localThreshold * m->thresholdPct / 100;
  g.threshold =
      m->thresholdPct < 0 ? 0 : (int)std::max<long long>(1, std::min<long long>(th, 1 << 30));
  if (c->N >= (1 << 30)) g.threshold = 0;  // outside-vertex encoding limit
  g.backoff = m->backoff;
  g.heavy = heavy_lockstep(m);
  g.startIter = 0;
  g.stopIter = g.maxIter;
  g.sliceTicks = 0;
  g.exitIter = m->dExit + gi;
  return g;
}

/* batch b of every graph that has one, up to and including the propagation
 * launch, on the shared stream */
int enqueue_batch_pre(dlp_multi *m, int b) {
  int rc;
  for (int g = 0; g < m->G; ++g) {
    dlp_context *c = m->ctx[g];
    if (b >= c->B) continue;
    const DBatch bt = dlp_ctx_make_batch(c, b);
    int tag = 0;
    if ((rc = dlp_ctx_batch_pre(c, b, bt, &tag)) != DLP_OK) return rc;
  }
  return dlp_launch_propagate_multi(m->dProps[b], m->counts[b], m->shape, m->ldsWords,
                                    m->stream);
}

/* the batch ends of batch b */
int enqueue_batch_post(dlp_multi *m, int b) {
  int rc;
  for (int g = 0; g < m->G; ++g) {
    dlp_context *c = m->ctx[g];
    if (b >= c->B) continue;
    if ((rc = dlp_ctx_batch_post(c, b, dlp_ctx_make_batch(c, b))) != DLP_OK) return rc;
  }
  return DLP_OK;
}

/* after a propagation launch over `graphs` (count n): the graphs it handed
 * back, read once the launch has finished (none when hand-back is off, in
 * which case nothing waits) */
int handed_back(dlp_multi *m, int heavy, const int *graphs, int n, std::vector<int> *handed) {
  handed->clear();
  if (heavy <= 0 || n == 0) return DLP_OK;
  DLP_HIP(hipMemcpyAsync(m->hExit, m->dExit, sizeof(int) * m->G, hipMemcpyDeviceToHost,
                         m->stream));
  DLP_HIP(hipStreamSynchronize(m->stream));
  for (int i = 0; i < n; ++i) {
    const int g = graphs[i];
    const int k = m->hExit[g];
    if (k < 0) continue;
    if (k > m->ctx[g]->cfg.maxIterations) {
      dlp_set_errorf_("graph %d handed back at iteration %d", g, k);
      return DLP_ERR_INTERNAL;
    }
    handed->push_back(g);
  }
  return DLP_OK;
}

/* the rest of a handed-back graph's propagation, on the shared stream */
int continue_graph(dlp_multi *m, int g) { return dlp_ctx_run_propagation(m->ctx[g], m->hExit[g]); }

/* Finishes handed-back graphs beside the serving loop's later steps: each
 * of `workers` host threads takes handed-back graphs from a queue and runs
 * them on a stream of its own (several run at once; the GPU overlaps their
 * launches with each other and with the steps).  A handed-back graph is
 * busy (not admitted again) until its propagation and batch end are done;
 * its context uses the worker's stream meanwhile, and nothing else touches
 * that context until it is released. */
class HandBackLane {
 public:
  /* finishBatch: a task also runs the graph's batch end (serving); without
   * it the caller runs the batch ends after drain() (lockstep batches) */
  HandBackLane(dlp_multi *m, int workers, bool finishBatch,
               std::function<void(int graph)> finished)
      : m_(m), workers_(workers), finishBatch_(finishBatch), finished_(std::move(finished)),
        busy_(m->G, 0) {}
  ~HandBackLane() { stop(); }

  int start() {
    if (hipGetDevice(&device_) != hipSuccess) {
      (void)hipGetLastError();
      dlp_set_error_("hand-back lane: hipGetDevice failed");
      return DLP_ERR_HIP;
    }
    for (int i = 0; i < workers_; ++i) {
      hipStream_t s = nullptr;
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        dlp_set_error_("hand-back lane: hipStreamCreate failed");
        return DLP_ERR_HIP;
      }
      streams_.push_back(s);
    }
    for (int i = 0; i < workers_; ++i) threads_.emplace_back([this, i] { loop(streams_[i]); });
    return DLP_OK;
  }
  /* graph g resumes at iteration k on the lane */
  int push(int g, int k) {
    std::lock_guard<std::mutex> lock(mu_);
    busy_[g] = 1;
    ++inFlight_;
    queue_.push_back({g, k});
    cv_.notify_all();
    return DLP_OK;
  }
  bool busy(int g) {
    std::lock_guard<std::mutex> lock(mu_);
    return busy_[g] != 0;
  }
  int inFlight() {
    std::lock_guard<std::mutex> lock(mu_);
    return inFlight_;
  }
  int workers() const { return workers_; }
  int error() {
    std::lock_guard<std::mutex> lock(mu_);
    return rc_;
  }
  std::string message() {
    std::lock_guard<std::mutex> lock(mu_);
    return msg_;
  }
  /* waits until `deadline` (seconds on `now`) or until a graph is released */
  template <class Now>
  void wait(double deadline, Now now) {
    std::unique_lock<std::mutex> lock(mu_);
    const int seen = released_;
    const double left = deadline - now();
    if (left <= 0) return;
    cv_.wait_for(lock, std::chrono::duration<double>(left),
                 [&] { return released_ != seen || rc_ != DLP_OK; });
  }
  void drain() {
    std::unique_lock<std::mutex> lock(mu_);
    cv_.wait(lock, [&] { return inFlight_ == 0 || rc_ != DLP_OK; });
  }
  void stop() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      stop_ = true;
      cv_.notify_all();
    }
    for (std::thread &t : threads_)
      if (t.joinable()) t.join();
    threads_.clear();
    for (hipStream_t s : streams_) (void)hipStreamDestroy(s);
    streams_.clear();
  }

 private:
  struct Task {
    int g, k;
  };
  void loop(hipStream_t stream) {
    if (hipSetDevice(device_) != hipSuccess) {  // the device is per host thread
      std::lock_guard<std::mutex> lock(mu_);
      rc_ = DLP_ERR_HIP;
      msg_ = "hipSetDevice failed";
      cv_.notify_all();
      return;
    }
    while (true) {
      Task t;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] { return stop_ || !queue_.empty(); });
        if (queue_.empty()) return;  // stop requested and nothing left
        t = queue_.front();
        queue_.pop_front();
      }
      dlp_context *c = m_->ctx[t.g];
      hipStream_t own = c->stream;
      c->stream = stream;
      int rc = dlp_ctx_run_propagation(c, t.k);
      if (rc == DLP_OK && finishBatch_)
        rc = dlp_ctx_batch_post(c, c->nextBatch, dlp_ctx_make_batch(c, c->nextBatch));
      if (rc == DLP_OK && hipStreamSynchronize(stream) != hipSuccess) {
        (void)hipGetLastError();
        dlp_set_error_("hand-back lane: stream synchronisation failed");
        rc = DLP_ERR_HIP;
      }
      c->stream = own;
      if (rc == DLP_OK) finished_(t.g);
      const std::string msg = rc != DLP_OK ? dlp_last_error() : "";
      std::lock_guard<std::mutex> lock(mu_);
      if (rc != DLP_OK && rc_ == DLP_OK) {
        rc_ = rc;
        msg_ = msg;
      }
      busy_[t.g] = 0;
      --inFlight_;
      ++released_;
      cv_.notify_all();
    }
  }
  dlp_multi *m_;
  int workers_;
  bool finishBatch_;
  std::function<void(int)> finished_;
  std::vector<char> busy_;
  std::deque<Task> queue_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::vector<std::thread> threads_;
  std::vector<hipStream_t> streams_;
  int device_ = 0;
  std::string msg_;
  int inFlight_ = 0, released_ = 0, rc_ = DLP_OK;
  bool stop_ = false;
};


int enqueue_reset(dlp_multi *m) {
  for (dlp_context *c : m->ctx) {
    const int rc = dlp_reset_state(c->d, m->stream);
    if (rc != DLP_OK) return rc;
  }
  return DLP_OK;
}

/* capture what `fn` enqueues into an executable graph */
template <class Fn>
int capture(dlp_multi *m, hipGraphExec_t *out, Fn fn) {
  DLP_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  const int rc = fn();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(m->stream, &graph);
  if (rc != DLP_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return dlp_hip_fail_(e, "hipStreamEndCapture");
  const hipError_t ie = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) return dlp_hip_fail_(ie, "hipGraphInstantiate");
  return DLP_OK;
}

bool pooled(const dlp_multi *m) { return m->poolBase != nullptr; }

/* Lends graph g the scratch of its next batch from the pool (nothing to do
 * without a pool or with a lease held).  *granted = false, and nothing
 * changed, when no free range fits now; DLP_ERR_NOMEM when the batch needs
 * more than the whole pool. */
int lease_scratch(dlp_multi *m, int g, bool *granted) {
  *granted = true;
  if (!pooled(m) || m->lease[g] >= 0) return DLP_OK;
  dlp_context *c = m->ctx[g];
  const int b = c->nextBatch;
  const int64_t need = dlp_ctx_batch_scratch_bytes(c, b);
  if (RangePool::rounded(need) > m->pool.capacity()) {
    dlp_set_errorf_("graph %d: batch %d needs %lld bytes of scratch, the pool holds %lld", g, b,
                    (long long)need, (long long)m->pool.capacity());
    return DLP_ERR_NOMEM;
  }
  int64_t off = 0;
  if (!m->pool.acquire(need, &off)) {
    *granted = false;
    return DLP_OK;
  }
  m->lease[g] = off;
  dlp_ctx_scratch_layout(c, c->insOff[b + 1] - c->insOff[b], c->eOff[b + 1] - c->eOff[b],
                         (char *)m->poolBase + off, true);
  return DLP_OK;
}

/* returns graph g's lease (its batch's work on the device is complete) */
void return_scratch(dlp_multi *m, int g) {
  if (!pooled(m) || m->lease[g] < 0) return;
  m->pool.release(m->lease[g]);
  m->lease[g] = -1;
  dlp_ctx_scratch_layout(m->ctx[g], 0, 0, nullptr, true);
}

/* Leases for every graph of a step, or none: DLP_ERR_NOMEM when the pool
 * cannot hold them all at once.  The leases go back when the guard ends
 * (after the stream is idle). */
class StepLeases {
 public:
  explicit StepLeases(dlp_multi *m) : m_(m) {}
  ~StepLeases() {
    if (graphs_.empty()) return;
    (void)hipStreamSynchronize(m_->stream);
    for (int g : graphs_) return_scratch(m_, g);
  }
  int take(const int *graphs, int n) {
    if (!pooled(m_)) return DLP_OK;
    for (int i = 0; i < n; ++i) {
      bool ok = false;
      const int rc = lease_scratch(m_, graphs[i], &ok);
      if (rc != DLP_OK) return rc;
      if (!ok) {
        dlp_set_errorf_("the scratch pool (%lld bytes, largest free range %lld) cannot hold the "
                        "batches of these %d graphs at once",
                        (long long)m_->pool.capacity(), (long long)m_->pool.largestFree(), n);
        return DLP_ERR_NOMEM;
      }
      graphs_.push_back(graphs[i]);
    }
    return DLP_OK;
  }

 private:
  dlp_multi *m_;
  std::vector<int> graphs_;
};

int run_next_impl(dlp_multi *m, const int *graphs, int count, int heavy,
                  const std::function<void(const std::vector<int> &)> *done,
                  const std::function<int(int, int)> *handOff = nullptr, bool lockstep = false);

void host_reset(dlp_multi *m) {
  for (dlp_context *c : m->ctx) {
    c->nextBatch = 0;
    c->stampBound = 1;  // the reset kernels set stampBase to 1
    c->lastCCBatch = -1;
    std::fill(c->ran.begin(), c->ran.end(), 0);
  }
  m->nextBatch = 0;
  std::fill(m->ran.begin(), m->ran.end(), 0);
}

}  // namespace

extern "C" int dlp_multi_create(dlp_multi **out, int numGraphs, const dlp_config *cfgs,
                                const dlp_graph_view *graphs, const float *const *trueLabels,
                                const dlp_batch_view *batches) {
  if (!out || numGraphs <= 0 || !cfgs || !graphs || !trueLabels || !batches) {
    dlp_set_error_("dlp_multi_create: bad argument");
    return DLP_ERR_INVALID;
  }
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    dlp_set_error_("no GPU visible to HIP");
    return DLP_ERR_NODEVICE;
  }
  dlp_multi *m = new dlp_multi();
  m->G = numGraphs;
  int rc = DLP_OK;
  int ldsBytes = 0;
#define MTRY(x)           \
  do {                    \
    rc = (x);             \
    if (rc != DLP_OK) {   \
      m->release();       \
      delete m;           \
      return rc;          \
    }                     \
  } while (0)
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    dlp_set_error_("hipStreamCreate failed");
    return DLP_ERR_HIP;
  }
  MTRY(dlp_local_lds_bytes(&ldsBytes));
  m->ldsWords = ldsBytes / 4;
  for (int g = 0; g < numGraphs; ++g) {
    dlp_context *c = nullptr;
    rc = dlp_context_create(&c, &cfgs[g], &graphs[g], trueLabels[g], &batches[g], m->stream);
    if (rc != DLP_OK) {
      const std::string msg = dlp_last_error();
      m->release();
      delete m;
      dlp_set_errorf_("graph %d: %s", g, msg.c_str());
      return rc;
    }
    m->ctx.push_back(c);
    m->maxB = std::max(m->maxB, c->B);
  }
  m->dProps.assign(m->maxB, nullptr);
  m->counts.assign(m->maxB, 0);
  m->exec.assign(m->maxB, nullptr);
  m->execPost.assign(m->maxB, nullptr);
  m->ran.assign(m->maxB, 0);
  m->evStart.resize(m->maxB);
  m->evEnd.resize(m->maxB);
  for (int b = 0; b < m->maxB; ++b) {
    if (hipEventCreate(&m->evStart[b]) != hipSuccess ||
        hipEventCreate(&m->evEnd[b]) != hipSuccess) {
      m->release();
      delete m;
      dlp_set_error_("hipEventCreate failed");
      return DLP_ERR_HIP;
    }
  }
  {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(int) * m->G) != hipSuccess ||
        hipHostMalloc((void **)&m->hExit, sizeof(int) * m->G, hipHostMallocDefault) !=
            hipSuccess) {
      if (p) (void)hipFree(p);
      m->release();
      delete m;
      dlp_set_error_("dlp_multi_create: allocation of the hand-back slots failed");
      return DLP_ERR_NOMEM;
    }
    m->dExit = (int *)p;
  }
  MTRY(upload_props(m));
  m->lease.assign(m->G, -1);
  {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&m->wallRateKHz, hipDeviceAttributeWallClockRate, dev) !=
            hipSuccess) {
      (void)hipGetLastError();
      m->wallRateKHz = 0;  // no time slices
    }
  }
  {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(GraphProp) * m->G) != hipSuccess ||
        hipHostMalloc((void **)&m->hStep, sizeof(GraphProp) * m->G, hipHostMallocDefault) !=
            hipSuccess) {
      if (p) (void)hipFree(p);
      m->release();
      delete m;
      dlp_set_error_("dlp_multi_create: allocation of the step descriptors failed");
      return DLP_ERR_NOMEM;
    }
    m->dStep = (GraphProp *)p;
  }
#undef MTRY
  host_reset(m);
  {
    long long total = 0;
    for (dlp_context *c : m->ctx) total += c->N;
    dlp_multi_shape(m->G, total, &m->logGraphs, &m->logNodes);
    dlp_tuning_table_autoload();
    dlp_multi_tuning t{};
    if (dlp_tuning_lookup_multi(m->logGraphs, m->logNodes, &t) &&
        dlp_multi_set_tuning(m, &t) == DLP_OK)
      m->tuningSource = 1;
  }
  *out = m;
  return DLP_OK;
}

extern "C" int dlp_multi_destroy(dlp_multi *m) {
  if (!m) return DLP_OK;
  m->release();
  delete m;
  return DLP_OK;
}

extern "C" int dlp_multi_set_tuning(dlp_multi *m, const dlp_multi_tuning *t) {
  if (!m || !t) {
    dlp_set_error_("dlp_multi_set_tuning: NULL argument");
    return DLP_ERR_INVALID;
  }
  if (t->localShape && (t->localShape < 1 || t->localShape > dlp_local_shape_count())) {
    dlp_set_errorf_("localShape must be in [1, %d]", dlp_local_shape_count());
    return DLP_ERR_INVALID;
  }
  if (t->thresholdPercent < -1 || t->backoff < 0 || t->useGraphs < 0 || t->useGraphs > 2 ||
      t->heavyFrontier < -1) {
    dlp_set_error_("dlp_multi_set_tuning: thresholdPercent < -1, negative backoff, "
                   "useGraphs not in {0,1,2} or heavyFrontier < -1");
    return DLP_ERR_INVALID;
  }
  DLP_HIP(hipStreamSynchronize(m->stream));
  if (t->localShape) m->shape = t->localShape - 1;
  if (t->thresholdPercent) m->thresholdPct = t->thresholdPercent;
  if (t->backoff) m->backoff = t->backoff;
  if (t->useGraphs) m->useGraphs = t->useGraphs == 1;
  if (t->heavyFrontier) m->heavy = t->heavyFrontier;
  m->drop_graphs();  // captured launches embed the shape
  m->exec.assign(m->maxB, nullptr);
  m->execPost.assign(m->maxB, nullptr);
  const int rc = upload_props(m);
  if (rc != DLP_OK) return rc;
  host_reset(m);
  for (dlp_context *c : m->ctx) dlp_reset_state(c->d, m->stream);
  return DLP_OK;
}

extern "C" int dlp_multi_get_tuning(const dlp_multi *m, dlp_multi_tuning *t) {
  if (!m || !t) {
    dlp_set_error_("dlp_multi_get_tuning: NULL argument");
    return DLP_ERR_INVALID;
  }
  t->localShape = m->shape + 1;
  t->thresholdPercent = m->thresholdPct;
  t->backoff = m->backoff;
  t->useGraphs = m->useGraphs ? 1 : 2;
  t->heavyFrontier = m->heavy;
  return DLP_OK;
}

extern "C" int dlp_multi_reset(dlp_multi *m) {
  if (!m) {
    dlp_set_error_("dlp_multi_reset: NULL handle");
    return DLP_ERR_STATE;
  }
  int rc;
  if (m->useGraphs) {
    if (!m->resetExec &&
        (rc = capture(m, &m->resetExec, [&] { return enqueue_reset(m); })) != DLP_OK)
      return rc;
    DLP_HIP(hipGraphLaunch(m->resetExec, m->stream));
  } else if ((rc = enqueue_reset(m)) != DLP_OK) {
    return rc;
  }
  host_reset(m);
  return DLP_OK;
}

extern "C" int dlp_multi_run_batch(dlp_multi *m, int b) {
  if (!m) {
    dlp_set_error_("dlp_multi_run_batch: NULL handle");
    return DLP_ERR_STATE;
  }
  if (m->nextBatch < 0) {
    dlp_set_error_("dlp_multi_run_batch: graphs were advanced one by one (run_next / serve); "
                   "reset to run lockstep batches");
    return DLP_ERR_STATE;
  }
  if (b != m->nextBatch || b < 0 || b >= m->maxB) {
    dlp_set_errorf_("batch %d requested, next batch is %d (reset to start over)", b,
                    m->nextBatch);
    return DLP_ERR_STATE;
  }
  for (int g = 0; g < m->G; ++g)
    if (b < m->ctx[g]->B && m->ctx[g]->nextBatch != b) {
      dlp_set_errorf_("dlp_multi_run_batch: graph %d is at batch %d, not %d", g,
                      m->ctx[g]->nextBatch, b);
      return DLP_ERR_STATE;
    }
  int rc;
  if (pooled(m)) {  // scratch lent per batch: no captured graphs (their pointers would move)
    std::vector<int> members;
    for (int g = 0; g < m->G; ++g)
      if (b < m->ctx[g]->B) members.push_back(g);
    StepLeases leases(m);
    if ((rc = leases.take(members.data(), (int)members.size())) != DLP_OK) return rc;
    DLP_HIP(hipEventRecord(m->evStart[b], m->stream));
    if ((rc = run_next_impl(m, members.data(), (int)members.size(), heavy_lockstep(m), nullptr,
                            nullptr, true)) != DLP_OK)
      return rc;
    DLP_HIP(hipEventRecord(m->evEnd[b], m->stream));
    m->ran[b] = 1;
    m->nextBatch = b + 1;
    return DLP_OK;
  }
  /* wrap-around guards: outside the captured graph, decided per run */
  for (int g = 0; g < m->G; ++g)
    if (b < m->ctx[g]->B && (rc = dlp_ctx_counter_guard(m->ctx[g], m->stream)) != DLP_OK)
      return rc;
  bool justCaptured = false;
  if (m->useGraphs && !m->exec[b]) {
    /* the pipelines read host bookkeeping (mark tags, last CC batch) that
     * evolves the same way on every run from a reset, so the capture of the
     * first run is valid for every later one */
    if ((rc = capture(m, &m->exec[b], [&] { return enqueue_batch_pre(m, b); })) != DLP_OK)
      return rc;
    if ((rc = capture(m, &m->execPost[b], [&] { return enqueue_batch_post(m, b); })) != DLP_OK)
      return rc;
    justCaptured = true;  // the captures ran the host bookkeeping already
  }
  DLP_HIP(hipEventRecord(m->evStart[b], m->stream));
  if (m->useGraphs) {
    DLP_HIP(hipGraphLaunch(m->exec[b], m->stream));
  } else if ((rc = enqueue_batch_pre(m, b)) != DLP_OK) {
    return rc;
  }
  std::vector<int> members, handed;
  for (int g = 0; g < m->G; ++g)
    if (b < m->ctx[g]->B) members.push_back(g);
  if ((rc = handed_back(m, heavy_lockstep(m), members.data(), (int)members.size(), &handed)) !=
      DLP_OK)
    return rc;
  if (handed.size() == 1) {
    if ((rc = continue_graph(m, handed[0])) != DLP_OK) return rc;
  } else if (!handed.empty()) {  // several: side by side, each on a stream of its own
    HandBackLane lane(m, std::min<int>((int)handed.size(), m->lanes), false, [](int) {});
    if ((rc = lane.start()) != DLP_OK) return rc;
    for (int g : handed) lane.push(g, m->hExit[g]);
    lane.drain();
    if ((rc = lane.error()) != DLP_OK) {
      dlp_set_errorf_("graph continuation: %s", lane.message().c_str());
      return rc;
    }
  }
  if (m->useGraphs) {
    DLP_HIP(hipGraphLaunch(m->execPost[b], m->stream));
    if (!justCaptured)
      for (dlp_context *c : m->ctx)
        if (b < c->B) dlp_ctx_advance_host(c, b, dlp_ctx_make_batch(c, b));
  } else if ((rc = enqueue_batch_post(m, b)) != DLP_OK) {
    return rc;
  }
  DLP_HIP(hipEventRecord(m->evEnd[b], m->stream));
  m->ran[b] = 1;
  m->nextBatch = b + 1;
  return DLP_OK;
}

int dlp_ctx_set_counter_bases(dlp_context *c, int stampBase, int localGen);
int dlp_ctx_counter_bases(dlp_context *c, int *stampBase, int *localGen);

extern "C" int dlp_multi_set_counter_bases(dlp_multi *m, int g, int stampBase, int localGen) {
  if (!m || g < 0 || g >= m->G) {
    dlp_set_error_("dlp_multi_set_counter_bases: bad argument");
    return DLP_ERR_INVALID;
  }
  DLP_HIP(hipStreamSynchronize(m->stream));
  return dlp_ctx_set_counter_bases(m->ctx[g], stampBase, localGen);
}

extern "C" int dlp_multi_counter_bases(dlp_multi *m, int g, int *stampBase, int *localGen) {
  if (!m || g < 0 || g >= m->G) {
    dlp_set_error_("dlp_multi_counter_bases: bad argument");
    return DLP_ERR_INVALID;
  }
  DLP_HIP(hipStreamSynchronize(m->stream));
  return dlp_ctx_counter_bases(m->ctx[g], stampBase, localGen);
}

extern "C" int dlp_multi_tuning_source(const dlp_multi *m, int *source) {
  if (!m || !source) {
    dlp_set_error_("dlp_multi_tuning_source: NULL argument");
    return DLP_ERR_INVALID;
  }
  *source = m->tuningSource;
  return DLP_OK;
}

extern "C" int dlp_multi_autotune(dlp_multi *m, int trialBatches, dlp_multi_tuning *chosen,
                                  float *bestMs) {
  if (!m) {
    dlp_set_error_("dlp_multi_autotune: NULL handle");
    return DLP_ERR_STATE;
  }
  const int nb = trialBatches <= 0 ? m->maxB : std::min(trialBatches, m->maxB);
  dlp_multi_tuning base{};
  int rc = dlp_multi_get_tuning(m, &base);
  if (rc != DLP_OK) return rc;
  /* heavyFrontier 0 keeps the modes' defaults (lockstep: never) */
  std::vector<dlp_multi_tuning> cand;
  for (int shape : {3, 1})
    for (int heavy : {0, 65536, 16384}) {
      dlp_multi_tuning t = base;
      t.localShape = shape;
      t.heavyFrontier = heavy;
      t.useGraphs = 1;
      cand.push_back(t);
    }
  hipEvent_t e0, e1;
  DLP_HIP(hipEventCreate(&e0));
  DLP_HIP(hipEventCreate(&e1));
  int best = -1;
  float bestT = 0.0f;
  auto pass = [&](float *ms) -> int {
    int r = dlp_multi_reset(m);
    if (r != DLP_OK) return r;
    DLP_HIP(hipEventRecord(e0, m->stream));
    for (int b = 0; b < nb; ++b)
      if ((r = dlp_multi_run_batch(m, b)) != DLP_OK) return r;
    DLP_HIP(hipEventRecord(e1, m->stream));
    DLP_HIP(hipEventSynchronize(e1));
    DLP_HIP(hipEventElapsedTime(ms, e0, e1));
    return dlp_multi_synchronize(m);
  };
  for (size_t i = 0; i < cand.size() && rc == DLP_OK; ++i) {
    /* set_tuning resets the heavy field only when non-zero: apply the
     * default explicitly through the stored value */
    m->heavy = cand[i].heavyFrontier;
    if ((rc = dlp_multi_set_tuning(m, &cand[i])) != DLP_OK) break;
    float ms = 0.0f;
    if ((rc = pass(&ms)) != DLP_OK) break;  // captures the batch graphs
    if ((rc = pass(&ms)) != DLP_OK) break;
    if (best < 0 || ms < bestT) {
      best = (int)i;
      bestT = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != DLP_OK) return rc;
  m->heavy = cand[best].heavyFrontier;
  if ((rc = dlp_multi_set_tuning(m, &cand[best])) != DLP_OK) return rc;
  if ((rc = dlp_multi_reset(m)) != DLP_OK) return rc;
  dlp_multi_tuning got{};
  if ((rc = dlp_multi_get_tuning(m, &got)) != DLP_OK) return rc;
  dlp_tuning_put_multi(m->logGraphs, m->logNodes, got, bestT);
  m->tuningSource = 2;
  if (chosen) *chosen = got;
  if (bestMs) *bestMs = bestT;
  return DLP_OK;
}

extern "C" int dlp_multi_run_all(dlp_multi *m) {
  int rc = dlp_multi_reset(m);
  if (rc != DLP_OK) return rc;
  for (int b = 0; b < m->maxB; ++b)
    if ((rc = dlp_multi_run_batch(m, b)) != DLP_OK) return rc;
  return DLP_OK;
}

extern "C" int dlp_multi_synchronize(dlp_multi *m) {
  if (!m) {
    dlp_set_error_("dlp_multi_synchronize: NULL handle");
    return DLP_ERR_STATE;
  }
  DLP_HIP(hipStreamSynchronize(m->stream));
  for (dlp_context *c : m->ctx) {
    DevCounters ctr;
    DLP_HIP(hipMemcpy(&ctr, c->d.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
    if (ctr.errorFlag) {
      dlp_set_errorf_("device kernels reported a violated precondition (flags 0x%x)",
                      ctr.errorFlag);
      return DLP_ERR_INTERNAL;
    }
  }
  return DLP_OK;
}

extern "C" int dlp_multi_read_labels(dlp_multi *m, int g, float *labels) {
  if (!m || g < 0 || g >= m->G) {
    dlp_set_error_("dlp_multi_read_labels: bad argument");
    return DLP_ERR_INVALID;
  }
  return dlp_context_read_labels(m->ctx[g], labels);
}

extern "C" int dlp_multi_batch_stats(dlp_multi *m, int g, int b, dlp_batch_stats *out) {
  if (!m || g < 0 || g >= m->G || !out) {
    dlp_set_error_("dlp_multi_batch_stats: bad argument");
    return DLP_ERR_INVALID;
  }
  dlp_context *c = m->ctx[g];
  if (b < 0 || b >= c->B || !c->ran[b]) {
    dlp_set_errorf_("graph %d: batch %d has not run since the last reset", g, b);
    return DLP_ERR_STATE;
  }
  int rc = dlp_ctx_fill_stats(c, b, out);
  if (rc != DLP_OK) return rc;
  /* the graph's batch ran inside the all-graph batch b (no timing when it
   * ran through dlp_multi_run_next / dlp_multi_serve) */
  float ms = 0.0f;
  if (b < m->maxB && m->ran[b]) DLP_HIP(hipEventElapsedTime(&ms, m->evStart[b], m->evEnd[b]));
  out->timeMs = ms;
  return DLP_OK;
}

extern "C" int dlp_multi_batch_time(dlp_multi *m, int b, float *ms) {
  if (!m || !ms || b < 0 || b >= m->maxB || !m->ran[b]) {
    dlp_set_error_("dlp_multi_batch_time: bad argument or batch not run");
    return DLP_ERR_INVALID;
  }
  DLP_HIP(hipEventSynchronize(m->evEnd[b]));
  DLP_HIP(hipEventElapsedTime(ms, m->evStart[b], m->evEnd[b]));
  return DLP_OK;
}

extern "C" int dlp_multi_device_bytes(const dlp_multi *m, int64_t *bytes) {
  if (!m || !bytes) {
    dlp_set_error_("dlp_multi_device_bytes: NULL argument");
    return DLP_ERR_INVALID;
  }
  int64_t t = 0;
  for (dlp_context *c : m->ctx) t += c->deviceBytes;
  *bytes = t + (int64_t)sizeof(GraphProp) * m->G * m->maxB + (m->poolBase ? m->pool.capacity() : 0);
  return DLP_OK;
}

extern "C" int dlp_multi_set_scratch_pool(dlp_multi *m, int64_t poolBytes) {
  if (!m || poolBytes < 0) {
    dlp_set_error_("dlp_multi_set_scratch_pool: NULL handle or negative size");
    return DLP_ERR_INVALID;
  }
  DLP_HIP(hipStreamSynchronize(m->stream));
  for (int64_t l : m->lease)
    if (l >= 0) {
      dlp_set_error_("dlp_multi_set_scratch_pool: a batch holds a lease");
      return DLP_ERR_STATE;
    }
  int rc = DLP_OK;
  if (poolBytes > 0) {
    const int64_t cap = poolBytes / kPoolAlign * kPoolAlign;
    if (cap <= 0) {
      dlp_set_errorf_("dlp_multi_set_scratch_pool: %lld bytes is below one %lld-byte range",
                      (long long)poolBytes, (long long)kPoolAlign);
      return DLP_ERR_INVALID;
    }
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, (size_t)cap);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      dlp_set_errorf_("hipMalloc of a %lld-byte scratch pool failed (%s)", (long long)cap,
                      hipGetErrorString(e));
      return DLP_ERR_NOMEM;
    }
    for (dlp_context *c : m->ctx)
      if ((rc = dlp_ctx_own_scratch(c, false)) != DLP_OK) {
        (void)hipFree(p);
        return rc;
      }
    if (m->poolBase) (void)hipFree(m->poolBase);
    m->poolBase = p;
    m->pool.reset(cap);
  } else {
    for (size_t i = 0; i < m->ctx.size(); ++i)
      if ((rc = dlp_ctx_own_scratch(m->ctx[i], true)) != DLP_OK) {
        for (size_t j = 0; j < i; ++j) (void)dlp_ctx_own_scratch(m->ctx[j], false);
        return rc;
      }
    if (m->poolBase) (void)hipFree(m->poolBase);
    m->poolBase = nullptr;
    m->pool.reset(0);
  }
  m->deferrals = 0;
  m->drop_graphs();  // captured launches embed the scratch pointers
  m->exec.assign(m->maxB, nullptr);
  m->execPost.assign(m->maxB, nullptr);
  return upload_props(m);
}

extern "C" int dlp_multi_pool_stats(dlp_multi *m, dlp_pool_stats *out) {
  if (!m || !out) {
    dlp_set_error_("dlp_multi_pool_stats: NULL argument");
    return DLP_ERR_INVALID;
  }
  out->capacity = m->poolBase ? m->pool.capacity() : 0;
  out->inUse = m->pool.inUse();
  out->peak = m->pool.peak();
  out->leases = m->pool.leases();
  out->deferrals = m->deferrals;
  return DLP_OK;
}

extern "C" int dlp_multi_batch_scratch_bytes(const dlp_multi *m, int g, int b, int64_t *bytes) {
  if (!m || !bytes || g < 0 || g >= m->G || b < 0 || b >= m->ctx[g]->B) {
    dlp_set_error_("dlp_multi_batch_scratch_bytes: bad argument");
    return DLP_ERR_INVALID;
  }
  *bytes = RangePool::rounded(dlp_ctx_batch_scratch_bytes(m->ctx[g], b));
  return DLP_OK;
}

/* ------------------------------------------------------------------------ */
/* serving: the next batch of any subset of graphs, and an arrival-driven   */
/* scheduler over it                                                        */
/* ------------------------------------------------------------------------ */
namespace {

/* the next batch of `graphs` (validated) together.  With `done` set, it is
 * called with the stream idle for the graphs whose batch has completed:
 * first all that the propagation launch finished, then each handed-back
 * graph on its own.  With `handOff` set, a handed-back graph is passed to
 * it (graph, resume iteration) instead of being finished here. */
int run_next_impl(dlp_multi *m, const int *graphs, int count, int heavy,
                  const std::function<void(const std::vector<int> &)> *done,
                  const std::function<int(int, int)> *handOff, bool lockstep) {
  DLP_HIP(hipStreamSynchronize(m->stream));  // the staging buffer is free
  int rc;
  if (!lockstep) m->nextBatch = -1;  // graphs advance one by one: no lockstep batch until a reset
  for (int i = 0; i < count; ++i) {
    dlp_context *c = m->ctx[graphs[i]];
    const int b = c->nextBatch;
    int tag = 0;
    if ((rc = dlp_ctx_counter_guard(c, m->stream)) != DLP_OK) return rc;
    if ((rc = dlp_ctx_batch_pre(c, b, dlp_ctx_make_batch(c, b), &tag)) != DLP_OK) return rc;
    m->hStep[i] = graph_prop(m, graphs[i]);  // the scratch pointers of a lease move
    m->hStep[i].heavy = heavy;
  }
  DLP_HIP(hipMemcpyAsync(m->dStep, m->hStep, sizeof(GraphProp) * count, hipMemcpyHostToDevice,
                         m->stream));
  if ((rc = dlp_launch_propagate_multi(m->dStep, count, m->shape, m->ldsWords, m->stream)) !=
      DLP_OK)
    return rc;
  std::vector<int> handed, finished;
  if ((rc = handed_back(m, heavy, graphs, count, &handed)) != DLP_OK) return rc;
  auto post = [&](int g) {
    dlp_context *c = m->ctx[g];
    return dlp_ctx_batch_post(c, c->nextBatch, dlp_ctx_make_batch(c, c->nextBatch));
  };
  for (int i = 0; i < count; ++i) {
    if (std::find(handed.begin(), handed.end(), graphs[i]) != handed.end()) continue;
    if ((rc = post(graphs[i])) != DLP_OK) return rc;
    finished.push_back(graphs[i]);
  }
  if (done && !finished.empty()) {
    DLP_HIP(hipStreamSynchronize(m->stream));
    (*done)(finished);
  }
  for (int g : handed) {
    if (handOff) {
      if ((rc = (*handOff)(g, m->hExit[g])) != DLP_OK) return rc;
      continue;
    }
    if ((rc = continue_graph(m, g)) != DLP_OK || (rc = post(g)) != DLP_OK) return rc;
    if (done) {
      DLP_HIP(hipStreamSynchronize(m->stream));
      (*done)(std::vector<int>{g});
    }
  }
  return DLP_OK;
}

}  // namespace

extern "C" int dlp_multi_run_next(dlp_multi *m, const int *graphs, int count) {
  if (!m || count < 0 || (count > 0 && !graphs)) {
    dlp_set_error_("dlp_multi_run_next: bad argument");
    return DLP_ERR_INVALID;
  }
  if (count == 0) return DLP_OK;
  std::vector<char> seen(m->G, 0);
  for (int i = 0; i < count; ++i) {
    const int g = graphs[i];
    if (g < 0 || g >= m->G || seen[g]) {
      dlp_set_errorf_("dlp_multi_run_next: graph %d out of range or listed twice", g);
      return DLP_ERR_INVALID;
    }
    seen[g] = 1;
    if (m->ctx[g]->nextBatch >= m->ctx[g]->B) {
      dlp_set_errorf_("dlp_multi_run_next: graph %d has no batch left", g);
      return DLP_ERR_STATE;
    }
  }
  StepLeases leases(m);
  const int rc = leases.take(graphs, count);
  if (rc != DLP_OK) return rc;
  return run_next_impl(m, graphs, count, heavy_lockstep(m), nullptr);
}

namespace {

/* iteration-level admission of dlp_multi_serve: a step stops each graph at
 * the first iteration boundary after kServeSliceIterations iterations, or
 * after kServeSliceMs of work-group HBM iterations, and the next step
 * resumes it beside the batches admitted in between */
constexpr int kServeSliceIterations = 2048;
constexpr double kServeSliceMs = 2.0;
/* default dlp_serve_config::handBackAfter: a batch still running after this
 * many iterations is in a long-lived frontier (a limit cycle runs to the
 * 100,000-iteration cap), which the single-graph loop's owner-mode LDS
 * kernels iterate several times faster than a work-group of the shared
 * launch in queue mode */
constexpr int kServeHandBackAfter = 1024;
/* default dlp_serve_config::handBackLanes: long batches (limit cycles) keep
 * a lane each for their whole run, so serving wants more lanes than the
 * lockstep runner */
constexpr int kServeLanes = 16;

/* a batch in the shared propagation launches of dlp_multi_serve */
struct Flight {
  int graph, req;
  int k; /* iteration the next launch starts it at */
};

}  // namespace

extern "C" int dlp_multi_serve(dlp_multi *m, const dlp_serve_request *reqs, int numReqs,
                               const dlp_serve_config *cfg, double *completion, int *steps) {
  if (!m || numReqs < 0 || (numReqs > 0 && (!reqs || !completion)) || !cfg ||
      cfg->maxSlots < 1 || cfg->hbmBudgetBytes < 0 || cfg->sliceIterations < -1 ||
      cfg->handBackAfter < -1) {
    dlp_set_error_("dlp_multi_serve: bad argument");
    return DLP_ERR_INVALID;
  }
  /* requests of a graph are its next batches, in the given (arrival) order */
  std::vector<int> need(m->G, 0);
  for (int i = 0; i < numReqs; ++i) {
    const int g = reqs[i].graph;
    if (g < 0 || g >= m->G || (i > 0 && reqs[i].arrival < reqs[i - 1].arrival)) {
      dlp_set_errorf_("dlp_multi_serve: request %d: bad graph or arrivals not sorted", i);
      return DLP_ERR_INVALID;
    }
    if (m->ctx[g]->nextBatch + ++need[g] > m->ctx[g]->B) {
      dlp_set_errorf_("dlp_multi_serve: graph %d has fewer batches left than requests", g);
      return DLP_ERR_INVALID;
    }
  }
  int rc;
  /* the HBM budget is the size of the scratch pool the batches draw from */
  if (cfg->hbmBudgetBytes > 0 &&
      (!pooled(m) || m->pool.capacity() != cfg->hbmBudgetBytes / kPoolAlign * kPoolAlign) &&
      (rc = dlp_multi_set_scratch_pool(m, cfg->hbmBudgetBytes)) != DLP_OK)
    return rc;
  m->nextBatch = -1;  // graphs advance one by one: no lockstep batch until a reset
  using clock = std::chrono::steady_clock;
  const clock::time_point t0 = clock::now();
  auto now = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
  const int slots = cfg->plain ? 1 : cfg->maxSlots;
  const int slice = cfg->sliceIterations == 0 ? kServeSliceIterations : cfg->sliceIterations;
  const double sliceMs = cfg->sliceMs == 0.0 ? kServeSliceMs : cfg->sliceMs;
  const long long sliceTicks =
      sliceMs > 0.0 && m->wallRateKHz > 0 ? (long long)(sliceMs * m->wallRateKHz) : 0;
  const int longAfter = cfg->handBackAfter == 0 ? kServeHandBackAfter : cfg->handBackAfter;
  std::vector<int> pending;  // request indices not yet admitted, in arrival order
  pending.reserve(numReqs);
  int nextArrival = 0, nSteps = 0;
  std::vector<char> blocked(m->G, 0), flying(m->G, 0);
  std::vector<Flight> flight, fresh, carried;
  /* handed-back graphs finish beside the following steps; their lease goes
   * back when they are done */
  std::vector<int> laneReq(m->G, -1);  // request a graph on the lane serves
  std::unique_ptr<HandBackLane> lane;
  const int heavy = cfg->plain ? 0 : heavy_serving(m);
  if (heavy > 0) {
    const int workers = cfg->handBackLanes > 0 ? cfg->handBackLanes : kServeLanes;
    lane.reset(new HandBackLane(m, workers, true, [&](int g) {
      completion[laneReq[g]] = now();
      return_scratch(m, g);
    }));
    if ((rc = lane->start()) != DLP_OK) return rc;
  }
  auto laneFailed = [&]() {
    rc = lane ? lane->error() : DLP_OK;
    if (rc != DLP_OK) dlp_set_errorf_("dlp_multi_serve: hand-back lane: %s", lane->message().c_str());
    return rc != DLP_OK;
  };
  /* returns every lease still held when an error ends the call */
  struct LeaseSweep {
    dlp_multi *m;
    std::unique_ptr<HandBackLane> *lane;
    ~LeaseSweep() {
      if (*lane) (*lane)->stop();
      (void)hipStreamSynchronize(m->stream);
      for (int g = 0; g < m->G; ++g) return_scratch(m, g);
    }
  } sweep{m, &lane};
  auto post = [&](int g) {
    dlp_context *c = m->ctx[g];
    return dlp_ctx_batch_post(c, c->nextBatch, dlp_ctx_make_batch(c, c->nextBatch));
  };
  while (nextArrival < numReqs || !pending.empty() || !flight.empty()) {
    if (laneFailed()) return rc;
    const double t = now();
    while (nextArrival < numReqs && reqs[nextArrival].arrival <= t) pending.push_back(nextArrival++);
    const double nextAt = nextArrival < numReqs ? reqs[nextArrival].arrival : t + 1.0;
    /* admission, in arrival order: one batch per graph at a time, a graph's
     * batches in order, at most `slots` batches in the shared launch, and
     * scratch from the pool (a batch that does not fit now is deferred; a
     * later, smaller one may pass it) */
    fresh.clear();
    std::fill(blocked.begin(), blocked.end(), 0);
    std::vector<int> rest;
    rest.reserve(pending.size());
    for (int idx : pending) {
      const int g = reqs[idx].graph;
      bool admit = !blocked[g] && !flying[g] && !(lane && lane->busy(g)) &&
                   (int)(flight.size() + fresh.size()) < slots;
      blocked[g] = 1;  // later requests of g wait for this one
      if (admit) {
        bool granted = false;
        if ((rc = lease_scratch(m, g, &granted)) != DLP_OK) return rc;
        if (!granted) ++m->deferrals;
        admit = granted;
      }
      if (admit) fresh.push_back({g, idx, 0});
      else rest.push_back(idx);
    }
    pending.swap(rest);
    if (flight.empty() && fresh.empty()) {
      if (pending.empty() && nextArrival >= numReqs) break;  // only hand-back work left
      if (lane && lane->inFlight() > 0) {
        lane->wait(nextAt, now);  // a graph or its lease comes back, or a request arrives
      } else if (pending.empty()) {
        std::this_thread::sleep_for(std::chrono::duration<double>(nextAt - t));
      } else {  // cannot happen: an idle runner admits the first pending batch
        dlp_set_error_("dlp_multi_serve: no batch admitted with the runner idle");
        return DLP_ERR_INTERNAL;
      }
      continue;
    }
    if (cfg->plain) {  // one batch at a time through the single-graph path
      const Flight f = fresh[0];
      dlp_context *c = m->ctx[f.graph];
      if ((rc = dlp_context_run_batch(c, c->nextBatch)) != DLP_OK) return rc;
      DLP_HIP(hipStreamSynchronize(c->stream));
      completion[f.req] = now();
      return_scratch(m, f.graph);
      ++nSteps;
      continue;
    }
    /* the new batches' pipelines up to their first frontier */
    for (const Flight &f : fresh) {
      dlp_context *c = m->ctx[f.graph];
      const int b = c->nextBatch;
      int tag = 0;
      if ((rc = dlp_ctx_counter_guard(c, m->stream)) != DLP_OK) return rc;
      if ((rc = dlp_ctx_batch_pre(c, b, dlp_ctx_make_batch(c, b), &tag)) != DLP_OK) return rc;
      flying[f.graph] = 1;
      flight.push_back(f);
    }
    /* one shared launch: resumed and new graphs, each until its slice ends */
    const int count = (int)flight.size();
    for (int i = 0; i < count; ++i) {
      const Flight &f = flight[i];
      GraphProp gp = graph_prop(m, f.graph);
      gp.heavy = heavy;
      gp.startIter = f.k;
      if (slice > 0)
        gp.stopIter = (int)std::min<long long>(gp.maxIter, (long long)f.k + slice);
      gp.sliceTicks = sliceTicks;
      m->hStep[i] = gp;
    }
    DLP_HIP(hipMemcpyAsync(m->dStep, m->hStep, sizeof(GraphProp) * count, hipMemcpyHostToDevice,
                           m->stream));
    if ((rc = dlp_launch_propagate_multi(m->dStep, count, m->shape, m->ldsWords, m->stream)) !=
        DLP_OK)
      return rc;
    DLP_HIP(hipMemcpyAsync(m->hExit, m->dExit, sizeof(int) * m->G, hipMemcpyDeviceToHost,
                           m->stream));
    DLP_HIP(hipStreamSynchronize(m->stream));
    ++nSteps;
    /* finished graphs: batch end; handed back: to a lane; sliced: carried */
    std::vector<int> done;
    carried.clear();
    for (const Flight &f : flight) {
      const int e = m->hExit[f.graph];
      const int maxIter = m->ctx[f.graph]->cfg.maxIterations;
      if (e == -1) {
        if ((rc = post(f.graph)) != DLP_OK) return rc;
        done.push_back(f.graph);
      } else if (e >= 0 && e <= maxIter && heavy > 0) {
        flying[f.graph] = 0;
        laneReq[f.graph] = f.req;
        if ((rc = lane->push(f.graph, e)) != DLP_OK) return rc;
      } else if (e <= -2 && -2 - e >= f.k && -2 - e < maxIter) {
        const int k = -2 - e;
        if (heavy > 0 && longAfter >= 0 && k >= longAfter && lane->inFlight() < lane->workers()) {
          /* a long-lived frontier, and a lane is free: the lane's owner-mode
           * kernels finish it (a slice stop leaves the state a hand-back
           * stop does); with every lane busy it goes on in the shared
           * launches and may leave at a later slice stop */
          flying[f.graph] = 0;
          laneReq[f.graph] = f.req;
          if ((rc = lane->push(f.graph, k)) != DLP_OK) return rc;
        } else {
          carried.push_back({f.graph, f.req, k});
        }
      } else {
        dlp_set_errorf_("graph %d: propagation launch exit %d from iteration %d", f.graph, e, f.k);
        return DLP_ERR_INTERNAL;
      }
    }
    if (!done.empty()) {
      DLP_HIP(hipStreamSynchronize(m->stream));
      const double tDone = now();
      for (const Flight &f : flight)
        if (std::find(done.begin(), done.end(), f.graph) != done.end()) {
          completion[f.req] = tDone;
          flying[f.graph] = 0;
          return_scratch(m, f.graph);
        }
    }
    flight.swap(carried);
  }
  if (lane) {
    lane->drain();
    if (laneFailed()) return rc;
  }
  if (steps) *steps = nSteps;
  return DLP_OK;
}
