/*
 * dlp_proploop.hip -- host driver of the propagation loop.
 *
 * The reference launches three kernels per iteration and reads two flags
 * back with a stream synchronisation each (iterativePropagation.cu:274-327).
 * Here the loop state lives on the device and the host only decides, from
 * the frontier size it reads back, which engine runs next:
 *   - frontier <= localThreshold: the LDS-resident kernel, which runs many
 *     iterations in one launch and returns when the frontier empties, the
 *     iteration cap is reached, or the working set no longer fits;
 *   - otherwise: a chunk of multi-block launches, one per iteration, whose
 *     length doubles up to maxChunk while the frontier stays large.
 */
#include <algorithm>

#include "dlp_kernels.h"

namespace {

int read_ctr(DevCounters *dev, DevCounters *pin, hipStream_t s) {
  DLP_HIP(hipMemcpyAsync(pin, dev, sizeof(DevCounters), hipMemcpyDeviceToHost, s));
  DLP_HIP(hipStreamSynchronize(s));
  if (pin->errorFlag) {
    dlp_set_errorf_("device kernels reported a violated precondition (flags 0x%x)",
                    pin->errorFlag);
    return DLP_ERR_INTERNAL;
  }
  return DLP_OK;
}

int begin_timed(PropTiming *t, hipStream_t s, bool local, int launches) {
  if (!t) return DLP_OK;
  while ((int)t->ev.size() < t->used + 2) {
    hipEvent_t e;
    DLP_HIP(hipEventCreate(&e));
    t->ev.push_back(e);
  }
  const int pair = t->used / 2;
  if ((int)t->isLocal.size() < pair + 1) {
    t->isLocal.resize(pair + 1);
    t->count.resize(pair + 1);
  }
  t->isLocal[pair] = local ? 1 : 0;
  t->count[pair] = launches;
  DLP_HIP(hipEventRecord(t->ev[t->used], s));
  return DLP_OK;
}

int end_timed(PropTiming *t, hipStream_t s) {
  if (!t) return DLP_OK;
  DLP_HIP(hipEventRecord(t->ev[t->used + 1], s));
  t->used += 2;
  return DLP_OK;
}

/* the stream is idle: fold the recorded launch times into the totals */
int collect_timed(PropTiming *t) {
  if (!t) return DLP_OK;
  for (int i = 0; i < t->used; i += 2) {
    float ms = 0.0f;
    DLP_HIP(hipEventElapsedTime(&ms, t->ev[i], t->ev[i + 1]));
    if (t->isLocal[i / 2]) {
      t->localMs += ms;
      t->localLaunches += t->count[i / 2];
    } else {
      t->ms += ms;
      t->launches += t->count[i / 2];
    }
  }
  t->used = 0;
  return DLP_OK;
}

}  // namespace

int dlp_run_propagation(const PropArgs &a, const LocalArgs *l, int maxIter,
                        const PropLoopConfig &cfg, DevCounters *pin, hipStream_t s,
                        PropTiming *timing, int startK) {
  int rc;
  int k = startK;
  int chunk = 2;
  int localAfter = 0; /* no local launch before this iteration (after an overflow) */
  bool owner = false;  /* the last queue launch of the LDS kernel ran long */
  int maskMin = 0;     /* smallest register-mask owner kernel to try (a kernel
                          that found the core too wide moves it up) */
  const int maxChunk = std::max(1, cfg.maxChunk);
  if ((rc = read_ctr(a.ctr, pin, s)) != DLP_OK) return rc;
  while (k < maxIter) {
    const int n = pin->qcount[k % 3];
    if (n == 0) break;
    if (l && n <= cfg.localThreshold && k >= localAfter) {
      /* restart the core generations before a launch could carry localGen
       * past kGenLimit (one launch builds at most dlp_gens_per_batch) */
      if ((long long)pin->localGen + dlp_gens_per_batch(maxIter) >= kGenLimit &&
          (rc = dlp_restart_local_gens(*l, a.ctr, s)) != DLP_OK)
        return rc;
      /* owner mode on a small frontier: a register-mask kernel, sized for a
       * core of about twice the frontier (a limit cycle alternates between
       * two frontiers), at least maskMin */
      LocalArgs la = *l;
      int level = kMaskLevels;
      if (owner && n <= std::min(cfg.maskOwnerMax, kMaskOwnerMax)) {
        level = maskMin;
        while (level < kMaskLevels && 2 * n > kMaskLevelThreads[level]) ++level;
        /* the multi-wave mask kernels read the next frontier off symmetric
         * rows: not for an asymmetric CSR */
        if (!a.symmetric && level > 0) level = kMaskLevels;
        if (level < kMaskLevels) la.shape = kLocalMaskShape + level;
      }
      if ((rc = begin_timed(timing, s, true, 1)) != DLP_OK) return rc;
      if ((rc = dlp_launch_propagate_local(a, la, k, maxIter, owner, s)) != DLP_OK) return rc;
      if ((rc = end_timed(timing, s)) != DLP_OK) return rc;
      if ((rc = read_ctr(a.ctr, pin, s)) != DLP_OK) return rc;
      const int next = pin->nextIter;
      if (next < k || next > maxIter) {
        dlp_set_errorf_("local propagation returned iteration %d from %d", next, k);
        return DLP_ERR_INTERNAL;
      }
      if (pin->localReason == DLP_LOCAL_TOO_BIG) {
        localAfter = next + std::max(2, maxChunk);
        chunk = 2;
        owner = false;
      }
      if (pin->localReason == DLP_LOCAL_LONG) {
        owner = true;
        maskMin = 0;
      }
      /* the core did not fit the mask kernel: the next larger one, or the
       * flag-based owner kernel */
      if (pin->localReason == DLP_LOCAL_WIDE) maskMin = level + 1;
      k = next;
      if (pin->localReason == DLP_LOCAL_CONVERGED || pin->localReason == DLP_LOCAL_CAPPED)
        break;
      continue;
    }
    chunk = std::min(chunk, maxChunk);
    const int m = std::min(chunk, maxIter - k);
    /* grid: one lane group per vertex of twice the chunk's first frontier
     * (frontiers grow within a chunk), at least kMinPropagateBlocks (a small
     * frontier spread over many compute units hides latency better than
     * the same waves packed onto a few) and at most cfg.blocks */
    const long long want =
        (2LL * n * cfg.groupLanes + kPropagateThreads - 1) / kPropagateThreads;
    const int blocks = (int)std::max<long long>(
        1, std::min<long long>(cfg.blocks, std::max<long long>(want, kMinPropagateBlocks)));
    PropTiming *perLaunch = (timing && !timing->perChunk) ? timing : nullptr;
    if (timing && timing->perChunk && (rc = begin_timed(timing, s, false, m)) != DLP_OK)
      return rc;
    for (int j = 0; j < m; ++j) {
      if ((rc = begin_timed(perLaunch, s, false, 1)) != DLP_OK) return rc;
      if ((rc = dlp_launch_propagate(a, k + j, cfg.groupLanes, blocks, s)) != DLP_OK)
        return rc;
      if ((rc = end_timed(perLaunch, s)) != DLP_OK) return rc;
    }
    if (timing && timing->perChunk && (rc = end_timed(timing, s)) != DLP_OK) return rc;
    k += m;
    if ((rc = read_ctr(a.ctr, pin, s)) != DLP_OK) return rc;
    chunk *= 2;
  }
  return collect_timed(timing);
}
