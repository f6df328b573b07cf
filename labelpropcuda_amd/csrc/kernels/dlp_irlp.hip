/*
 * dlp_irlp.hip -- IrLP, the full-recompute baseline (validation.cu:47-205).
 *
 * Every iteration recomputes the label of every active vertex that is not
 * ground truth from all of its neighbours, double-buffered, with ground
 * truth enforced to 0 / 1, until no label moves by more than delta.
 *
 * The reference runs one 256-thread block per vertex and one enforcement
 * kernel per iteration, and copies a changed flag to the host after every
 * iteration.  Here:
 *   - before the first iteration, one pass replaces every column that
 *     points at a DELETED vertex by -1 in a copy of colInd (the neighbour
 *     status never changes during IrLP), so a sweep reads no neighbour
 *     status: the entry adds nothing either way;
 *   - one sweep kernel per iteration over consecutive vertices, with ground
 *     truth enforcement fused into the label store.  Default: one lane per
 *     vertex (64 consecutive vertices per wave, so the per-vertex arrays are
 *     read in coalesced runs and every lane's row loads and label gathers
 *     are independent), rows of more than 16 entries taken by the whole wave
 *     one at a time; or G lanes (4 .. 64) per vertex;
 *   - the row sums are the reference block's 256-slot tree
 *     (dlp::lane_row_sums / dlp::group_row_sums), so every label is the same
 *     float;
 *   - the changed flag of iteration i is word i+1 of a small per-chunk
 *     array; the kernel of iteration i returns at once when word i is zero
 *     (iteration i-1 changed nothing), so a chunk of iterations is queued
 *     without a host round trip and the host reads the words once per chunk.
 *     The first zero word gives the reference's iterationsUsed exactly;
 *     kernels queued after it write nothing.
 *
 * Per iteration over N vertices and E stored entries the algorithmic
 * traffic is 20 B per vertex (row pointer, ground-truth class, status, old
 * and new label) and 12 B per entry (column, weight, the neighbour's label).
 */
#include <algorithm>

#include "../dlp_kernels.h"
#include "../dlp_rowsum.h"

namespace {

constexpr int kIrlpThreads = 256;

/* word 0 <- 1 (the iteration before the chunk changed something), words
 * 1..C <- 0 */
__global__ void k_irlp_flags(int *flags, int C) {
  for (int i = threadIdx.x; i <= C; i += blockDim.x) flags[i] = i == 0 ? 1 : 0;
}

/* colMasked[e] = colInd[e], or -1 where that is negative or a DELETED
 * vertex */
__global__ void k_irlp_mask(const int *colInd, const int *status, int *colMasked, long long E) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (long long)gridDim.x * blockDim.x) {
    const int v = colInd[e];
    colMasked[e] = (v >= 0 && status[v] != DLP_VERTEX_DELETED) ? v : -1;
  }
}

/* the label update of vertex u from its row sums, with ground truth
 * enforced (validation.cu:56-70, :107-126, :131-137) */
__device__ __forceinline__ float irlp_label(int gt, bool live, float W, float S, float old,
                                            float delta, bool &changed) {
  if (gt == 0) return 0.0f;
  if (gt == 1) return 1.0f;
  if (!live) return old;
  const float nl = dlp::new_label(W, S, old);
  changed = changed || fabsf(nl - old) > delta;
  return nl;
}

constexpr int kIrlpLaneRowMax = 16;
constexpr int kIrlpStage = 1024; /* entries of a wave's rows staged in LDS */

/* row position p of a staged row: weight and neighbour label copied to LDS,
 * both 0 for an entry that adds nothing (the product is then +0.0 too) */
struct StagedPos {
  const float *w;
  const float *f;
  int off;
  __device__ void operator()(int p, float &ww, float &ss) const {
    ww = w[off + p];
    ss = ww * f[off + p];
  }
};

/* the labels of a wave's 64 consecutive vertices (lane = vertex): rows up to
 * 16 entries by their own lane, longer rows by the whole wave one at a time;
 * `at(b)` is the position functor of a row starting at entry b */
template <class At>
__device__ __forceinline__ void irlp_wave_rows(const At &at, long long base, int lane, bool act,
                                               int gt, bool live, int beg, int len,
                                               const float *Fold, float *Fnew, float delta,
                                               bool &changed) {
  const long long u = base + lane;
  const bool laneRow = live && len <= kIrlpLaneRowMax;
  const int lenS = laneRow ? len : 0;
  float W, S;
  if (!__any(lenS > 4))
    dlp::lane_row_sums<4>(lenS, at(beg), W, S);
  else if (!__any(lenS > 8))
    dlp::lane_row_sums<8>(lenS, at(beg), W, S);
  else
    dlp::lane_row_sums<kIrlpLaneRowMax>(lenS, at(beg), W, S);
  if (act && (laneRow || !live)) Fnew[u] = irlp_label(gt, live, W, S, Fold[u], delta, changed);
  unsigned long long longMask = dlp_ballot(live && !laneRow);
  while (longMask) {  // wave-uniform
    const int src = __ffsll((long long)longMask) - 1;
    longMask &= longMask - 1;
    const int lbeg = __shfl(beg, src, 64), llen = __shfl(len, src, 64);
    float LW, LS;
    dlp::group_row_sums<64>(lane, llen, at(lbeg), LW, LS);
    if (lane == 0) {
      const long long lu = base + src;
      Fnew[lu] = irlp_label(-1, true, LW, LS, Fold[lu], delta, changed);
    }
  }
}

/* one lane per vertex, 64 consecutive vertices per wave; iteration `iter`
 * as word j of the chunk.  When the wave's rows hold at most kIrlpStage
 * entries, their weights and neighbour labels are first copied to LDS with
 * coalesced loads (one label gather per entry, all independent), and the
 * rows are summed from there. */
__global__ __launch_bounds__(kIrlpThreads) void k_irlp_lanes(IrlpArgs a, int iter, int j) {
  __shared__ float stW[kIrlpThreads / 64][kIrlpStage];
  __shared__ float stF[kIrlpThreads / 64][kIrlpStage];
  if (a.flags[j] == 0) return;  // converged before this iteration (grid-uniform)
  const float *Fold = (iter & 1) ? a.lab1 : a.lab0;
  float *Fnew = (iter & 1) ? a.lab0 : a.lab1;
  const int *col = a.colMasked ? a.colMasked : a.colInd;
  const int *nbStatus = a.colMasked ? nullptr : a.status;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float *sw = stW[wv], *sf = stF[wv];
  const long long gwave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  const long long N = a.numNodes;
  bool changed = false;
  for (long long base = gwave * 64; base < N; base += nwaves * 64) {  // wave-uniform
    const long long u = base + lane;
    const bool act = u < N;
    int gt = -1, st = DLP_VERTEX_DELETED, beg = 0, len = 0;
    if (act) {
      gt = a.gtClass[u];
      st = a.status[u];
    }
    const bool live = act && gt < 0 && st == DLP_VERTEX_ACTIVE;
    if (live) {
      beg = a.rowPtr[u];
      len = a.rowPtr[u + 1] - beg;
    }
    const int e0 = a.rowPtr[base];
    const int e1 = a.rowPtr[base + 64 < N ? base + 64 : N];
    if (e1 - e0 <= kIrlpStage) {  // wave-uniform
      for (int i = lane; i < e1 - e0; i += 64) {
        const int v = col[e0 + i];
        const bool ok = v >= 0 && !(nbStatus && nbStatus[v] == DLP_VERTEX_DELETED);
        sw[i] = ok ? a.values[e0 + i] : 0.0f;
        sf[i] = ok ? Fold[v] : 0.0f;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // stage written before read
      irlp_wave_rows([&](int b) { return StagedPos{sw, sf, b - e0}; }, base, lane, act, gt, live,
                     beg, len, Fold, Fnew, a.delta, changed);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // read before the next stage
    } else {
      irlp_wave_rows([&](int b) { return dlp::GlobalPos{col, a.values, Fold, nbStatus, b}; },
                     base, lane, act, gt, live, beg, len, Fold, Fnew, a.delta, changed);
    }
  }
  if (__any(changed) && lane == 0) a.flags[j + 1] = 1;
}

/* iteration `iter` (global index, picks the buffers) as word j of the
 * chunk */
template <int G>
__global__ __launch_bounds__(kIrlpThreads) void k_irlp(IrlpArgs a, int iter, int j) {
  if (a.flags[j] == 0) return;  // converged before this iteration (grid-uniform)
  constexpr int GPW = 64 / G;   /* vertices per wave */
  const float *Fold = (iter & 1) ? a.lab1 : a.lab0;
  float *Fnew = (iter & 1) ? a.lab0 : a.lab1;
  const int lane = threadIdx.x & 63;
  const int gl = lane & (G - 1);
  const int grp = lane / G;
  const long long gwave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  const long long N = a.numNodes;
  bool changed = false;
  for (long long base = gwave * GPW; base < N; base += nwaves * GPW) {  // wave-uniform
    const long long u = base + grp;
    const bool act = u < N;
    int gt = -1, st = DLP_VERTEX_DELETED, beg = 0, len = 0;
    if (act) {
      gt = a.gtClass[u];
      st = a.status[u];
    }
    /* validation.cu:56-70: ground truth and non-active vertices keep their
     * label (then :131-137 enforces ground truth) */
    const bool live = act && gt < 0 && st == DLP_VERTEX_ACTIVE;
    if (live) {
      beg = a.rowPtr[u];
      len = a.rowPtr[u + 1] - beg;
    }
    float W, S;
    dlp::group_row_sums<G>(gl, len,
                           dlp::GlobalPos{a.colMasked ? a.colMasked : a.colInd, a.values, Fold,
                                          a.colMasked ? nullptr : a.status, beg},
                           W, S);
    if (act && gl == 0) Fnew[u] = irlp_label(gt, live, W, S, Fold[u], a.delta, changed);
  }
  if (__any(changed) && lane == 0) a.flags[j + 1] = 1;
}

template <int G>
int launch_irlp(const IrlpArgs &a, int iter, int j, int blocks, hipStream_t s) {
  if (G == 1)
    k_irlp_lanes<<<blocks, kIrlpThreads, 0, s>>>(a, iter, j);
  else
    k_irlp<G><<<blocks, kIrlpThreads, 0, s>>>(a, iter, j);
  return DLP_OK;
}

}  // namespace

int dlp_irlp_blocks_default(int numNodes, int groupLanes) {
  const long long waves = ((long long)numNodes * groupLanes + 63) / 64;
  const long long blocks = (waves + kIrlpThreads / 64 - 1) / (kIrlpThreads / 64);
  return (int)std::max<long long>(1, std::min<long long>(blocks, 8192));
}

int dlp_run_irlp(const IrlpArgs &a, int maxIterations, const IrlpLoopConfig &cfg, int *hostFlags,
                 hipStream_t s, int *iterationsUsed, float *kernelMs) {
  *iterationsUsed = 0;
  if (kernelMs) *kernelMs = 0.0f;
  if (maxIterations <= 0 || a.numNodes <= 0) return DLP_OK;
  const int G = cfg.groupLanes;
  if (G != 1 && G != 4 && G != 8 && G != 16 && G != 32 && G != 64) {
    dlp_set_errorf_("IrLP: group lanes %d not in {1,4,8,16,32,64}", G);
    return DLP_ERR_INVALID;
  }
  if (cfg.firstChunk < 1 || cfg.maxChunk < cfg.firstChunk || cfg.maxChunk > kIrlpMaxChunk) {
    dlp_set_error_("IrLP: bad chunk lengths");
    return DLP_ERR_INVALID;
  }
  const int blocks = cfg.blocks > 0 ? cfg.blocks : dlp_irlp_blocks_default(a.numNodes, G);
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (kernelMs) {
    DLP_HIP(hipEventCreate(&ev[0]));
    DLP_HIP(hipEventCreate(&ev[1]));
    DLP_HIP(hipEventRecord(ev[0], s));
  }
  if (a.colMasked && a.numEntries > 0) {
    const long long blocksM = std::min<long long>((a.numEntries + 255) / 256, 8192);
    k_irlp_mask<<<(int)blocksM, 256, 0, s>>>(a.colInd, a.status, a.colMasked, a.numEntries);
  }
  int used = -1, it = 0, chunk = cfg.firstChunk;
  int rc = DLP_OK;
  while (it < maxIterations && used < 0 && rc == DLP_OK) {
    const int C = std::min(chunk, maxIterations - it);
    k_irlp_flags<<<1, 64, 0, s>>>(a.flags, C);
    for (int j = 0; j < C && rc == DLP_OK; ++j) {
      switch (G) {
        case 1: rc = launch_irlp<1>(a, it + j, j, blocks, s); break;
        case 4: rc = launch_irlp<4>(a, it + j, j, blocks, s); break;
        case 8: rc = launch_irlp<8>(a, it + j, j, blocks, s); break;
        case 16: rc = launch_irlp<16>(a, it + j, j, blocks, s); break;
        case 32: rc = launch_irlp<32>(a, it + j, j, blocks, s); break;
        default: rc = launch_irlp<64>(a, it + j, j, blocks, s); break;
      }
    }
    if (rc != DLP_OK) break;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(hostFlags, a.flags, sizeof(int) * (size_t)(C + 1), hipMemcpyDeviceToHost,
                         s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      dlp_set_errorf_("IrLP: HIP error %d (%s)", (int)e, hipGetErrorString(e));
      rc = DLP_ERR_HIP;
      break;
    }
    for (int j = 0; j < C; ++j)
      if (hostFlags[j + 1] == 0) {  // iteration it+j changed nothing (validation.cu:192-196)
        used = it + j + 1;
        break;
      }
    it += C;
    chunk = std::min(2 * chunk, cfg.maxChunk);
  }
  if (rc == DLP_OK) {
    if (used < 0) used = maxIterations;
    /* the last iteration wrote lab1 when its index is even
     * (validation.cu:199-203) */
    if (used & 1) {
      hipError_t e = hipMemcpyAsync(a.lab0, a.lab1, sizeof(float) * (size_t)a.numNodes,
                                    hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        dlp_set_errorf_("IrLP: HIP error %d (%s)", (int)e, hipGetErrorString(e));
        rc = DLP_ERR_HIP;
      }
    }
    *iterationsUsed = used;
  }
  if (ev[0]) {
    if (rc == DLP_OK) {
      (void)hipEventRecord(ev[1], s);
      (void)hipEventSynchronize(ev[1]);
      (void)hipEventElapsedTime(kernelMs, ev[0], ev[1]);
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
  if (rc == DLP_OK) DLP_HIP(hipStreamSynchronize(s));
  return rc;
}
