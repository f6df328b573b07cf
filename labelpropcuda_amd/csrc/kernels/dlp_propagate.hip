/*
 * dlp_propagate.hip -- frontier propagation (the hot loop).
 *
 * Replaces, per iteration, propagateLabelsKernel + markNeighborsKernel +
 * the next iteration's collectAffectedKernel + the isAffected copy
 * (iterativePropagation.cu:63-217, :274-327) with work proportional to the
 * frontier, not to N:
 *
 *   for u in queue_k:
 *     W, S = weighted sums over row(u) reading F_old (buffer k & 1)
 *     F_new[u] = clamp(S / W)  (F_old[u] if W <= 1e-10f)   -> buffer (k+1) & 1
 *     if |F_new[u] - F_old[u]| > delta:
 *       every eligible neighbour not yet in queue_{k+1}, and u itself, is
 *       appended to queue_{k+1}.
 *
 * Deduplication is an atomicMax on a per-vertex iteration stamp that also
 * carries eligibility (INT_MAX = may never enter a frontier).  Row sums use
 * dlp::group_row_sums / lane_row_sums, bit-identical to the reference's
 * block reduction.
 *
 * Kernels implementing the same iteration:
 *   k_propagate        one launch per iteration, many work-groups, G lanes
 *                      per frontier vertex, queues / labels / stamps in HBM;
 *                      for large frontiers.
 *   k_propagate_local  one work-group that runs many iterations back to back
 *                      with the frontier's rows, their labels and the labels
 *                      of their neighbours staged in LDS; the limit cycles
 *                      and convergence tails of the reference's loop (most
 *                      iterations, with small frontiers) run without a
 *                      launch or a host round trip per iteration.  Two
 *                      variants: queue iterations (<..., false>) and owner
 *                      mode (<..., true>: one thread per core vertex, the
 *                      frontier as flags), which the host launches once a
 *                      queue launch has run LocalArgs::ownerAfter iterations.
 *   k_propagate_multi  one work-group per graph of a batched step (LDS
 *                      passes and work-group HBM iterations).
 */
#include "../dlp_kernels.h"
#include "../dlp_rowsum.h"

namespace {

/* ------------------------------------------------------------------------ */
/* multi-block iteration                                                    */
/* ------------------------------------------------------------------------ */
using dlp::GlobalPos;

constexpr int kPT = kPropagateThreads; /* threads of the multi-block kernel */
constexpr int kStageQ = 4096;  /* per-block staged next-frontier appends     */
constexpr int kStageT = 2048;  /* per-block staged touched-list appends      */

/* Block-staged append: lanes with `pred` get a slot in the block's LDS stage
 * (one LDS atomic per wave); entries past its capacity go straight to the
 * global list.  The stage is flushed with one global atomic per block, so a
 * launch issues about one same-address global atomic per block instead of
 * one per wave and mark round. */
template <int CAP>
__device__ __forceinline__ void stage_append(int *stage, int *stageCount, int *gCount,
                                             int *gList, bool pred, int v) {
  const unsigned long long m = dlp_ballot(pred);
  if (m == 0ull) return;
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (dlp_lane() == leader) base = atomicAdd(stageCount, __popcll(m));
  base = __shfl(base, leader, 64);
  const int idx = base + dlp_rank_below(m);
  const bool spill = pred && idx >= CAP;
  if (pred && !spill) stage[idx] = v;
  const int slot = dlp_wave_append(gCount, spill);
  if (slot >= 0) gList[slot] = v;
}

/* copies the block's stage [0, min(count, CAP)) to a range of the global
 * list reserved with one atomic; all threads of the block, after a barrier */
template <int CAP>
__device__ __forceinline__ void stage_flush(const int *stage, int count, int *gCount, int *gList,
                                            int *baseSlot) {
  const int nb = min(count, CAP);
  if (threadIdx.x == 0) *baseSlot = nb ? atomicAdd(gCount, nb) : 0;
  __syncthreads();
  const int b = *baseSlot;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) gList[b + i] = stage[i];
}

/* Claims vertex v (stamp word `stamp`) for the frontier of the iteration
 * whose stamp is stampNext and returns the stamp it held before; the caller
 * appends v when that is below stampNext.  With `filter`, an agent-scope
 * load first skips the atomic when v is ineligible (INT_MAX) or already
 * claimed in this iteration: on large frontiers a vertex is usually marked
 * by several neighbours and returning atomics are the throughput limit
 * (about half the rate of random loads); on small ones the extra round trip
 * only adds latency. */
__device__ __forceinline__ int claim_stamp(int *stamp, int stampNext, bool filter) {
  if (filter) {
    const int cur = __hip_atomic_load(stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur >= stampNext) return cur;
  }
  return atomicMax(stamp, stampNext);
}

/* frontier size from which k_propagate filters its claims */
constexpr int kClaimFilterMin = 16384;

/* hub rows (longer than PropArgs::blockRowMin) one work-group of the
 * multi-block kernel sums together after its lane-group and wave work; more
 * than this many in one work-group go to the wave path */
constexpr int kHubCap = 64;

/* HUB: the work-group tier for rows longer than blockRowMin is compiled in
 * (a separate instantiation, launched only when the tier is on, so the
 * common kernel carries none of its LDS or code) */
template <int G, bool HUB>
__global__ __launch_bounds__(kPT) void k_propagate(PropArgs a, int k) {
  constexpr int GPW = 64 / G; /* frontier vertices per wave */
  __shared__ int stageQ[kStageQ];
  __shared__ int stageT[kStageT];
  __shared__ int stageCount[2], stageBase[2];
  __shared__ int hubList[HUB ? kHubCap : 1];
  __shared__ int hubCount, hubChanged;
  __shared__ float hubW[HUB ? 256 : 1], hubS[HUB ? 256 : 1];
  DevCounters *c = a.ctr;
  const int *in = (k & 1) ? a.q1 : a.q0;
  const int lane = threadIdx.x & 63;
  const int gl = lane & (G - 1);
  const int grp = lane / G;
  const long long gwave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  /* the lane group's first queue entry, loaded together with the frontier
   * size instead of after it (one dependent round trip less); used only if
   * that entry is below the size */
  const long long item0 = gwave * GPW + grp;
  int u0 = 0;
  if (item0 < a.queueCap) u0 = in[item0];
  const int n = c->qcount[k % 3];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c->qcount[(k + 2) % 3] = 0; /* consumed by iteration k-1; next iteration appends here */
    if (k == 0) c->frontier0 = n;
    if (n > 0) {
      if (a.iterSizes && k < a.iterSizesCap) a.iterSizes[k] = n;
      c->iterUsed = k + 1;
      c->updates += (unsigned long long)n;
    }
  }
  if (n == 0) return;  // block-uniform
  if (threadIdx.x == 0) stageCount[0] = stageCount[1] = hubCount = 0;
  __syncthreads();

  const int stampBase = c->stampBase;
  const int stampNext = stampBase + k + 1;
  int *out = (k & 1) ? a.q0 : a.q1;
  int *outCount = &c->qcount[(k + 1) % 3];
  const float *Fold = (k & 1) ? a.lab1 : a.lab0;
  float *Fnew = (k & 1) ? a.lab0 : a.lab1;
  unsigned profE = 0, profC = 0, profA = 0; /* lane-local work counters */

  for (long long base = gwave * GPW; base < n; base += nwaves * GPW) {  // wave-uniform
    const long long item = base + grp;
    bool act = item < n;
    int u = 0, beg = 0, len = 0;
    float oldLabel = 0.0f;
    if (act) {
      u = (item == item0 && a.queueCap > 0) ? u0 : in[item];
      if (gl == 0) oldLabel = Fold[u];  // in flight with the row loads
      if (a.rowInfo) {
        const int2 r = a.rowInfo[u];
        beg = r.x;
        len = r.y;
      } else {
        beg = a.rowBeg[u];
        len = a.rowLen ? a.rowLen[u] : a.rowEnd[u] - beg;
      }
    }
    /* hub rows (longer than blockRowMin) wait for the whole work-group
     * after this loop; a full hub list leaves them to the wave path.  Off
     * (blockRowMin 0, uniform) it costs nothing: the group learns its
     * leader's decision by a ballot, not a shuffle in the row's chain. */
    if (HUB) {
      bool deferred = false;
      if (act && gl == 0 && len > a.blockRowMin) {
        const int slot = atomicAdd(&hubCount, 1);
        if (slot < kHubCap) {
          hubList[slot] = u;
          deferred = true;
        }
      }
      if ((dlp_ballot(deferred) >> (grp * G)) & 1ull) {
        act = false;
        len = 0;
      }
    }
    /* rows of up to waveRowMin (<= 64) entries by the lane group; longer
     * ones (rare at the degrees of these graphs) by the whole wave, one
     * after another, so the group path needs no registers for the 256-slot
     * form and the kernel keeps a high occupancy */
    const bool longRow = act && len > a.waveRowMin;
    /* the lane's first column, read by the sum below and kept for the marks */
    const int col0 = (act && !longRow && gl < len) ? a.colInd[beg + gl] : -1;
    float W, S;
    dlp::group_short_row_sums<G>(gl, longRow ? 0 : len,
                                 GlobalPos{a.colInd, a.values, Fold, a.status, beg}, W, S);
    unsigned long long longMask = dlp_ballot(longRow && gl == 0);
    while (longMask) {  // wave-uniform
      const int src = __ffsll((long long)longMask) - 1;
      longMask &= longMask - 1;
      const int lbeg = __shfl(beg, src, 64), llen = __shfl(len, src, 64);
      float LW, LS;
      dlp::group_row_sums<64>(lane, llen, GlobalPos{a.colInd, a.values, Fold, a.status, lbeg},
                              LW, LS);
      LW = __shfl(LW, 0, 64);
      LS = __shfl(LS, 0, 64);
      if (lane == src) {
        W = LW;
        S = LS;
      }
    }
    int changed = 0;
    if (act && gl == 0) {
      const float nl = dlp::new_label(W, S, oldLabel);
      Fnew[u] = nl;
      changed = fabsf(nl - oldLabel) > a.delta;
      profE += (unsigned)len;
      if (changed) profC += (unsigned)len + 1u;
    }
    changed = __shfl(changed, grp * G, 64);

    /* next frontier: neighbours of changed vertices at row positions
     * 0..len-1, and the vertex itself as position len */
    for (int p0 = 0; __any(changed && p0 <= len); p0 += G) {  // wave-uniform
      const int p = p0 + gl;
      int v = -1;
      bool pred = false;
      if (changed && p < len) {
        v = (p0 == 0 && !longRow) ? col0 : a.colInd[beg + p];
        pred = v >= 0;
      } else if (changed && p == len) {
        v = u;
        pred = true;
      }
      const int old = pred ? claim_stamp(&a.iterStamp[v], stampNext, n >= kClaimFilterMin)
                           : DLP_INT_MAX;
      const bool app = pred && old < stampNext;
      profA += app ? 1u : 0u;
      stage_append<kStageQ>(stageQ, &stageCount[0], outCount, out, app, v);
      if (a.touched)
        stage_append<kStageT>(stageT, &stageCount[1], &c->touched, a.touched,
                              app && old < stampBase, v);
    }
  }
  __syncthreads();
  /* hub rows: the whole work-group sums each (the reference's 256-slot
   * order) and claims its neighbours */
  const int nHub = HUB ? min(hubCount, kHubCap) : 0;
  for (int h = 0; h < nHub; ++h) {  // block-uniform
    const int u = hubList[h];
    int beg, len;
    if (a.rowInfo) {
      const int2 r = a.rowInfo[u];
      beg = r.x;
      len = r.y;
    } else {
      beg = a.rowBeg[u];
      len = a.rowLen ? a.rowLen[u] : a.rowEnd[u] - beg;
    }
    float W = 0.0f, S = 0.0f;
    if constexpr (HUB)
      dlp::block_row_sums(len, GlobalPos{a.colInd, a.values, Fold, a.status, beg}, hubW, hubS, W,
                          S);
    if (threadIdx.x == 0) {
      const float oldLabel = Fold[u];
      const float nl = dlp::new_label(W, S, oldLabel);
      Fnew[u] = nl;
      hubChanged = fabsf(nl - oldLabel) > a.delta;
      profE += (unsigned)len;
      if (hubChanged) profC += (unsigned)len + 1u;
    }
    __syncthreads();
    if (hubChanged) {  // block-uniform
      for (int p0 = 0; p0 <= len; p0 += kPT) {  // block-uniform
        const int p = p0 + (int)threadIdx.x;
        int v = -1;
        bool pred = false;
        if (p < len) {
          v = a.colInd[beg + p];
          pred = v >= 0 && !(a.status && a.status[v] == DLP_VERTEX_DELETED);
        } else if (p == len) {
          v = u;
          pred = true;
        }
        const int old = pred ? claim_stamp(&a.iterStamp[v], stampNext, n >= kClaimFilterMin)
                             : DLP_INT_MAX;
        const bool app = pred && old < stampNext;
        profA += app ? 1u : 0u;
        stage_append<kStageQ>(stageQ, &stageCount[0], outCount, out, app, v);
        if (a.touched)
          stage_append<kStageT>(stageT, &stageCount[1], &c->touched, a.touched,
                                app && old < stampBase, v);
      }
    }
    __syncthreads();  // hubChanged is read by every thread before the next hub
  }
  stage_flush<kStageQ>(stageQ, stageCount[0], outCount, out, &stageBase[0]);
  if (a.touched) stage_flush<kStageT>(stageT, stageCount[1], &c->touched, a.touched, &stageBase[1]);
  if (a.profile) {  // LDS sums per block, then one global atomic per counter per block
    __shared__ unsigned long long blockProf[3];
    if (threadIdx.x < 3) blockProf[threadIdx.x] = 0ull;
    __syncthreads();
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      profE += __shfl_xor(profE, off, 64);
      profC += __shfl_xor(profC, off, 64);
      profA += __shfl_xor(profA, off, 64);
    }
    if (lane == 0) {
      atomicAdd(&blockProf[0], (unsigned long long)profE);
      atomicAdd(&blockProf[1], (unsigned long long)profC);
      atomicAdd(&blockProf[2], (unsigned long long)profA);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (blockProf[0]) atomicAdd(&c->profEdges, blockProf[0]);
      if (blockProf[1]) atomicAdd(&c->profChangedEdges, blockProf[1]);
      if (blockProf[2]) atomicAdd(&c->profAppended, blockProf[2]);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* LDS-resident iterations (one work-group)                                 */
/* ------------------------------------------------------------------------ */
/* Work-group shapes of the local kernel: LT threads; rows of up to SR
 * entries are stored padded to SR records, so one lane reads them without
 * per-entry masks.  Picked per context (tuning). */
struct LocalShape {
  int threads, stride;
};
constexpr LocalShape kLocalShapes[] = {{1024, 8}, {1024, 16}, {512, 8}, {256, 8}, {1024, 4}};
constexpr int kNumLocalShapes = sizeof(kLocalShapes) / sizeof(kLocalShapes[0]);

/* Rows up to this length are summed by one lane (canonical tree over 16
 * slots); longer rows by a wave. */
constexpr int kLaneRowMax = 16;

/* The working set of a pass.  Label slots: [0, nCore) core vertices,
 * [nCore, nCore + nHE) eligible vertices outside the core (halo), then the
 * ineligible halo, then one null slot whose labels are 0.  An LDS row entry
 * is a record {slot, weight bits}; holes and entries pointing at deleted
 * vertices are {null, 0}, so every entry is summed as w * F[slot] with no
 * branch (0 * 0 adds +0.0, which changes no partial sum).  Halo labels do
 * not change while a vertex is outside the core. */
struct LocalPos {
  const int2 *rec;
  const float *Fo;
  int rs;
  __device__ void operator()(int p, float &w, float &s) const {
    const int2 r = rec[rs + p];
    w = __int_as_float(r.y);
    s = w * Fo[r.x];
  }
};

__device__ __forceinline__ int ld_agent(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* global writes of this work-group become visible to its own later loads */
__device__ __forceinline__ void wg_global_sync() {
  __threadfence();
  __syncthreads();
  __threadfence();
}

/* exclusive scan of a[0..n) in LDS into out (may alias a), returns the
 * total; all LT threads */
template <int LT>
__device__ int block_scan_lds(const int *a, int *out, int n, int *tmp /* >= 17 ints */) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int per = (n + LT - 1) / LT;
  const int lo = min(tid * per, n), hi = min(lo + per, n);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += a[i];
  int v = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) tmp[wave] = v;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int w = 0; w < LT / 64; ++w) {
      int t = tmp[w];
      tmp[w] = run;
      run += t;
    }
    tmp[16] = run;
  }
  __syncthreads();
  int run = tmp[wave] + v - s;
  for (int i = lo; i < hi; ++i) {
    int x = a[i];
    out[i] = run;
    run += x;
  }
  const int total = tmp[16];
  __syncthreads();
  return total;
}

enum {
  SH_GEN, SH_CC, SH_NCAND, SH_NCORE, SH_BIG, SH_REBUILD, SH_NOVF, SH_HASLONG,
  SH_NHE, SH_NHI, SH_QC0, SH_QC1, SH_QC2, SH_NLONG0, SH_NLONG1, SH_INELIG, SH_MAXLEN,
  SH_COUNT
};

/* state of one local iteration shared by its row paths */
struct LocalIter {
  int *stampL;
  int *qout;
  int *qcOut;
  int *sh;
  const int *haloE;  /* gid of each eligible halo slot (slot - nCore) */
  int *dummyStamp;   /* 64 words, INT_MAX: the mark target of idle lanes */
  int *dummyOut;     /* 64 words: the queue write target of idle lanes   */
  int nCore, nHE;
  int stampNextL, stampNextG, stampBase;
};

/* an eligible vertex outside the core was marked: global stamp, overflow
 * list (it joins the core at the rebuild after this iteration), touched */
__device__ __forceinline__ unsigned mark_outside(const PropArgs &a, const LocalArgs &l,
                                                 const LocalIter &it, int gid) {
  const int old = atomicMax(&a.iterStamp[gid], it.stampNextG);
  if (old >= it.stampNextG) return 0u;
  const int os = atomicAdd(&it.sh[SH_NOVF], 1);
  if (os < l.ovfCap) l.ovf[os] = gid;
  it.sh[SH_REBUILD] = 1;
  if (a.touched && old < it.stampBase) a.touched[atomicAdd(&a.ctr->touched, 1)] = gid;
  return 1u;
}

/* One frontier vertex per lane (all 64 lanes call it; `act` lanes hold a
 * vertex whose row has at most L entries).  The row's records start at
 * `row`; without MASKED every position 0..L-1 holds a record (padding is
 * {null, 0}), with MASKED positions >= len are skipped.  Marks and queue
 * writes are branch-free: a lane with nothing to mark or write at a
 * position targets its own dummy LDS word (a stamp that stays INT_MAX, a
 * scratch slot), so a lane waits for a few LDS round trips, not one per
 * entry, and no exec-mask juggling is needed. */
template <int L, bool MASKED>
__device__ __forceinline__ void lane_rows(const PropArgs &a, const LocalArgs &l,
                                          const LocalIter &it, const int2 *row, int nullSlot,
                                          const float *Fo, float *Fn, bool act, int ci, int len,
                                          bool hasOut, unsigned &profE, unsigned &profC,
                                          unsigned &profA) {
  static_assert(L + 1 <= 32, "hit mask is 32 bits");
  int tt[L];
  float ww[L], ff[L];
#pragma unroll
  for (int j = 0; j < L; ++j) {
    int2 r = make_int2(nullSlot, 0);
    if (!MASKED || j < len) r = row[j];
    tt[j] = r.x;
    ww[j] = __int_as_float(r.y);
  }
#pragma unroll
  for (int j = 0; j < L; ++j) ff[j] = Fo[tt[j]];
  const float oldLabel = Fo[ci];
  float W, S;
  dlp::lane_row_sums<L>(
      L,
      [&](int j, float &w, float &s) {
        w = ww[j];
        s = ww[j] * ff[j];
      },
      W, S);
  const float nl = dlp::new_label(W, S, oldLabel);
  const bool changed = act && fabsf(nl - oldLabel) > a.delta;
  if (act) {
    Fn[ci] = nl;
    profE += (unsigned)len;
    if (changed) profC += (unsigned)len + 1u;
  }
  /* marks: the row's core targets (bits 0..L-1) and the vertex itself (bit
   * L) */
  const int lane = dlp_lane();
  int *dumStamp = it.dummyStamp + lane;
  const int nCore = it.nCore, next = it.stampNextL;
  int olds[L + 1];
#pragma unroll
  for (int j = 0; j < L; ++j)
    olds[j] = atomicMax((changed && tt[j] < nCore) ? &it.stampL[tt[j]] : dumStamp, next);
  olds[L] = atomicMax(changed ? &it.stampL[ci] : dumStamp, next);
  unsigned hit = 0u;
#pragma unroll
  for (int j = 0; j <= L; ++j) hit |= (olds[j] < next ? 1u : 0u) << j;
  /* eligible vertices outside the core (rows flagged at the build) */
  if (__any(changed && hasOut)) {  // wave-uniform
    if (changed && hasOut) {
#pragma unroll
      for (int j = 0; j < L; ++j) {
        const int h = tt[j] - nCore;
        if ((unsigned)h < (unsigned)it.nHE) profA += mark_outside(a, l, it, it.haloE[h]);
      }
    }
  }
  /* one LDS atomic per wave reserves the wave's core appends */
  const int cnt = __popc(hit);
  profA += (unsigned)cnt;
  const int incl = dlp_wave_inclusive_sum(cnt);
  const int total = __builtin_amdgcn_readlane(incl, 63);
  if (total > 0) {  // wave-uniform
    int wbase = 0;
    if (lane == 63) wbase = atomicAdd(it.qcOut, total);
    wbase = __builtin_amdgcn_readlane(wbase, 63);
    int pos = wbase + incl - cnt;
    int *dumOut = it.dummyOut + lane;
#pragma unroll
    for (int j = 0; j <= L; ++j) {
      const bool h = (hit >> j) & 1u;
      *(h ? &it.qout[pos] : dumOut) = j < L ? tt[j < L ? j : 0] : ci;
      pos += h ? 1 : 0;
    }
  }
}

/* rlL word of a core vertex: row length, and a flag for rows that name an
 * eligible vertex outside the core */
constexpr int kRowLenMask = 0x3fffffff;
constexpr int kRowHasOut = 0x40000000;

/* LDS words past the working set (rounding of the aligned sections) */
constexpr int kLdsSlack = 64;

/* most waves of a register-mask owner work-group (512 threads) */
constexpr int kMaskWavesMax = 8;

/* static LDS of the local passes */
struct LocalShared {
  int sh[SH_COUNT];
  int misc[4];        /* frontier size broadcast of the work-group drivers */
  int tmp[17];
  int dummyStamp[64]; /* INT_MAX, set by the kernel before the first pass */
  int dummyOut[64];
  int ownerCount;     /* queue rebuilt from the owner-mode flags */
  unsigned long long ownerSum; /* frontier sizes of the owner-mode iterations */
  /* register-mask owner mode with several waves: per wave and iteration
   * parity, the changed-vertex mask, whether a mark left the core, and the
   * size of the wave's next frontier */
  unsigned long long xChanged[2][kMaskWavesMax];
  unsigned char xOut[2][kMaskWavesMax];
  unsigned char xCount[2][kMaskWavesMax]; /* 0 .. 64 */
};

/* lane-local work counters of the local passes */
struct LocalProf {
  unsigned U = 0, E = 0, C = 0, A = 0, D = 0; /* D: updates run in owner mode */
  unsigned Wv = 0;                            /* of D: in the single-wave owner mode */
};

/* One owned vertex per lane in owner mode (all 64 lanes call it; `act`
 * lanes hold a flagged vertex whose row has at most L entries, stored from
 * `row`, padded to L records unless MASKED; `row` points at a stored row
 * for every lane).  The label arithmetic is lane_rows'; the next frontier is
 * set with plain flag stores (every core vertex is eligible in owner mode,
 * and so is the vertex itself).  Branch-free like lane_rows: loads come from
 * stored rows only (a masked position past the row reads the row's first
 * record and is then treated as padding), and a lane with nothing to store
 * at a step stores to its dummy LDS word. */
template <int L, bool MASKED>
__device__ __forceinline__ void owner_rows(const PropArgs &a, const LocalArgs &l,
                                           const LocalIter &it, const int2 *row, int nullSlot,
                                           const float *Fo, float *Fn, int *nxt, int *dumOut,
                                           bool act, int ci, int len, bool hasOut,
                                           float oldLabel, float &nl, bool &changed,
                                           LocalProf &prof) {
  /* all record loads, then all label loads, then the sums: two LDS round
   * trips for the row (a weight sum taken once per pass, as the mask
   * kernels do, costs this kernel more spilled registers than it saves) */
  int tt[L];
  float ww[L], ff[L];
#pragma unroll
  for (int j = 0; j < L; ++j) {
    const bool in = act && (!MASKED || j < len);
    const int2 r = row[(!MASKED || in) ? j : 0];
    tt[j] = in ? r.x : nullSlot;
    ww[j] = in ? __int_as_float(r.y) : 0.0f;
  }
#pragma unroll
  for (int j = 0; j < L; ++j) ff[j] = Fo[tt[j]];
  float W, S;
  dlp::lane_row_sums<L>(
      L,
      [&](int j, float &w, float &s) {
        w = ww[j];
        s = ww[j] * ff[j];
      },
      W, S);
  nl = dlp::new_label(W, S, oldLabel);
  changed = act && fabsf(nl - oldLabel) > a.delta;
  *(act ? &Fn[ci] : (float *)dumOut) = nl;
  prof.E += act ? (unsigned)len : 0u;
  prof.C += changed ? (unsigned)len + 1u : 0u;
#pragma unroll
  for (int j = 0; j < L; ++j) *((changed && tt[j] < it.nCore) ? &nxt[tt[j]] : dumOut) = 1;
  *(changed ? &nxt[ci] : dumOut) = 1;
  if (__any(changed && hasOut)) {  // wave-uniform
    if (changed && hasOut) {
#pragma unroll
      for (int j = 0; j < L; ++j) {
        const int h = tt[j] - it.nCore;
        if ((unsigned)h < (unsigned)it.nHE) prof.A += mark_outside(a, l, it, it.haloE[h]);
      }
    }
  }
}

/* Owner mode of the LDS-resident iterations, for a working set whose core
 * has at most V vertices per thread and holds eligible vertices only (as it
 * does when it is made of frontier vertices).  Thread tid owns core vertices
 * tid + v*LT and keeps in registers their row length and place and both
 * label slots; the frontier is a flag
 * word per core vertex (the two queue arrays serve as the flags of
 * iterations k and k+1).  An iteration reads the owned flags, loads the rows
 * and neighbour labels of flagged vertices, writes the new labels and, for
 * changed vertices, sets the next flags with plain stores (a vertex flagged
 * by several neighbours is still one vertex); rows longer than kLaneRowMax
 * are done by the whole wave, one after another.  The frontier size is a
 * per-wave count added before the iteration's one barrier.  No returning
 * atomics, appends or frontier-list loads: a few LDS round trips per
 * iteration.  Iterations and per-vertex arithmetic are the queue
 * iterations' (same sums, same Jacobi buffers), so labels, frontiers and
 * counts are identical.  On exit the flags of iteration k become its queue
 * again and the stamps of the vertices that entered a frontier in the pass
 * are set, as the queue iterations leave them.  Returns the DLP_LOCAL_*
 * reason, or 0 when an eligible vertex outside the core was marked
 * (rebuild). */
template <int LT, int SR, int V>
__device__ __forceinline__ int owner_iterations(
    const PropArgs &a, const LocalArgs &l, LocalShared &S, int &k, int maxIter, int kBuild,
    int nCore, int nHE, int nullSlot, int stampBase, const int *rlL, const int *rsL,
    const int2 *recS, const int2 *recL, int *stampL, int *q0L, int *q1L, float *lab0L,
    float *lab1L, const int *haloE, unsigned long long &passUpdates, LocalProf &prof) {
  constexpr int RS = SR + 1;
  int *sh = S.sh;
  DevCounters *c = a.ctr;
  const int tid = threadIdx.x, lane = tid & 63;
  int *dumOut = S.dummyOut + lane;
  /* the queue of iteration k becomes flags */
  int qv[V];
  const int nq = sh[SH_QC0 + k % 3];
  {
    const int *qin = (k & 1) ? q1L : q0L;
#pragma unroll
    for (int v = 0; v < V; ++v) qv[v] = tid + v * LT < nq ? qin[tid + v * LT] : -1;
  }
  __syncthreads();  // the queue and its count are read
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int ci = tid + v * LT;
    if (ci < nCore) {
      q0L[ci] = 0;
      q1L[ci] = 0;
    }
  }
  if (tid == 0) {
    sh[SH_QC0] = sh[SH_QC0 + 1] = sh[SH_QC0 + 2] = 0;
    S.ownerSum = 0ull;
  }
  __syncthreads();
  {
    int *cur = (k & 1) ? q1L : q0L;
#pragma unroll
    for (int v = 0; v < V; ++v)
      if (qv[v] >= 0) cur[qv[v]] = 1;
  }
  /* the owned vertices: row length and record offset (in recS for rows of
   * at most SR entries, else in recL), label slots */
  float l0[V], l1[V];
  int len[V], ro[V];
  bool own[V], hasOut[V], entered[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int ci = tid + v * LT;
    own[v] = ci < nCore;
    len[v] = ro[v] = 0;
    hasOut[v] = entered[v] = false;
    l0[v] = l1[v] = 0.0f;
    if (own[v]) {
      const int rl = rlL[ci];
      len[v] = rl & kRowLenMask;
      hasOut[v] = (rl & kRowHasOut) != 0;
      ro[v] = len[v] <= SR ? ci * RS : rsL[ci];
      l0[v] = lab0L[ci];
      l1[v] = lab1L[ci];
    }
  }
  __syncthreads();  // flags of iteration k are set

  int reason = 0;
  while (true) {
    const int kb = k & 1;
    int *cur = kb ? q1L : q0L;
    int *nxt = kb ? q0L : q1L;
    const float *Fo = kb ? lab1L : lab0L;
    float *Fn = kb ? lab0L : lab1L;
    bool f[V];
    int fl[V];
    int cnt = 0;
#pragma unroll
    for (int v = 0; v < V; ++v) fl[v] = cur[own[v] ? tid + v * LT : 0];  // loads in flight together
#pragma unroll
    for (int v = 0; v < V; ++v) {
      f[v] = own[v] && fl[v] != 0;
      cnt += __popcll(dlp_ballot(f[v]));
    }
    if (lane == 0 && cnt) atomicAdd(&sh[SH_QC0 + k % 3], cnt);
    if (k >= maxIter) {  // uniform: the size of frontier k decides the reason
      __syncthreads();
      reason = sh[SH_QC0 + k % 3] ? DLP_LOCAL_CAPPED : DLP_LOCAL_CONVERGED;
      break;
    }
    /* slot (k+1) % 3 was last read after the barrier of iteration k-2 */
    if (tid == 0) sh[SH_QC0 + (k + 1) % 3] = 0;
    const LocalIter it{stampL, nullptr, nullptr, sh, haloE, nullptr, dumOut, nCore, nHE,
                       k + 1, stampBase + k + 1, stampBase};
#pragma unroll
    for (int v = 0; v < V; ++v) {
      if (!__any(f[v])) continue;  // wave-uniform
      const int ci = own[v] ? tid + v * LT : 0;
      const float oldLabel = kb ? l1[v] : l0[v];
      /* rows up to kLaneRowMax: one lane per vertex */
      const bool lanePath = f[v] && len[v] <= kLaneRowMax;
      if (__any(lanePath)) {  // wave-uniform
        float nl = 0.0f;
        bool changed = false;
        /* a stored row for every lane: the lane's own short row, else row 0 */
        const int2 *rowS = recS + (len[v] <= SR ? ro[v] : 0);
        if (!__any(lanePath && len[v] > 4))
          owner_rows<(SR > 4 ? 4 : SR), false>(a, l, it, rowS, nullSlot, Fo, Fn, nxt, dumOut,
                                               lanePath, ci, len[v], hasOut[v], oldLabel,
                                               nl, changed, prof);
        else if (!__any(lanePath && len[v] > SR))
          owner_rows<SR, false>(a, l, it, rowS, nullSlot, Fo, Fn, nxt, dumOut, lanePath, ci,
                                len[v], hasOut[v], oldLabel, nl, changed, prof);
        else
          owner_rows<kLaneRowMax, true>(a, l, it, len[v] <= SR ? rowS : recL + ro[v], nullSlot,
                                        Fo, Fn, nxt, dumOut, lanePath, ci, len[v], hasOut[v],
                                        oldLabel, nl, changed, prof);
        l0[v] = (lanePath && kb) ? nl : l0[v];
        l1[v] = (lanePath && !kb) ? nl : l1[v];
      }
      /* longer rows: the wave takes them one by one */
      unsigned long long longMask = dlp_ballot(f[v] && len[v] > kLaneRowMax);
      while (longMask) {  // wave-uniform
        const int src = __ffsll((long long)longMask) - 1;
        longMask &= longMask - 1;
        const int lci = __shfl(ci, src, 64), llen = __shfl(len[v], src, 64),
                  lro = __shfl(ro[v], src, 64);
        const float lold = __shfl(oldLabel, src, 64);
        float W, Sm;
        dlp::group_row_sums<64>(lane, llen, LocalPos{recL, Fo, lro}, W, Sm);
        float lnl = 0.0f;
        int changed = 0;
        if (lane == 0) {
          lnl = dlp::new_label(W, Sm, lold);
          Fn[lci] = lnl;
          changed = fabsf(lnl - lold) > a.delta;
          prof.E += (unsigned)llen;
          if (changed) prof.C += (unsigned)llen + 1u;
        }
        lnl = __shfl(lnl, 0, 64);
        changed = __shfl(changed, 0, 64);
        if (lane == src) {
          if (kb)
            l0[v] = lnl;
          else
            l1[v] = lnl;
        }
        if (!changed) continue;  // wave-uniform
        for (int p0 = 0; p0 <= llen; p0 += 64) {  // wave-uniform
          const int p = p0 + lane;
          if (p < llen) {
            const int t = recL[lro + p].x;
            if (t < nCore) {
              nxt[t] = 1;
            } else if (t < nCore + nHE) {
              prof.A += mark_outside(a, l, it, haloE[t - nCore]);
            }
          } else if (p == llen) {
            nxt[lci] = 1;  // itself: flagged, hence eligible
          }
        }
      }
      *(f[v] ? &cur[ci] : dumOut) = 0;
      entered[v] = entered[v] || (f[v] && k > kBuild);
    }
    __syncthreads();
    const int n = sh[SH_QC0 + k % 3];
    if (n == 0) {  // uniform
      reason = DLP_LOCAL_CONVERGED;
      break;
    }
    if (tid == 0) {  // sums kept in LDS: no registers live across the loop
      if (k == 0) c->frontier0 = n;
      if (a.iterSizes && k < a.iterSizesCap) a.iterSizes[k] = n;
      atomicAdd(&S.ownerSum, (unsigned long long)n);
    }
    ++k;
    if (sh[SH_REBUILD]) break;  // uniform after the barrier
  }

  /* flags of iteration k -> its queue; stamps of the vertices that entered */
  {
    int *cur = (k & 1) ? q1L : q0L;
    bool f[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      f[v] = own[v] && cur[tid + v * LT] != 0;
      entered[v] = entered[v] || (f[v] && k > kBuild);
    }
    if (tid == 0) S.ownerCount = 0;
    __syncthreads();  // every flag is read
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const unsigned long long m = dlp_ballot(f[v]);
      int base = 0;
      if (lane == 0 && m) base = atomicAdd(&S.ownerCount, __popcll(m));
      base = __shfl(base, 0, 64);
      if (f[v]) cur[base + dlp_rank_below(m)] = tid + v * LT;
      if (entered[v]) stampL[tid + v * LT] = k;  // k > kBuild
    }
    __syncthreads();
    if (tid == 0) {
      const int nOut = S.ownerCount;
      sh[SH_QC0 + k % 3] = nOut;
      /* when owner iterations ran, the first processed the nq entries of the
       * queue it started from (counted where they were appended, or not
       * appended in this pass); every later frontier and the one left for
       * iteration k were appended here */
      const unsigned long long sum = S.ownerSum;
      passUpdates += sum;
      prof.U += (unsigned)sum;
      prof.D += (unsigned)sum;
      if (sum > 0ull) prof.A += (unsigned)(sum - (unsigned long long)nq) + (unsigned)nOut;
    }
    __syncthreads();
  }
  return reason;
}

/* the OR of a 64-bit mask over the 64 lanes of the wave, in every lane */
__device__ __forceinline__ unsigned long long wave_or64(unsigned long long m) {
  unsigned lo = (unsigned)m, hi = (unsigned)(m >> 32);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    lo |= (unsigned)__shfl_xor((int)lo, off, 64);
    hi |= (unsigned)__shfl_xor((int)hi, off, 64);
  }
  return ((unsigned long long)hi << 32) | lo;
}

/* New label of a register-held row (its first LL <= L records used; rows
 * are padded with null records of weight 0) whose weight sum W was taken
 * once per pass (the same tree over the same weights): the neighbour labels
 * come from the LDS label array Fo.  Lanes without a flagged vertex read
 * the null slot (one address: a broadcast, no bank traffic) and their
 * result is not used. */
template <int LL, int L>
__device__ __forceinline__ float mask_row_label(const int (&src)[L], const float (&w)[L],
                                                float W, const float *Fo, int nullSlot, bool f,
                                                float old) {
  float ff[LL];
#pragma unroll
  for (int p = 0; p < LL; ++p) ff[p] = Fo[f ? src[p] : nullSlot];
  const float Sm = dlp::lane_row_sum<LL>(LL, [&](int p) { return w[p] * ff[p]; });
  return dlp::new_label(W, Sm, old);
}

/* the weight sum of a register-held row of at most L records, in the
 * reference's order */
template <int L>
__device__ __forceinline__ float mask_row_weight(const float (&w)[L]) {
  return dlp::lane_row_sum<L>(L, [&](int p) { return w[p]; });
}

/* mask_row_label with the narrowest row width that covers the wave's
 * flagged rows (wave-uniform choice); all lanes of the wave call it */
template <int L>
__device__ __forceinline__ float mask_row_label_any(const int (&src)[L], const float (&w)[L],
                                                    float W, const float *Fo, int nullSlot,
                                                    bool f, int len, float old) {
  static_assert(L == 8 || L == 16, "register rows of 8 or 16 records");
  if (!__any(f && len > 4)) return mask_row_label<4, L>(src, w, W, Fo, nullSlot, f, old);
  if (L == 8 || !__any(f && len > 8))
    return mask_row_label<8, L>(src, w, W, Fo, nullSlot, f, old);
  return mask_row_label<L, L>(src, w, W, Fo, nullSlot, f, old);
}

/* Single-wave owner mode: the work-group is one wave (64 threads) and the
 * core at most 64 * V vertices, every core row at most L <= SR entries
 * (stored padded in recS), every core vertex eligible.  Lane `lane` owns
 * core vertices lane + 64 r (r < V) and keeps in registers their row
 * records, both labels, and for each of them the core vertices its row
 * names, as V 64-bit masks.  The frontier of an iteration is V 64-bit masks
 * (uniform).  The next one is the changed vertices plus every core vertex
 * some changed vertex's row names: on a symmetric CSR (PropArgs::symmetric:
 * a row names v iff v's row names it) that is a vertex whose own row mask
 * meets the changed mask -- register operations only; otherwise the
 * changed vertices' row masks are OR-reduced over the wave.  No flag stores, no atomics, no barrier between
 * waves: an iteration is a short chain of register and LDS operations.
 * New labels go to the LDS label array, where neighbours read them, and to
 * the registers.  Iterations, arithmetic and the state left behind (labels,
 * the queue of iteration k and its count, stamps of the vertices that
 * entered a frontier, the counters) are those of owner_iterations; marks of
 * eligible vertices outside the core go through mark_outside and end the
 * pass (rebuild).  Returns the DLP_LOCAL_* reason or 0 (rebuild). */
template <int SR, int V, int L>
__device__ __forceinline__ int wave_owner_iterations(
    const PropArgs &a, const LocalArgs &l, LocalShared &S, int &k, int maxIter, int kBuild,
    int nCore, int nHE, int nullSlot, int stampBase, const int *rlL, const int2 *recS,
    int *stampL, int *q0L, int *q1L, float *lab0L, float *lab1L, const int *haloE,
    unsigned long long &passUpdates, LocalProf &prof) {
  static_assert(L <= SR, "rows are read from the padded short-row records");
  constexpr int RS = SR + 1;
  int *sh = S.sh;
  DevCounters *c = a.ctr;
  const int lane = threadIdx.x;  // the work-group is one wave
  int *dumOut = S.dummyOut + lane;
  const bool symmetric = a.symmetric != 0;
  /* frontier k: the queue's core indices as flags in the other queue
   * array, then as masks */
  const int nq = sh[SH_QC0 + k % 3];
  {
    const int *qin = (k & 1) ? q1L : q0L;
    int *flags = (k & 1) ? q0L : q1L;
    for (int i = lane; i < nCore; i += 64) flags[i] = 0;
    __syncthreads();
    for (int i = lane; i < nq; i += 64) flags[qin[i]] = 1;
    __syncthreads();
  }
  const int *flags0 = (k & 1) ? q0L : q1L;
  unsigned long long fm[V], ownMask[V];
  int src[V][L];
  float w[V][L];
  unsigned long long out[V][V]; /* out[r][q]: core vertices q*64+j named by row r */
  float l0[V], l1[V];
  int len[V];
  unsigned outPos[V]; /* row positions naming eligible vertices outside the core */
  float Wr[V];        /* weight sums of the rows */
  bool entered[V];
#pragma unroll
  for (int r = 0; r < V; ++r) {
    const int ci = lane + 64 * r;
    const bool own = ci < nCore;
    ownMask[r] = dlp_ballot(own);
    entered[r] = false;
    fm[r] = dlp_ballot(own && flags0[ci] != 0);
    len[r] = own ? (rlL[ci] & kRowLenMask) : 0;
    l0[r] = own ? lab0L[ci] : 0.0f;
    l1[r] = own ? lab1L[ci] : 0.0f;
    outPos[r] = 0u;
#pragma unroll
    for (int q = 0; q < V; ++q) out[r][q] = 0ull;
#pragma unroll
    for (int p = 0; p < L; ++p) {
      const int2 rec = own ? recS[ci * RS + p] : make_int2(nullSlot, 0);
      src[r][p] = rec.x;
      w[r][p] = __int_as_float(rec.y);
      if (rec.x < nCore) {
#pragma unroll
        for (int q = 0; q < V; ++q)
          if ((rec.x >> 6) == q) out[r][q] |= 1ull << (rec.x & 63);
      } else if (rec.x < nCore + nHE) {
        outPos[r] |= 1u << p;
      }
    }
    Wr[r] = mask_row_weight<L>(w[r]);
  }
  __syncthreads();  // the flags are read

  int reason = 0;
  unsigned long long ownerSum = 0ull;
  while (true) {
    int n = 0;
#pragma unroll
    for (int r = 0; r < V; ++r) n += __popcll(fm[r]);
    if (k >= maxIter) {  // uniform: the size of frontier k decides the reason
      reason = n ? DLP_LOCAL_CAPPED : DLP_LOCAL_CONVERGED;
      break;
    }
    if (n == 0) {
      reason = DLP_LOCAL_CONVERGED;
      break;
    }
    const int kb = k & 1;
    const float *Fo = kb ? lab1L : lab0L;
    float *Fn = kb ? lab0L : lab1L;
    unsigned long long cm[V];
    bool ch[V], anyOut = false;
#pragma unroll
    for (int r = 0; r < V; ++r) {
      const bool f = (fm[r] >> lane) & 1ull;
      const float old = kb ? l1[r] : l0[r];
      const float nl = mask_row_label_any<L>(src[r], w[r], Wr[r], Fo, nullSlot, f, len[r], old);
      ch[r] = f && fabsf(nl - old) > a.delta;
      *(f ? &Fn[lane + 64 * r] : (float *)dumOut) = nl;
      l0[r] = (f && kb) ? nl : l0[r];
      l1[r] = (f && !kb) ? nl : l1[r];
      cm[r] = dlp_ballot(ch[r]);
      prof.E += f ? (unsigned)len[r] : 0u;
      prof.C += ch[r] ? (unsigned)len[r] + 1u : 0u;
      anyOut = anyOut || (ch[r] && outPos[r] != 0u);
      entered[r] = entered[r] || (f && k > kBuild);
    }
    const bool rebuild = __any(anyOut);  // uniform
    if (rebuild) {  // marks of eligible vertices outside the core
      const LocalIter it{stampL, nullptr, nullptr, sh, haloE, nullptr, dumOut, nCore, nHE,
                         k + 1, stampBase + k + 1, stampBase};
#pragma unroll
      for (int r = 0; r < V; ++r) {
        if (!(ch[r] && outPos[r])) continue;
#pragma unroll
        for (int p = 0; p < L; ++p)
          if ((outPos[r] >> p) & 1u) prof.A += mark_outside(a, l, it, haloE[src[r][p] - nCore]);
      }
    }
    /* next frontier: every changed vertex and the core vertices their rows
     * name */
    if (symmetric) {
#pragma unroll
      for (int r = 0; r < V; ++r) {
        unsigned long long meet = 0ull;
#pragma unroll
        for (int q = 0; q < V; ++q) meet |= out[r][q] & cm[q];
        fm[r] = dlp_ballot(ch[r] || meet != 0ull) & ownMask[r];
      }
    } else {
#pragma unroll
      for (int q = 0; q < V; ++q) {
        unsigned long long reach = 0ull;
#pragma unroll
        for (int r = 0; r < V; ++r) reach |= ch[r] ? out[r][q] : 0ull;
        fm[q] = (wave_or64(reach) | cm[q]) & ownMask[q];
      }
    }
    if (lane == 0) {
      if (k == 0) c->frontier0 = n;
      if (a.iterSizes && k < a.iterSizesCap) a.iterSizes[k] = n;
    }
    ownerSum += (unsigned long long)n;
    ++k;
    __syncthreads();  // new labels (and mark_outside's writes) visible to every lane
    if (rebuild) break;
  }

  /* the masks of iteration k -> its queue (by core index); stamps of the
   * vertices that entered a frontier */
  {
    int *cur = (k & 1) ? q1L : q0L;
    int base = 0, nOut = 0;
#pragma unroll
    for (int r = 0; r < V; ++r) nOut += __popcll(fm[r]);
#pragma unroll
    for (int r = 0; r < V; ++r) {
      const bool f = (fm[r] >> lane) & 1ull;
      entered[r] = entered[r] || (f && k > kBuild);
      if (f) cur[base + dlp_rank_below(fm[r])] = lane + 64 * r;
      base += __popcll(fm[r]);
      if (entered[r]) stampL[lane + 64 * r] = k;  // k > kBuild
    }
    if (lane == 0) {
      sh[SH_QC0 + k % 3] = nOut;
      passUpdates += ownerSum;
      prof.U += (unsigned)ownerSum;
      prof.D += (unsigned)ownerSum;
      prof.Wv += (unsigned)ownerSum;
      if (ownerSum > 0ull) prof.A += (unsigned)(ownerSum - (unsigned long long)nq) + (unsigned)nOut;
    }
    __syncthreads();
  }
  return reason;
}

/* Register-mask owner mode with W = LT / 64 waves (2 .. kMaskWavesMax):
 * thread tid owns core vertex tid (core at most LT vertices, rows at most
 * L <= SR entries, every core vertex eligible) and keeps in registers its
 * row records, both labels and, as W 64-bit masks, the core vertices its
 * row names.  Each wave holds its part of the frontier as one 64-bit mask.
 * An iteration: flagged threads compute their labels (neighbour labels from
 * the LDS label array) and store them; every wave publishes its changed
 * mask, and whether a change reached an eligible vertex outside the core,
 * in LDS; one barrier; then each thread's next flag is `changed`, or its row
 * mask meeting some wave's changed mask (the CSR is symmetric: an
 * asymmetric one -- PropArgs::symmetric 0 -- never comes here).  The size of
 * the next frontier is published with the masks of the following
 * iteration, so one barrier per iteration suffices: an iteration whose
 * frontier turns out empty has run with no flagged thread and changed
 * nothing.  State left behind as owner_iterations. */
template <int LT, int SR, int L>
__device__ __forceinline__ int mask_owner_iterations(
    const PropArgs &a, const LocalArgs &l, LocalShared &S, int &k, int maxIter, int kBuild,
    int nCore, int nHE, int nullSlot, int stampBase, const int *rlL, const int2 *recS,
    int *stampL, int *q0L, int *q1L, float *lab0L, float *lab1L, const int *haloE,
    unsigned long long &passUpdates, LocalProf &prof) {
  constexpr int W = LT / 64;
  static_assert(W >= 2 && W <= kMaskWavesMax && L <= SR, "mask owner shape");
  constexpr int RS = SR + 1;
  int *sh = S.sh;
  DevCounters *c = a.ctr;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int *dumOut = S.dummyOut + lane;
  const int nq = sh[SH_QC0 + k % 3];
  {
    const int *qin = (k & 1) ? q1L : q0L;
    int *flags = (k & 1) ? q0L : q1L;
    for (int i = tid; i < nCore; i += LT) flags[i] = 0;
    __syncthreads();
    for (int i = tid; i < nq; i += LT) flags[qin[i]] = 1;
    __syncthreads();
  }
  const int *flags0 = (k & 1) ? q0L : q1L;
  const bool own = tid < nCore;
  const unsigned long long ownMask = dlp_ballot(own);
  unsigned long long fm = dlp_ballot(own && flags0[tid] != 0);
  int src[L];
  float w[L];
  unsigned long long out[W];
  const int len = own ? (rlL[tid] & kRowLenMask) : 0;
  float l0 = own ? lab0L[tid] : 0.0f, l1 = own ? lab1L[tid] : 0.0f;
  unsigned outPos = 0u;
  bool entered = false;
#pragma unroll
  for (int q = 0; q < W; ++q) out[q] = 0ull;
#pragma unroll
  for (int p = 0; p < L; ++p) {
    const int2 rec = own ? recS[tid * RS + p] : make_int2(nullSlot, 0);
    src[p] = rec.x;
    w[p] = __int_as_float(rec.y);
    if (rec.x < nCore) {
#pragma unroll
      for (int q = 0; q < W; ++q)
        if ((rec.x >> 6) == q) out[q] |= 1ull << (rec.x & 63);
    } else if (rec.x < nCore + nHE) {
      outPos |= 1u << p;
    }
  }
  const float Wrow = mask_row_weight<L>(w);
  int n = nq;  // size of frontier k, known for the first iteration
  __syncthreads();  // the flags are read

  int reason = 0;
  unsigned long long ownerSum = 0ull;
  bool first = true;
  while (true) {
    if (k >= maxIter) {
      if (!first) {  // the size of frontier k was published in the last iteration
        __syncthreads();
        n = 0;
#pragma unroll
        for (int q = 0; q < W; ++q) n += S.xCount[(k - 1) & 1][q];
      }
      reason = n ? DLP_LOCAL_CAPPED : DLP_LOCAL_CONVERGED;
      break;
    }
    const int kb = k & 1;
    const float *Fo = kb ? lab1L : lab0L;
    float *Fn = kb ? lab0L : lab1L;
    const bool f = (fm >> lane) & 1ull;
    const float old = kb ? l1 : l0;
    float nl = old;
    if (fm != 0ull)  // wave-uniform: a wave with no flagged vertex only syncs
      nl = mask_row_label_any<L>(src, w, Wrow, Fo, nullSlot, f, len, old);
    const bool ch = f && fabsf(nl - old) > a.delta;
    *(f ? &Fn[tid] : (float *)dumOut) = nl;
    l0 = (f && kb) ? nl : l0;
    l1 = (f && !kb) ? nl : l1;
    const unsigned long long cmW = dlp_ballot(ch);
    const bool outW = __any(ch && outPos != 0u);
    if (outW) {  // wave-uniform: marks of eligible vertices outside the core
      const LocalIter it{stampL, nullptr, nullptr, sh, haloE, nullptr, dumOut, nCore, nHE,
                         k + 1, stampBase + k + 1, stampBase};
      if (ch && outPos) {
#pragma unroll
        for (int p = 0; p < L; ++p)
          if ((outPos >> p) & 1u) prof.A += mark_outside(a, l, it, haloE[src[p] - nCore]);
      }
    }
    if (lane == 0) {
      S.xChanged[kb][wave] = cmW;
      S.xOut[kb][wave] = outW ? 1u : 0u;
    }
    __syncthreads();
    if (!first) {  // the size of frontier k, published in iteration k - 1
      int cnt[W];
#pragma unroll
      for (int q = 0; q < W; ++q) cnt[q] = S.xCount[kb ^ 1][q];
      n = 0;
#pragma unroll
      for (int q = 0; q < W; ++q) n += cnt[q];
    }
    first = false;
    if (n == 0) {  // frontier k was empty: nothing above changed anything
      reason = DLP_LOCAL_CONVERGED;
      break;
    }
    prof.E += f ? (unsigned)len : 0u;
    prof.C += ch ? (unsigned)len + 1u : 0u;
    entered = entered || (f && k > kBuild);
    /* every wave's words first (independent loads), then branch-free ORs */
    int outs = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) outs |= S.xOut[kb][q];
    unsigned long long xc[W], meet = 0ull;
#pragma unroll
    for (int q = 0; q < W; ++q) xc[q] = S.xChanged[kb][q];
#pragma unroll
    for (int q = 0; q < W; ++q) meet |= out[q] & xc[q];
    const bool rebuild = outs != 0;
    fm = dlp_ballot(ch || meet != 0ull) & ownMask;
    if (lane == 0) S.xCount[kb][wave] = (unsigned char)__popcll(fm);
    if (tid == 0) {
      if (k == 0) c->frontier0 = n;
      if (a.iterSizes && k < a.iterSizesCap) a.iterSizes[k] = n;
    }
    ownerSum += (unsigned long long)n;
    ++k;
    if (rebuild) break;  // uniform
  }

  /* the masks of iteration k -> its queue (by core index); stamps of the
   * vertices that entered a frontier */
  {
    int *cur = (k & 1) ? q1L : q0L;
    const bool f = (fm >> lane) & 1ull;
    entered = entered || (f && k > kBuild);
    __syncthreads();  // every wave's mask of iteration k is final
    if (lane == 0) S.xCount[0][wave] = (unsigned char)__popcll(fm);
    __syncthreads();
    int base = 0, nOut = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      base += q < wave ? S.xCount[0][q] : 0;
      nOut += S.xCount[0][q];
    }
    if (f) cur[base + dlp_rank_below(fm)] = tid;
    if (entered) stampL[tid] = k;  // k > kBuild
    if (tid == 0) {
      sh[SH_QC0 + k % 3] = nOut;
      passUpdates += ownerSum;
      prof.U += (unsigned)ownerSum;
      prof.D += (unsigned)ownerSum;
      prof.Wv += (unsigned)ownerSum;
      if (ownerSum > 0ull) prof.A += (unsigned)(ownerSum - (unsigned long long)nq) + (unsigned)nOut;
    }
    __syncthreads();
  }
  return reason;
}

/* the multi-wave register-mask owner iterations for the core's longest row */
template <int LT, int SR>
__device__ __forceinline__ int mask_owner_dispatch(
    const PropArgs &a, const LocalArgs &l, LocalShared &S, int &k, int maxIter, int kBuild,
    int nCore, int nHE, int nullSlot, int stampBase, const int *rlL, const int2 *recS,
    int *stampL, int *q0L, int *q1L, float *lab0L, float *lab1L, const int *haloE,
    unsigned long long &passUpdates, LocalProf &prof, int maxLen) {
  if (maxLen <= 8)  // 8-record rows in registers: fewer registers
    return mask_owner_iterations<LT, SR, 8>(a, l, S, k, maxIter, kBuild, nCore, nHE, nullSlot,
                                            stampBase, rlL, recS, stampL, q0L, q1L, lab0L, lab1L,
                                            haloE, passUpdates, prof);
  return mask_owner_iterations<LT, SR, SR>(a, l, S, k, maxIter, kBuild, nCore, nHE, nullSlot,
                                           stampBase, rlL, recS, stampL, q0L, q1L, lab0L, lab1L,
                                           haloE, passUpdates, prof);
}

/* the single-wave owner iterations for the core's size and longest row */
template <int SR>
__device__ __forceinline__ int wave_owner_dispatch(
    const PropArgs &a, const LocalArgs &l, LocalShared &S, int &k, int maxIter, int kBuild,
    int nCore, int nHE, int nullSlot, int stampBase, const int *rlL, const int2 *recS,
    int *stampL, int *q0L, int *q1L, float *lab0L, float *lab1L, const int *haloE,
    unsigned long long &passUpdates, LocalProf &prof, int maxLen) {
#define DLP_WAVE_OWNER(V, LL)                                                                  \
  return wave_owner_iterations<SR, V, LL>(a, l, S, k, maxIter, kBuild, nCore, nHE, nullSlot,  \
                                          stampBase, rlL, recS, stampL, q0L, q1L, lab0L, lab1L, \
                                          haloE, passUpdates, prof)
  (void)nCore;  // at most 64: one vertex per lane
  if (maxLen <= 8) DLP_WAVE_OWNER(1, 8);
  DLP_WAVE_OWNER(1, SR);
#undef DLP_WAVE_OWNER
}

/* Runs LDS-resident passes from iteration k (updated) until the frontier
 * empties, k reaches maxIter, or the working set outgrows the LDS budget;
 * returns the DLP_LOCAL_* reason.  All LT threads of the work-group call
 * it; `pool` is the dynamic LDS (l.ldsWords words). */
template <int LT, int SR, bool OWNER, bool MASK = false>
__device__ __forceinline__ int local_passes(const PropArgs &a, const LocalArgs &l, int &k, int maxIter,
                            int *pool, LocalShared &S, LocalProf &prof, int longAfter) {
  static_assert(SR <= kLaneRowMax, "padded stride above the lane row limit");
  /* padded rows are SR + 1 records apart: with 2*(SR+1) words (SR a power
   * of two) the rows of neighbouring core vertices start in different LDS
   * banks, so a wave's lanes reading position j of their rows do not all
   * collide in a few banks */
  constexpr int RS = SR + 1;
  int *sh = S.sh, *tmp = S.tmp, *dummyStamp = S.dummyStamp, *dummyOut = S.dummyOut;
  DevCounters *c = a.ctr;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int P = l.ldsWords;
  const int stampBase = c->stampBase;
  int reason = 0;
  /* the owner kernel starts from the frontier alone: a long-lived frontier
   * then grows the core (rebuilds) only by the vertices it reaches, which
   * keeps it within one vertex per thread more often */
  bool useCore = !OWNER;
  unsigned &profU = prof.U, &profE = prof.E, &profC = prof.C, &profA = prof.A;
  unsigned long long passUpdates = 0; /* tid 0: frontier sizes of this pass */
  const int kLaunch = k;

  while (true) {  // one pass = build the LDS working set, iterate, write back
    /* every thread has read the previous pass's sh[] values */
    __syncthreads();
    /* ---- candidates: saved core list + frontier of iteration k ---------- */
    if (tid == 0) {
      const int gen = c->localGen + 1;
      c->localGen = gen;
      if (a.profile) c->profLocalBuilds += 1ull;
      const int cc = useCore ? min(c->coreCount, l.coreListCap) : 0;
      sh[SH_GEN] = gen;
      sh[SH_CC] = cc;
      sh[SH_NCAND] = cc + c->qcount[k % 3];
      sh[SH_NCORE] = 0;
      sh[SH_BIG] = 0;
      sh[SH_REBUILD] = 0;
      sh[SH_NOVF] = 0;
      sh[SH_HASLONG] = 0;
      sh[SH_NHE] = 0;
      sh[SH_NHI] = 0;
      sh[SH_NLONG0] = 0;
      sh[SH_NLONG1] = 0;
      sh[SH_INELIG] = 0;
    }
    __syncthreads();
    const int gen = sh[SH_GEN], cc = sh[SH_CC], nCand = sh[SH_NCAND];
    const int *qk = (k & 1) ? a.q1 : a.q0;
    if (8LL * nCand > P) {  // uniform
      if (useCore && cc > 0) {
        useCore = false;
        continue;
      }
      reason = DLP_LOCAL_TOO_BIG;
      break;
    }
    int *gidL = pool;  // [nCand]
    for (int i = tid; i < nCand; i += LT) {
      const int g = i < cc ? l.coreList[i] : qk[i - cc];
      if (atomicMax(&l.coreGen[g], gen) < gen) {
        const int idx = atomicAdd(&sh[SH_NCORE], 1);
        gidL[idx] = g;
        l.coreIdx[g] = idx;
      }
    }
    wg_global_sync();
    const int nCore = sh[SH_NCORE];
    int *rlL = gidL + nCand;  // row length
    int *rsL = rlL + nCore;   // offset of a row longer than SR in the long records
    for (int i = tid; i < nCore; i += LT) {
      const int g = gidL[i];
      const int rl = a.rowLen ? a.rowLen[g] : a.rowEnd[g] - a.rowBeg[g];
      rlL[i] = rl;
      rsL[i] = rl > SR ? rl : 0;
      if (rl > kLaneRowMax) sh[SH_HASLONG] = 1;
    }
    __syncthreads();
    const int Elong = block_scan_lds<LT>(rsL, rsL, nCore, tmp);  // ends with a barrier
    const bool hasLong = sh[SH_HASLONG] != 0;
    /* ---- halo: distinct eligible / ineligible neighbours outside the core */
    for (int i = tid; i < nCore; i += LT) {
      const int beg = a.rowBeg[gidL[i]], len = rlL[i];
      for (int p = 0; p < len; ++p) {
        const int v = a.colInd[beg + p];
        if (v < 0 || (a.status && a.status[v] == DLP_VERTEX_DELETED)) continue;
        if (ld_agent(&l.coreGen[v]) == gen) continue;
        if (atomicMax(&l.haloGen[v], gen) < gen) {
          const bool elig = a.iterStamp[v] != DLP_INT_MAX;
          const int idx = atomicAdd(&sh[elig ? SH_NHE : SH_NHI], 1);
          l.haloIdx[v] = elig ? idx : (idx | 0x40000000);
        }
      }
    }
    wg_global_sync();
    const int nHE = sh[SH_NHE], nHI = sh[SH_NHI];
    const int nullSlot = nCore + nHE + nHI;
    const int nLab = nullSlot + 1;
    /* layout after gidL, rlL, rsL */
    const int offStamp = nCand + 2 * nCore;
    const int offQ0 = offStamp + nCore, offQ1 = offQ0 + nCore, offLong = offQ1 + nCore;
    const int offLab0 = offLong + nCore, offLab1 = offLab0 + nLab;
    const int offHalo = offLab1 + nLab;
    const int offRecS = (offHalo + nHE + 1) & ~1;                  // int2 aligned
    const int offRecL = offRecS + 2 * RS * nCore;
    const long long need = (long long)offRecL + 2LL * Elong + kLdsSlack;
    if (need > P) {  // uniform
      if (useCore && cc > 0) {
        useCore = false;
        continue;
      }
      reason = DLP_LOCAL_TOO_BIG;
      break;
    }
    int *stampL = pool + offStamp;
    int *q0L = pool + offQ0;
    int *q1L = pool + offQ1;
    int *longL = pool + offLong;
    float *lab0L = (float *)(pool + offLab0);
    float *lab1L = (float *)(pool + offLab1);
    int *haloE = pool + offHalo;
    int2 *recS = (int2 *)(pool + offRecS);
    int2 *recL = (int2 *)(pool + offRecL);
    /* ---- stage labels, stamps and row records ---------------------------- */
    for (int i = tid; i < nCore; i += LT) {
      const int g = gidL[i];
      lab0L[i] = a.lab0[g];
      lab1L[i] = a.lab1[g];
      const bool inelig = a.iterStamp[g] == DLP_INT_MAX;
      stampL[i] = inelig ? DLP_INT_MAX : -1;
      if (inelig) sh[SH_INELIG] = 1;
      const int beg = a.rowBeg[g], len = rlL[i];
      const bool shortRow = len <= SR;
      int2 *out = shortRow ? recS + i * RS : recL + rsL[i];
      bool hasOut = false;
      for (int p = 0; p < len; ++p) {
        const int v = a.colInd[beg + p];
        int2 r = make_int2(nullSlot, 0);
        if (v >= 0 && !(a.status && a.status[v] == DLP_VERTEX_DELETED)) {
          r.y = __float_as_int(a.values[beg + p]);
          if (ld_agent(&l.coreGen[v]) == gen) {
            r.x = ld_agent(&l.coreIdx[v]);
          } else {
            const int h = ld_agent(&l.haloIdx[v]);
            const int hi = h & 0x3fffffff;
            r.x = (h & 0x40000000) ? nCore + nHE + hi : nCore + hi;
            lab0L[r.x] = a.lab0[v];  // the same values from every entry naming v
            lab1L[r.x] = a.lab1[v];
            if (!(h & 0x40000000)) {
              haloE[hi] = v;
              hasOut = true;
            }
          }
        }
        out[p] = r;
      }
      if (hasOut) rlL[i] = len | kRowHasOut;
      /* padding of a short row; the unused short slots of a long row */
      for (int p = shortRow ? len : 0; p < SR; ++p) recS[i * RS + p] = make_int2(nullSlot, 0);
    }
    if (tid == 0) {
      lab0L[nullSlot] = 0.0f;
      lab1L[nullSlot] = 0.0f;
    }
    __syncthreads();
    /* ---- frontier of iteration k ----------------------------------------- */
    const int nq = nCand - cc;
    {
      int *qin = (k & 1) ? q1L : q0L;
      for (int i = tid; i < nq; i += LT) {
        const int idx = ld_agent(&l.coreIdx[qk[i]]);
        qin[i] = idx;
        stampL[idx] = k;
      }
      if (tid == 0) {
        sh[SH_QC0 + k % 3] = nq;
        sh[SH_QC0 + (k + 1) % 3] = 0;
        sh[SH_QC0 + (k + 2) % 3] = 0;
      }
    }
    __syncthreads();
    const int kBuild = k;

    /* ---- iterations ------------------------------------------------------- */
    /* Owner mode (the OWNER kernels) when the core fits it, else queue
     * iterations; the queue kernels stop with DLP_LOCAL_LONG after
     * longAfter iterations of a launch (>= 0), so that the host goes on
     * with the owner kernel on a long-lived frontier such as a limit cycle.
     * The two modes are separate kernels: owner mode's registers would
     * otherwise cost the queue iterations spills. */
    constexpr int OV = 2048 / LT;  // owned vertices per thread in owner mode: cores up to 2048
    const bool ownerOK = OWNER && nCore <= OV * LT && !sh[SH_INELIG];  // uniform
    if constexpr (OWNER && MASK) {
      /* the register-mask owner kernels: one core vertex per thread (cores
       * of at most LT vertices: with more per thread the serial work
       * outweighs the barriers saved), rows of at most SR entries, a
       * symmetric CSR; anything else goes back to the host
       * for a larger mask kernel or the flag-based owner kernel */
      if (!ownerOK || nCore > LT || hasLong || Elong > 0 || (LT > 64 && !a.symmetric)) {
        reason = DLP_LOCAL_WIDE;
      } else {
        if (tid == 0) sh[SH_MAXLEN] = 0;
        __syncthreads();
        for (int i = tid; i < nCore; i += LT) atomicMax(&sh[SH_MAXLEN], rlL[i] & kRowLenMask);
        __syncthreads();
        if constexpr (LT == 64)
          reason = wave_owner_dispatch<SR>(a, l, S, k, maxIter, kBuild, nCore, nHE, nullSlot,
                                           stampBase, rlL, recS, stampL, q0L, q1L, lab0L, lab1L,
                                           haloE, passUpdates, prof, sh[SH_MAXLEN]);
        else
          reason = mask_owner_dispatch<LT, SR>(a, l, S, k, maxIter, kBuild, nCore, nHE, nullSlot,
                                               stampBase, rlL, recS, stampL, q0L, q1L, lab0L,
                                               lab1L, haloE, passUpdates, prof, sh[SH_MAXLEN]);
      }
    } else if (ownerOK) {
      if (nCore <= LT)
        reason = owner_iterations<LT, SR, 1>(a, l, S, k, maxIter, kBuild, nCore, nHE, nullSlot,
                                             stampBase, rlL, rsL, recS, recL, stampL, q0L, q1L,
                                             lab0L, lab1L, haloE, passUpdates, prof);
      else
        reason = owner_iterations<LT, SR, OV>(a, l, S, k, maxIter, kBuild, nCore, nHE,
                                              nullSlot, stampBase, rlL, rsL, recS, recL, stampL,
                                              q0L, q1L, lab0L, lab1L, haloE, passUpdates, prof);
    } else {  // queue iterations
      while (true) {
        const int n = sh[SH_QC0 + k % 3];
        if (n == 0) {
          reason = DLP_LOCAL_CONVERGED;
          break;
        }
        if (k >= maxIter) {
          reason = DLP_LOCAL_CAPPED;
          break;
        }
        if (!OWNER && longAfter >= 0 && k - kLaunch >= longAfter) {  // uniform
          reason = DLP_LOCAL_LONG;
          break;
        }
        const int kb = k & 1;
        if (tid == 0) {
          sh[SH_QC0 + (k + 2) % 3] = 0;    // appended to by iteration k+1
          sh[SH_NLONG0 + (kb ^ 1)] = 0;    // long-row list of iteration k+1
          if (k == 0) c->frontier0 = n;
          if (a.iterSizes && k < a.iterSizesCap) a.iterSizes[k] = n;
          passUpdates += (unsigned long long)n;
          profU += (unsigned)n;
        }
        const int *qin = kb ? q1L : q0L;
        const float *Fo = kb ? lab1L : lab0L;
        float *Fn = kb ? lab0L : lab1L;
        int *nLongCtr = &sh[SH_NLONG0 + kb];
        const LocalIter it{stampL, kb ? q0L : q1L, &sh[SH_QC0 + (k + 1) % 3], sh, haloE,
                           dummyStamp, dummyOut, nCore, nHE, k + 1, stampBase + k + 1,
                           stampBase};

        /* rows up to kLaneRowMax: one lane per frontier vertex */
        for (int base = 0; base < n; base += LT) {  // block-uniform
          if (base + (tid & ~63) >= n) continue;     // wave-uniform: no vertex for this wave
          const int item = base + tid;
          const bool act = item < n;
          const int ci = act ? qin[item] : 0;
          const int rl = act ? rlL[ci] : 0;
          const int len = rl & kRowLenMask;
          const bool hasOut = (rl & kRowHasOut) != 0;
          const bool shortRow = act && len <= SR;
          const bool midRow = act && len > SR && len <= kLaneRowMax;
          if (hasLong) {
            const int slot = dlp_wave_append(nLongCtr, act && len > kLaneRowMax);
            if (slot >= 0) longL[slot] = ci;
          }
          if (__any(shortRow)) {  // wave-uniform
            const int2 *row = recS + ci * RS;
            if (SR > 4 && !__any(shortRow && len > 4))
              lane_rows<(SR > 4 ? 4 : SR), false>(a, l, it, row, nullSlot, Fo, Fn, shortRow, ci,
                                                  len, hasOut, profE, profC, profA);
            else
              lane_rows<SR, false>(a, l, it, row, nullSlot, Fo, Fn, shortRow, ci, len, hasOut,
                                   profE, profC, profA);
          }
          if (SR < kLaneRowMax && __any(midRow)) {  // wave-uniform
            const int2 *row = recL + (midRow ? rsL[ci] : 0);
            lane_rows<kLaneRowMax, true>(a, l, it, row, nullSlot, Fo, Fn, midRow, ci,
                                         midRow ? len : 0, hasOut, profE, profC, profA);
          }
        }
        /* longer rows: one wave per frontier vertex */
        if (hasLong) {  // uniform
          __syncthreads();
          const int nLong = *nLongCtr;
          for (int i = wave; i < nLong; i += LT / 64) {  // wave-uniform
            const int ci = longL[i];
            const int rs = rsL[ci], len = rlL[ci] & kRowLenMask;
            float W, S;
            dlp::group_row_sums<64>(lane, len, LocalPos{recL, Fo, rs}, W, S);
            int changed = 0;
            if (lane == 0) {
              const float oldLabel = Fo[ci];
              const float nl = dlp::new_label(W, S, oldLabel);
              Fn[ci] = nl;
              changed = fabsf(nl - oldLabel) > a.delta;
              profE += (unsigned)len;
              if (changed) profC += (unsigned)len + 1u;
            }
            changed = __shfl(changed, 0, 64);
            if (!changed) continue;  // wave-uniform
            for (int p0 = 0; p0 <= len; p0 += 64) {  // wave-uniform
              const int p = p0 + lane;
              int t = nullSlot;
              if (p < len) t = recL[rs + p].x;
              else if (p == len) t = ci;  // the vertex itself
              bool appL = false;
              if (t < nCore) {
                appL = atomicMax(&stampL[t], it.stampNextL) < it.stampNextL;
              } else if (t < nCore + nHE) {
                profA += mark_outside(a, l, it, haloE[t - nCore]);
              }
              profA += appL ? 1u : 0u;
              const int slot = dlp_wave_append(it.qcOut, appL);
              if (slot >= 0) it.qout[slot] = t;
            }
          }
        }
        __syncthreads();
        ++k;
        if (sh[SH_REBUILD]) break;  // uniform after the barrier
      }
    }

    /* ---- write back: labels, touched, next frontier, core list ---------- */
    for (int base = 0; base < nCore; base += LT) {  // block-uniform
      const int i = base + tid;
      bool first = false;
      int g = 0;
      if (i < nCore) {
        g = gidL[i];
        a.lab0[g] = lab0L[i];
        a.lab1[g] = lab1L[i];
        const int st = stampL[i];
        if (st != DLP_INT_MAX && st > kBuild) {  // entered a frontier here
          const int old = atomicMax(&a.iterStamp[g], stampBase + k);
          first = old < stampBase;
        }
        if (i < l.coreListCap) l.coreList[i] = g;
      }
      if (a.touched) {
        const int ts = dlp_wave_append(&c->touched, first);
        if (ts >= 0) a.touched[ts] = g;
      }
    }
    {
      const int nl = sh[SH_QC0 + k % 3];
      const int no = min(sh[SH_NOVF], l.ovfCap);
      const int *qloc = (k & 1) ? q1L : q0L;
      int *gq = (k & 1) ? a.q1 : a.q0;
      for (int i = tid; i < nl; i += LT) gq[i] = gidL[qloc[i]];
      for (int i = tid; i < no; i += LT) gq[nl + i] = l.ovf[i];
      if (tid == 0) {
        c->qcount[k % 3] = nl + no;
        c->qcount[(k + 1) % 3] = 0;
        c->coreCount = min(nCore, l.coreListCap);
        if (k > kBuild) {  // iterations kBuild .. k-1 ran
          c->iterUsed = k;
          c->updates += passUpdates;
        }
        passUpdates = 0;
        /* cannot happen with ovfCap >= ldsWords / 4 (outside targets <= E) */
        if (sh[SH_NOVF] > l.ovfCap) {
          atomicOr(&c->errorFlag, 8);
          sh[SH_BIG] = 1;
        }
      }
    }
    wg_global_sync();
    if (reason != 0) break;
    if (sh[SH_BIG]) {
      reason = DLP_LOCAL_TOO_BIG;
      break;
    }
    useCore = true;  // rebuild with the grown core at iteration k
  }
  return reason;
}

/* adds the work counters of the local passes to the graph's totals */
__device__ __forceinline__ void local_prof_flush(const PropArgs &a, LocalProf &prof) {
  if (!a.profile) return;
  DevCounters *c = a.ctr;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    prof.E += __shfl_xor(prof.E, off, 64);
    prof.C += __shfl_xor(prof.C, off, 64);
    prof.A += __shfl_xor(prof.A, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (prof.E) atomicAdd(&c->profLocalEdges, (unsigned long long)prof.E);
    if (prof.C) atomicAdd(&c->profLocalChangedEdges, (unsigned long long)prof.C);
    if (prof.A) atomicAdd(&c->profLocalAppended, (unsigned long long)prof.A);
  }
  if (threadIdx.x == 0 && prof.U) atomicAdd(&c->profLocalUpdates, (unsigned long long)prof.U);
  if (threadIdx.x == 0 && prof.D)
    atomicAdd(&c->profLocalOwnerUpdates, (unsigned long long)prof.D);
  if (threadIdx.x == 0 && prof.Wv)
    atomicAdd(&c->profLocalWaveUpdates, (unsigned long long)prof.Wv);
  prof = LocalProf{};
}

template <int LT, int SR, bool OWNER, bool MASK = false>
__global__ __launch_bounds__(LT) void k_propagate_local(PropArgs a, LocalArgs l, int k0,
                                                        int maxIter) {
  extern __shared__ int4 pool4[];
  __shared__ LocalShared S;
  if (threadIdx.x < 64) S.dummyStamp[threadIdx.x] = DLP_INT_MAX;  // first pass starts with a barrier
  LocalProf prof;
  int k = k0;
  const int reason = local_passes<LT, SR, OWNER, MASK>(a, l, k, maxIter, (int *)pool4, S, prof,
                                                 OWNER ? -1 : l.ownerAfter);
  if (threadIdx.x == 0) {
    a.ctr->nextIter = k;
    a.ctr->localReason = reason;
  }
  local_prof_flush(a, prof);
}

/* ------------------------------------------------------------------------ */
/* one work-group iterating over HBM (graphs whose working set does not fit */
/* the LDS), and the multi-graph driver                                     */
/* ------------------------------------------------------------------------ */

/* Frontier vertices with rows of at most L entries, one per lane, reading
 * the row, the labels and the stamps in HBM; appends with one atomic per
 * wave.  All 64 lanes call it. */
template <int L>
__device__ __forceinline__ void wg_lane_rows(const PropArgs &a, const float *Fold, float *Fnew,
                                             bool act, int u, int beg, int len, int stampNext,
                                             int stampBase, int *out, int *outCount,
                                             unsigned &profE, unsigned &profC,
                                             unsigned &profA) {
  int vv[L];
  float ww[L], ff[L];
#pragma unroll
  for (int j = 0; j < L; ++j) {
    int v = -1;
    float w = 0.0f;
    if (j < len) {
      v = a.colInd[beg + j];
      w = a.values[beg + j];
    }
    if (v >= 0 && a.status && a.status[v] == DLP_VERTEX_DELETED) v = -1;
    vv[j] = v;
    ww[j] = v >= 0 ? w : 0.0f;
  }
#pragma unroll
  for (int j = 0; j < L; ++j) ff[j] = vv[j] >= 0 ? Fold[vv[j]] : 0.0f;
  const float oldLabel = Fold[u];
  float W, S;
  dlp::lane_row_sums<L>(
      L,
      [&](int j, float &w, float &s) {
        w = ww[j];
        s = ww[j] * ff[j];
      },
      W, S);
  const float nl = dlp::new_label(W, S, oldLabel);
  const bool changed = act && fabsf(nl - oldLabel) > a.delta;
  if (act) {
    Fnew[u] = nl;
    profE += (unsigned)len;
    if (changed) profC += (unsigned)len + 1u;
  }
  int olds[L + 1];
#pragma unroll
  for (int j = 0; j < L; ++j) {
    int o = DLP_INT_MAX;
    if (changed && vv[j] >= 0) o = atomicMax(&a.iterStamp[vv[j]], stampNext);
    olds[j] = o;
  }
  {
    int o = DLP_INT_MAX;
    if (changed) o = atomicMax(&a.iterStamp[u], stampNext);
    olds[L] = o;
  }
  unsigned hit = 0u, first = 0u;
#pragma unroll
  for (int j = 0; j <= L; ++j) {
    hit |= (olds[j] < stampNext ? 1u : 0u) << j;
    first |= (olds[j] < stampBase ? 1u : 0u) << j;
  }
  const int lane = dlp_lane();
  const int cnt = __popc(hit);
  profA += (unsigned)cnt;
  {
    const int incl = dlp_wave_inclusive_sum(cnt);
    const int total = __builtin_amdgcn_readlane(incl, 63);
    if (total > 0) {  // wave-uniform
      int wbase = 0;
      if (lane == 63) wbase = atomicAdd(outCount, total);
      wbase = __builtin_amdgcn_readlane(wbase, 63);
      int pos = wbase + incl - cnt;
#pragma unroll
      for (int j = 0; j <= L; ++j)
        if ((hit >> j) & 1u) out[pos++] = j < L ? vv[j < L ? j : 0] : u;
    }
  }
  if (a.touched) {
    const int tc = __popc(first);
    const int incl = dlp_wave_inclusive_sum(tc);
    const int total = __builtin_amdgcn_readlane(incl, 63);
    if (total > 0) {  // wave-uniform
      int wbase = 0;
      if (lane == 63) wbase = atomicAdd(&a.ctr->touched, total);
      wbase = __builtin_amdgcn_readlane(wbase, 63);
      int pos = wbase + incl - tc;
#pragma unroll
      for (int j = 0; j <= L; ++j)
        if ((first >> j) & 1u) a.touched[pos++] = j < L ? vv[j < L ? j : 0] : u;
    }
  }
}

/* Iterations k, k+1, ... by one work-group over HBM, while the frontier is
 * non-empty, k < maxIter, not (frontier <= threshold and k >= resumeAt), not
 * (heavy > 0 and frontier > heavy) and not past the deadline (wall-clock
 * ticks, 0 = none; checked once k > kFirst): the same global state
 * transitions as one k_propagate launch each.  On return S.misc[0] holds
 * the frontier size of iteration k and S.misc[1] is 1 if the deadline
 * ended the iterations. */
template <int LT>
__device__ __forceinline__ void wg_iterations(const PropArgs &a, int &k, int maxIter, int threshold,
                              int resumeAt, int heavy, long long deadline, int kFirst,
                              LocalShared &S, LocalProf &prof) {
  constexpr int kWgLaneRowMax = 16;
  DevCounters *c = a.ctr;
  const int tid = threadIdx.x, lane = tid & 63;
  const int stampBase = c->stampBase;
  while (true) {
    __syncthreads();  // every thread has read the previous S.misc[0..1]
    if (tid == 0) {
      S.misc[0] = ld_agent(&c->qcount[k % 3]);
      S.misc[1] = (deadline > 0 && k > kFirst && wall_clock64() >= deadline) ? 1 : 0;
    }
    __syncthreads();
    const int n = S.misc[0];
    if (n == 0 || k >= maxIter || (n <= threshold && k >= resumeAt) || (heavy > 0 && n > heavy) ||
        S.misc[1])
      break;  // uniform
    if (tid == 0) {
      c->qcount[(k + 2) % 3] = 0;
      if (k == 0) c->frontier0 = n;
      if (a.iterSizes && k < a.iterSizesCap) a.iterSizes[k] = n;
      c->iterUsed = k + 1;
      c->updates += (unsigned long long)n;
      prof.U += (unsigned)n;
    }
    const int stampNext = stampBase + k + 1;
    const int *in = (k & 1) ? a.q1 : a.q0;
    int *out = (k & 1) ? a.q0 : a.q1;
    int *outCount = &c->qcount[(k + 1) % 3];
    const float *Fold = (k & 1) ? a.lab1 : a.lab0;
    float *Fnew = (k & 1) ? a.lab0 : a.lab1;
    for (int base = 0; base < n; base += LT) {  // block-uniform
      if (base + (tid & ~63) >= n) continue;     // wave-uniform: no vertex for this wave
      const int item = base + tid;
      const bool act = item < n;
      const int u = act ? in[item] : 0;
      int beg = 0, len = 0;
      if (act && a.rowInfo) {
        const int2 r = a.rowInfo[u];
        beg = r.x;
        len = r.y;
      } else if (act) {
        beg = a.rowBeg[u];
        len = a.rowLen ? a.rowLen[u] : a.rowEnd[u] - beg;
      }
      const bool laneRow = act && len <= kWgLaneRowMax;
      const int lenS = laneRow ? len : 0;
      if (!__any(lenS > 4))
        wg_lane_rows<4>(a, Fold, Fnew, laneRow, u, beg, lenS, stampNext, stampBase, out,
                        outCount, prof.E, prof.C, prof.A);
      else if (!__any(lenS > 8))
        wg_lane_rows<8>(a, Fold, Fnew, laneRow, u, beg, lenS, stampNext, stampBase, out,
                        outCount, prof.E, prof.C, prof.A);
      else
        wg_lane_rows<kWgLaneRowMax>(a, Fold, Fnew, laneRow, u, beg, lenS, stampNext,
                                    stampBase, out, outCount, prof.E, prof.C, prof.A);
      /* longer rows: the wave takes them one by one */
      unsigned long long longMask = dlp_ballot(act && !laneRow);
      while (longMask) {  // wave-uniform
        const int src = __ffsll((long long)longMask) - 1;
        longMask &= longMask - 1;
        const int lu = __shfl(u, src, 64), lbeg = __shfl(beg, src, 64),
                  llen = __shfl(len, src, 64);
        float W, Sm;
        dlp::group_row_sums<64>(lane, llen, GlobalPos{a.colInd, a.values, Fold, a.status, lbeg},
                                W, Sm);
        int changed = 0;
        if (lane == 0) {
          const float oldLabel = Fold[lu];
          const float nl = dlp::new_label(W, Sm, oldLabel);
          Fnew[lu] = nl;
          changed = fabsf(nl - oldLabel) > a.delta;
          prof.E += (unsigned)llen;
          if (changed) prof.C += (unsigned)llen + 1u;
        }
        changed = __shfl(changed, 0, 64);
        if (!changed) continue;  // wave-uniform
        for (int p0 = 0; p0 <= llen; p0 += 64) {  // wave-uniform
          const int p = p0 + lane;
          int v = -1;
          bool pred = false;
          if (p < llen) {
            v = a.colInd[lbeg + p];
            pred = v >= 0 && !(a.status && a.status[v] == DLP_VERTEX_DELETED);
          } else if (p == llen) {
            v = lu;
            pred = true;
          }
          const int old = pred ? atomicMax(&a.iterStamp[v], stampNext) : DLP_INT_MAX;
          const bool app = pred && old < stampNext;
          prof.A += app ? 1u : 0u;
          const int slot = dlp_wave_append(outCount, app);
          if (slot >= 0) out[slot] = v;
          if (a.touched) {
            const int ts = dlp_wave_append(&c->touched, app && old < stampBase);
            if (ts >= 0) a.touched[ts] = v;
          }
        }
      }
    }
    wg_global_sync();  // this iteration's labels, stamps and queue for every wave
    ++k;
  }
}

/* the HBM-iteration work counters of one work-group */
__device__ __forceinline__ void wg_prof_flush(const PropArgs &a, LocalProf &prof) {
  if (!a.profile) return;
  DevCounters *c = a.ctr;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    prof.E += __shfl_xor(prof.E, off, 64);
    prof.C += __shfl_xor(prof.C, off, 64);
    prof.A += __shfl_xor(prof.A, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (prof.E) atomicAdd(&c->profEdges, (unsigned long long)prof.E);
    if (prof.C) atomicAdd(&c->profChangedEdges, (unsigned long long)prof.C);
    if (prof.A) atomicAdd(&c->profAppended, (unsigned long long)prof.A);
  }
  prof = LocalProf{};
}

/* One work-group per graph runs the propagation of its graph's batch from
 * iteration startIter: LDS-resident passes while the frontier is at most the
 * threshold (and no overflow in the last `backoff` iterations), HBM
 * iterations otherwise.  It stops at the start of an iteration when
 *   - the frontier is empty or maxIter is reached: *exitIter = -1;
 *   - the frontier is above `heavy` (when > 0): *exitIter = k, with the
 *     state a multi-block launch would leave (resume on the full GPU);
 *   - k reaches stopIter < maxIter, or the launch has run sliceTicks
 *     wall-clock ticks and the graph at least one iteration:
 *     *exitIter = dlp_slice_exit(k) (resume in a later launch).
 * Every stop leaves the queue, counters and labels of iteration k on the
 * device, so the iterations and results do not depend on where a graph was
 * stopped and resumed. */
template <int LT, int SR>
__global__ __launch_bounds__(LT) void k_propagate_multi(const GraphProp *graphs) {
  extern __shared__ int4 pool4[];
  __shared__ LocalShared S;
  const GraphProp &gp = graphs[blockIdx.x];
  const PropArgs a = gp.a;
  const LocalArgs l = gp.l;
  const int maxIter = gp.maxIter, threshold = gp.threshold, backoff = gp.backoff;
  const int kFirst = gp.startIter;
  const int stopAt = min(gp.stopIter, maxIter);
  const long long deadline = gp.sliceTicks > 0 ? wall_clock64() + gp.sliceTicks : 0;
  if (threadIdx.x < 64) S.dummyStamp[threadIdx.x] = DLP_INT_MAX;
  LocalProf lprof, gprof;
  const int heavy = gp.heavy;
  int k = kFirst, resumeAt = kFirst, exitCode = -1;
  bool more = false; /* frontier of iteration k non-empty at the stop */
  if (l.ldsWords > 0 && threshold > 0) {
    while (true) {
      wg_iterations<LT>(a, k, stopAt, threshold, resumeAt, heavy, deadline, kFirst, S, gprof);
      const int n = S.misc[0];
      if (n == 0) break;  // uniform
      if (k >= stopAt || S.misc[1]) {
        more = true;
        break;
      }
      if (heavy > 0 && n > heavy) {
        exitCode = k;
        break;
      }
      const int reason = local_passes<LT, SR, false>(a, l, k, stopAt, (int *)pool4, S, lprof, -1);
      if (reason == DLP_LOCAL_CAPPED) more = true;
      if (reason != DLP_LOCAL_TOO_BIG) break;
      resumeAt = k + backoff;
      wg_global_sync();
    }
  } else {
    wg_iterations<LT>(a, k, stopAt, -1, stopAt, heavy, deadline, kFirst, S, gprof);
    const int n = S.misc[0];
    if (n > 0 && k < stopAt && !S.misc[1] && heavy > 0 && n > heavy) exitCode = k;
    else if (n > 0) more = true;
  }
  /* a non-empty frontier below maxIter that no hand-back took: resume later */
  if (exitCode < 0 && more && k < maxIter) exitCode = dlp_slice_exit(k);
  local_prof_flush(a, lprof);
  wg_prof_flush(a, gprof);
  if (threadIdx.x == 0 && gp.exitIter) *gp.exitIter = exitCode;
}

}  // namespace

int dlp_launch_propagate(const PropArgs &a, int k, int groupLanes, int blocks, hipStream_t s) {
  dim3 grid(blocks > 0 ? blocks : 1), block(kPT);
  const bool hub = a.blockRowMin > 0;
#define DLP_LAUNCH_PROP(Gl)                                                           \
  case Gl:                                                                            \
    if (hub)                                                                          \
      hipLaunchKernelGGL((k_propagate<Gl, true>), grid, block, 0, s, a, k);           \
    else                                                                              \
      hipLaunchKernelGGL((k_propagate<Gl, false>), grid, block, 0, s, a, k);          \
    break;
  switch (groupLanes) {
    DLP_LAUNCH_PROP(4)
    DLP_LAUNCH_PROP(8)
    DLP_LAUNCH_PROP(16)
    DLP_LAUNCH_PROP(32)
    DLP_LAUNCH_PROP(64)
#undef DLP_LAUNCH_PROP
    default:
      dlp_set_errorf_("propagation group size %d not in {4,8,16,32,64}", groupLanes);
      return DLP_ERR_INVALID;
  }
  DLP_LAUNCHED("k_propagate");
  return DLP_OK;
}

#define DLP_LOCAL_SHAPES(X) X(0, 1024, 8) X(1, 1024, 16) X(2, 512, 8) X(3, 256, 8) X(4, 1024, 4)

int dlp_local_shape_count() { return kNumLocalShapes; }

int dlp_local_lds_bytes(int *bytes) {
  int dev = 0, v = 0;
  DLP_HIP(hipGetDevice(&dev));
  DLP_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  v -= 1024; /* the kernel's static LDS */
  if (v < 16384) {
    dlp_set_errorf_("only %d bytes of LDS per work-group", v);
    return DLP_ERR_INVALID;
  }
  v &= ~255;
#define DLP_SET_LDS(I, T, R)                                                       \
  DLP_HIP(hipFuncSetAttribute((const void *)k_propagate_local<T, R, false>,       \
                              hipFuncAttributeMaxDynamicSharedMemorySize, v));     \
  DLP_HIP(hipFuncSetAttribute((const void *)k_propagate_local<T, R, true>,        \
                              hipFuncAttributeMaxDynamicSharedMemorySize, v));     \
  DLP_HIP(hipFuncSetAttribute((const void *)k_propagate_multi<T, R>,              \
                              hipFuncAttributeMaxDynamicSharedMemorySize, v));
  DLP_LOCAL_SHAPES(DLP_SET_LDS)
#undef DLP_SET_LDS
  DLP_HIP(hipFuncSetAttribute((const void *)k_propagate_local<64, 16, true, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, v));
  DLP_HIP(hipFuncSetAttribute((const void *)k_propagate_local<256, 16, true, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, v));
  DLP_HIP(hipFuncSetAttribute((const void *)k_propagate_local<512, 16, true, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, v));

  *bytes = v;
  return DLP_OK;
}

int dlp_launch_propagate_local(const PropArgs &a, const LocalArgs &l, int k, int maxIter,
                               bool owner, hipStream_t s) {
  const size_t lds = (size_t)l.ldsWords * 4;
  if (l.shape >= kLocalMaskShape && l.shape < kLocalMaskShape + kMaskLevels) {
    if (!owner) {
      dlp_set_error_("the register-mask local kernels run owner mode only");
      return DLP_ERR_INVALID;
    }
    switch (l.shape - kLocalMaskShape) {
      case 0:
        hipLaunchKernelGGL((k_propagate_local<64, 16, true, true>), dim3(1), dim3(64), lds, s, a,
                           l, k, maxIter);
        break;
      case 1:
        hipLaunchKernelGGL((k_propagate_local<256, 16, true, true>), dim3(1), dim3(256), lds, s,
                           a, l, k, maxIter);
        break;
      default:
        hipLaunchKernelGGL((k_propagate_local<512, 16, true, true>), dim3(1), dim3(512), lds, s,
                           a, l, k, maxIter);
        break;
    }
    DLP_LAUNCHED("k_propagate_local (register-mask owner)");
    return DLP_OK;
  }
  switch (l.shape) {
#define DLP_LAUNCH_LOCAL(I, T, R)                                                       \
  case I:                                                                              \
    static_assert(kLocalShapes[I].threads == T && kLocalShapes[I].stride == R, "");  \
    if (owner)                                                                         \
      hipLaunchKernelGGL((k_propagate_local<T, R, true>), dim3(1), dim3(T), lds, s, a, l, k, \
                         maxIter);                                                     \
    else                                                                               \
      hipLaunchKernelGGL((k_propagate_local<T, R, false>), dim3(1), dim3(T), lds, s, a, l, \
                         k, maxIter);                                                  \
    break;
    DLP_LOCAL_SHAPES(DLP_LAUNCH_LOCAL)
#undef DLP_LAUNCH_LOCAL
    default:
      dlp_set_errorf_("local propagation shape %d not in [0, %d)", l.shape, kNumLocalShapes);
      return DLP_ERR_INVALID;
  }
  DLP_LAUNCHED("k_propagate_local");
  return DLP_OK;
}

int dlp_launch_propagate_multi(const GraphProp *graphs, int numGraphs, int shape, int ldsWords,
                               hipStream_t s) {
  if (numGraphs <= 0) return DLP_OK;
  const size_t lds = (size_t)ldsWords * 4;
  switch (shape) {
#define DLP_LAUNCH_MULTI(I, T, R)                                                        \
  case I:                                                                               \
    hipLaunchKernelGGL((k_propagate_multi<T, R>), dim3(numGraphs), dim3(T), lds, s, graphs); \
    break;
    DLP_LOCAL_SHAPES(DLP_LAUNCH_MULTI)
#undef DLP_LAUNCH_MULTI
    default:
      dlp_set_errorf_("propagation shape %d not in [0, %d)", shape, kNumLocalShapes);
      return DLP_ERR_INVALID;
  }
  DLP_LAUNCHED("k_propagate_multi");
  return DLP_OK;
}
