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
 * dlp::group_row_sums, bit-identical to the reference's block reduction.
 *
 * Two kernels implement the same iteration:
 *   k_propagate        one launch per iteration, many work-groups, G lanes
 *                      per frontier vertex, queues / labels / stamps in HBM;
 *                      for large frontiers.
 *   k_propagate_local  one work-group that runs many iterations back to back
 *                      with the frontier's rows, their labels and the labels
 *                      of their neighbours staged in LDS; the limit cycles
 *                      and convergence tails of the reference's loop (most
 *                      iterations, with small frontiers) run without a
 *                      launch or a host round trip per iteration.
 */
#include "../dlp_kernels.h"
#include "../dlp_rowsum.h"

namespace {

/* ------------------------------------------------------------------------ */
/* multi-block iteration                                                    */
/* ------------------------------------------------------------------------ */
struct GlobalPos {
  const int *col;
  const float *val;
  const float *F;
  const int *status;
  int beg;
  __device__ void operator()(int p, float &w, float &s) const {
    const int v = col[beg + p];
    if (v >= 0 && !(status && status[v] == DLP_VERTEX_DELETED)) {
      w = val[beg + p];
      s = w * F[v];
    } else {
      w = 0.0f;
      s = 0.0f;
    }
  }
};

template <int G>
__global__ __launch_bounds__(256) void k_propagate(PropArgs a, int k) {
  constexpr int GPW = 64 / G; /* frontier vertices per wave */
  DevCounters *c = a.ctr;
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
  if (n == 0) return;

  const int stampBase = c->stampBase;
  const int stampNext = stampBase + k + 1;
  const int *in = (k & 1) ? a.q1 : a.q0;
  int *out = (k & 1) ? a.q0 : a.q1;
  int *outCount = &c->qcount[(k + 1) % 3];
  const float *Fold = (k & 1) ? a.lab1 : a.lab0;
  float *Fnew = (k & 1) ? a.lab0 : a.lab1;

  const int lane = threadIdx.x & 63;
  const int gl = lane & (G - 1);
  const int grp = lane / G;
  const long long gwave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  unsigned profE = 0, profC = 0, profA = 0; /* lane-local work counters */

  for (long long base = gwave * GPW; base < n; base += nwaves * GPW) {  // wave-uniform
    const long long item = base + grp;
    const bool act = item < n;
    int u = 0, beg = 0, len = 0;
    if (act) {
      u = in[item];
      beg = a.rowBeg[u];
      len = a.rowLen ? a.rowLen[u] : a.rowEnd[u] - beg;
    }
    float W, S;
    dlp::group_row_sums<G>(gl, len, GlobalPos{a.colInd, a.values, Fold, a.status, beg}, W, S);
    int changed = 0;
    if (act && gl == 0) {
      const float oldLabel = Fold[u];
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
        v = a.colInd[beg + p];
        pred = v >= 0;
      } else if (changed && p == len) {
        v = u;
        pred = true;
      }
      const int old = pred ? atomicMax(&a.iterStamp[v], stampNext) : DLP_INT_MAX;
      const bool app = pred && old < stampNext;
      profA += app ? 1u : 0u;
      const int slot = dlp_wave_append(outCount, app);
      if (slot >= 0) out[slot] = v;
      if (a.touched) {
        const int ts = dlp_wave_append(&c->touched, app && old < stampBase);
        if (ts >= 0) a.touched[ts] = v;
      }
    }
  }
  if (a.profile) {  // one atomic per counter per wave
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      profE += __shfl_xor(profE, off, 64);
      profC += __shfl_xor(profC, off, 64);
      profA += __shfl_xor(profA, off, 64);
    }
    if (lane == 0) {
      if (profE) atomicAdd(&c->profEdges, (unsigned long long)profE);
      if (profC) atomicAdd(&c->profChangedEdges, (unsigned long long)profC);
      if (profA) atomicAdd(&c->profAppended, (unsigned long long)profA);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* LDS-resident iterations (one work-group)                                 */
/* ------------------------------------------------------------------------ */
constexpr int kLT = 1024; /* threads of the local kernel             */
constexpr int kLG = 4;    /* lanes per frontier vertex               */
constexpr int kLGroups = kLT / kLG;

/* LDS edge record: target t (>= 0 core index, -1 contributes nothing,
 * <= -2 a vertex outside the core: -2 - (2*gid + eligible)), weight, and the
 * outside vertex's labels in both buffers (they cannot change while it is
 * outside the core) */
struct LocalPos {
  const int *etgt;
  const float *ew;
  const float *eh;  /* outside-vertex label in the buffer being read */
  const float *Fo;  /* core labels in the buffer being read          */
  int rs;
  __device__ void operator()(int p, float &w, float &s) const {
    const int e = rs + p;
    const int t = etgt[e];
    if (t == -1) {
      w = 0.0f;
      s = 0.0f;
      return;
    }
    w = ew[e];
    const float F = t >= 0 ? Fo[t] : eh[e];
    s = w * F;
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

/* exclusive scan of a[0..n) in LDS into out, returns the total; all threads */
__device__ int block_scan_lds(const int *a, int *out, int n, int *tmp /* >= 17 ints */) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int per = (n + kLT - 1) / kLT;
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
    for (int w = 0; w < kLT / 64; ++w) {
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
  SH_GEN, SH_CC, SH_NCAND, SH_NCORE, SH_BIG, SH_REBUILD, SH_NOVF,
  SH_QC0, SH_QC1, SH_QC2, SH_COUNT
};

__global__ __launch_bounds__(kLT) void k_propagate_local(PropArgs a, LocalArgs l, int k0,
                                                         int maxIter) {
  extern __shared__ int pool[];
  __shared__ int sh[SH_COUNT];
  __shared__ int tmp[17];
  DevCounters *c = a.ctr;
  const int tid = threadIdx.x, lane = tid & 63;
  const int gl = tid & (kLG - 1);
  const int P = l.ldsWords;
  const int stampBase = c->stampBase;
  int k = k0;
  int reason = 0;
  bool useCore = true;
  unsigned profU = 0, profE = 0, profC = 0, profA = 0;

  while (true) {  // one pass = build the LDS working set, iterate, write back
    /* every thread has read the previous pass's sh[] values */
    __syncthreads();
    /* ---- candidates: saved core list + frontier of iteration k ---------- */
    if (tid == 0) {
      const int gen = c->localGen + 1;
      c->localGen = gen;
      const int cc = useCore ? min(c->coreCount, l.coreListCap) : 0;
      sh[SH_GEN] = gen;
      sh[SH_CC] = cc;
      sh[SH_NCAND] = cc + c->qcount[k % 3];
      sh[SH_NCORE] = 0;
      sh[SH_BIG] = 0;
      sh[SH_REBUILD] = 0;
      sh[SH_NOVF] = 0;
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
    for (int i = tid; i < nCand; i += kLT) {
      const int g = i < cc ? l.coreList[i] : qk[i - cc];
      if (atomicMax(&l.coreGen[g], gen) < gen) {
        const int idx = atomicAdd(&sh[SH_NCORE], 1);
        gidL[idx] = g;
        l.coreIdx[g] = idx;
      }
    }
    wg_global_sync();
    const int nCore = sh[SH_NCORE];
    int *rlL = gidL + nCand;
    int *rsL = rlL + nCore;
    for (int i = tid; i < nCore; i += kLT) {
      const int g = gidL[i];
      rlL[i] = a.rowLen ? a.rowLen[g] : a.rowEnd[g] - a.rowBeg[g];
    }
    __syncthreads();
    const int E = block_scan_lds(rlL, rsL, nCore, tmp);
    if ((long long)nCand + 7LL * nCore + 4LL * E > P) {  // uniform
      if (useCore && cc > 0) {
        useCore = false;
        continue;
      }
      reason = DLP_LOCAL_TOO_BIG;
      break;
    }
    float *lab0L = (float *)(rsL + nCore);
    float *lab1L = lab0L + nCore;
    int *stampL = (int *)(lab1L + nCore);
    int *q0L = stampL + nCore;
    int *q1L = q0L + nCore;
    int *etgt = q1L + nCore;
    float *ew = (float *)(etgt + E);
    float *ehA = ew + E;
    float *ehB = ehA + E;
    /* ---- stage core labels, stamps and rows ------------------------------ */
    for (int i = tid; i < nCore; i += kLT) {
      const int g = gidL[i];
      lab0L[i] = a.lab0[g];
      lab1L[i] = a.lab1[g];
      stampL[i] = (a.iterStamp[g] == DLP_INT_MAX) ? DLP_INT_MAX : -1;
      const int beg = a.rowBeg[g], len = rlL[i], out = rsL[i];
      for (int p = 0; p < len; ++p) {
        const int v = a.colInd[beg + p];
        const int e = out + p;
        ew[e] = a.values[beg + p];
        if (v < 0 || (a.status && a.status[v] == DLP_VERTEX_DELETED)) {
          etgt[e] = -1;
        } else if (ld_agent(&l.coreGen[v]) == gen) {
          etgt[e] = ld_agent(&l.coreIdx[v]);
        } else {
          const int elig = a.iterStamp[v] != DLP_INT_MAX ? 1 : 0;
          etgt[e] = -2 - (2 * v + elig);
          ehA[e] = a.lab0[v];
          ehB[e] = a.lab1[v];
        }
      }
    }
    __syncthreads();
    /* ---- frontier of iteration k ----------------------------------------- */
    const int nq = nCand - cc;
    {
      int *qin = (k & 1) ? q1L : q0L;
      for (int i = tid; i < nq; i += kLT) {
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
      if (tid == 0) {
        sh[SH_QC0 + (k + 2) % 3] = 0;
        if (k == 0) c->frontier0 = n;
        if (a.iterSizes && k < a.iterSizesCap) a.iterSizes[k] = n;
        c->iterUsed = k + 1;
        c->updates += (unsigned long long)n;
        profU += (unsigned)n;
      }
      const int kb = k & 1;
      const int *qin = kb ? q1L : q0L;
      int *qout = kb ? q0L : q1L;
      const float *Fo = kb ? lab1L : lab0L;
      float *Fn = kb ? lab0L : lab1L;
      const float *eh = kb ? ehB : ehA;
      int *qcOut = &sh[SH_QC0 + (k + 1) % 3];
      const int stampNextL = k + 1;
      const int stampNextG = stampBase + k + 1;
      for (int base = 0; base < n; base += kLGroups) {  // block-uniform
        const int item = base + (tid >> 2);
        const bool act = item < n;
        const int ci = act ? qin[item] : 0;
        const int rs = act ? rsL[ci] : 0;
        const int len = act ? rlL[ci] : 0;
        float W, S;
        dlp::group_row_sums<kLG>(gl, len, LocalPos{etgt, ew, eh, Fo, rs}, W, S);
        int changed = 0;
        if (act && gl == 0) {
          const float oldLabel = Fo[ci];
          const float nl = dlp::new_label(W, S, oldLabel);
          Fn[ci] = nl;
          changed = fabsf(nl - oldLabel) > a.delta;
          profE += (unsigned)len;
          if (changed) profC += (unsigned)len + 1u;
        }
        changed = __shfl(changed, lane & ~(kLG - 1), 64);
        for (int p0 = 0; __any(changed && p0 <= len); p0 += kLG) {  // wave-uniform
          const int p = p0 + gl;
          int t = -1;
          if (changed && p < len) t = etgt[rs + p];
          else if (changed && p == len) t = ci;  // the vertex itself
          bool appL = false, appG = false, firstG = false;
          int gid = 0;
          if (t >= 0) {
            appL = atomicMax(&stampL[t], stampNextL) < stampNextL;
          } else if (t <= -2) {
            const int x = -2 - t;
            if (x & 1) {  // eligible vertex outside the core: rebuild after this iteration
              gid = x >> 1;
              const int old = atomicMax(&a.iterStamp[gid], stampNextG);
              appG = old < stampNextG;
              firstG = appG && old < stampBase;
            }
          }
          profA += (appL || appG) ? 1u : 0u;
          const int slot = dlp_wave_append(qcOut, appL);
          if (slot >= 0) qout[slot] = t;
          const int os = dlp_wave_append(&sh[SH_NOVF], appG);
          if (os >= 0) {
            if (os < l.ovfCap) l.ovf[os] = gid;
            sh[SH_REBUILD] = 1;
          }
          if (a.touched) {
            const int ts = dlp_wave_append(&c->touched, firstG);
            if (ts >= 0) a.touched[ts] = gid;
          }
        }
      }
      __syncthreads();
      ++k;
      if (sh[SH_REBUILD]) break;  // uniform after the barrier
    }

    /* ---- write back: labels, touched, next frontier, core list ---------- */
    for (int base = 0; base < nCore; base += kLT) {  // block-uniform
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
      for (int i = tid; i < nl; i += kLT) gq[i] = gidL[qloc[i]];
      for (int i = tid; i < no; i += kLT) gq[nl + i] = l.ovf[i];
      if (tid == 0) {
        c->qcount[k % 3] = nl + no;
        c->qcount[(k + 1) % 3] = 0;
        c->coreCount = min(nCore, l.coreListCap);
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
  if (tid == 0) {
    c->nextIter = k;
    c->localReason = reason;
  }
  if (a.profile) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      profE += __shfl_xor(profE, off, 64);
      profC += __shfl_xor(profC, off, 64);
      profA += __shfl_xor(profA, off, 64);
    }
    if (lane == 0) {
      if (profE) atomicAdd(&c->profLocalEdges, (unsigned long long)profE);
      if (profC) atomicAdd(&c->profLocalChangedEdges, (unsigned long long)profC);
      if (profA) atomicAdd(&c->profLocalAppended, (unsigned long long)profA);
    }
    if (tid == 0 && profU) atomicAdd(&c->profLocalUpdates, (unsigned long long)profU);
  }
}

}  // namespace

int dlp_launch_propagate(const PropArgs &a, int k, int groupLanes, int blocks, hipStream_t s) {
  dim3 grid(blocks > 0 ? blocks : 1), block(256);
  switch (groupLanes) {
    case 4: hipLaunchKernelGGL(k_propagate<4>, grid, block, 0, s, a, k); break;
    case 8: hipLaunchKernelGGL(k_propagate<8>, grid, block, 0, s, a, k); break;
    case 16: hipLaunchKernelGGL(k_propagate<16>, grid, block, 0, s, a, k); break;
    case 32: hipLaunchKernelGGL(k_propagate<32>, grid, block, 0, s, a, k); break;
    case 64: hipLaunchKernelGGL(k_propagate<64>, grid, block, 0, s, a, k); break;
    default:
      dlp_set_errorf_("propagation group size %d not in {4,8,16,32,64}", groupLanes);
      return DLP_ERR_INVALID;
  }
  DLP_LAUNCHED("k_propagate");
  return DLP_OK;
}

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
  DLP_HIP(hipFuncSetAttribute((const void *)k_propagate_local,
                              hipFuncAttributeMaxDynamicSharedMemorySize, v));
  *bytes = v;
  return DLP_OK;
}

int dlp_launch_propagate_local(const PropArgs &a, const LocalArgs &l, int k, int maxIter,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_propagate_local, dim3(1), dim3(kLT), (size_t)l.ldsWords * 4, s, a, l, k,
                     maxIter);
  DLP_LAUNCHED("k_propagate_local");
  return DLP_OK;
}
