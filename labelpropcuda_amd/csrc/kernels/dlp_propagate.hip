/*
 * dlp_propagate.hip -- one frontier-propagation iteration (the hot loop).
 *
 * Replaces, per iteration, propagateLabelsKernel + markNeighborsKernel +
 * the next iteration's collectAffectedKernel + the isAffected copy
 * (iterativePropagation.cu:63-217, :274-327) with a single launch whose work
 * is proportional to the frontier, not to N:
 *
 *   for u in queue_k (size qcount[k % 3]):
 *     W, S = weighted sums over row(u) reading F_old (buffer k & 1)
 *     F_new[u] = clamp(S / W)  (F_old[u] if W <= 1e-10f)   -> buffer (k+1) & 1
 *     if |F_new[u] - F_old[u]| > delta:
 *       every neighbour v whose stamp is below stampBase+k+1 and that may
 *       enter a frontier (stamp != INT_MAX) is appended once to queue_{k+1},
 *       and u itself.
 *
 * Work layout (gfx950, 64-lane waves): G lanes per frontier vertex, 64/G
 * vertices per wave.  Lane gl of a group reads row positions gl, gl+G, ...
 * (coalesced), so the row's 4-byte column ids and weights arrive in G-lane
 * segments; the label gather F_old[v] is the only random access per edge.
 *
 * Bit-exact reduction order: the reference sums a row with a 256-thread
 * block -- slot t accumulates positions t, t+256, ... in order, then
 * s[t] += s[t+stride] for stride = 128, 64, ..., 1
 * (iterativePropagation.cu:126-150).  Here slot t in [0,64) of the group is
 * owned by lane t % G as its register j = t / G, holding
 * (s[t] + s[t+128]) + (s[t+64] + s[t+192])   (strides 128 and 64);
 * strides 32 .. G are the in-register halving over j, strides G/2 .. 1 are
 * cross-lane shuffles.  Every addition pairs the same two partial sums as
 * the reference block, so the result is the same float.
 *
 * Queue appends and the touched list use wave-aggregated atomics (one
 * atomic per wave per list).  Deduplication is an atomicMax on a per-vertex
 * iteration stamp, which also carries eligibility (INT_MAX = never).
 */
#include "../dlp_kernels.h"

namespace {

struct Acc {
  float w, s;
};

/* slot x of the reference's 256-slot block: positions x, x+256, ... */
__device__ __forceinline__ Acc slot_sum(const PropArgs &a, const float *Fold, int beg, int len,
                                        int x) {
  Acc r{0.0f, 0.0f};
  for (int p = x; p < len; p += 256) {
    int v = a.colInd[beg + p];
    if (v < 0) continue;
    if (a.status && a.status[v] == DLP_VERTEX_DELETED) continue;
    float w = a.values[beg + p];
    float prod = w * Fold[v];
    r.w = r.w + w;
    r.s = r.s + prod;
  }
  return r;
}

template <int G>
__global__ __launch_bounds__(256) void k_propagate(PropArgs a, int k) {
  constexpr int J = 64 / G;    /* slots per lane (slots 0..63)   */
  constexpr int GPW = 64 / G;  /* frontier vertices per wave     */
  DevCounters *c = a.ctr;
  const int n = c->qcount[k % 3];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c->qcount[(k + 2) % 3] = 0;  /* consumed by iteration k-1; next iteration appends here */
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

  unsigned profE = 0, profC = 0, profA = 0;  /* lane-local work counters */
  for (long long base = gwave * GPW; base < n; base += nwaves * GPW) {  // wave-uniform
    const long long item = base + grp;
    const bool act = item < n;
    int u = 0, beg = 0, len = 0;
    if (act) {
      u = in[item];
      beg = a.rowBeg[u];
      len = a.rowLen ? a.rowLen[u] : a.rowEnd[u] - beg;
    }

    float ws[J], ss[J];
    int col[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int p = gl + G * j;
      float w = 0.0f, sv = 0.0f;
      int v = -1;
      if (p < len) {
        v = a.colInd[beg + p];
        if (v >= 0 && !(a.status && a.status[v] == DLP_VERTEX_DELETED)) {
          w = a.values[beg + p];
          sv = w * Fold[v];
        }
      }
      col[j] = v;
      ws[j] = w;
      ss[j] = sv;
    }
    if (len > 64) {  // rows longer than 64: full 256-slot form (rare)
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int t = gl + G * j;
        Acc s0 = slot_sum(a, Fold, beg, len, t);
        Acc s1 = slot_sum(a, Fold, beg, len, t + 64);
        Acc s2 = slot_sum(a, Fold, beg, len, t + 128);
        Acc s3 = slot_sum(a, Fold, beg, len, t + 192);
        ws[j] = (s0.w + s2.w) + (s1.w + s3.w);
        ss[j] = (s0.s + s2.s) + (s1.s + s3.s);
      }
    }
    /* strides 32 .. G: in-register halving over j */
#pragma unroll
    for (int h = J / 2; h >= 1; h >>= 1) {
#pragma unroll
      for (int j = 0; j < h; ++j) {
        ws[j] = ws[j] + ws[j + h];
        ss[j] = ss[j] + ss[j + h];
      }
    }
    float W = ws[0], S = ss[0];
    /* strides G/2 .. 1: across the group's lanes */
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) {
      W = W + __shfl_down(W, off, G);
      S = S + __shfl_down(S, off, G);
    }

    int changed = 0;
    if (act && gl == 0) {
      const float oldLabel = Fold[u];
      float nl = (W > 1e-10f) ? S / W : oldLabel;
      nl = fminf(fmaxf(nl, 0.0f), 1.0f);
      Fnew[u] = nl;
      changed = fabsf(nl - oldLabel) > a.delta;
    }
    changed = __shfl(changed, grp * G, 64);
    if (act && gl == 0) {
      profE += (unsigned)len;
      if (changed) profC += (unsigned)len + 1u;
    }

    /* next frontier: neighbours of changed vertices, and the vertex itself */
#pragma unroll
    for (int j = 0; j < J + 1; ++j) {
      int v;
      bool pred;
      if (j < J) {
        const int p = gl + G * j;
        v = col[j];
        pred = changed && p < len && v >= 0;
      } else {
        v = u;
        pred = changed && gl == 0;
      }
      int old = pred ? atomicMax(&a.iterStamp[v], stampNext) : DLP_INT_MAX;
      const bool app = pred && old < stampNext;
      profA += app ? 1u : 0u;
      int slot = dlp_wave_append(outCount, app);
      if (slot >= 0) out[slot] = v;
      if (a.touched) {
        int ts = dlp_wave_append(&c->touched, app && old < stampBase);
        if (ts >= 0) a.touched[ts] = v;
      }
    }
    if (__any(changed && len > 64)) {  // wave-uniform
      for (int p0 = 64;; p0 += G) {
        const int p = p0 + gl;
        const bool more = changed && p < len;
        if (!__any(more)) break;
        const int v = more ? a.colInd[beg + p] : -1;
        const bool pred = more && v >= 0;
        int old = pred ? atomicMax(&a.iterStamp[v], stampNext) : DLP_INT_MAX;
        const bool app = pred && old < stampNext;
        profA += app ? 1u : 0u;
        int slot = dlp_wave_append(outCount, app);
        if (slot >= 0) out[slot] = v;
        if (a.touched) {
          int ts = dlp_wave_append(&c->touched, app && old < stampBase);
          if (ts >= 0) a.touched[ts] = v;
        }
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
