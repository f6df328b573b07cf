/*
 * dlp_batch.hip -- per-batch kernels of the DynLP update, before and after
 * the propagation loop, for the device-resident pipeline and for the step
 * API (standard CSR input).
 *
 * Reference steps and where they land here:
 *   updateHostCSR status order, GT setup     dynLP.cu:118-130, :301-305
 *       -> k_clear_prev, k_apply_*, k_finalize (work proportional to the
 *          batch's lists, not to N)
 *   host CSR rebuild + markDeletions + cleanIncomingEdges
 *                                            dynLP.cu:134-166,
 *                                            sparsifySubgraph.cu:58-111
 *       -> k_csr_update: one pass over the full graph, 8 lanes per row,
 *          ballot + mbcnt compaction into the row's fixed slot; entries that
 *          point at vertices deleted in this batch become -1 holes at the
 *          position the reference leaves them
 *   sparsify                                 sparsifySubgraph.cu:119-179
 *       -> order-preserving scan compaction (G' in input-edge order)
 *   Shiloach-Vishkin CC                      connectedComponents.cu:64-290
 *       -> atomicMin hooking of roots with retry (larger root under smaller,
 *          so every root is its component's minimum id) + pointer jumping to
 *          the root; ranks by a scan over the ascending inserted list
 *   label initialisation                     labelInitialization.cu:59-187
 *       -> per-member row sums, stable radix grouping by component, and a
 *          wave-level bottom-up pairwise tree per component
 *   collect / GT enforcement / accuracy      iterativePropagation.cu:63-79,
 *                                            :224-235, dynLP.cu:65-85
 */
#include "../dlp_kernels.h"
#include "../dlp_scan.h"

namespace {

constexpr int kT = 256;

/* ------------------------------------------------------------------------ */
/* shared device helpers                                                    */
/* ------------------------------------------------------------------------ */
struct RowsSlack {
  const int *beg, *len;
  __device__ void get(int u, int &b, int &l) const {
    b = beg[u];
    l = len[u];
  }
};
struct RowsStd {
  const int *rowPtr;
  __device__ void get(int u, int &b, int &l) const {
    b = rowPtr[u];
    l = rowPtr[u + 1] - b;
  }
};
struct GtFlags {
  const uint8_t *f;
  __device__ int cls(int v) const {
    uint8_t x = f[v];
    return (x & VF_GT0) ? 0 : ((x & VF_GT1) ? 1 : -1);
  }
};
struct GtInts {
  const int *g;
  __device__ int cls(int v) const { return g[v]; }
};

__device__ __forceinline__ bool bit_test(const uint32_t *bits, int v) {
  return (bits[v >> 5] >> (v & 31)) & 1u;
}

__device__ __forceinline__ int par_load(const int *par, int x) {
  return __hip_atomic_load(par + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* root of x: parents only decrease along a path (par[x] <= x), so the walk
 * ends; a parent out of range flags an error */
__device__ __forceinline__ int uf_find(const int *par, int x, int *err) {
  int p = par_load(par, x);
  while (p != x) {
    if (p < 0 || p > x) {
      atomicOr(err, 1);
      return x;
    }
    x = p;
    p = par_load(par, x);
  }
  return x;
}

/* hook the two trees: atomicMin(par[larger root], smaller root); if the
 * larger root was already hooked elsewhere, continue with its parent.  The
 * larger root strictly decreases every retry. */
__device__ void uf_union(int *par, int a, int b, int *err) {
  for (int guard = 0; guard < (1 << 26); ++guard) {
    int ra = uf_find(par, a, err), rb = uf_find(par, b, err);
    if (ra == rb) return;
    int hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
    int old = atomicMin(&par[hi], lo);
    if (old == hi) return;
    a = old;
    b = lo;
  }
  atomicOr(err, 2);
}

/* ------------------------------------------------------------------------ */
/* pipeline: state reset and per-batch status                              */
/* ------------------------------------------------------------------------ */
__global__ void k_reset(DState d) {
  long long i = (long long)blockIdx.x * kT + threadIdx.x;
  long long stride = (long long)gridDim.x * kT;
  for (long long v = i; v < d.N; v += stride) {
    d.hStatus[v] = DLP_VERTEX_DELETED;
    d.dStatus[v] = DLP_VERTEX_DELETED;
    d.gtClass[v] = -1;
    d.vflags[v] = 0;
    d.lab0[v] = 0.5f;
    d.lab1[v] = 0.5f;
    d.iterStamp[v] = DLP_INT_MAX;
    d.markStamp[v] = 0;
    d.compId[v] = -1;
    d.aLen[v] = 0;
    d.aRow[v] = make_int2(d.fRow[v], 0);
    d.rowStamp[v] = 0;
    d.accState[v] = 0;
  }
  for (long long w = i; w < d.numWords; w += stride) d.newBits[w] = 0u;
  if (i == 0) {
    DevCounters z = {};
    z.stampBase = 1;
    z.profEdges = d.ctr->profEdges;  // work counters survive resets
    z.profChangedEdges = d.ctr->profChangedEdges;
    z.profAppended = d.ctr->profAppended;
    z.profLocalUpdates = d.ctr->profLocalUpdates;
    z.profLocalEdges = d.ctr->profLocalEdges;
    z.profLocalChangedEdges = d.ctr->profLocalChangedEdges;
    z.profLocalAppended = d.ctr->profLocalAppended;
    z.profLocalOwnerUpdates = d.ctr->profLocalOwnerUpdates;
    z.profLocalBuilds = d.ctr->profLocalBuilds;
    z.profLocalWaveUpdates = d.ctr->profLocalWaveUpdates;
    z.localGen = d.ctr->localGen;  // the core generations stored per vertex stay valid
    *d.ctr = z;
  }
}

/* wrap-around guards: eligible stamps (not INT_MAX) back to 0 under a
 * stampBase of 1; every core / halo generation back to 0 with localGen */
__global__ void k_rebase_stamps(int *iterStamp, long long n, DevCounters *ctr) {
  const long long i = (long long)blockIdx.x * kT + threadIdx.x;
  for (long long v = i; v < n; v += (long long)gridDim.x * kT)
    if (iterStamp[v] != DLP_INT_MAX) iterStamp[v] = 0;
  if (i == 0) ctr->stampBase = 1;
}

__global__ void k_restart_gens(int *coreGen, int *haloGen, long long n, DevCounters *ctr) {
  const long long i = (long long)blockIdx.x * kT + threadIdx.x;
  for (long long v = i; v < n; v += (long long)gridDim.x * kT) {
    coreGen[v] = 0;
    haloGen[v] = 0;
  }
  if (i == 0) ctr->localGen = 0;
}

/* zero per-batch counters; clear the previous batch's per-batch flags */
__global__ void k_clear_prev(DState d, DBatch bt) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i == 0) {
    DevCounters *c = d.ctr;
    c->qcount[0] = c->qcount[1] = c->qcount[2] = 0;
    c->numCand = 0;
    c->touched = 0;
    c->iterUsed = 0;
    c->subPairs = 0;
    c->numComponents = 0;
    c->numNew = 0;
    c->frontier0 = 0;
    c->updates = 0;
    c->correct = 0;
    c->total = 0;
    c->coreCount = 0;
    c->rowCount = 0;
  }
  const int n1 = bt.nPrevIns, n2 = n1 + bt.nPrevDel, n3 = n2 + bt.nPrevGt;
  if (i >= n3) return;
  int v = i < n1 ? bt.prevIns[i] : (i < n2 ? bt.prevDel[i - n1] : bt.prevGt[i - n2]);
  d.vflags[v] &= (uint8_t)~(VF_DDEL | VF_NEW | VF_GTNOW);
  if (i < n1) atomicAnd(&d.newBits[v >> 5], ~(1u << (v & 31)));
}

__global__ void k_apply_deleted(DState d, DBatch bt) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= bt.nDel) return;
  int v = bt.del[i];
  d.hStatus[v] = DLP_VERTEX_DELETED;
  d.vflags[v] |= VF_DDEL;
}

__global__ void k_apply_inserted(DState d, DBatch bt) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= bt.nIns) return;
  int v = bt.ins[i];
  d.hStatus[v] = DLP_VERTEX_ACTIVE;
  d.vflags[v] |= VF_NEW;
  atomicOr(&d.newBits[v >> 5], 1u << (v & 31));
}

__global__ void k_apply_gt(DState d, DBatch bt) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= bt.nGt) return;
  int v = bt.gtIds[i];
  float lab = bt.gtLab[i];
  d.hStatus[v] = DLP_VERTEX_GROUND_TRUTH;
  d.gtClass[v] = (lab < 0.5f) ? 0 : 1;
  d.lab0[v] = lab;
  d.lab1[v] = lab;
  d.vflags[v] |= VF_GTNOW;
}

/* device status, CSR membership, eligibility and stamps of every vertex a
 * list of this batch or the previous batch's deletion list names */
__global__ void k_finalize(DState d, DBatch bt) {
  const int i = blockIdx.x * kT + threadIdx.x;
  const int n1 = bt.nPrevDel, n2 = n1 + bt.nDel, n3 = n2 + bt.nIns, n4 = n3 + bt.nGt;
  if (i >= n4) return;
  int v = i < n1 ? bt.prevDel[i]
                 : (i < n2 ? bt.del[i - n1] : (i < n3 ? bt.ins[i - n2] : bt.gtIds[i - n3]));
  const int hs = d.hStatus[v];
  const uint8_t f = d.vflags[v];
  const int ds = (f & VF_DDEL) ? DLP_VERTEX_DELETED : hs;
  d.dStatus[v] = ds;
  const int g = d.gtClass[v];
  const int bm = d.batchMap[v];
  const bool hact = hs != DLP_VERTEX_DELETED && bm >= 0 && bm <= bt.b;
  const bool elig = ds == DLP_VERTEX_ACTIVE && g < 0;
  uint8_t nf = (uint8_t)(f & (VF_DDEL | VF_NEW | VF_GTNOW));
  if (hact) nf |= VF_HACT;
  if (elig) nf |= VF_ELIG;
  if (g == 0) nf |= VF_GT0;
  if (g == 1) nf |= VF_GT1;
  if (hs == DLP_VERTEX_ACTIVE) nf |= VF_HSACT;
  d.vflags[v] = nf;
  if (!elig) d.iterStamp[v] = DLP_INT_MAX;
  else if (d.iterStamp[v] == DLP_INT_MAX) d.iterStamp[v] = 0;
}

/* ------------------------------------------------------------------------ */
/* pipeline: CSR update with ballot compaction (8 lanes per row)            */
/* ------------------------------------------------------------------------ */
/* Row u of the active CSR (dynLP.cu:134-166) rebuilt from its full row by a
 * group of G = 8 lanes: entries naming a vertex that is present (host
 * status not DELETED, arrived) are compacted in order into the row's fixed
 * slot; entries naming a vertex deleted in this batch, and every entry of
 * the row of such a vertex, become -1 holes (cleanIncomingEdges,
 * sparsifySubgraph.cu:58-111), and the neighbours named by such a row are
 * marked (markDeletions).  All 64 lanes of the wave call it; act and u are
 * group-uniform. */
__device__ __forceinline__ void csr_rebuild_row(const DState &d, int markTag, bool act, int u,
                                                int gl, int grp) {
  constexpr int G = 8;
  const uint8_t fu = act ? d.vflags[u] : 0;
  const bool q1 = (fu & VF_DDEL) != 0;  // row of a vertex deleted in this batch
  const int fb = act ? d.fRow[u] : 0;
  const int flen = (act && (fu & VF_HACT)) ? d.fRow[u + 1] - fb : 0;
  int written = 0;
  for (int c0 = 0; __any(c0 < flen); c0 += G) {
    const int p = c0 + gl;
    const bool valid = p < flen;
    const int v = valid ? d.fCol[fb + p] : -1;
    const uint8_t fv = (v >= 0) ? d.vflags[v] : 0;
    const bool keep = valid && v >= 0 && (fv & VF_HACT);
    const unsigned long long m = dlp_ballot(keep);
    const unsigned gm = (unsigned)(m >> (grp * G)) & ((1u << G) - 1u);
    if (keep) {
      const int pos = fb + written + __popc(gm & ((1u << gl) - 1u));
      d.aCol[pos] = (q1 || (fv & VF_DDEL)) ? -1 : v;
      d.aVal[pos] = d.fVal[fb + p];
    }
    dlp_mark_candidate(d.markStamp, markTag, d.cand, &d.ctr->numCand, v, keep && q1);
    written += __popc(gm);
  }
  if (act && gl == 0) {
    d.aLen[u] = written;
    d.aRow[u] = make_int2(fb, written);
  }
}

/* every row of the graph, 8 rows per wave */
__global__ __launch_bounds__(kT) void k_csr_update(DState d, int markTag) {
  constexpr int G = 8;
  const int lane = threadIdx.x & 63, gl = lane & (G - 1), grp = lane / G;
  const long long gwave = ((long long)blockIdx.x * kT + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * kT) >> 6;
  for (long long rb = gwave * (64 / G); rb < d.N; rb += nwaves * (64 / G)) {  // wave-uniform
    const long long uu = rb + grp;
    const bool act = uu < d.N;
    csr_rebuild_row(d, markTag, act, act ? (int)uu : 0, gl, grp);
  }
}

/* The rows a batch changes.  A row's content depends on whether its vertex
 * and the vertices it names are present and whether they are deleted in
 * this batch; those change only for the vertices of the batch's lists
 * (deleted, inserted, ground truth) and of the previous batch's deletion
 * list.  On a symmetric full graph the rows naming such a vertex x are the
 * rows of x's full-graph neighbours, so the rows to rebuild are the listed
 * vertices' own rows and their neighbours' rows.  8 lanes per listed
 * vertex; each row once per batch (rowStamp = the batch's tag) into
 * rows[0 .. ctr->rowCount). */
__global__ __launch_bounds__(kT) void k_csr_rows(DState d, DBatch bt, int tag, int *rows) {
  constexpr int G = 8;
  const int lane = threadIdx.x & 63, gl = lane & (G - 1), grp = lane / G;
  const long long n1 = bt.nPrevDel, n2 = n1 + bt.nDel, n3 = n2 + bt.nIns, n4 = n3 + bt.nGt;
  const long long gwave = ((long long)blockIdx.x * kT + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * kT) >> 6;
  for (long long xb = gwave * (64 / G); xb < n4; xb += nwaves * (64 / G)) {  // wave-uniform
    const long long i = xb + grp;
    const bool act = i < n4;
    int x = 0;
    if (act)
      x = i < n1 ? bt.prevDel[i]
                 : (i < n2 ? bt.del[i - n1] : (i < n3 ? bt.ins[i - n2] : bt.gtIds[i - n3]));
    const int fb = act ? d.fRow[x] : 0;
    const int flen = act ? d.fRow[x + 1] - fb : 0;
    /* position -1 is the vertex itself, 0 .. flen-1 its full row */
    for (int c0 = -1; __any(act && c0 < flen); c0 += G) {
      const int p = c0 + gl;
      int v = -1;
      if (act && p < flen) v = p < 0 ? x : d.fCol[fb + p];
      const bool fresh = v >= 0 && atomicMax(&d.rowStamp[v], tag) < tag;
      const int slot = dlp_wave_append(&d.ctr->rowCount, fresh);
      if (slot >= 0) rows[slot] = v;
    }
  }
}

/* the listed rows rows[0 .. ctr->rowCount), 8 rows per wave */
__global__ __launch_bounds__(kT) void k_csr_update_rows(DState d, int markTag, const int *rows) {
  constexpr int G = 8;
  const int lane = threadIdx.x & 63, gl = lane & (G - 1), grp = lane / G;
  const long long n = d.ctr->rowCount;
  const long long gwave = ((long long)blockIdx.x * kT + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * kT) >> 6;
  for (long long rb = gwave * (64 / G); rb < n; rb += nwaves * (64 / G)) {  // wave-uniform
    const long long i = rb + grp;
    const bool act = i < n;
    csr_rebuild_row(d, markTag, act, act ? rows[i] : 0, gl, grp);
  }
}

/* ------------------------------------------------------------------------ */
/* sparsification (both paths)                                              */
/* ------------------------------------------------------------------------ */
struct NewBits {
  const uint32_t *bits;
  __device__ bool operator()(int v) const { return bit_test(bits, v); }
};
struct NewInts {
  const int *flags;
  __device__ bool operator()(int v) const { return flags[v] != 0; }
};

/* One pass of LDS-staged stream compaction: tiles of kSpTile edges, one
 * work-group each, taken in launch order from a device counter.  Row j of a
 * tile is edges tileBase + 256 j + t (coalesced loads); a wave's kept edges
 * of a row are packed by ballot and rank, the counts of every (row, wave)
 * are exchanged and scanned in LDS, so each kept edge knows its place in
 * the tile in input order, and the kept edges are staged there in that
 * order.  The tile's offset among all tiles comes from a decoupled
 * look-back over per-tile status words (aggregate, then inclusive prefix):
 * a tile waits only on tiles that took their number before it and are
 * running, and it gives up after kSpSpinMax polls (device error flag)
 * rather than wait for ever.  The staged edges then leave as one contiguous
 * run per tile, each as its two mirrored entries; every heavy edge's
 * endpoints are marked; the last tile stores the number of kept edges. */
constexpr int kSpRows = 8;
constexpr int kSpTile = kSpRows * kT; /* 2048 edges */
constexpr int kSpWaves = kT / 64;
constexpr int kSpSpinMax = 1 << 22;

__device__ __forceinline__ unsigned long long sp_status(unsigned flag, int value) {
  return ((unsigned long long)flag << 62) | (unsigned)value;
}

/* pipeline output: mirrored pairs as 8-byte stores, candidate marks */
struct SpOutPipe {
  int *subSrc, *subDst;
  float *subW;
  int numInserted;
  int *markStamp, *cand, *numCand;
  int markTag;
  __device__ void write(int q, int u, int v, float w) const {
    reinterpret_cast<int2 *>(subSrc)[q] = make_int2(u, v);
    reinterpret_cast<int2 *>(subDst)[q] = make_int2(v, u);
    reinterpret_cast<float2 *>(subW)[q] = make_float2(w, w);
  }
  /* every lane of the wave calls it (wave-aggregated appends) */
  __device__ void mark(int x, bool heavy) const {
    dlp_mark_candidate(markStamp, markTag, cand, numCand, x,
                       heavy && x >= 0 && x < numInserted);
  }
};

/* step-API output (the reference's arrays: destinations and weights, sources
 * optional; caller buffers, so 4-byte stores) and isAffected marks */
struct SpOutStd {
  int *subSrc, *subDst;
  float *subW;
  int numInserted;
  int *isAffected;
  __device__ void write(int q, int u, int v, float w) const {
    if (subSrc) {
      subSrc[2 * q] = u;
      subSrc[2 * q + 1] = v;
    }
    subDst[2 * q] = v;
    subDst[2 * q + 1] = u;
    subW[2 * q] = w;
    subW[2 * q + 1] = w;
  }
  __device__ void mark(int x, bool heavy) const {
    if (heavy && x >= 0 && x < numInserted) atomicMax(&isAffected[x], 1);
  }
};

template <class IsNew, class Out>
__global__ __launch_bounds__(kT) void k_sparsify_pass(const int *src, const int *dst,
                                                      const float *wt, int n, float tau,
                                                      IsNew isNew, Out out,
                                                      unsigned long long *tileState,
                                                      int *tileCounter, int *total, int *err) {
  __shared__ int cnt[kSpRows * kSpWaves]; /* kept edges per (row, wave), then offsets */
  __shared__ int tileInfo[3];             /* tile number, tile offset, kept edges */
  __shared__ int stU[kSpTile], stV[kSpTile];
  __shared__ float stW[kSpTile];          /* the tile's kept edges, in input order */
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (t == 0) tileInfo[0] = atomicAdd(tileCounter, 1);
  __syncthreads();
  const int tile = tileInfo[0];
  const int tiles = (int)(((long long)n + kSpTile - 1) / kSpTile);
  const long long base = (long long)tile * kSpTile;
  int u[kSpRows], v[kSpRows];
  float w[kSpRows];
  unsigned long long keep[kSpRows];
  bool heavy[kSpRows];
#pragma unroll
  for (int j = 0; j < kSpRows; ++j) {
    const long long i = base + (long long)j * kT + t;
    const bool valid = i < n;
    u[j] = valid ? src[i] : -1;
    v[j] = valid ? dst[i] : -1;
    w[j] = valid ? wt[i] : 0.0f;
    /* the reference's test, !(w <= tau) (sparsifySubgraph.cu:163), keeps a
     * NaN weight too */
    heavy[j] = valid && !(w[j] <= tau);
    keep[j] = dlp_ballot(heavy[j] && isNew(u[j]) && isNew(v[j]));
    if (lane == 0) cnt[j * kSpWaves + wave] = __popcll(keep[j]);
  }
  __syncthreads();
  if (wave == 0) {
    /* exclusive offsets of the (row, wave) groups in input order */
    constexpr int G = kSpRows * kSpWaves; /* 32 */
    const int x = lane < G ? cnt[lane] : 0;
    int incl = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    if (lane < G) cnt[lane] = incl - x;
    const int agg = __shfl(incl, G - 1, 64);
    /* publish the aggregate, look back for the prefix, publish the prefix */
    if (lane == 0)
      __hip_atomic_store(&tileState[tile], sp_status(tile == 0 ? 2u : 1u, agg), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
    int excl = 0;
    int p = tile - 1; /* newest predecessor not yet summed */
    int spins = 0;
    while (p >= 0) {  // wave-uniform
      const int q = p - lane;
      const unsigned long long st =
          q >= 0 ? __hip_atomic_load(&tileState[q], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)
                 : sp_status(2u, 0);
      const unsigned flag = (unsigned)(st >> 62);
      const unsigned long long incM = dlp_ballot(flag == 2u);
      const int first = incM ? __ffsll((long long)incM) - 1 : 64; /* nearest prefix */
      const unsigned long long upTo = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
      if (dlp_ballot(flag == 0u) & upTo) {  // a predecessor has published nothing yet
        if (++spins > kSpSpinMax) {
          if (lane == 0) atomicOr(err, 16);
          break;
        }
        continue;
      }
      int val = lane <= first ? (int)(unsigned)st : 0;
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) val += __shfl_xor(val, o, 64);
      excl += val;
      if (first < 64) break;
      p -= 64;
    }
    if (lane == 0) {
      if (tile > 0)
        __hip_atomic_store(&tileState[tile], sp_status(2u, excl + agg), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
      if (tile == tiles - 1) *total = excl + agg;
      tileInfo[1] = excl;
      tileInfo[2] = agg;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSpRows; ++j) {
    if ((keep[j] >> lane) & 1ull) {
      const int pos = cnt[j * kSpWaves + wave] + dlp_rank_below(keep[j]);
      stU[pos] = u[j];
      stV[pos] = v[j];
      stW[pos] = w[j];
    }
    out.mark(u[j], heavy[j]);
    out.mark(v[j], heavy[j]);
  }
  __syncthreads();
  const int off = tileInfo[1], kept = tileInfo[2];
  for (int p = t; p < kept; p += kT) out.write(off + p, stU[p], stV[p], stW[p]);
}

/* ------------------------------------------------------------------------ */
/* connected components (both paths)                                        */
/* ------------------------------------------------------------------------ */
__global__ void k_cc_init(int *par, const int *list, int n) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i < n) par[list[i]] = list[i];
}

/* pairs == true: edges come as (2p, 2p+1) mirrored pairs, hook once per pair;
 * the count is read on the device */
__global__ void k_cc_hook(int *par, const int *src, const int *dst, const int *countDev,
                          int countHost, int pairs, int *err) {
  const int n = countDev ? *countDev : countHost;
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n;
       i += (long long)gridDim.x * kT) {
    long long e = pairs ? 2 * i : i;
    int u = src[e], v = dst[e];
    if (par_load(par, u) < 0 || par_load(par, v) < 0) {
      atomicOr(err, 4);  // endpoint that is not an inserted vertex
      continue;
    }
    uf_union(par, u, v, err);
  }
}

__global__ void k_cc_compress(int *par, const int *list, int n, int *err) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  int v = list[i];
  par[v] = uf_find(par, v, err);
}

struct RootGet {
  const int *par, *list;
  __device__ int operator()(int i) const { return par[list[i]] == list[i] ? 1 : 0; }
};
struct RootEmit {
  const int *par, *list;
  int *compId;
  __device__ void operator()(int i, int prefix, bool valid) const {
    if (valid && par[list[i]] == list[i]) compId[list[i]] = prefix;
  }
};

/* non-roots take their root's rank; with an empty G' the pipeline numbers
 * components by position in the batch's inserted list (dynLP.cu:441-452) */
__global__ void k_cc_assign(const int *par, const int *sorted, const int *ins, int n,
                            int *compId, const int *subPairsDev, int *numComponents) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  if (subPairsDev && *subPairsDev == 0) {
    compId[ins[i]] = i;
    if (i == 0) *numComponents = n;
    return;
  }
  int v = sorted[i];
  int r = par[v];
  if (r != v) compId[v] = compId[r];
}

__global__ void k_fill_int(int *p, int value, long long n) {
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n;
       i += (long long)gridDim.x * kT)
    p[i] = value;
}
__global__ void k_fill_float(float *p, float value, long long n) {
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n;
       i += (long long)gridDim.x * kT)
    p[i] = value;
}
__global__ void k_copy_float(float *dst, const float *src, long long n) {
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n;
       i += (long long)gridDim.x * kT)
    dst[i] = src[i];
}
__global__ void k_scatter_int(int *a, const int *list, int n, int value) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i < n) a[list[i]] = value;
}
__global__ void k_set_bits(uint32_t *bits, const int *list, int n) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i < n) atomicOr(&bits[list[i] >> 5], 1u << (list[i] & 31));
}

/* ------------------------------------------------------------------------ */
/* label initialisation (both paths)                                        */
/* ------------------------------------------------------------------------ */

/* member i of the ascending inserted list: left-to-right row sums of the
 * weights to GT class 0 / 1 (soft-deleted entries skipped,
 * labelInitialization.cu:76-87); key = component (numComponents when the
 * vertex has none, so it sorts after every real component) */
template <class Rows, class Gt>
__global__ void k_li_partials(Rows rows, const int *colInd, const float *values, Gt gt,
                              const int *sorted, int n, const int *compId, int numCompHost,
                              const int *numCompDev, float *p0, float *p1, int *key,
                              int *val) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const int numComp = numCompDev ? *numCompDev : numCompHost;
  const int u = sorted[i];
  const int c = compId[u];
  float a0 = 0.0f, a1 = 0.0f;
  if (c >= 0) {
    int b, l;
    rows.get(u, b, l);
    for (int e = b; e < b + l; ++e) {
      int v = colInd[e];
      if (v < 0) continue;
      int g = gt.cls(v);
      if (g == 0) a0 = a0 + values[e];
      else if (g == 1) a1 = a1 + values[e];
    }
  }
  p0[i] = a0;
  p1[i] = a1;
  key[i] = (c >= 0 && c < numComp) ? c : numComp;
  val[i] = i;
}

__global__ void k_li_segments(const int *key, int n, const int *numCompDev, int numCompHost,
                              int *segStart, int *segEnd) {
  const int j = blockIdx.x * kT + threadIdx.x;
  if (j >= n) return;
  const int numComp = numCompDev ? *numCompDev : numCompHost;
  const int c = key[j];
  if (c >= numComp) return;
  if (j == 0 || key[j - 1] != c) segStart[c] = j;
  if (j == n - 1 || key[j + 1] != c) segEnd[c] = j + 1;
}

/* one wave per component: bottom-up pairwise tree (stride 1, 2, 4, ...)
 * over the members' sums in ascending vertex id.  Within a 64-member chunk
 * the strides 1..32 are shuffles; chunk results form the next level, which
 * is the same tree (every stride >= 64 pairs chunk results). */
__global__ __launch_bounds__(kT) void k_li_reduce(const int *segStart, const int *segEnd,
                                                  const int *numCompDev, int numCompHost,
                                                  const int *order, const float *p0,
                                                  const float *p1, float *red0, float *red1,
                                                  float *W0, float *W1) {
  const int numComp = numCompDev ? *numCompDev : numCompHost;
  const int lane = threadIdx.x & 63;
  const long long gwave = ((long long)blockIdx.x * kT + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * kT) >> 6;
  for (long long c = gwave; c < numComp; c += nwaves) {
    const int s = segStart[c], n = segEnd[c] - s;
    int nl = n;
    bool lvl0 = true;
    float x0 = 0.0f, x1 = 0.0f;
    while (true) {
      const int nch = (nl + 63) / 64;
      for (int ch = 0; ch < nch; ++ch) {
        const int i = ch * 64 + lane;
        x0 = 0.0f;
        x1 = 0.0f;
        if (i < nl) {
          if (lvl0) {
            int m = order[s + i];
            x0 = p0[m];
            x1 = p1[m];
          } else {
            x0 = red0[s + i];
            x1 = red1[s + i];
          }
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          x0 = x0 + __shfl_down(x0, off, 64);
          x1 = x1 + __shfl_down(x1, off, 64);
        }
        if (nch > 1 && lane == 0) {
          red0[s + ch] = x0;
          red1[s + ch] = x1;
        }
      }
      if (nch <= 1) break;
      __threadfence();
      nl = nch;
      lvl0 = false;
    }
    if (lane == 0) {
      W0[c] = x0;
      W1[c] = x1;
    }
  }
}

/* initializeLabelsKernel (labelInitialization.cu:99-142); writes every
 * buffer given (the pipeline keeps two equal buffers between batches) */
template <class Gt>
__global__ void k_li_labels(const int *ins, int n, Gt gt, const int *compId, const float *W0,
                            const float *W1, float *lab0, float *lab1) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const int u = ins[i];
  const int g = gt.cls(u);
  float lab;
  if (g == 0) lab = 0.0f;
  else if (g == 1) lab = 1.0f;
  else {
    const int c = compId[u];
    if (c < 0) lab = 0.5f;
    else {
      const float den = W0[c] + W1[c];
      lab = (den > 1e-10f) ? W1[c] / den : 0.5f;
    }
  }
  lab0[u] = lab;
  if (lab1) lab1[u] = lab;
}

/* new vertices without new edges (dynLP.cu:482-504) */
__global__ void k_labels_no_edges(DState d, DBatch bt) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= bt.nIns) return;
  const int v = bt.ins[i];
  float lab = 0.5f;
  if (d.gtClass[v] >= 0 && (d.vflags[v] & VF_GTNOW)) lab = d.lab0[v];
  d.lab0[v] = lab;
  d.lab1[v] = lab;
}

/* ------------------------------------------------------------------------ */
/* frontier seeding, batch end                                              */
/* ------------------------------------------------------------------------ */
__global__ __launch_bounds__(kT) void k_frontier0(DState d) {
  DevCounters *c = d.ctr;
  const int n = c->numCand;
  const int base = c->stampBase;
  const long long gwave = ((long long)blockIdx.x * kT + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * kT) >> 6;
  const int lane = threadIdx.x & 63;
  for (long long i0 = gwave * 64; i0 < n; i0 += nwaves * 64) {  // wave-uniform
    const long long i = i0 + lane;
    const int v = i < n ? d.cand[i] : 0;
    const bool ok = i < n && (d.vflags[v] & VF_ELIG);
    if (ok) d.iterStamp[v] = base;
    int slot = dlp_wave_append(&c->qcount[0], ok);
    if (slot >= 0) d.q0[slot] = v;
    int ts = dlp_wave_append(&c->touched, ok);
    if (ts >= 0) d.touched[ts] = v;
  }
}

/* make the other label buffer equal to the final one where propagation
 * wrote (both were equal on entry) */
__global__ void k_sync_touched(DState d) {
  const DevCounters *c = d.ctr;
  const int n = c->touched;
  const bool odd = c->iterUsed & 1;
  const float *fin = odd ? d.lab1 : d.lab0;
  float *oth = odd ? d.lab0 : d.lab1;
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n;
       i += (long long)gridDim.x * kT) {
    int v = d.touched[i];
    oth[v] = fin[v];
  }
}

/* enforceGroundTruthKernel (iterativePropagation.cu:224-235) on the only
 * vertices whose labels can differ from 0/1: this batch's ground truth */
__global__ void k_enforce_gt(DState d, DBatch bt) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= bt.nGt) return;
  const int v = bt.gtIds[i];
  const int g = d.gtClass[v];
  if (g == 0 || g == 1) {
    d.lab0[v] = (float)g;
    d.lab1[v] = (float)g;
  }
}

/* computeAccuracy (dynLP.cu:65-85): over the vertices whose host status is
 * ACTIVE, the share whose label (>= 0.5) gives the true class.  Each
 * vertex's contribution is kept (accState) with running totals over all
 * vertices.  A contribution changes only when the vertex's host status or
 * its label does, i.e. for the batch's listed vertices and those
 * propagation wrote (the touched list); when those are few (at most N/16)
 * only they are visited and their changes added to the totals, atomicExch
 * making a vertex listed twice count once.  Otherwise a coalesced pass over
 * all vertices is cheaper than that many scattered visits: it rewrites
 * every state and counts the batch's totals afresh (k_batch_record then
 * takes them as the running totals). */
__device__ __forceinline__ int acc_state(const DState &d, int v) {
  if (!(d.vflags[v] & VF_HSACT)) return 0;
  const int pred = d.lab0[v] >= 0.5f ? 1 : 0;
  const int truth = d.trueLab[v] >= 0.5f ? 1 : 0;
  return 1 | (pred == truth ? 2 : 0);
}

__global__ __launch_bounds__(kT) void k_accuracy_delta(DState d, DBatch bt) {
  __shared__ int lds[4];
  DevCounters *c = d.ctr;
  const long long n1 = bt.nIns, n2 = n1 + bt.nDel, n3 = n2 + bt.nGt, n4 = n3 + c->touched;
  const bool full = 16 * n4 > (long long)d.N;  // uniform
  int corr = 0, tot = 0;
  if (full) {
    for (long long v = (long long)blockIdx.x * kT + threadIdx.x; v < d.N;
         v += (long long)gridDim.x * kT) {
      const int st = acc_state(d, (int)v);
      d.accState[v] = st;
      corr += st == 3;
      tot += st & 1;
    }
  } else {
    for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n4;
         i += (long long)gridDim.x * kT) {
      const int v = i < n1 ? bt.ins[i]
                           : (i < n2 ? bt.del[i - n1]
                                     : (i < n3 ? bt.gtIds[i - n2] : d.touched[i - n3]));
      const int st = acc_state(d, v);
      const int old = atomicExch(&d.accState[v], st);
      corr += (st == 3) - (old == 3);
      tot += (st & 1) - (old & 1);
    }
  }
  corr = dlp::block_sum(corr, lds);
  tot = dlp::block_sum(tot, lds);
  if (threadIdx.x == 0) {
    if (blockIdx.x == 0) c->accFull = full ? 1 : 0;
    /* full: this batch's counts (zeroed per batch); else changes of the
     * running totals (a negative change adds as its two's complement) */
    unsigned long long *cc = full ? &c->correct : &c->accCorrect;
    unsigned long long *tt = full ? &c->total : &c->accTotal;
    if (corr) atomicAdd(cc, (unsigned long long)(long long)corr);
    if (tot) atomicAdd(tt, (unsigned long long)(long long)tot);
  }
}

__global__ void k_batch_record(DState d, int b) {
  DevCounters *c = d.ctr;
  DevBatchRecord r;
  r.numSub = 2 * c->subPairs;
  r.numComponents = c->numComponents;
  r.numCand = c->numCand;
  r.frontier0 = c->frontier0;
  r.iterations = c->iterUsed;
  r.touched = c->touched;
  r.updates = c->updates;
  if (c->accFull) {  // the whole-graph pass counted afresh
    c->accCorrect = c->correct;
    c->accTotal = c->total;
  } else {
    c->correct = c->accCorrect;
    c->total = c->accTotal;
  }
  r.correct = c->accCorrect;
  r.total = c->accTotal;
  d.records[b] = r;
  c->stampBase += c->iterUsed + 1;
}

/* ------------------------------------------------------------------------ */
/* trace export                                                             */
/* ------------------------------------------------------------------------ */
__global__ __launch_bounds__(kT) void k_export_rows(DState d, const int *rowPtr, int *colInd,
                                                    float *values) {
  constexpr int G = 8;
  const int lane = threadIdx.x & 63, gl = lane & (G - 1), grp = lane / G;
  const long long gwave = ((long long)blockIdx.x * kT + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * kT) >> 6;
  for (long long u = gwave * (64 / G) + grp; u < d.N; u += nwaves * (64 / G)) {
    const int src = d.fRow[u], dst = rowPtr[u], len = d.aLen[u];
    for (int p = gl; p < len; p += G) {
      colInd[dst + p] = d.aCol[src + p];
      values[dst + p] = d.aVal[src + p];
    }
  }
}

__global__ void k_export_marks(DState d, int tag, int *out) {
  for (long long v = (long long)blockIdx.x * kT + threadIdx.x; v < d.N;
       v += (long long)gridDim.x * kT)
    out[v] = d.markStamp[v] == tag ? 1 : 0;
}

__global__ void k_last_row_ptr(int *rowPtr, int N, const int *total) { rowPtr[N] = *total; }

/* ------------------------------------------------------------------------ */
/* step API: deletions on a standard CSR                                    */
/* ------------------------------------------------------------------------ */
__global__ void k_md_rows(const int *rowPtr, int *colInd, int *status, int *isAffected,
                          const int *deleted, int numDeleted, int numNodes) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= numDeleted) return;
  const int u = deleted[i];
  if (u < 0 || u >= numNodes) return;
  for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
    int v = colInd[e];
    if (v >= 0 && v < numNodes) atomicMax(&isAffected[v], 1);
    colInd[e] = -1;
  }
  status[u] = DLP_VERTEX_DELETED;
}

__global__ __launch_bounds__(kT) void k_md_clean(const int *rowPtr, int *colInd,
                                                 const int *status, int numNodes) {
  constexpr int G = 8;
  const int lane = threadIdx.x & 63, gl = lane & (G - 1), grp = lane / G;
  const long long gwave = ((long long)blockIdx.x * kT + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * kT) >> 6;
  for (long long u = gwave * (64 / G) + grp; u < numNodes; u += nwaves * (64 / G)) {
    if (status[u] == DLP_VERTEX_DELETED) continue;
    for (int e = rowPtr[u] + gl; e < rowPtr[u + 1]; e += G) {
      int v = colInd[e];
      if (v >= 0 && v < numNodes && status[v] == DLP_VERTEX_DELETED) colInd[e] = -1;
    }
  }
}

/* step API propagation set-up: stamps carry eligibility, first frontier */
__global__ __launch_bounds__(kT) void k_prop_prepare(const int *isAffected, const int *status,
                                                     const int *gtClass, int numNodes,
                                                     int *iterStamp, int *q0, DevCounters *c) {
  const long long gwave = ((long long)blockIdx.x * kT + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * kT) >> 6;
  const int lane = threadIdx.x & 63;
  for (long long v0 = gwave * 64; v0 < numNodes; v0 += nwaves * 64) {
    const long long v = v0 + lane;
    bool in = false;
    if (v < numNodes) {
      const bool elig = status[v] == DLP_VERTEX_ACTIVE && gtClass[v] < 0;
      in = elig && isAffected[v] == 1;
      iterStamp[v] = elig ? (in ? 1 : 0) : DLP_INT_MAX;
    }
    int slot = dlp_wave_append(&c->qcount[0], in);
    if (slot >= 0) q0[slot] = (int)v;
  }
}

__global__ void k_prop_finish_gt(float *labels, const int *gtClass, int numNodes) {
  for (long long v = (long long)blockIdx.x * kT + threadIdx.x; v < numNodes;
       v += (long long)gridDim.x * kT) {
    if (gtClass[v] == 0) labels[v] = 0.0f;
    else if (gtClass[v] == 1) labels[v] = 1.0f;
  }
}

__global__ void k_prop_pending(int *isAffected, const int *pending, const int *countDev) {
  const int n = *countDev;
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n;
       i += (long long)gridDim.x * kT)
    isAffected[pending[i]] = 1;
}

int grid_for(long long n) { return dlp_blocks(n, kT, 65536); }

}  // namespace

/* ========================================================================== */
/* pipeline launchers                                                         */
/* ========================================================================== */
int dlp_rebase_stamps(const DState &d, hipStream_t s) {
  hipLaunchKernelGGL(k_rebase_stamps, dim3(grid_for(d.N)), dim3(kT), 0, s, d.iterStamp,
                     (long long)d.N, d.ctr);
  DLP_LAUNCHED("k_rebase_stamps");
  return DLP_OK;
}

int dlp_restart_local_gens(const LocalArgs &l, DevCounters *ctr, hipStream_t s) {
  hipLaunchKernelGGL(k_restart_gens, dim3(grid_for(l.numNodes)), dim3(kT), 0, s, l.coreGen,
                     l.haloGen, (long long)l.numNodes, ctr);
  DLP_LAUNCHED("k_restart_gens");
  return DLP_OK;
}

int dlp_reset_state(const DState &d, hipStream_t s) {
  hipLaunchKernelGGL(k_reset, dim3(grid_for(d.N)), dim3(kT), 0, s, d);
  DLP_LAUNCHED("k_reset");
  return DLP_OK;
}

int dlp_batch_begin(const DState &d, const DBatch &bt, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_prev, dim3(dlp_blocks((long long)bt.nPrevIns + bt.nPrevDel + bt.nPrevGt, kT)),
                     dim3(kT), 0, s, d, bt);
  DLP_LAUNCHED("k_clear_prev");
  if (bt.nDel > 0) {
    hipLaunchKernelGGL(k_apply_deleted, dim3(dlp_blocks(bt.nDel, kT)), dim3(kT), 0, s, d, bt);
    DLP_LAUNCHED("k_apply_deleted");
  }
  if (bt.nIns > 0) {
    hipLaunchKernelGGL(k_apply_inserted, dim3(dlp_blocks(bt.nIns, kT)), dim3(kT), 0, s, d, bt);
    DLP_LAUNCHED("k_apply_inserted");
  }
  if (bt.nGt > 0) {
    hipLaunchKernelGGL(k_apply_gt, dim3(dlp_blocks(bt.nGt, kT)), dim3(kT), 0, s, d, bt);
    DLP_LAUNCHED("k_apply_gt");
  }
  long long nf = (long long)bt.nPrevDel + bt.nDel + bt.nIns + bt.nGt;
  if (nf > 0) {
    hipLaunchKernelGGL(k_finalize, dim3(dlp_blocks(nf, kT)), dim3(kT), 0, s, d, bt);
    DLP_LAUNCHED("k_finalize");
  }
  return DLP_OK;
}

int dlp_csr_update(const DState &d, const DBatch &bt, int markTag, long long rowBound,
                   hipStream_t s) {
  /* 8 rows (listed vertices) per wave, 32 per block; the row list lives in
   * the odd-iteration queue, which propagation first writes after this */
  if (rowBound >= 0) {
    const long long listed = (long long)bt.nPrevDel + bt.nDel + bt.nIns + bt.nGt;
    if (listed == 0) return DLP_OK;  // no row changes
    hipLaunchKernelGGL(k_csr_rows, dim3(dlp_blocks(listed, 32, 65536)), dim3(kT), 0, s, d, bt,
                       markTag, d.q1);
    DLP_LAUNCHED("k_csr_rows");
    hipLaunchKernelGGL(k_csr_update_rows,
                       dim3(dlp_blocks(std::min<long long>(rowBound, d.N), 32, 65536)), dim3(kT),
                       0, s, d, markTag, (const int *)d.q1);
    DLP_LAUNCHED("k_csr_update_rows");
    return DLP_OK;
  }
  hipLaunchKernelGGL(k_csr_update, dim3(dlp_blocks(d.N, 32)), dim3(kT), 0, s, d, markTag);
  DLP_LAUNCHED("k_csr_update");
  return DLP_OK;
}

int dlp_sorted_inserted(const DState &d, hipStream_t s) {
  return dlp_bitmap_to_list(d.newBits, d.numWords, d.S, &d.ctr->numNew, d.scratch, s);
}

/* the one-pass compaction over n edges; scratch holds
 * dlp_sparsify_scratch_ints(n) ints (tile status words, tile counter) */
template <class IsNew, class Out>
int sparsify_launch(const int *src, const int *dst, const float *w, int n, float tau,
                    IsNew isNew, Out out, int *total, int *scratch, int *err, hipStream_t s) {
  const int tiles = (int)std::max<long long>(1, ((long long)n + kSpTile - 1) / kSpTile);
  unsigned long long *state = reinterpret_cast<unsigned long long *>(scratch);
  int *counter = scratch + 2 * tiles;
  hipLaunchKernelGGL(k_fill_int, dim3(dlp_blocks(2LL * tiles + 1, kT)), dim3(kT), 0, s, scratch,
                     0, 2LL * tiles + 1);
  DLP_LAUNCHED("k_fill_int");
  hipLaunchKernelGGL((k_sparsify_pass<IsNew, Out>), dim3(tiles), dim3(kT), 0, s, src, dst, w, n,
                     tau, isNew, out, state, counter, total, err);
  DLP_LAUNCHED("k_sparsify_pass");
  return DLP_OK;
}

int dlp_sparsify(const DState &d, const DBatch &bt, float tau, int markTag, hipStream_t s) {
  SpOutPipe out{d.subSrc, d.subDst, d.subW, bt.nIns, d.markStamp, d.cand, &d.ctr->numCand,
                markTag};
  return sparsify_launch(bt.eSrc, bt.eDst, bt.eW, bt.nE, tau, NewBits{d.newBits}, out,
                         &d.ctr->subPairs, d.scratch, &d.ctr->errorFlag, s);
}

int dlp_components(const DState &d, const DBatch &bt, hipStream_t s) {
  const int n = bt.nIns;
  hipLaunchKernelGGL(k_cc_init, dim3(dlp_blocks(n, kT)), dim3(kT), 0, s, d.par, (const int *)d.S,
                     n);
  DLP_LAUNCHED("k_cc_init");
  hipLaunchKernelGGL(k_cc_hook, dim3(dlp_blocks(bt.nE, kT, 4096)), dim3(kT), 0, s, d.par,
                     (const int *)d.subSrc, (const int *)d.subDst, (const int *)&d.ctr->subPairs,
                     0, 1, &d.ctr->errorFlag);
  DLP_LAUNCHED("k_cc_hook");
  hipLaunchKernelGGL(k_cc_compress, dim3(dlp_blocks(n, kT)), dim3(kT), 0, s, d.par,
                     (const int *)d.S, n, &d.ctr->errorFlag);
  DLP_LAUNCHED("k_cc_compress");
  int rc = dlp::scan_launch(RootGet{d.par, d.S}, RootEmit{d.par, d.S, d.compId}, n,
                            &d.ctr->numComponents, d.scratch, s);
  if (rc != DLP_OK) return rc;
  hipLaunchKernelGGL(k_cc_assign, dim3(dlp_blocks(n, kT)), dim3(kT), 0, s, (const int *)d.par,
                     (const int *)d.S, bt.ins, n, d.compId, (const int *)&d.ctr->subPairs,
                     &d.ctr->numComponents);
  DLP_LAUNCHED("k_cc_assign");
  return DLP_OK;
}

int dlp_clear_component_ids(const DState &d, const int *ins, int n, hipStream_t s) {
  if (n <= 0) return DLP_OK;
  hipLaunchKernelGGL(k_scatter_int, dim3(dlp_blocks(n, kT)), dim3(kT), 0, s, d.compId, ins, n, -1);
  DLP_LAUNCHED("k_scatter_int");
  return DLP_OK;
}

static int bits_for(long long maxValue) {
  int b = 1;
  while (b < 31 && (1ll << b) <= maxValue) ++b;
  return b;
}

int dlp_label_init(const DState &d, const DBatch &bt, hipStream_t s) {
  const int n = bt.nIns;
  hipLaunchKernelGGL((k_li_partials<RowsSlack, GtFlags>), dim3(dlp_blocks(n, kT)), dim3(kT), 0,
                     s, RowsSlack{d.fRow, d.aLen}, (const int *)d.aCol, (const float *)d.aVal,
                     GtFlags{d.vflags}, (const int *)d.S, n, (const int *)d.compId, 0,
                     (const int *)&d.ctr->numComponents, d.p0, d.p1, d.keyA, d.valA);
  DLP_LAUNCHED("k_li_partials");
  int rc = dlp_radix_sort_pairs(d.keyA, d.valA, d.keyC, d.valC, d.keyB, d.valB, n, bits_for(n),
                                d.scratch, s);
  if (rc != DLP_OK) return rc;
  hipLaunchKernelGGL(k_li_segments, dim3(dlp_blocks(n, kT)), dim3(kT), 0, s, (const int *)d.keyC,
                     n, (const int *)&d.ctr->numComponents, 0, d.segStart, d.segEnd);
  DLP_LAUNCHED("k_li_segments");
  hipLaunchKernelGGL(k_li_reduce, dim3(dlp_blocks(n, 4, 16384)), dim3(kT), 0, s,
                     (const int *)d.segStart, (const int *)d.segEnd,
                     (const int *)&d.ctr->numComponents, 0, (const int *)d.valC,
                     (const float *)d.p0, (const float *)d.p1, d.red0, d.red1, d.W0, d.W1);
  DLP_LAUNCHED("k_li_reduce");
  hipLaunchKernelGGL((k_li_labels<GtFlags>), dim3(dlp_blocks(n, kT)), dim3(kT), 0, s, bt.ins, n,
                     GtFlags{d.vflags}, (const int *)d.compId, (const float *)d.W0,
                     (const float *)d.W1, d.lab0, d.lab1);
  DLP_LAUNCHED("k_li_labels");
  return DLP_OK;
}

int dlp_labels_no_edges(const DState &d, const DBatch &bt, hipStream_t s) {
  if (bt.nIns <= 0) return DLP_OK;
  hipLaunchKernelGGL(k_labels_no_edges, dim3(dlp_blocks(bt.nIns, kT)), dim3(kT), 0, s, d, bt);
  DLP_LAUNCHED("k_labels_no_edges");
  return DLP_OK;
}

int dlp_frontier0(const DState &d, int bound, hipStream_t s) {
  hipLaunchKernelGGL(k_frontier0, dim3(dlp_blocks(bound, kT, 2048)), dim3(kT), 0, s, d);
  DLP_LAUNCHED("k_frontier0");
  return DLP_OK;
}

int dlp_batch_end(const DState &d, const DBatch &bt, hipStream_t s) {
  hipLaunchKernelGGL(k_sync_touched, dim3(dlp_blocks(d.N, kT, 2048)), dim3(kT), 0, s, d);
  DLP_LAUNCHED("k_sync_touched");
  if (bt.nGt > 0) {
    hipLaunchKernelGGL(k_enforce_gt, dim3(dlp_blocks(bt.nGt, kT)), dim3(kT), 0, s, d, bt);
    DLP_LAUNCHED("k_enforce_gt");
  }
  hipLaunchKernelGGL(k_accuracy_delta, dim3(dlp_blocks(d.N, kT, 2048)), dim3(kT), 0, s, d, bt);
  DLP_LAUNCHED("k_accuracy_delta");
  hipLaunchKernelGGL(k_batch_record, dim3(1), dim3(1), 0, s, d, bt.b);
  DLP_LAUNCHED("k_batch_record");
  return DLP_OK;
}

int dlp_export_csr(const DState &d, int *rowPtr, int *colInd, float *values, hipStream_t s) {
  int rc = dlp_scan_exclusive(d.aLen, rowPtr, d.N, &d.ctr->scanTotal, d.scratch, s);
  if (rc != DLP_OK) return rc;
  hipLaunchKernelGGL(k_last_row_ptr, dim3(1), dim3(1), 0, s, rowPtr, d.N,
                     (const int *)&d.ctr->scanTotal);
  DLP_LAUNCHED("k_last_row_ptr");
  hipLaunchKernelGGL(k_export_rows, dim3(dlp_blocks(d.N, 32)), dim3(kT), 0, s, d,
                     (const int *)rowPtr, colInd, values);
  DLP_LAUNCHED("k_export_rows");
  return DLP_OK;
}

int dlp_export_marks(const DState &d, int tag, int *out, hipStream_t s) {
  hipLaunchKernelGGL(k_export_marks, dim3(grid_for(d.N)), dim3(kT), 0, s, d, tag, out);
  DLP_LAUNCHED("k_export_marks");
  return DLP_OK;
}

/* ========================================================================== */
/* step API launchers                                                         */
/* ========================================================================== */
size_t dlp_sparsify_scratch_ints(long long n) {
  const long long tiles = std::max<long long>(1, (n + kSpTile - 1) / kSpTile);
  return (size_t)(2 * tiles + 64);
}

int dlp_k_fill_int(int *p, int value, long long n, hipStream_t s) {
  if (n <= 0) return DLP_OK;
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(n)), dim3(kT), 0, s, p, value, n);
  DLP_LAUNCHED("k_fill_int");
  return DLP_OK;
}

int dlp_k_copy_float(float *dst, const float *src, long long n, hipStream_t s) {
  if (n <= 0) return DLP_OK;
  hipLaunchKernelGGL(k_copy_float, dim3(grid_for(n)), dim3(kT), 0, s, dst, src, n);
  DLP_LAUNCHED("k_copy_float");
  return DLP_OK;
}

int dlp_k_mark_deletions(int *rowPtr, int *colInd, int *status, int *isAffected,
                         const int *deleted, int numDeleted, int numNodes, hipStream_t s) {
  hipLaunchKernelGGL(k_md_rows, dim3(dlp_blocks(numDeleted, kT)), dim3(kT), 0, s,
                     (const int *)rowPtr, colInd, status, isAffected, deleted, numDeleted,
                     numNodes);
  DLP_LAUNCHED("k_md_rows");
  hipLaunchKernelGGL(k_md_clean, dim3(dlp_blocks(numNodes, 32)), dim3(kT), 0, s,
                     (const int *)rowPtr, colInd, (const int *)status, numNodes);
  DLP_LAUNCHED("k_md_clean");
  return DLP_OK;
}

int dlp_k_sparsify(const int *src, const int *dst, const float *w, int numNewEdges,
                   int numInserted, const int *isNew, float tau, int *isAffected, int *subSrc,
                   int *subDst, float *subW, int *countDev, int *scratch, int *errFlag,
                   hipStream_t s) {
  SpOutStd out{subSrc, subDst, subW, numInserted, isAffected};
  return sparsify_launch(src, dst, w, numNewEdges, tau, NewInts{isNew}, out, countDev, scratch,
                         errFlag, s);
}

int dlp_k_sorted_list(const int *inserted, int numInserted, int numNodes, uint32_t *bits,
                      int *sorted, DevCounters *ctr, int *scratch, hipStream_t s) {
  const int numWords = (numNodes + 31) / 32;
  int rc;
  if ((rc = dlp_k_fill_int((int *)bits, 0, numWords, s)) != DLP_OK) return rc;
  hipLaunchKernelGGL(k_set_bits, dim3(dlp_blocks(numInserted, kT)), dim3(kT), 0, s, bits,
                     inserted, numInserted);
  DLP_LAUNCHED("k_set_bits");
  return dlp_bitmap_to_list(bits, numWords, sorted, &ctr->numNew, scratch, s);
}

int dlp_k_cc(const int *src, const int *dst, int numEdges, const int *inserted, int numInserted,
             int *componentId, int numNodes, int *par, const int *sorted, DevCounters *ctr,
             int *scratch, hipStream_t s) {
  int rc;
  if ((rc = dlp_k_fill_int(componentId, -1, numNodes, s)) != DLP_OK) return rc;
  if ((rc = dlp_k_fill_int(par, -1, numNodes, s)) != DLP_OK) return rc;
  hipLaunchKernelGGL(k_cc_init, dim3(dlp_blocks(numInserted, kT)), dim3(kT), 0, s, par,
                     (const int *)sorted, numInserted);
  DLP_LAUNCHED("k_cc_init");
  hipLaunchKernelGGL(k_cc_hook, dim3(dlp_blocks(numEdges, kT, 4096)), dim3(kT), 0, s, par, src,
                     dst, (const int *)nullptr, numEdges, 0, &ctr->errorFlag);
  DLP_LAUNCHED("k_cc_hook");
  hipLaunchKernelGGL(k_cc_compress, dim3(dlp_blocks(numInserted, kT)), dim3(kT), 0, s, par,
                     (const int *)sorted, numInserted, &ctr->errorFlag);
  DLP_LAUNCHED("k_cc_compress");
  rc = dlp::scan_launch(RootGet{par, sorted}, RootEmit{par, sorted, componentId}, numInserted,
                        &ctr->numComponents, scratch, s);
  if (rc != DLP_OK) return rc;
  hipLaunchKernelGGL(k_cc_assign, dim3(dlp_blocks(numInserted, kT)), dim3(kT), 0, s,
                     (const int *)par, (const int *)sorted, inserted, numInserted, componentId,
                     (const int *)nullptr, &ctr->numComponents);
  DLP_LAUNCHED("k_cc_assign");
  return DLP_OK;
}

int dlp_k_label_init(const int *rowPtr, const int *colInd, const float *values, float *labels,
                     const int *componentId, int numComponents, const int *inserted,
                     int numInserted, const int *gtClass, const int *sorted, float *p0,
                     float *p1, int *keyA, int *valA, int *keyB, int *valB, int *keyC, int *valC,
                     int *segStart, int *segEnd, float *red0, float *red1, float *W0, float *W1,
                     int *scratch, hipStream_t s) {
  const int numComp = numComponents > 1 ? numComponents : 1;
  int rc;
  hipLaunchKernelGGL((k_li_partials<RowsStd, GtInts>), dim3(dlp_blocks(numInserted, kT)),
                     dim3(kT), 0, s, RowsStd{rowPtr}, colInd, values, GtInts{gtClass},
                     (const int *)sorted, numInserted, componentId, numComponents,
                     (const int *)nullptr, p0, p1, keyA, valA);
  DLP_LAUNCHED("k_li_partials");
  rc = dlp_radix_sort_pairs(keyA, valA, keyC, valC, keyB, valB, numInserted,
                            bits_for(numComponents), scratch, s);
  if (rc != DLP_OK) return rc;
  hipLaunchKernelGGL(k_fill_float, dim3(grid_for(numComp)), dim3(kT), 0, s, W0, 0.0f,
                     (long long)numComp);
  hipLaunchKernelGGL(k_fill_float, dim3(grid_for(numComp)), dim3(kT), 0, s, W1, 0.0f,
                     (long long)numComp);
  DLP_LAUNCHED("k_fill_float");
  /* components without members keep W = 0: segments default to empty */
  if ((rc = dlp_k_fill_int(segStart, 0, numComp, s)) != DLP_OK) return rc;
  if ((rc = dlp_k_fill_int(segEnd, 0, numComp, s)) != DLP_OK) return rc;
  hipLaunchKernelGGL(k_li_segments, dim3(dlp_blocks(numInserted, kT)), dim3(kT), 0, s,
                     (const int *)keyC, numInserted, (const int *)nullptr, numComponents,
                     segStart, segEnd);
  DLP_LAUNCHED("k_li_segments");
  hipLaunchKernelGGL(k_li_reduce, dim3(dlp_blocks(numComponents, 4, 16384)), dim3(kT), 0, s,
                     (const int *)segStart, (const int *)segEnd, (const int *)nullptr,
                     numComponents, (const int *)valC, (const float *)p0, (const float *)p1, red0,
                     red1, W0, W1);
  DLP_LAUNCHED("k_li_reduce");
  hipLaunchKernelGGL((k_li_labels<GtInts>), dim3(dlp_blocks(numInserted, kT)), dim3(kT), 0, s,
                     inserted, numInserted, GtInts{gtClass}, componentId, (const float *)W0,
                     (const float *)W1, labels, (float *)nullptr);
  DLP_LAUNCHED("k_li_labels");
  return DLP_OK;
}

int dlp_k_prop_prepare(const int *isAffected, const int *status, const int *gtClass,
                       int numNodes, int *iterStamp, int *q0, DevCounters *ctr, hipStream_t s) {
  hipLaunchKernelGGL(k_prop_prepare, dim3(dlp_blocks(numNodes, kT, 4096)), dim3(kT), 0, s,
                     isAffected, status, gtClass, numNodes, iterStamp, q0, ctr);
  DLP_LAUNCHED("k_prop_prepare");
  return DLP_OK;
}

int dlp_k_prop_finish(float *labels, const int *gtClass, int numNodes, int *isAffected,
                      const int *pending, DevCounters *ctr, int pendingIndex, int capped,
                      hipStream_t s) {
  int rc;
  if (isAffected) {
    if ((rc = dlp_k_fill_int(isAffected, 0, numNodes, s)) != DLP_OK) return rc;
    if (capped) {
      hipLaunchKernelGGL(k_prop_pending, dim3(dlp_blocks(numNodes, kT, 4096)), dim3(kT), 0, s,
                         isAffected, pending, (const int *)&ctr->qcount[pendingIndex]);
      DLP_LAUNCHED("k_prop_pending");
    }
  }
  hipLaunchKernelGGL(k_prop_finish_gt, dim3(grid_for(numNodes)), dim3(kT), 0, s, labels,
                     gtClass, numNodes);
  DLP_LAUNCHED("k_prop_finish_gt");
  return DLP_OK;
}
