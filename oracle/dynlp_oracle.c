/*
 * dynlp_oracle.c -- CPU oracle for the DynLP batch-update path.
 *
 * TEST INFRASTRUCTURE (see dynlp_oracle.h): a literal single-threaded
 * restatement of the reference.  Compiled twice, with
 *   -DORC_REAL=float  -DORC_PFX=orc32_   (parity target)
 *   -DORC_REAL=double -DORC_PFX=orc64_   (FP64 twin for error statistics)
 * and always with -ffp-contract=off so that every multiply and add rounds
 * separately, as the restated kernels do.
 */
#include "dynlp_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL float
#define ORC_PFX orc32_
#endif

#define ORC_CAT_(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT_(a, b)
#define FN(name) ORC_CAT(ORC_PFX, name)
#define STATE ORC_CAT(ORC_PFX, state)

typedef ORC_REAL real;

enum { V_ACTIVE = 0, V_DELETED = 1, V_GT = 2 };

/* clamp to [0,1] with the reference's fminf(fmaxf(x, 0), 1)
 * (iterativePropagation.cu:168); double twin uses fmin/fmax. */
static real clamp01(real x) {
  if (sizeof(real) == sizeof(float))
    return (real)fminf(fmaxf((float)x, 0.0f), 1.0f);
  return (real)fmin(fmax((double)x, 0.0), 1.0);
}

static real absr(real x) {
  if (sizeof(real) == sizeof(float)) return (real)fabsf((float)x);
  return (real)fabs((double)x);
}

/* ---------------------------------------------------------------------- */
/* Step 1a: markDeletionsKernel + cleanIncomingEdgesKernel                */
/* (sparsifySubgraph.cu:58-111, launcher :185-203).  The kernels are     */
/* order-independent: one thread per deleted vertex writes only its own  */
/* row, and every mark writes the same value 1.                          */
/* ---------------------------------------------------------------------- */
void FN(mark_deletions)(const int *rowPtr, int *colInd, int *status,
                        int *isAffected, const int *deleted, int numDeleted,
                        int numNodes) {
  if (numDeleted == 0) return; /* :188 */
  for (int i = 0; i < numDeleted; ++i) {
    int u = deleted[i];
    if (u < 0 || u >= numNodes) continue; /* :66 */
    for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
      int v = colInd[e];
      if (v >= 0 && v < numNodes && isAffected[v] < 1) isAffected[v] = 1; /* :73-76 */
      colInd[e] = -1;                                                     /* :78 */
    }
    status[u] = V_DELETED; /* :82 */
  }
  /* cleanIncomingEdgesKernel over every row (:92-111) */
  for (int u = 0; u < numNodes; ++u) {
    if (status[u] == V_DELETED) continue; /* :99 */
    for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
      int v = colInd[e];
      if (v >= 0 && v < numNodes && status[v] == V_DELETED) colInd[e] = -1;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Step 1b: countSubgraphEdgesKernel + buildSubgraphKernel               */
/* (sparsifySubgraph.cu:119-179, launcher :205-287).  G' is emitted in   */
/* input-edge order (the reference's atomicAdd order is arbitrary, so    */
/* only the set is defined there).  Marks are bounded by numInserted:    */
/* the launcher passes numSubNodes == numInserted as the kernel's        */
/* numNodes (:219, :259-263, :177-178).                                   */
/* Returns the number of directed G' entries.                            */
/* ---------------------------------------------------------------------- */
int FN(sparsify)(const int *src, const int *dst, const float *w,
                 int numNewEdges, int numInserted, const int *isNew, float tau,
                 int *isAffected, int *subSrc, int *subDst, float *subW) {
  if (numNewEdges == 0) return 0; /* :213-217 */
  int count = 0;
  for (int i = 0; i < numNewEdges; ++i) {
    int u = src[i], v = dst[i];
    float wt = w[i];
    if (wt <= tau) continue; /* :163 */
    if (isNew[u] && isNew[v]) { /* :165-174 */
      if (subSrc) {
        subSrc[count] = u; subDst[count] = v; subW[count] = wt;
        subSrc[count + 1] = v; subDst[count + 1] = u; subW[count + 1] = wt;
      }
      count += 2;
    }
    if (u >= 0 && u < numInserted && isAffected[u] < 1) isAffected[u] = 1;
    if (v >= 0 && v < numInserted && isAffected[v] < 1) isAffected[v] = 1;
  }
  return count;
}

/* ---------------------------------------------------------------------- */
/* Step 1c: connected components (connectedComponents.cu:162-290)        */
/* The SV hook/jump loop is restated by its result: hooking always       */
/* attaches the larger root to the smaller one, so each component's root */
/* is its minimum vertex id; componentId = rank of that root among all   */
/* roots in ascending id order (:248-283).  Computed with union-find by  */
/* minimum id.  Returns 0, or -1 if an edge endpoint is not inserted     */
/* (the reference reads par[-1] there).                                  */
/* ---------------------------------------------------------------------- */
static int uf_find(int *par, int x) {
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}

int FN(connected_components)(const int *src, const int *dst, int numEdges,
                             const int *inserted, int numInserted,
                             int *componentId, int numNodes,
                             int *numComponents) {
  if (numInserted == 0) { /* :167-170 */
    *numComponents = 0;
    return 0;
  }
  if (numEdges == 0) { /* :191-205 */
    for (int v = 0; v < numNodes; ++v) componentId[v] = -1;
    *numComponents = numInserted;
    return 0;
  }
  int *par = (int *)malloc(sizeof(int) * (size_t)numNodes);
  for (int v = 0; v < numNodes; ++v) par[v] = -1; /* :176-178 */
  for (int i = 0; i < numInserted; ++i) par[inserted[i]] = inserted[i];
  for (int i = 0; i < numEdges; ++i) {
    int u = src[i], v = dst[i];
    if (u < 0 || u >= numNodes || v < 0 || v >= numNodes || par[u] < 0 ||
        par[v] < 0) {
      free(par);
      return -1;
    }
    int ru = uf_find(par, u), rv = uf_find(par, v);
    if (ru == rv) continue;
    if (ru < rv) par[rv] = ru; else par[ru] = rv;
  }
  int *rootToId = (int *)malloc(sizeof(int) * (size_t)numNodes);
  for (int v = 0; v < numNodes; ++v) rootToId[v] = 0;
  for (int i = 0; i < numInserted; ++i) rootToId[uf_find(par, inserted[i])] = 1;
  int c = 0;
  for (int v = 0; v < numNodes; ++v) rootToId[v] = rootToId[v] ? c++ : -1; /* :264-270 */
  for (int v = 0; v < numNodes; ++v) componentId[v] = -1; /* :279-280 */
  for (int i = 0; i < numInserted; ++i) {
    int v = inserted[i];
    componentId[v] = rootToId[uf_find(par, v)];
  }
  *numComponents = c;
  free(rootToId);
  free(par);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Step 2: label initialisation (labelInitialization.cu:59-187)          */
/* ---------------------------------------------------------------------- */

/* bottom-up pairwise tree: stride 1, 2, 4, ... a[i] += a[i + stride] for
 * i multiple of 2*stride.  The canonical combination order of component
 * members (replaces the reference's float atomicAdd order). */
static real pairwise_tree(real *a, int n) {
  if (n <= 0) return (real)0;
  for (int stride = 1; stride < n; stride *= 2)
    for (int i = 0; i + stride < n; i += 2 * stride) a[i] = a[i] + a[i + stride];
  return a[0];
}

typedef struct {
  int comp, id, idx;
} member_t;

static int member_cmp(const void *pa, const void *pb) {
  const member_t *a = (const member_t *)pa, *b = (const member_t *)pb;
  if (a->comp != b->comp) return a->comp < b->comp ? -1 : 1;
  if (a->id != b->id) return a->id < b->id ? -1 : 1;
  return 0;
}

int FN(label_initialization)(const int *rowPtr, const int *colInd,
                             const float *values, real *labels,
                             const int *componentId, int numComponents,
                             const int *inserted, int numInserted,
                             const int *gtClass) {
  if (numInserted == 0) return 0; /* :155 */
  int numComp = numComponents > 1 ? numComponents : 1; /* :164 */
  real *W0 = (real *)calloc((size_t)numComp, sizeof(real));
  real *W1 = (real *)calloc((size_t)numComp, sizeof(real));
  real *p0 = (real *)malloc(sizeof(real) * (size_t)numInserted);
  real *p1 = (real *)malloc(sizeof(real) * (size_t)numInserted);
  member_t *m = (member_t *)malloc(sizeof(member_t) * (size_t)numInserted);
  int nm = 0;
  /* computeSupernodeWeightsKernel (:59-88): per member, its row left to
   * right, skipping soft-deleted entries, weights of GT-class-0 / -1
   * neighbours */
  for (int i = 0; i < numInserted; ++i) {
    int u = inserted[i];
    int c = componentId[u];
    p0[i] = 0; p1[i] = 0;
    if (c < 0) continue; /* :70 */
    if (c >= numComp) goto bad;
    for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
      int v = colInd[e];
      if (v < 0) continue; /* :78 */
      int g = gtClass[v];
      if (g == 0) p0[i] = p0[i] + (real)values[e];
      else if (g == 1) p1[i] = p1[i] + (real)values[e];
    }
    m[nm].comp = c; m[nm].id = u; m[nm].idx = i; ++nm;
  }
  qsort(m, (size_t)nm, sizeof(member_t), member_cmp);
  {
    real *a0 = (real *)malloc(sizeof(real) * (size_t)(nm ? nm : 1));
    real *a1 = (real *)malloc(sizeof(real) * (size_t)(nm ? nm : 1));
    for (int s = 0; s < nm;) {
      int e = s;
      while (e < nm && m[e].comp == m[s].comp) ++e;
      for (int k = s; k < e; ++k) { a0[k - s] = p0[m[k].idx]; a1[k - s] = p1[m[k].idx]; }
      W0[m[s].comp] = pairwise_tree(a0, e - s);
      W1[m[s].comp] = pairwise_tree(a1, e - s);
      s = e;
    }
    free(a0); free(a1);
  }
  /* initializeLabelsKernel (:99-142) */
  for (int i = 0; i < numInserted; ++i) {
    int u = inserted[i];
    int g = gtClass[u];
    if (g == 0) { labels[u] = (real)0; continue; }
    if (g == 1) { labels[u] = (real)1; continue; }
    int c = componentId[u];
    if (c < 0) { labels[u] = (real)0.5f; continue; }
    real den = W0[c] + W1[c];
    labels[u] = (den > (real)1e-10f) ? W1[c] / den : (real)0.5f;
  }
  free(W0); free(W1); free(p0); free(p1); free(m);
  return 0;
bad:
  free(W0); free(W1); free(p0); free(p1); free(m);
  return -1;
}

/* ---------------------------------------------------------------------- */
/* Step 3: iterative propagation (iterativePropagation.cu:241-346)       */
/* ---------------------------------------------------------------------- */

/* propagateLabelsKernel's sums for one row (:113-150): a 256-thread block
 * where slot t accumulates row positions t, t+256, ... in order (soft-
 * deleted and DELETED neighbours add nothing), then s[t] += s[t+stride]
 * for stride = 128 .. 1.  Slots beyond pow2ceil(len) only ever hold +0.0f,
 * and x + 0.0f == x for every x that can arise here, so the tree is run
 * over the first min(256, pow2ceil(len)) slots only -- the same sums,
 * bit for bit. */
static void row_sums(const int *colInd, const float *values, int beg, int end,
                     const int *status, const real *F, real *outW,
                     real *outS) {
  real sw[256], ss[256];
  int len = end - beg;
  int S = 1;
  while (S < len && S < 256) S <<= 1;
  for (int t = 0; t < S; ++t) { sw[t] = 0; ss[t] = 0; }
  for (int p = 0; p < len; ++p) {
    int v = colInd[beg + p];
    if (v < 0) continue;                             /* :128 */
    if (status && status[v] == V_DELETED) continue;  /* :129 */
    real w = (real)values[beg + p];
    real prod = w * F[v];
    int t = p & 255;
    sw[t] = sw[t] + w;
    ss[t] = ss[t] + prod;
  }
  for (int stride = S >> 1; stride > 0; stride >>= 1)
    for (int t = 0; t < stride; ++t) {
      sw[t] = sw[t] + sw[t + stride];
      ss[t] = ss[t] + ss[t + stride];
    }
  *outW = sw[0];
  *outS = ss[0];
}

/* launchIterativePropagation.  The reference collects the frontier with an
 * O(N) scan of the flag array each iteration; the flags it scans are exactly
 * the marks of the previous iteration, so here the marks are kept as a
 * deduplicated list and the next frontier is that list filtered by the same
 * predicate.  Frontier order does not matter: every update reads only the
 * old buffer and every mark writes the same value.  Two label buffers are
 * swapped by pointer as in the reference.  Returns iterationsUsed;
 * frontierSizes (nullable) receives numAffected of every processed
 * iteration; isAffected ends as the reference leaves it (the last
 * iteration's marks, or unchanged when no iteration ran). */
/* Worker threads of the frontier-parallel variant of propagate (1 = the
 * literal sequential loop).  With more threads the label updates of a
 * frontier run in parallel (each reads only the old buffer, so the floats
 * are the same) and the next frontier is collected per thread and
 * concatenated: the same vertex set in another order, so labels, frontier
 * sizes, iteration counts and isAffected are unchanged. */
static int g_threads = 1;
static int g_parMin = 1024; /* smallest frontier run in parallel */
void FN(set_threads)(int threads, int minFrontier) {
  g_threads = threads < 1 ? 1 : threads;
  g_parMin = minFrontier < 1 ? 1 : minFrontier;
}

/* thread-local mark buffers of the parallel variant */
typedef struct {
  int *v;
  int n, cap;
} MarkBuf;

static void markbuf_push(MarkBuf *b, int v) {
  if (b->n == b->cap) {
    b->cap = b->cap ? 2 * b->cap : 1024;
    b->v = (int *)realloc(b->v, sizeof(int) * (size_t)b->cap);
  }
  b->v[b->n++] = v;
}

/* marks of the changed frontier vertices, in parallel (markNeighborsKernel,
 * iterativePropagation.cu:187-217): each thread collects the vertices it
 * marked first, then the per-thread lists are concatenated into `marks`.
 * bufs and offs hold g_threads and g_threads + 1 entries.  Returns the
 * number of marks. */
static int mark_parallel(const int *rowPtr, const int *colInd, const int *status,
                         const int *gtClass, const int *list, int n, const real *nw,
                         const real *old, real d, char *marked, int *marks, MarkBuf *bufs,
                         int *offs) {
  int total = 0;
#pragma omp parallel num_threads(g_threads)
  {
    const int t = omp_get_thread_num();
    MarkBuf *b = &bufs[t];
    b->n = 0;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      const int u = list[i];
      if (!(absr(nw[u] - old[u]) > d)) continue;
      for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
        const int v = colInd[e];
        if (v >= 0 && status[v] != V_DELETED && gtClass[v] < 0 &&
            !__atomic_exchange_n(&marked[v], 1, __ATOMIC_RELAXED))
          markbuf_push(b, v);
      }
      if (gtClass[u] < 0 && !__atomic_exchange_n(&marked[u], 1, __ATOMIC_RELAXED))
        markbuf_push(b, u);
    }
#pragma omp single
    {
      const int T = omp_get_num_threads();
      offs[0] = 0;
      for (int k = 0; k < T; ++k) offs[k + 1] = offs[k] + bufs[k].n;
      total = offs[T];
    }
    if (b->n) memcpy(marks + offs[t], b->v, sizeof(int) * (size_t)b->n);
  }
  return total;
}

int FN(propagate)(const int *rowPtr, const int *colInd, const float *values,
                  real *labels, real *labelsBuffer, const int *status,
                  const int *gtClass, int *isAffected, int numNodes,
                  float delta, int maxIterations, int *frontierSizes,
                  int maxFrontierSizes, int *numFrontierSizes) {
  const size_t N1 = (size_t)(numNodes ? numNodes : 1);
  int *list = (int *)malloc(sizeof(int) * N1);
  int *marks = (int *)malloc(sizeof(int) * N1);
  char *marked = (char *)calloc(N1, 1);
  MarkBuf *bufs = (MarkBuf *)calloc((size_t)g_threads, sizeof(MarkBuf));
  int *offs = (int *)malloc(sizeof(int) * (size_t)(g_threads + 1));
  int numMarks = 0;
  real *old = labels, *nw = labelsBuffer;
  const real d = (real)delta;
  memcpy(nw, old, sizeof(real) * (size_t)numNodes); /* :268-270 */
  int iterationsUsed = 0, nfs = 0;
  for (int iter = 0; iter < maxIterations; ++iter) {
    /* collectAffectedKernel (:63-79) */
    int n = 0;
    if (iter == 0) {
      for (int v = 0; v < numNodes; ++v)
        if (isAffected[v] == 1 && status[v] == V_ACTIVE && gtClass[v] < 0) list[n++] = v;
    } else {
      for (int i = 0; i < numMarks; ++i) {
        int v = marks[i];
        if (status[v] == V_ACTIVE && gtClass[v] < 0) list[n++] = v;
      }
    }
    if (n == 0) break; /* :287 */
    if (frontierSizes && nfs < maxFrontierSizes) frontierSizes[nfs] = n;
    ++nfs;
    int changed = 0;
    const int par = g_threads > 1 && n >= g_parMin;
    /* propagateLabelsKernel (:98-179) */
#pragma omp parallel for num_threads(g_threads) if (par) reduction(| : changed) \
    schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      int u = list[i];
      real W, S;
      row_sums(colInd, values, rowPtr[u], rowPtr[u + 1], status, old, &W, &S);
      real oldLabel = old[u];
      real nl = (W > (real)1e-10f) ? S / W : oldLabel;
      nl = clamp01(nl);
      nw[u] = nl;
      if (absr(nl - oldLabel) > d) changed = 1;
    }
    /* markNeighborsKernel (:187-217) into cleared flags (:299-305) */
    for (int i = 0; i < numMarks; ++i) marked[marks[i]] = 0;
    numMarks = 0;
    if (par) {
      numMarks = mark_parallel(rowPtr, colInd, status, gtClass, list, n, nw, old, d, marked,
                               marks, bufs, offs);
      real *tmp = old; old = nw; nw = tmp;
      iterationsUsed = iter + 1;
      if (!changed) break;
      continue;
    }
    for (int i = 0; i < n; ++i) {
      int u = list[i];
      if (absr(nw[u] - old[u]) > d) {
        for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
          int v = colInd[e];
          if (v >= 0 && status[v] != V_DELETED && gtClass[v] < 0 && !marked[v]) {
            marked[v] = 1;
            marks[numMarks++] = v;
          }
        }
        if (gtClass[u] < 0 && !marked[u]) {
          marked[u] = 1;
          marks[numMarks++] = u;
        }
      }
    }
    real *tmp = old; old = nw; nw = tmp; /* :308-310 */
    iterationsUsed = iter + 1;
    if (!changed) break;
  }
  if (iterationsUsed > 0) { /* isAffected <- last marks (:314-316) */
    memset(isAffected, 0, sizeof(int) * (size_t)numNodes);
    for (int i = 0; i < numMarks; ++i) isAffected[marks[i]] = 1;
  }
  if (old != labels) memcpy(labels, old, sizeof(real) * (size_t)numNodes); /* :330-334 */
  for (int v = 0; v < numNodes; ++v) { /* enforceGroundTruthKernel (:224-235) */
    if (gtClass[v] == 0) labels[v] = (real)0;
    else if (gtClass[v] == 1) labels[v] = (real)1;
  }
  if (numFrontierSizes) *numFrontierSizes = nfs;
  free(list);
  free(marks);
  free(marked);
  for (int t = 0; t < g_threads; ++t) free(bufs[t].v);
  free(bufs);
  free(offs);
  return iterationsUsed;
}

/* ---------------------------------------------------------------------- */
/* runDynLP (dynLP.cu:175-587), one batch per call                        */
/* ---------------------------------------------------------------------- */
struct STATE {
  int N, B;
  const int *fRow, *fCol;
  const float *fVal;
  const float *trueLabels;
  const int64_t *insOff, *gtOff, *delOff, *eOff;
  const int *ins, *gtIds, *del, *eSrc, *eDst;
  const float *gtLab, *eW;
  float tau, delta;
  int maxIter;
  int nextBatch;

  int *batchMap, *hStatus, *gtClass, *dStatus;
  real *hLabels, *labA, *labB, *initLabels;
  int *rowPtr, *colInd;
  float *values;
  int nnz;
  int *isAffected, *affectedSnap, *isNew, *compId;
  int *subSrc, *subDst;
  float *subW;
  int subCap, numSub, numComponents;
  int *frontier;
  int frontierCap, numFrontier, iterations;
  int64_t updates;
  float accuracy;
};

STATE *FN(create)(int numNodes, const int *rowPtr, const int *colInd,
                  const float *values, const float *trueLabels,
                  int numBatches, const int64_t *insOff, const int *ins,
                  const int64_t *gtOff, const int *gtIds,
                  const float *gtLabels, const int64_t *delOff,
                  const int *del, const int64_t *edgeOff, const int *edgeSrc,
                  const int *edgeDst, const float *edgeW, float tau,
                  float delta, int maxIterations) {
  STATE *st = (STATE *)calloc(1, sizeof(STATE));
  size_t N = (size_t)numNodes;
  int nnzFull = rowPtr[numNodes];
  st->N = numNodes; st->B = numBatches;
  st->fRow = rowPtr; st->fCol = colInd; st->fVal = values;
  st->trueLabels = trueLabels;
  st->insOff = insOff; st->ins = ins; st->gtOff = gtOff; st->gtIds = gtIds;
  st->gtLab = gtLabels; st->delOff = delOff; st->del = del;
  st->eOff = edgeOff; st->eSrc = edgeSrc; st->eDst = edgeDst; st->eW = edgeW;
  st->delta = delta; st->maxIter = maxIterations;
  /* auto tau: computeAvgEdgeWeight (dynLP.cu:90-102, :186-190) */
  if (tau < 0.0f) {
    if (nnzFull == 0) tau = 0.5f;
    else {
      double sum = 0.0;
      int count = 0;
      for (int i = 0; i < nnzFull; ++i)
        if (colInd[i] >= 0) { sum += values[i]; ++count; }
      tau = count > 0 ? (float)(sum / count) : 0.5f;
    }
  }
  st->tau = tau;
  st->batchMap = (int *)malloc(sizeof(int) * N);
  st->hStatus = (int *)malloc(sizeof(int) * N);
  st->gtClass = (int *)malloc(sizeof(int) * N);
  st->dStatus = (int *)malloc(sizeof(int) * N);
  st->hLabels = (real *)malloc(sizeof(real) * N);
  st->labA = (real *)malloc(sizeof(real) * N);
  st->labB = (real *)malloc(sizeof(real) * N);
  st->initLabels = (real *)malloc(sizeof(real) * N);
  st->rowPtr = (int *)malloc(sizeof(int) * (N + 1));
  st->colInd = (int *)malloc(sizeof(int) * (size_t)(nnzFull ? nnzFull : 1));
  st->values = (float *)malloc(sizeof(float) * (size_t)(nnzFull ? nnzFull : 1));
  st->isAffected = (int *)malloc(sizeof(int) * N);
  st->affectedSnap = (int *)malloc(sizeof(int) * N);
  st->isNew = (int *)malloc(sizeof(int) * N);
  st->compId = (int *)malloc(sizeof(int) * N);
  int64_t maxE = 0;
  for (int b = 0; b < numBatches; ++b)
    if (edgeOff[b + 1] - edgeOff[b] > maxE) maxE = edgeOff[b + 1] - edgeOff[b];
  st->subCap = (int)(2 * maxE + 1);
  st->subSrc = (int *)malloc(sizeof(int) * (size_t)st->subCap);
  st->subDst = (int *)malloc(sizeof(int) * (size_t)st->subCap);
  st->subW = (float *)malloc(sizeof(float) * (size_t)st->subCap);
  st->frontierCap = maxIterations < 1000000 ? maxIterations : 1000000;
  st->frontier = (int *)malloc(sizeof(int) * (size_t)(st->frontierCap + 1));
  /* vertex -> batch map (:193-198) */
  for (size_t v = 0; v < N; ++v) st->batchMap[v] = -1;
  for (int b = 0; b < numBatches; ++b)
    for (int64_t i = insOff[b]; i < insOff[b + 1]; ++i) st->batchMap[ins[i]] = b;
  /* host state (:201-203), device init (:238-266) */
  for (size_t v = 0; v < N; ++v) {
    st->hStatus[v] = V_DELETED;
    st->gtClass[v] = -1;
    st->hLabels[v] = (real)0.5f;
    st->compId[v] = -1;
    st->isAffected[v] = 0;
  }
  st->nextBatch = 0;
  return st;
}

void FN(destroy)(STATE *st) {
  if (!st) return;
  free(st->batchMap); free(st->hStatus); free(st->gtClass); free(st->dStatus);
  free(st->hLabels); free(st->labA); free(st->labB); free(st->initLabels);
  free(st->rowPtr); free(st->colInd); free(st->values);
  free(st->isAffected); free(st->affectedSnap); free(st->isNew); free(st->compId);
  free(st->subSrc); free(st->subDst); free(st->subW); free(st->frontier);
  free(st);
}

float FN(tau)(const STATE *st) { return st->tau; }

int FN(run_batch)(STATE *st, int b) {
  if (b != st->nextBatch || b < 0 || b >= st->B) return -1;
  const int N = st->N;
  const int *ins = st->ins + st->insOff[b];
  const int numIns = (int)(st->insOff[b + 1] - st->insOff[b]);
  const int *gtIds = st->gtIds + st->gtOff[b];
  const float *gtLab = st->gtLab + st->gtOff[b];
  const int numGt = (int)(st->gtOff[b + 1] - st->gtOff[b]);
  const int *del = st->del + st->delOff[b];
  const int numDel = (int)(st->delOff[b + 1] - st->delOff[b]);
  const int *eSrc = st->eSrc + st->eOff[b];
  const int *eDst = st->eDst + st->eOff[b];
  const float *eW = st->eW + st->eOff[b];
  const int numE = (int)(st->eOff[b + 1] - st->eOff[b]);

  /* GT class and labels (:301-305) */
  for (int i = 0; i < numGt; ++i) {
    int v = gtIds[i];
    st->gtClass[v] = (gtLab[i] < 0.5f) ? 0 : 1;
    st->hLabels[v] = (real)gtLab[i];
  }
  /* updateHostCSR (:111-167): status order deleted -> active -> GT */
  for (int i = 0; i < numDel; ++i) st->hStatus[del[i]] = V_DELETED;
  for (int i = 0; i < numIns; ++i) st->hStatus[ins[i]] = V_ACTIVE;
  for (int i = 0; i < numGt; ++i) st->hStatus[gtIds[i]] = V_GT;
  {
    int nnz = 0;
    st->rowPtr[0] = 0;
    for (int u = 0; u < N; ++u) {
      if (st->hStatus[u] == V_DELETED || st->batchMap[u] > b || st->batchMap[u] < 0) {
        st->rowPtr[u + 1] = nnz;
        continue;
      }
      for (int e = st->fRow[u]; e < st->fRow[u + 1]; ++e) {
        int v = st->fCol[e];
        if (v < 0) continue;
        if (st->hStatus[v] == V_DELETED) continue;
        if (st->batchMap[v] > b || st->batchMap[v] < 0) continue;
        st->colInd[nnz] = v;
        st->values[nnz] = st->fVal[e];
        ++nnz;
      }
      st->rowPtr[u + 1] = nnz;
    }
    st->nnz = nnz;
  }
  /* upload (:315-335) */
  memcpy(st->dStatus, st->hStatus, sizeof(int) * (size_t)N);
  memcpy(st->labA, st->hLabels, sizeof(real) * (size_t)N);
  /* reset flags, isNewVertex (:342-358) */
  for (int v = 0; v < N; ++v) { st->isNew[v] = 0; st->isAffected[v] = 0; }
  for (int i = 0; i < numIns; ++i) st->isNew[ins[i]] = 1;
  /* Step 1a (:364-377) */
  if (numDel > 0)
    FN(mark_deletions)(st->rowPtr, st->colInd, st->dStatus, st->isAffected, del,
                       numDel, N);
  /* Step 1b, 1c, 2 (:383-504) */
  st->numSub = 0;
  st->numComponents = 0;
  if (numIns > 0 && numE > 0) {
    st->numSub = FN(sparsify)(eSrc, eDst, eW, numE, numIns, st->isNew, st->tau,
                              st->isAffected, st->subSrc, st->subDst, st->subW);
    if (st->numSub > 0) {
      if (FN(connected_components)(st->subSrc, st->subDst, st->numSub, ins, numIns,
                                   st->compId, N, &st->numComponents) != 0)
        return -2;
    } else { /* :441-452 */
      st->numComponents = numIns;
      for (int v = 0; v < N; ++v) st->compId[v] = -1;
      for (int i = 0; i < numIns; ++i) st->compId[ins[i]] = i;
    }
    if (FN(label_initialization)(st->rowPtr, st->colInd, st->values, st->labA,
                                 st->compId, st->numComponents, ins, numIns,
                                 st->gtClass) != 0)
      return -3;
  } else if (numIns > 0) { /* :482-504 */
    st->numComponents = numIns;
    real *batchLabels = (real *)malloc(sizeof(real) * (size_t)N);
    for (int v = 0; v < N; ++v) batchLabels[v] = (real)0.5f;
    for (int k = 0; k < numGt; ++k) batchLabels[gtIds[k]] = (real)gtLab[k];
    for (int i = 0; i < numIns; ++i) {
      int v = ins[i];
      st->hLabels[v] = (st->gtClass[v] >= 0) ? batchLabels[v] : (real)0.5f;
    }
    free(batchLabels);
    memcpy(st->labA, st->hLabels, sizeof(real) * (size_t)N);
  }
  memcpy(st->affectedSnap, st->isAffected, sizeof(int) * (size_t)N);
  memcpy(st->initLabels, st->labA, sizeof(real) * (size_t)N);
  /* Step 3 (:518-528) */
  memcpy(st->labB, st->labA, sizeof(real) * (size_t)N);
  st->iterations = FN(propagate)(st->rowPtr, st->colInd, st->values, st->labA,
                                 st->labB, st->dStatus, st->gtClass,
                                 st->isAffected, N, st->delta, st->maxIter,
                                 st->frontier, st->frontierCap, &st->numFrontier);
  st->updates = 0;
  for (int k = 0; k < st->numFrontier && k < st->frontierCap; ++k)
    st->updates += st->frontier[k];
  /* download (:538-541) and accuracy over host status (:65-85, :543-545) */
  memcpy(st->hLabels, st->labA, sizeof(real) * (size_t)N);
  {
    int correct = 0, total = 0;
    for (int i = 0; i < N; ++i) {
      if (st->hStatus[i] == V_DELETED || st->hStatus[i] == V_GT) continue;
      int predicted = (st->hLabels[i] >= (real)0.5f) ? 1 : 0;
      int truth = (st->trueLabels[i] >= 0.5f) ? 1 : 0;
      if (predicted == truth) correct++;
      total++;
    }
    st->accuracy = total > 0 ? (float)correct / (float)total : 1.0f;
  }
  st->nextBatch = b + 1;
  return 0;
}

/* IrLP full recompute (validation.cu:47-213): every ACTIVE non-GT vertex
 * takes the weighted average of its non-deleted neighbours each iteration
 * (same 256-slot row sum), GT vertices are set to their class, until no
 * label moves by more than delta; final labels in `labels`.  Returns
 * iterationsUsed. */
int FN(irlp)(const int *rowPtr, const int *colInd, const float *values, real *labels,
             const int *status, const int *gtClass, int numNodes, float delta,
             int maxIterations) {
  const size_t N1 = (size_t)(numNodes ? numNodes : 1);
  real *buf = (real *)malloc(sizeof(real) * N1);
  memcpy(buf, labels, sizeof(real) * (size_t)numNodes); /* :157-159 */
  real *old = labels, *nw = buf;
  const real d = (real)delta;
  int iterationsUsed = 0;
  for (int iter = 0; iter < maxIterations; ++iter) {
    int changed = 0;
    for (int u = 0; u < numNodes; ++u) { /* irLPKernel (:47-126) */
      if (gtClass[u] >= 0 || status[u] != V_ACTIVE) {
        nw[u] = old[u];
        continue;
      }
      real W, S;
      row_sums(colInd, values, rowPtr[u], rowPtr[u + 1], status, old, &W, &S);
      real nl = (W > (real)1e-10f) ? S / W : old[u];
      nl = clamp01(nl);
      nw[u] = nl;
      if (absr(nl - old[u]) > d) changed = 1;
    }
    for (int v = 0; v < numNodes; ++v) { /* irLPEnforceGTKernel (:131-137) */
      if (gtClass[v] == 0) nw[v] = (real)0;
      else if (gtClass[v] == 1) nw[v] = (real)1;
    }
    real *tmp = old; old = nw; nw = tmp; /* :186-189 */
    iterationsUsed = iter + 1;
    if (!changed) break; /* :192-197 */
  }
  if (old != labels) memcpy(labels, old, sizeof(real) * (size_t)numNodes); /* :200-205 */
  free(buf);
  return iterationsUsed;
}

int FN(nnz)(const STATE *st) { return st->nnz; }

void FN(get_csr)(const STATE *st, int *rowPtr, int *colInd, float *values) {
  memcpy(rowPtr, st->rowPtr, sizeof(int) * (size_t)(st->N + 1));
  memcpy(colInd, st->colInd, sizeof(int) * (size_t)st->nnz);
  memcpy(values, st->values, sizeof(float) * (size_t)st->nnz);
}

void FN(get_affected)(const STATE *st, int *isAffected) {
  memcpy(isAffected, st->affectedSnap, sizeof(int) * (size_t)st->N);
}

int FN(get_sub)(const STATE *st, int *src, int *dst, float *w) {
  if (src) memcpy(src, st->subSrc, sizeof(int) * (size_t)st->numSub);
  if (dst) memcpy(dst, st->subDst, sizeof(int) * (size_t)st->numSub);
  if (w) memcpy(w, st->subW, sizeof(float) * (size_t)st->numSub);
  return st->numSub;
}

int FN(get_component_ids)(const STATE *st, int *componentId) {
  memcpy(componentId, st->compId, sizeof(int) * (size_t)st->N);
  return st->numComponents;
}

void FN(get_init_labels)(const STATE *st, real *labels) {
  memcpy(labels, st->initLabels, sizeof(real) * (size_t)st->N);
}

void FN(get_labels)(const STATE *st, real *labels) {
  memcpy(labels, st->hLabels, sizeof(real) * (size_t)st->N);
}

int FN(get_frontier_sizes)(const STATE *st, int *sizes, int maxSizes) {
  int n = st->numFrontier < st->frontierCap ? st->numFrontier : st->frontierCap;
  if (n > maxSizes) n = maxSizes;
  memcpy(sizes, st->frontier, sizeof(int) * (size_t)n);
  return st->numFrontier;
}

void FN(get_status)(const STATE *st, int *dStatus, int *hStatus, int *gtClass) {
  if (dStatus) memcpy(dStatus, st->dStatus, sizeof(int) * (size_t)st->N);
  if (hStatus) memcpy(hStatus, st->hStatus, sizeof(int) * (size_t)st->N);
  if (gtClass) memcpy(gtClass, st->gtClass, sizeof(int) * (size_t)st->N);
}

void FN(get_stats)(const STATE *st, int *stats5, int64_t *updates,
                   float *accuracy) {
  int marked = 0;
  for (int v = 0; v < st->N; ++v) marked += st->affectedSnap[v] == 1;
  stats5[0] = st->iterations;
  stats5[1] = st->numSub;
  stats5[2] = st->numComponents;
  stats5[3] = marked;
  stats5[4] = st->numFrontier > 0 && st->frontierCap > 0 ? st->frontier[0] : 0;
  *updates = st->updates;
  *accuracy = st->accuracy;
}
