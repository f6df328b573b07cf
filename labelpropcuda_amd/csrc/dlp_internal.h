/*
 * dlp_internal.h -- internal declarations of libdynlp (not installed).
 *
 * Device state layout, wave-level helpers for gfx950 (64-lane wavefronts)
 * and the host-side launchers that the runtime sequences.  Everything here
 * is written for CDNA directly: 64-bit ballots, mbcnt-based lane ranks,
 * wave-aggregated atomics for queue appends.
 */
#ifndef DLP_INTERNAL_H
#define DLP_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dynlp/dynlp_c.h"

#define DLP_INT_MAX 0x7fffffff

/* ------------------------------------------------------------------------ */
/* vertex flag byte (one per vertex, rewritten for the vertices a batch     */
/* touches)                                                                  */
/* ------------------------------------------------------------------------ */
enum : uint8_t {
  VF_HACT = 1,   /* row present in this batch's CSR: host status != DELETED
                    and the vertex has arrived (dynLP.cu:143-157)          */
  VF_DDEL = 2,   /* in this batch's deletion list (device status DELETED)  */
  VF_NEW = 4,    /* inserted in this batch                                  */
  VF_ELIG = 8,   /* device status ACTIVE and not ground truth: may enter a
                    frontier (iterativePropagation.cu:73-75, :208)         */
  VF_GT0 = 16,   /* ground truth, class 0                                   */
  VF_GT1 = 32,   /* ground truth, class 1                                   */
  VF_HSACT = 64, /* host status ACTIVE: counted by accuracy (dynLP.cu:73-74) */
  VF_GTNOW = 128 /* in this batch's ground-truth list                       */
};

/* ------------------------------------------------------------------------ */
/* device counters of one context (zeroed per batch where noted)            */
/* ------------------------------------------------------------------------ */
struct DevCounters {
  int qcount[3];       /* frontier queue sizes, ring indexed by iteration % 3 */
  int numCand;         /* marked candidates this batch                        */
  int touched;         /* distinct vertices written by propagation            */
  int iterUsed;        /* iterations with a non-empty frontier                */
  int subPairs;        /* undirected G' edges (directed entries = 2x)         */
  int numComponents;
  int numNew;          /* sorted inserted list length                         */
  int frontier0;
  int stampBase;       /* iteration stamps of this batch are > stampBase      */
  int errorFlag;       /* set by kernels on violated preconditions            */
  unsigned long long updates;  /* sum of frontier sizes                       */
  unsigned long long correct;  /* accuracy numerator                          */
  unsigned long long total;    /* accuracy denominator                        */
  int scanTotal;       /* scratch total of the last scan                      */
  int pad[3];
  /* LDS-resident propagation (single work-group) */
  int localGen;        /* generation of the last local build (never reset)   */
  int coreCount;       /* vertices in the saved core list (reset per batch)  */
  int nextIter;        /* iteration the local kernel stopped at              */
  int localReason;     /* why it stopped (DLP_LOCAL_*)                        */
  /* propagation work counters, accumulated while profiling is on */
  unsigned long long profEdges;        /* row entries read by frontier vertices */
  unsigned long long profChangedEdges; /* (row length + 1) of changed vertices  */
  unsigned long long profAppended;     /* queue appends                         */
  /* the same counters for iterations run by the LDS-resident kernel */
  unsigned long long profLocalUpdates;
  unsigned long long profLocalEdges;
  unsigned long long profLocalChangedEdges;
  unsigned long long profLocalAppended;
  unsigned long long profLocalOwnerUpdates; /* of profLocalUpdates: owner mode */
  unsigned long long profLocalBuilds;       /* LDS working sets built         */
  unsigned long long profLocalWaveUpdates;  /* of profLocalOwnerUpdates: in the
                                               single-wave owner kernel      */
  int rowCount;        /* rows listed for the batch's CSR rebuild           */
  int accFull;         /* the batch's accuracy came from a whole-graph pass */
  /* accuracy numerator / denominator over all vertices, kept up to date
   * from the vertices each batch changes (zeroed by a reset only) */
  unsigned long long accCorrect;
  unsigned long long accTotal;
};

/* exit reasons of the LDS-resident propagation kernel */
enum {
  DLP_LOCAL_CONVERGED = 1, /* next frontier empty                            */
  DLP_LOCAL_CAPPED = 2,    /* maxIterations reached                          */
  DLP_LOCAL_TOO_BIG = 3,   /* working set does not fit; continue multi-block */
  DLP_LOCAL_LONG = 4,      /* queue kernel ran LocalArgs::ownerAfter iterations:
                              continue with the owner kernel                 */
  DLP_LOCAL_WIDE = 5       /* single-wave owner kernel: the core is too large
                              (or has long rows or ineligible vertices) for
                              it; continue with the work-group owner kernel  */
};

/* per-batch record kept on the device, read back once per run */
struct DevBatchRecord {
  int numSub, numComponents, numCand, frontier0, iterations, touched;
  unsigned long long updates, correct, total;
};

/* ------------------------------------------------------------------------ */
/* wave helpers                                                             */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ int dlp_lane() { return (int)(threadIdx.x & 63u); }

/* number of set bits of `mask` below this lane */
__device__ __forceinline__ int dlp_rank_below(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ unsigned long long dlp_ballot(bool p) { return __ballot(p ? 1 : 0); }

/* inclusive prefix sum over the 64 lanes of a wave, all lanes active: DPP
 * row shifts by 1, 2, 4, 8 scan each row of 16 lanes, then row_bcast:15
 * adds row 0's total to row 1 (and row 2's to row 3) and row_bcast:31 adds
 * rows 0+1 to rows 2 and 3 */
__device__ __forceinline__ int dlp_wave_inclusive_sum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

/* Wave-aggregated append: every lane of the wave that is active calls this
 * (converged); lanes with `pred` get a distinct slot of `counter`.  One
 * atomic per wave.  Returns the slot or -1. */
__device__ __forceinline__ int dlp_wave_append(int *counter, bool pred) {
  unsigned long long m = dlp_ballot(pred);
  if (m == 0ull) return -1;
  int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (dlp_lane() == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader, 64);
  return pred ? base + dlp_rank_below(m) : -1;
}

/* mark candidate: first marker of this batch appends v to the candidate
 * list (dedup through an N-sized stamp array) */
__device__ __forceinline__ void dlp_mark_candidate(int *markStamp, int tag, int *cand,
                                                   int *numCand, int v, bool pred) {
  bool first = false;
  if (pred) first = atomicMax(&markStamp[v], tag) < tag;
  int slot = dlp_wave_append(numCand, first);
  if (slot >= 0) cand[slot] = v;
}

/* ------------------------------------------------------------------------ */
/* error helpers (host)                                                     */
/* ------------------------------------------------------------------------ */
extern "C" void dlp_set_error_(const char *msg);
void dlp_set_errorf_(const char *fmt, ...);
int dlp_hip_fail_(hipError_t e, const char *what);

#define DLP_HIP(call)                                            \
  do {                                                           \
    hipError_t dlp_e_ = (call);                                  \
    if (dlp_e_ != hipSuccess) return dlp_hip_fail_(dlp_e_, #call); \
  } while (0)

#define DLP_LAUNCHED(name)                                        \
  do {                                                            \
    hipError_t dlp_e_ = hipGetLastError();                        \
    if (dlp_e_ != hipSuccess) return dlp_hip_fail_(dlp_e_, name); \
  } while (0)

static inline int dlp_blocks(long long n, int per, int cap = 1 << 20) {
  long long b = (n + per - 1) / per;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

/* ------------------------------------------------------------------------ */
/* scan / sort primitives (dlp_scan.hip, dlp_radix.hip)                     */
/* ------------------------------------------------------------------------ */
/* scratch sizes for n items */
size_t dlp_scan_scratch_ints(long long n);
/* scratch of the one-pass sparsification of n new edges */
size_t dlp_sparsify_scratch_ints(long long n);
/* exclusive scan of in[0..n) into out, total into *totalDev (device) */
int dlp_scan_exclusive(const int *in, int *out, int n, int *totalDev, int *scratch,
                       hipStream_t s);
/* ascending list of the set bits of bitmap[0..numWords) into list; count
 * into *countDev */
int dlp_bitmap_to_list(const uint32_t *bitmap, int numWords, int *list, int *countDev,
                       int *scratch, hipStream_t s);
/* stable LSD radix sort of (key, val) by key, keys < 2^keyBits; n items;
 * result in keysOut/valsOut; tmp buffers of n each */
size_t dlp_radix_scratch_ints(long long n);
int dlp_radix_sort_pairs(const int *keysIn, const int *valsIn, int *keysOut, int *valsOut,
                         int *keysTmp, int *valsTmp, int n, int keyBits, int *scratch,
                         hipStream_t s);

#endif /* DLP_INTERNAL_H */
