/*
 * dlp_rowsum.h -- the propagation row sum, bit-identical to the reference's
 * 256-thread block reduction (iterativePropagation.cu:126-150).
 *
 * The reference block sums a row of length L like this: thread t (0..255)
 * accumulates row positions t, t+256, ... in order into slot t (positions
 * that are soft-deleted or point at a deleted vertex add nothing), then
 * s[t] += s[t+stride] for stride = 128, 64, ..., 1.
 *
 * Here G lanes (a power of two dividing 64) cooperate on one row.  Slot
 * t in [0,64) belongs to lane t % G as register j = t / G and holds
 * (s[t] + s[t+128]) + (s[t+64] + s[t+192]) -- the reference's first two tree
 * levels; strides 32 .. G are the in-register halving over j and strides
 * G/2 .. 1 are cross-lane shuffles.  Each addition pairs the same two
 * partial sums as the reference block, so the result is the same float.
 *
 * The caller supplies `pos(p, w, s)`: for row position p it sets the weight
 * w and the product s = w * F[v] of a contributing entry, or w = s = 0.
 * Adding +0.0f leaves every value that can arise here unchanged, so
 * non-contributing positions may be added as zeros.
 */
#ifndef DLP_ROWSUM_H
#define DLP_ROWSUM_H

#include <hip/hip_runtime.h>

#include "dynlp/dynlp_c.h"

namespace dlp {

template <class Pos>
__device__ __forceinline__ void slot_accumulate(const Pos &pos, int len, int x, float &w,
                                                float &s) {
  w = 0.0f;
  s = 0.0f;
  for (int p = x; p < len; p += 256) {  // positions x, x+256, ... in order
    float pw, ps;
    pos(p, pw, ps);
    w = w + pw;
    s = s + ps;
  }
}

/* strides 32 .. 1 of the tree over the 64 slot registers of a group of G
 * lanes (register j of lane gl is slot gl + G*j); the result is valid in the
 * group's lane 0 */
template <int G>
__device__ __forceinline__ void group_tree(float (&ws)[64 / G], float (&ss)[64 / G], float &W,
                                           float &S) {
  constexpr int J = 64 / G;
#pragma unroll
  for (int h = J / 2; h >= 1; h >>= 1) {  // strides 32 .. G
#pragma unroll
    for (int j = 0; j < h; ++j) {
      ws[j] = ws[j] + ws[j + h];
      ss[j] = ss[j] + ss[j + h];
    }
  }
  W = ws[0];
  S = ss[0];
#pragma unroll
  for (int off = G / 2; off >= 1; off >>= 1) {  // strides G/2 .. 1
    W = W + __shfl_down(W, off, G);
    S = S + __shfl_down(S, off, G);
  }
}

/* W and S of a row of at most 64 entries (slots 64 .. 255 of the reference
 * block hold zeros), valid in the group's lane 0.  Every lane of the group
 * must call it; `len` is the row length (group-uniform). */
template <int G, class Pos>
__device__ __forceinline__ void group_short_row_sums(int gl, int len, const Pos &pos, float &W,
                                                     float &S) {
  constexpr int J = 64 / G;
  float ws[J], ss[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int p = gl + G * j;
    float w = 0.0f, s = 0.0f;
    if (p < len) pos(p, w, s);
    ws[j] = w;
    ss[j] = s;
  }
  group_tree<G>(ws, ss, W, S);
}

/* W and S of one row of any length, valid in the group's lane 0 (gl == 0).
 * Every lane of the group must call it; `len` is the row length
 * (group-uniform). */
template <int G, class Pos>
__device__ __forceinline__ void group_row_sums(int gl, int len, const Pos &pos, float &W,
                                               float &S) {
  if (len <= 64) {
    group_short_row_sums<G>(gl, len, pos, W, S);
    return;
  }
  constexpr int J = 64 / G;
  float ws[J], ss[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {  // four slots per register (strides 128, 64)
    const int t = gl + G * j;
    float w0, s0, w1, s1, w2, s2, w3, s3;
    slot_accumulate(pos, len, t, w0, s0);
    slot_accumulate(pos, len, t + 64, w1, s1);
    slot_accumulate(pos, len, t + 128, w2, s2);
    slot_accumulate(pos, len, t + 192, w3, s3);
    ws[j] = (w0 + w2) + (w1 + w3);
    ss[j] = (s0 + s2) + (s1 + s3);
  }
  group_tree<G>(ws, ss, W, S);
}

/* W and S of one row of any length by a whole work-group (at least 256
 * threads): thread t < 256 accumulates slot t, then the reference's tree --
 * strides 128 and 64 through LDS, 32 .. 1 by shuffles in the first wave.
 * The result is valid in thread 0.  Every thread of the work-group must
 * call it (it synchronises the work-group); ldsW / ldsS hold 256 floats. */
template <class Pos>
__device__ __forceinline__ void block_row_sums(int len, const Pos &pos, float *ldsW,
                                               float *ldsS, float &W, float &S) {
  const int t = threadIdx.x;
  if (t < 256) {
    float w, s;
    slot_accumulate(pos, len, t, w, s);
    ldsW[t] = w;
    ldsS[t] = s;
  }
  __syncthreads();
  if (t < 128) {
    ldsW[t] = ldsW[t] + ldsW[t + 128];
    ldsS[t] = ldsS[t] + ldsS[t + 128];
  }
  __syncthreads();
  W = 0.0f;
  S = 0.0f;
  if (t < 64) {  // the whole first wave
    W = ldsW[t] + ldsW[t + 64];
    S = ldsS[t] + ldsS[t + 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      W = W + __shfl_down(W, off, 64);
      S = S + __shfl_down(S, off, 64);
    }
  }
  __syncthreads();  // ldsW / ldsS may be reused
}

constexpr int bit_reverse(int x, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r |= ((x >> b) & 1) << (bits - 1 - b);
  return r;
}

constexpr int log2_exact(int x) { return x <= 1 ? 0 : 1 + log2_exact(x / 2); }

/* W and S of a short row (len <= L, L a power of two <= 64) summed by one
 * lane.  Slots len .. 63 of the reference block hold zeros, so its strides
 * 128 .. L add zeros only, and strides L/2 .. 1 form a halving tree over
 * slots 0..L-1.  That tree splits first on the lowest bit of the slot index,
 * so it is the adjacent-pair tree over the slots in bit-reversed order,
 * which a binary counter reduces in one pass with log2(L)+1 partial sums
 * per quantity (float addition is commutative, so only the grouping
 * matters). */
template <int L, class Pos>
__device__ __forceinline__ void lane_row_sums(int len, const Pos &pos, float &W, float &S) {
  static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "L must be a power of two <= 64");
  constexpr int LOG = log2_exact(L);
  float sw[LOG + 1], ss[LOG + 1];
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int j = bit_reverse(i, LOG);
    float w = 0.0f, s = 0.0f;
    if (j < len) pos(j, w, s);
    int b = 0;
#pragma unroll
    for (; (i >> b) & 1; ++b) {
      w = sw[b] + w;
      s = ss[b] + s;
    }
    sw[b] = w;
    ss[b] = s;
  }
  W = sw[LOG];
  S = ss[LOG];
}

/* lane_row_sums for one quantity: the same tree over val(j), j < len (zero
 * beyond).  Used where the weight sum W of a row is known beforehand: it
 * is the same tree over the same weights every iteration. */
template <int L, class Val>
__device__ __forceinline__ float lane_row_sum(int len, const Val &val) {
  static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "L must be a power of two <= 64");
  constexpr int LOG = log2_exact(L);
  float sv[LOG + 1];
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int j = bit_reverse(i, LOG);
    float v = j < len ? val(j) : 0.0f;
    int b = 0;
#pragma unroll
    for (; (i >> b) & 1; ++b) v = sv[b] + v;
    sv[b] = v;
  }
  return sv[LOG];
}

/* the reference's label update from the row sums (iterativePropagation.cu:
 * 153-171): F = S / W if W > 1e-10f else F_old, clamped to [0,1] */
__device__ __forceinline__ float new_label(float W, float S, float oldLabel) {
  float nl = (W > 1e-10f) ? S / W : oldLabel;
  return fminf(fmaxf(nl, 0.0f), 1.0f);
}

/* row position p of the CSR row that starts at `beg`: a soft-deleted entry
 * (negative column) or one pointing at a deleted vertex adds nothing
 * (iterativePropagation.cu:126-137, validation.cu:82-91) */
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

}  // namespace dlp

#endif /* DLP_ROWSUM_H */
