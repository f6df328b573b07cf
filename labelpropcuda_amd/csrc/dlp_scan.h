/*
 * dlp_scan.h -- device-wide exclusive scan with a per-item callback.
 *
 * Reduce-then-scan in three launches, tiles of 2048 items (256 threads x 8
 * consecutive items):
 *   1. per-tile totals,
 *   2. one block scans the tile totals (writes the grand total to device),
 *   3. per-tile block scan; every thread calls emit(i, exclusivePrefix, valid)
 *      for its 8 items in lockstep, so emitters may use wave ballots.
 * No inter-block waiting anywhere, so no launch can hang.
 * Item values come from a getter functor: get(i) for i < n.
 */
#ifndef DLP_SCAN_H
#define DLP_SCAN_H

#include "dlp_internal.h"

namespace dlp {
namespace {  // kernels get internal linkage: every including unit has its own

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

/* exclusive scan of one value per thread across a 256-thread block */
__device__ __forceinline__ int block_exclusive_scan(int x, int *lds4, int *blockTotal) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) lds4[wave] = v;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    int t = lds4[w];
    off += (w < wave) ? t : 0;
    tot += t;
  }
  __syncthreads();
  *blockTotal = tot;
  return off + v - x;
}

__device__ __forceinline__ int block_sum(int x, int *lds4) {
  int tot;
  (void)block_exclusive_scan(x, lds4, &tot);
  return tot;
}

template <class Get>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(Get get, int n, int *tileSums) {
  __shared__ int lds[4];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    long long i = base + j;
    if (i < n) s += get((int)i);
  }
  int tot = block_sum(s, lds);
  if (threadIdx.x == 0) tileSums[blockIdx.x] = tot;
}

/* single block: exclusive scan of tileSums in place, grand total out */
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(int *tileSums, int numTiles,
                                                            int *totalDev) {
  __shared__ int lds[4];
  int carry = 0;
  for (int base = 0; base < numTiles; base += kScanTile) {
    int vals[kScanItems];
    int s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      int i = base + threadIdx.x * kScanItems + j;
      vals[j] = (i < numTiles) ? tileSums[i] : 0;
      s += vals[j];
    }
    int tot;
    int ex = block_exclusive_scan(s, lds, &tot);
    int run = carry + ex;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      int i = base + threadIdx.x * kScanItems + j;
      if (i < numTiles) tileSums[i] = run;
      run += vals[j];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && totalDev) *totalDev = carry;
}

template <class Get, class Emit>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(Get get, Emit emit, int n,
                                                            const int *tileOffsets) {
  __shared__ int lds[4];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
  int vals[kScanItems];
  int s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    long long i = base + j;
    vals[j] = (i < n) ? get((int)i) : 0;
    s += vals[j];
  }
  int tot;
  int run = block_exclusive_scan(s, lds, &tot) + tileOffsets[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    long long i = base + j;
    emit((int)i, run, i < n);
    run += vals[j];
  }
}

/* ints of scratch a scan over n items needs */
inline size_t scan_scratch_ints(long long n) {
  return (size_t)((n + kScanTile - 1) / kScanTile) + 16;
}

template <class Get, class Emit>
int scan_launch(Get get, Emit emit, int n, int *totalDev, int *scratch, hipStream_t s) {
  int tiles = (int)((n + kScanTile - 1) / kScanTile);
  if (tiles < 1) tiles = 1;
  hipLaunchKernelGGL(k_scan_reduce<Get>, dim3(tiles), dim3(kScanThreads), 0, s, get, n, scratch);
  DLP_LAUNCHED("k_scan_reduce");
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s, scratch, tiles, totalDev);
  DLP_LAUNCHED("k_scan_tiles");
  hipLaunchKernelGGL((k_scan_apply<Get, Emit>), dim3(tiles), dim3(kScanThreads), 0, s, get,
                     emit, n, (const int *)scratch);
  DLP_LAUNCHED("k_scan_apply");
  return DLP_OK;
}

}  // namespace
}  // namespace dlp

#endif /* DLP_SCAN_H */
