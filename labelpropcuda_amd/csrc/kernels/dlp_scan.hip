/*
 * dlp_scan.hip -- generic scans and the stable radix sort.
 *
 * Used by: sorted list of a batch's inserted vertices (bitmap -> ascending
 * id list), component ranks, and the label-initialisation grouping of
 * component members.
 */
#include "../dlp_internal.h"
#include "../dlp_scan.h"

namespace {

struct GetArray {
  const int *a;
  __device__ int operator()(int i) const { return a[i]; }
};
struct EmitArray {
  int *out;
  __device__ void operator()(int i, int prefix, bool valid) const {
    if (valid) out[i] = prefix;
  }
};

struct GetPopc {
  const uint32_t *bits;
  __device__ int operator()(int i) const { return __popc(bits[i]); }
};
struct EmitBits {
  const uint32_t *bits;
  int *list;
  __device__ void operator()(int i, int prefix, bool valid) const {
    if (!valid) return;
    uint32_t w = bits[i];
    int k = prefix;
    while (w) {
      int b = __ffs((int)w) - 1;
      list[k++] = i * 32 + b;
      w &= w - 1;
    }
  }
};

constexpr int kRadixThreads = 256;

__global__ __launch_bounds__(kRadixThreads) void k_radix_hist(const int *keys, int n, int shift,
                                                              int *hist, int nb) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int i = blockIdx.x * kRadixThreads + threadIdx.x;
  if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
  __syncthreads();
  hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

/* stable scatter: rank among equal digits = (earlier blocks, via the
 * digit-major scanned histogram) + (earlier waves of this block) + (lower
 * lanes of this wave with the same digit, found with 8 ballots) */
__global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(
    const int *keysIn, const int *valsIn, int n, int shift, const int *histScanned, int nb,
    int *keysOut, int *valsOut) {
  __shared__ int wc[4][256];
  for (int t = threadIdx.x; t < 1024; t += kRadixThreads) wc[t >> 8][t & 255] = 0;
  __syncthreads();
  const int i = blockIdx.x * kRadixThreads + threadIdx.x;
  const bool valid = i < n;
  const int key = valid ? keysIn[i] : 0;
  const int d = (key >> shift) & 255;
  unsigned long long m = dlp_ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1;
    unsigned long long bb = dlp_ballot(valid && bit);
    m &= bit ? bb : ~bb;
  }
  const int rank = dlp_rank_below(m);
  const int wave = threadIdx.x >> 6;
  if (valid && rank == 0) wc[wave][d] = __popcll(m);
  __syncthreads();
  if (valid) {
    int off = histScanned[d * nb + blockIdx.x];
    for (int w = 0; w < wave; ++w) off += wc[w][d];
    keysOut[off + rank] = key;
    valsOut[off + rank] = valsIn[i];
  }
}

}  // namespace

size_t dlp_scan_scratch_ints(long long n) { return dlp::scan_scratch_ints(n); }

int dlp_scan_exclusive(const int *in, int *out, int n, int *totalDev, int *scratch,
                       hipStream_t s) {
  return dlp::scan_launch(GetArray{in}, EmitArray{out}, n, totalDev, scratch, s);
}

int dlp_bitmap_to_list(const uint32_t *bitmap, int numWords, int *list, int *countDev,
                       int *scratch, hipStream_t s) {
  return dlp::scan_launch(GetPopc{bitmap}, EmitBits{bitmap, list}, numWords, countDev, scratch,
                          s);
}

size_t dlp_radix_scratch_ints(long long n) {
  long long nb = (n + kRadixThreads - 1) / kRadixThreads;
  if (nb < 1) nb = 1;
  return (size_t)(256 * nb) + dlp::scan_scratch_ints(256 * nb) + 16;
}

int dlp_radix_sort_pairs(const int *keysIn, const int *valsIn, int *keysOut, int *valsOut,
                         int *keysTmp, int *valsTmp, int n, int keyBits, int *scratch,
                         hipStream_t s) {
  if (n <= 0) return DLP_OK;
  int passes = (keyBits + 7) / 8;
  if (passes < 1) passes = 1;
  const int nb = (n + kRadixThreads - 1) / kRadixThreads;
  int *hist = scratch;
  int *scanScratch = scratch + (size_t)256 * nb;
  const int *srcK = keysIn, *srcV = valsIn;
  for (int p = 0; p < passes; ++p) {
    int *dstK = ((passes - p) & 1) ? keysOut : keysTmp;
    int *dstV = ((passes - p) & 1) ? valsOut : valsTmp;
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kRadixThreads), 0, s, srcK, n, 8 * p, hist,
                       nb);
    DLP_LAUNCHED("k_radix_hist");
    int rc = dlp_scan_exclusive(hist, hist, 256 * nb, nullptr, scanScratch, s);
    if (rc != DLP_OK) return rc;
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kRadixThreads), 0, s, srcK, srcV, n,
                       8 * p, (const int *)hist, nb, dstK, dstV);
    DLP_LAUNCHED("k_radix_scatter");
    srcK = dstK;
    srcV = dstV;
  }
  return DLP_OK;
}
