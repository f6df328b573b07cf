/*
 * dlp_ceilings.hip -- the device's random-access rates, the ceilings of the
 * gather- and atomic-bound propagation kernels.
 *
 * A propagation iteration on a large frontier is not limited by HBM bytes:
 * every vertex update is about twenty independent 4-byte accesses at random
 * addresses (row start and length, old and new label, the neighbours'
 * labels, a returning atomicMax per claimed neighbour).  This measures how
 * many such accesses per second the GPU sustains, with the same access
 * shapes: 4-byte random loads, 4-byte random stores and returning 32-bit
 * atomicMax, over an array of a given size (cache-resident or not).
 * Indices come from a fixed integer hash of the access number, so every
 * access is in bounds and the sequence is the same on every run.
 */
#include <hip/hip_runtime.h>

#include "../dlp_internal.h"

namespace {

__device__ __forceinline__ unsigned ra_hash(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__global__ void k_ra_atomic(int *a, unsigned size, long long n, int val, int *sink) {
  int acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    acc += atomicMax(&a[ra_hash((unsigned)i) % size], val) < val;
  if (acc == -1) sink[0] = acc;  // keeps the returned values live
}

__global__ void k_ra_load(const int *a, unsigned size, long long n, int *sink) {
  int acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    acc += a[ra_hash((unsigned)i) % size];
  if (acc == -1) sink[0] = acc;
}

__global__ void k_ra_store(int *a, unsigned size, long long n, int val) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    a[ra_hash((unsigned)i) % size] = val;
}

}  // namespace

extern "C" int dlp_random_access_rates(int64_t arrayInts, int64_t accesses, double *rates) {
  if (!rates || arrayInts < 1 || arrayInts > 0x7fffffffLL || accesses < 1) {
    dlp_set_error_("dlp_random_access_rates: bad argument");
    return DLP_ERR_INVALID;
  }
  int *a = nullptr, *sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, sizeof(int) * (size_t)arrayInts);
  if (e == hipSuccess) e = hipMalloc(&sink, 64);
  if (e == hipSuccess) e = hipMemset(a, 0, sizeof(int) * (size_t)arrayInts);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const unsigned size = (unsigned)arrayInts;
  const dim3 grid(8192), block(256);
  /* rates[0] atomicMax, [1] load, [2] store: the second of two runs each */
  for (int kind = 0; kind < 3 && e == hipSuccess; ++kind) {
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
      e = hipEventRecord(e0, nullptr);
      if (kind == 0) hipLaunchKernelGGL(k_ra_atomic, grid, block, 0, nullptr, a, size, accesses, rep + 1, sink);
      if (kind == 1) hipLaunchKernelGGL(k_ra_load, grid, block, 0, nullptr, a, size, accesses, sink);
      if (kind == 2) hipLaunchKernelGGL(k_ra_store, grid, block, 0, nullptr, a, size, accesses, rep);
      if (e == hipSuccess) e = hipGetLastError();
      if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e == hipSuccess) rates[kind] = (double)accesses / ((double)ms * 1e-3);
    }
  }
  int rc = DLP_OK;
  if (e != hipSuccess) rc = dlp_hip_fail_(e, "dlp_random_access_rates");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (sink) (void)hipFree(sink);
  return rc;
}
