/*
 * dlp_ceilings.hip -- the device's random-access rates, the ceilings of the
 * gather- and atomic-bound propagation kernels, and the cost of a
 * device-wide sync between dependent iterations (a launch boundary against
 * an in-kernel barrier).
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

#include <algorithm>
#include <vector>

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

/* Device-wide exchange rounds, the pattern of one propagation iteration per
 * round (Jacobi buffers): every thread reads, from buffer r & 1, the word
 * another block wrote in the previous round and writes its own (value + 1)
 * to buffer (r + 1) & 1.  After `rounds` rounds every word of buffer
 * rounds & 1 holds `rounds` exactly when each round saw all writes of the
 * one before. */
constexpr int kGsThreads = 512;

__device__ __forceinline__ void gs_exchange(int *x, int blocks, int r) {
  const size_t half = (size_t)blocks * kGsThreads;
  const int other = (int)((blockIdx.x + 1u + (unsigned)r) % (unsigned)blocks);
  const int v = x[(r & 1) * half + other * kGsThreads + threadIdx.x];
  x[((r + 1) & 1) * half + blockIdx.x * kGsThreads + threadIdx.x] = v + 1;
}

/* one round per launch: the kernel boundary is the device-wide sync */
__global__ __launch_bounds__(kGsThreads) void k_gs_launch(int *x, int blocks, int r) {
  gs_exchange(x, blocks, r);
}

/* all rounds in one launch, a device-wide barrier between rounds: a
 * sense-counting barrier on bar[0] (arrivals) / bar[1] (generation) with
 * agent-scope release before arriving and acquire after leaving.  Every
 * wait is bounded (and ends at once when another block gave up), so the
 * kernel always ends; a wait that gives up sets *fail. */
__global__ __launch_bounds__(kGsThreads) void k_gs_barrier(int *x, int blocks, int rounds,
                                                           unsigned *bar, int *fail) {
  __shared__ int gaveUp;
  unsigned gen = 0u;
  for (int r = 0; r < rounds; ++r) {
    gs_exchange(x, blocks, r);
    __syncthreads();
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this block's round-r writes
      if (atomicAdd(&bar[0], 1u) == (unsigned)blocks - 1u) {
        __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        for (int spins = 0;
             __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen;
             ++spins) {
          if (spins > (1 << 20) ||
              __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            atomicOr(fail, 1);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      ++gen;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the other blocks' round-r writes
      gaveUp = __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (gaveUp) return;  // block-uniform
  }
}

/* the same rounds with a barrier of per-block flags instead of one counter:
 * each block stores its generation to its own word (no same-address
 * atomics); block 0 polls all of them with one load per thread and then
 * publishes the generation in word `blocks`, which the others poll.  Waits
 * are bounded as in k_gs_barrier. */
__global__ __launch_bounds__(kGsThreads) void k_gs_flags(int *x, int blocks, int rounds,
                                                         unsigned *flags, int *fail) {
  __shared__ int gaveUp;
  for (int r = 0; r < rounds; ++r) {
    gs_exchange(x, blocks, r);
    const unsigned gen = (unsigned)r + 1u;
    __syncthreads();
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_store(&flags[blockIdx.x], gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (blockIdx.x == 0) {  // block-uniform
      for (int spins = 0;; ++spins) {
        bool mine = true;
        for (int i = threadIdx.x; i < blocks; i += kGsThreads)
          mine = mine && __hip_atomic_load(&flags[i], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT) >= gen;
        const bool stop = spins > (1 << 20) ||
                          __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__syncthreads_and(mine)) break;
        if (__syncthreads_or(stop)) {
          if (threadIdx.x == 0) atomicOr(fail, 1);
          break;
        }
      }
      if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
        __hip_atomic_store(&flags[blocks], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else if (threadIdx.x == 0) {
      for (int spins = 0;
           __hip_atomic_load(&flags[blocks], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < gen;
           ++spins) {
        if (spins > (1 << 20) ||
            __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          atomicOr(fail, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      gaveUp = __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (gaveUp) return;  // block-uniform
  }
}

}  // namespace

/* Device-wide synchronisation between two dependent rounds of work, three
 * ways, on `blocks` work-groups of 512 threads: us[0] per round as one
 * launch per round (back to back on one stream), us[1] per round with all
 * rounds in one cooperative launch and a counter barrier between them,
 * us[2] the same with the per-block flag barrier; ok[i] is 1 when every
 * word ended at `rounds` (each round saw all writes of the one before) and
 * no barrier wait gave up. */
extern "C" int dlp_grid_sync_cost(int blocks, int rounds, double *us, int *ok) {
  if (!us || !ok || blocks < 1 || rounds < 1 || rounds > (1 << 20)) {
    dlp_set_error_("dlp_grid_sync_cost: bad argument");
    return DLP_ERR_INVALID;
  }
  int dev = 0, cus = 0, perCU = 0, perCU2 = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_gs_barrier, kGsThreads, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU2, k_gs_flags, kGsThreads, 0);
  if (e != hipSuccess) return dlp_hip_fail_(e, "dlp_grid_sync_cost");
  perCU = std::min(perCU, perCU2);
  if ((long long)perCU * cus < blocks) {
    dlp_set_errorf_("dlp_grid_sync_cost: %d work-groups cannot all be resident (%d per CU x %d)",
                    blocks, perCU, cus);
    return DLP_ERR_INVALID;
  }
  const size_t words = (size_t)blocks * kGsThreads; /* one Jacobi buffer */
  int *x = nullptr, *fail = nullptr;
  unsigned *bar = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<int> host(words);
  e = hipMalloc(&x, 2 * words * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&bar, ((size_t)blocks + 1) * sizeof(unsigned));
  if (e == hipSuccess) e = hipMalloc(&fail, sizeof(int));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int way = 0; way < 3 && e == hipSuccess; ++way) {
    e = hipMemset(x, 0, 2 * words * sizeof(int));
    if (e == hipSuccess) e = hipMemset(bar, 0, ((size_t)blocks + 1) * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(fail, 0, sizeof(int));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (way == 0) {
      for (int r = 0; r < rounds && e == hipSuccess; ++r) {
        hipLaunchKernelGGL(k_gs_launch, dim3(blocks), dim3(kGsThreads), 0, nullptr, x, blocks, r);
        e = hipGetLastError();
      }
    } else if (e == hipSuccess) {
      void *args[] = {&x, &blocks, &rounds, &bar, &fail};
      e = hipLaunchCooperativeKernel(way == 1 ? (const void *)k_gs_barrier : (const void *)k_gs_flags,
                                     dim3(blocks), dim3(kGsThreads), args, 0, nullptr);
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) us[way] = 1e3 * (double)ms / rounds;
    if (e == hipSuccess)
      e = hipMemcpy(host.data(), x + (rounds & 1) * words, words * sizeof(int),
                    hipMemcpyDeviceToHost);
    int failed = 0;
    if (e == hipSuccess) e = hipMemcpy(&failed, fail, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
      bool all = failed == 0;
      for (size_t i = 0; i < words && all; ++i) all = host[i] == rounds;
      ok[way] = all ? 1 : 0;
    }
  }
  int rc = DLP_OK;
  if (e != hipSuccess) rc = dlp_hip_fail_(e, "dlp_grid_sync_cost");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (x) (void)hipFree(x);
  if (bar) (void)hipFree(bar);
  if (fail) (void)hipFree(fail);
  return rc;
}

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
