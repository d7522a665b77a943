// probe.hip - measurement kernels for the benchmark (libnmc_probe.so), not part
// of the transport library:
//   nmcp_atomic_rate  device-wide FP64 atomic-add throughput to scattered
//                     addresses, the ceiling of the fluence flushes that miss
//                     the LDS tile (one atomic per voxel visit, as in mcxyz);
//   nmcp_copy         a plain streaming copy of known size, for scaling the
//                     FETCH_SIZE / WRITE_SIZE counters to bytes;
//   nmcp_rcp_check    the transport kernels' fast reciprocal against IEEE
//                     division over every binary32 input.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nmc_device.h"

namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// Each lane adds 1.0 to `iters` pseudo-random slots of [0, slots).
__global__ void __launch_bounds__(256) atomic_kernel(double* d, uint32_t slots, int iters) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int i = 0; i < iters; ++i) {
    uint32_t idx = mix32(t * 7919u + (uint32_t)i * 104729u) % slots;
    unsafeAtomicAdd(&d[idx], 1.0);
  }
}

__global__ void __launch_bounds__(256) copy_kernel(const float4* __restrict__ a,
                                                   float4* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

constexpr int kGrid = 4096, kBlock = 256;

// Counts the inputs x (all 2^32 bit patterns) whose rcp_rn(x) differs in bits
// from 1.0f / x (two NaNs count as equal); keeps the smallest differing one.
__global__ void __launch_bounds__(256) rcp_check_kernel(unsigned long long* bad,
                                                        unsigned int* first_bad) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long n = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
    const float x = __uint_as_float((uint32_t)i);
    const float f = nmc::rcp_rn(x), q = 1.0f / x;
    const bool same = __float_as_uint(f) == __float_as_uint(q) || (f != f && q != q);
    if (!same) {
      ++n;
      atomicMin(first_bad, (uint32_t)i);
    }
  }
  if (n) atomicAdd(bad, n);
}

}  // namespace

// Returns 0 on success; *ms = the timed launch, *atomics = the atomics it made.
extern "C" int nmcp_atomic_rate(uint32_t slots, int iters, float* ms, double* atomics) {
  if (slots == 0 || iters <= 0 || !ms || !atomics) return -1;
  double* d = nullptr;
  if (hipMalloc(&d, (size_t)slots * sizeof(double)) != hipSuccess) return -4;
  hipEvent_t a, b;
  int rc = -3;
  if (hipMemset(d, 0, (size_t)slots * sizeof(double)) == hipSuccess &&
      hipEventCreate(&a) == hipSuccess) {
    if (hipEventCreate(&b) == hipSuccess) {
      hipLaunchKernelGGL(atomic_kernel, dim3(kGrid), dim3(kBlock), 0, 0, d, slots, iters);  // warm-up
      hipEventRecord(a, 0);
      hipLaunchKernelGGL(atomic_kernel, dim3(kGrid), dim3(kBlock), 0, 0, d, slots, iters);
      hipEventRecord(b, 0);
      if (hipEventSynchronize(b) == hipSuccess && hipGetLastError() == hipSuccess &&
          hipEventElapsedTime(ms, a, b) == hipSuccess) {
        *atomics = (double)kGrid * kBlock * iters;
        rc = 0;
      }
      hipEventDestroy(b);
    }
    hipEventDestroy(a);
  }
  hipFree(d);
  return rc;
}

// Copies `bytes` (a multiple of 16) `reps` times; *ms = time of the last copy.
extern "C" int nmcp_copy(int64_t bytes, int reps, float* ms) {
  if (bytes <= 0 || bytes % 16 != 0 || reps <= 0 || !ms) return -1;
  float4 *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, bytes) != hipSuccess) return -4;
  if (hipMalloc(&dst, bytes) != hipSuccess) {
    hipFree(src);
    return -4;
  }
  int rc = -3;
  hipEvent_t a, b;
  if (hipMemset(src, 1, bytes) == hipSuccess && hipEventCreate(&a) == hipSuccess) {
    if (hipEventCreate(&b) == hipSuccess) {
      for (int r = 0; r < reps; ++r) {
        hipEventRecord(a, 0);
        hipLaunchKernelGGL(copy_kernel, dim3(kGrid), dim3(kBlock), 0, 0, src, dst, bytes / 16);
        hipEventRecord(b, 0);
      }
      if (hipEventSynchronize(b) == hipSuccess && hipGetLastError() == hipSuccess &&
          hipEventElapsedTime(ms, a, b) == hipSuccess)
        rc = 0;
      hipEventDestroy(b);
    }
    hipEventDestroy(a);
  }
  hipFree(src);
  hipFree(dst);
  return rc;
}

// Exhaustive check of the fast reciprocal; returns 0 and fills *mismatches
// (inputs that differ from IEEE division) and *first (the smallest such bit
// pattern, 0xFFFFFFFF if none).
extern "C" int nmcp_rcp_check(unsigned long long* mismatches, unsigned int* first) {
  if (!mismatches || !first) return -1;
  unsigned long long* d_bad = nullptr;
  unsigned int* d_first = nullptr;
  if (hipMalloc(&d_bad, sizeof(*d_bad)) != hipSuccess) return -4;
  if (hipMalloc(&d_first, sizeof(*d_first)) != hipSuccess) {
    hipFree(d_bad);
    return -4;
  }
  int rc = -3;
  const unsigned int none = 0xFFFFFFFFu;
  if (hipMemset(d_bad, 0, sizeof(*d_bad)) == hipSuccess &&
      hipMemcpy(d_first, &none, sizeof(none), hipMemcpyHostToDevice) == hipSuccess) {
    hipLaunchKernelGGL(rcp_check_kernel, dim3(kGrid), dim3(kBlock), 0, 0, d_bad, d_first);
    if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
        hipMemcpy(mismatches, d_bad, sizeof(*d_bad), hipMemcpyDeviceToHost) == hipSuccess &&
        hipMemcpy(first, d_first, sizeof(*first), hipMemcpyDeviceToHost) == hipSuccess)
      rc = 0;
  }
  hipFree(d_bad);
  hipFree(d_first);
  return rc;
}
