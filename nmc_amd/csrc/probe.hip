// probe.hip - measurement kernels for the benchmark (libnmc_probe.so), not part
// of the transport library:
//   nmcp_atomic_rate  device-wide FP64 atomic-add throughput to scattered
//                     addresses, the ceiling of the fluence flushes that miss
//                     the LDS tile (one atomic per voxel visit, as in mcxyz);
//   nmcp_copy         a plain streaming copy of known size, for scaling the
//                     FETCH_SIZE / WRITE_SIZE counters to bytes;
//   nmcp_rcp_check    the transport kernels' fast reciprocal against IEEE
//                     division over every binary32 input;
//   nmcp_div_check    their short division against IEEE division on the
//                     operand domains of its call sites and on general pairs;
//   nmcp_sqrt_check   their short square root against sqrtf over every input
//                     of its domain;
//   nmcp_expm1_check  the one-path expm1 of the voxel walk against the
//                     branched one over every input <= 0;
//   nmcp_log_tab      the walks' table log of every uniform draw k 2^-24,
//                     k = 1..2^24, for comparison with the oracle;
//   nmcp_sincos_tab   the walks' table azimuth sine and cosine of every
//                     random word's top 24 bits, for the same.
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

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Operands (a, b) of case i of div_rn's check domain `dom`:
//   0  log_pos: every mantissa m of [sqrt2/2, sqrt2] as log_pos reduces it,
//      (m - 1, 2 + (m - 1))                                    i < 2^23
//   1  Henyey-Greenstein: g = -1 + (k + 1/2)/128, k = i >> 24 < 256, u = (i mod
//      2^24) 2^-24: (1 - g^2, (1 - g) + 2g u)                    i < 2^32
//   2  spin: (1 - c^2, 1 - w^2), c uniform in [-1, 1], w uniform in
//      [-(1 - 1e-6), 1 - 1e-6] (hashed)                       i < 2^32
//   3  general: sign, exponent in [-60, 60] and mantissa of a and b hashed
//                                                              i < 2^32
__device__ __forceinline__ void div_operands(int dom, uint64_t i, float* a, float* b) {
  if (dom == 0) {
    float m = __uint_as_float(0x3f800000u | (uint32_t)i);
    if (m > 0x1.6a09e6p+0f) m = m * 0.5f;
    const float f = m - 1.0f;
    *a = f;
    *b = 2.0f + f;
  } else if (dom == 1) {
    const float g = -1.0f + ((float)(i >> 24) + 0.5f) * 0x1p-7f;
    const float u = (float)(uint32_t)(i & 0xFFFFFFu) * 0x1p-24f;
    *a = 1.0f - g * g;
    *b = (1.0f - g) + (2.0f * g) * u;
  } else if (dom == 2) {
    // the rotation's 1 - cos^2 over ux^2 + uy^2: half the pairs take the
    // latter as 1 - w^2 of a unit vector (|w| < 1, down to about 1.2e-7), half
    // as any magnitude in [2^-100, 1) (the spin's lower bound for it)
    const uint64_t h = mix64(i);
    const float c = 2.0f * ((float)(uint32_t)(h >> 40) * 0x1p-24f) - 1.0f;
    *a = 1.0f - c * c;
    if (h & 1ull) {
      const float w = 2.0f * ((float)(uint32_t)((h >> 1) & 0xFFFFFFu) * 0x1p-24f) - 1.0f;
      const float r2 = 1.0f - w * w;
      *b = r2 >= 0x1p-100f ? r2 : 0x1p-100f;
    } else {
      const uint32_t e = 27u + (uint32_t)((h >> 1) % 100u);
      *b = __uint_as_float((e << 23) | (uint32_t)((h >> 8) & 0x7FFFFFu));
    }
  } else {
    const uint64_t h = mix64(i);
    const uint32_t ea = 67u + (uint32_t)((h >> 23) % 121u), eb = 67u + (uint32_t)((h >> 46) % 121u);
    *a = __uint_as_float(((uint32_t)(h >> 63) << 31) | (ea << 23) | (uint32_t)(h & 0x7FFFFFu));
    const uint64_t h2 = mix64(h);
    *b = __uint_as_float(((uint32_t)(h2 >> 63) << 31) | (eb << 23) | (uint32_t)(h2 & 0x7FFFFFu));
  }
}

// Counts the cases of domain `dom` whose div_rn(a, b) differs in bits from
// a / b; keeps the smallest differing case index.
__global__ void __launch_bounds__(256) div_check_kernel(int dom, uint64_t n, unsigned long long* bad,
                                                        unsigned long long* first_bad) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long cnt = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    float a, b;
    div_operands(dom, i, &a, &b);
    const float f = nmc::div_rn(a, b), q = a / b;
    if (__float_as_uint(f) != __float_as_uint(q)) {
      ++cnt;
      atomicMin(first_bad, (unsigned long long)i);
    }
  }
  if (cnt) atomicAdd(bad, cnt);
}

// Counts the inputs x = +0 or x in [2^-100, 2^126] (every such bit pattern)
// whose sqrt_rn(x) differs in bits from sqrtf(x); keeps the smallest.
__global__ void __launch_bounds__(256) sqrt_check_kernel(unsigned long long* bad,
                                                         unsigned int* first_bad) {
  const uint32_t lo = __float_as_uint(0x1p-100f), hi = __float_as_uint(0x1p126f);
  const uint64_t n = (uint64_t)(hi - lo) + 2u;   // [lo, hi] and +0
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long cnt = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t bits = i + 1 == n ? 0u : lo + (uint32_t)i;
    const float x = __uint_as_float(bits);
    const float f = nmc::sqrt_rn(x), q = sqrtf(x);
    if (__float_as_uint(f) != __float_as_uint(q)) {
      ++cnt;
      atomicMin(first_bad, bits);
    }
  }
  if (cnt) atomicAdd(bad, cnt);
}

// Counts the inputs x in [-inf, -0] (every such bit pattern) whose
// expm1_neg(x) differs in bits from expm1_f(x); keeps the smallest pattern.
__global__ void __launch_bounds__(256) expm1_check_kernel(unsigned long long* bad,
                                                          unsigned int* first_bad) {
  const uint32_t lo = 0x80000000u, hi = 0xFF800000u;   // -0 .. -inf
  const uint64_t n = (uint64_t)(hi - lo) + 1u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long cnt = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t bits = lo + (uint32_t)i;
    const float x = __uint_as_float(bits);
    if (__float_as_uint(nmc::expm1_neg(x)) != __float_as_uint(nmc::expm1_f(x))) {
      ++cnt;
      atomicMin(first_bad, bits);
    }
  }
  if (cnt) atomicAdd(bad, cnt);
}

// out[k - 1] = log_tab(k 2^-24) for k = 1..2^24, the table read from LDS as
// in the transport kernel.
__global__ void __launch_bounds__(256) log_tab_kernel(float* out) {
  __shared__ float4 lt[128];
  if (threadIdx.x < 128) {
    const float* row = nmc::kLogTable[threadIdx.x];
    lt[threadIdx.x] = make_float4(row[0], row[1], row[2], row[3]);
  }
  __syncthreads();
  const uint32_t n = 1u << 24;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = nmc::log_tab((float)(i + 1u) * 0x1p-24f, lt);
}

// s[k], c[k] = sincos2pi_bits(k << 8 | 0x5A) for k < 2^24, the table read
// from LDS as in the transport kernels.
__global__ void __launch_bounds__(256) sincos_tab_kernel(float* s, float* c) {
  __shared__ float2 sc[64];
  if (threadIdx.x < 64)
    sc[threadIdx.x] = make_float2(nmc::kSinCosTable[threadIdx.x][0], nmc::kSinCosTable[threadIdx.x][1]);
  __syncthreads();
  const uint32_t n = 1u << 24;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float si, ci;
    nmc::sincos2pi_bits(i << 8 | 0x5Au, sc, si, ci);
    s[i] = si;
    c[i] = ci;
  }
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

// div_rn against IEEE division on check domain `domain` (0..3, see
// div_operands); returns 0 and fills *cases, *mismatches and *first (the
// smallest differing case index, ~0 if none).
extern "C" int nmcp_div_check(int domain, unsigned long long* cases, unsigned long long* mismatches,
                              unsigned long long* first) {
  if (domain < 0 || domain > 3 || !cases || !mismatches || !first) return -1;
  const uint64_t n = domain == 0 ? (1ull << 23) : (1ull << 32);
  unsigned long long* d = nullptr;   // [mismatches, first]
  if (hipMalloc(&d, 2 * sizeof(*d)) != hipSuccess) return -4;
  int rc = -3;
  const unsigned long long init[2] = {0ull, ~0ull};
  if (hipMemcpy(d, init, sizeof(init), hipMemcpyHostToDevice) == hipSuccess) {
    hipLaunchKernelGGL(div_check_kernel, dim3(kGrid), dim3(kBlock), 0, 0, domain, n, d, d + 1);
    unsigned long long out[2];
    if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
        hipMemcpy(out, d, sizeof(out), hipMemcpyDeviceToHost) == hipSuccess) {
      *cases = n;
      *mismatches = out[0];
      *first = out[1];
      rc = 0;
    }
  }
  hipFree(d);
  return rc;
}

// sqrt_rn against sqrtf on every input of its domain; returns 0 and fills
// *mismatches and *first (the smallest differing bit pattern, 0xFFFFFFFF if
// none).
extern "C" int nmcp_sqrt_check(unsigned long long* mismatches, unsigned int* first) {
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
    hipLaunchKernelGGL(sqrt_check_kernel, dim3(kGrid), dim3(kBlock), 0, 0, d_bad, d_first);
    if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
        hipMemcpy(mismatches, d_bad, sizeof(*d_bad), hipMemcpyDeviceToHost) == hipSuccess &&
        hipMemcpy(first, d_first, sizeof(*first), hipMemcpyDeviceToHost) == hipSuccess)
      rc = 0;
  }
  hipFree(d_bad);
  hipFree(d_first);
  return rc;
}

// expm1_neg against expm1_f on every input <= 0; returns 0 and fills
// *mismatches and *first (the smallest differing bit pattern, 0xFFFFFFFF if
// none).
extern "C" int nmcp_expm1_check(unsigned long long* mismatches, unsigned int* first) {
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
    hipLaunchKernelGGL(expm1_check_kernel, dim3(kGrid), dim3(kBlock), 0, 0, d_bad, d_first);
    if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
        hipMemcpy(mismatches, d_bad, sizeof(*d_bad), hipMemcpyDeviceToHost) == hipSuccess &&
        hipMemcpy(first, d_first, sizeof(*first), hipMemcpyDeviceToHost) == hipSuccess)
      rc = 0;
  }
  hipFree(d_bad);
  hipFree(d_first);
  return rc;
}

// log_tab of every uniform draw k 2^-24, k = 1..2^24, into out[0..2^24)
// (host memory); returns 0 on success.
extern "C" int nmcp_log_tab(float* out) {
  if (!out) return -1;
  float* d = nullptr;
  const size_t bytes = sizeof(float) << 24;
  if (hipMalloc(&d, bytes) != hipSuccess) return -4;
  int rc = -3;
  hipLaunchKernelGGL(log_tab_kernel, dim3(kGrid), dim3(kBlock), 0, 0, d);
  if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
      hipMemcpy(out, d, bytes, hipMemcpyDeviceToHost) == hipSuccess)
    rc = 0;
  hipFree(d);
  return rc;
}

// sincos2pi_bits of every word k << 8 | 0x5A, k < 2^24, into s[k] and c[k]
// (host memory, 2^24 floats each); returns 0 on success.
extern "C" int nmcp_sincos_tab(float* s, float* c) {
  if (!s || !c) return -1;
  float* d = nullptr;
  const size_t bytes = sizeof(float) << 24;
  if (hipMalloc(&d, 2 * bytes) != hipSuccess) return -4;
  int rc = -3;
  hipLaunchKernelGGL(sincos_tab_kernel, dim3(kGrid), dim3(kBlock), 0, 0, d, d + (1u << 24));
  if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
      hipMemcpy(s, d, bytes, hipMemcpyDeviceToHost) == hipSuccess &&
      hipMemcpy(c, d + (1u << 24), bytes, hipMemcpyDeviceToHost) == hipSuccess)
    rc = 0;
  hipFree(d);
  return rc;
}
