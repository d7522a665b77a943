// mcsub.hip - semi-infinite homogeneous medium with cylindrical tallies
// (reference MCSubKernel, NMC/NMCCode/kernels.metal:59-319).
//
// Every loop iteration is one event of every live lane: hop, surface test
// (Fresnel all-or-none escape / mirror reflection), drop of W(1-albedo) into
// A(r,z), HG spin and roulette; lanes are refilled from a wave-private pool
// of photon ids as in mcxyz.hip.  Rd(r,alpha) and A(r,z) are privatised per
// workgroup in LDS as FP64 and written out once as block partials, which a
// second kernel sums over blocks in a fixed order into the FP64 tallies.
#include "nmc_internal.h"

namespace nmc {

namespace {

__device__ __forceinline__ int rbin(float q, int n) { return q >= (float)n ? n : (int)q; }

__device__ __forceinline__ void addf(nmc_totals* t, int k, double v) {
  unsafeAtomicAdd(&((double*)t)[k], v);
}
__device__ __forceinline__ void addi(nmc_totals* t, int k, unsigned long long v) {
  atomicAdd(&(&t->photons)[k], v);
}

template <bool LDS>
__global__ __launch_bounds__(kSubBlock) void mcsub_kernel(SubArgs a) {
  extern __shared__ double tal[];  // [(nr+1)*na] Rd(r,a) then [(nr+1)*(nz+1)] A(r,z)
  __shared__ nmc_totals st;
  __shared__ float cos_tab[64];
  const int n_ra = (a.nr + 1) * a.na;
  const int n_rz = a.tally_rz ? (a.nr + 1) * (a.nz + 1) : 0;
  double* t_ra = LDS ? tal : a.rd_ra;
  double* t_rz = LDS ? tal + n_ra : a.a_rz;
  if (LDS)
    for (int i = threadIdx.x; i < n_ra + n_rz; i += blockDim.x) tal[i] = 0.0;
  for (int i = threadIdx.x; i < a.na - 1; i += blockDim.x) cos_tab[i] = a.cos_tab[i];
  if (threadIdx.x < TF_COUNT) ((double*)&st)[threadIdx.x] = 0.0;
  if (threadIdx.x < TI_COUNT) (&st.photons)[threadIdx.x] = 0ull;
  __syncthreads();

  const int lane = lane_id();
  float x = 0.0f, y = 0.0f, z = 0.0f, ux = 0.0f, uy = 0.0f, uz = 1.0f, W = 0.0f;
  uint64_t pid = 0;
  uint32_t ev = 0, events = 0;
  bool alive = false, finished = false;
  uint64_t pool_next = 0, pool_end = 0;
  double absorbed = 0.0, specular = 0.0;
  uint32_t c_int = 0, c_phot = 0;

  for (;;) {
    uint64_t need = ballot(!alive && !finished);
    while (need) {
      if (pool_next >= pool_end) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(a.counter, (unsigned long long)a.chunk);
        base = __shfl(base, 0);
        if (base >= a.n) {
          finished = finished || !alive;
          break;
        }
        pool_next = base;
        pool_end = base + a.chunk < a.n ? base + a.chunk : a.n;
      }
      uint64_t avail = pool_end - pool_next;
      uint64_t cnt = (uint64_t)__popcll(need);
      uint32_t take = (uint32_t)(cnt < avail ? cnt : avail);
      uint32_t rank = (uint32_t)__popcll(need & lanemask_lt());
      if (((need >> lane) & 1ull) && rank < take) {
        pid = a.first + pool_next + rank;
        c_phot++;
        // LAUNCH (kernels.metal:106-166)
        U4 b = philox(pid, 0u, 0u, a.key0, a.key1);
        float rsp;
        if (a.mcflag == 0) {
          x = a.radius * sqrtf(u01(b.x));
          y = 0.0f;
          z = 0.0f;
          ux = 0.0f;
          uy = 0.0f;
          uz = 1.0f;
          rsp = a.rsp0;
        } else if (a.mcflag == 1) {
          x = a.radius * sqrtf(-log_pos(u01nz(b.x)));
          y = 0.0f;
          z = 0.0f;
          float xf = a.waist * sqrtf(-log_pos(u01nz(b.y)));
          float tmp = sqrtf((x - xf) * (x - xf) + a.zfocus * a.zfocus);
          float ct = a.zfocus / tmp;
          float uzn;
          rsp = fresnel(a.n2, a.n1, ct, &uzn);
          uz = uzn;
          ux = sqrtf(1.0f - uz * uz);
          uy = 0.0f;
        } else {
          float ct = 1.0f - 2.0f * u01(b.x);
          float stt = sqrtf(1.0f - ct * ct);
          float sp, cp;
          sincos2pi(u01(b.y), &sp, &cp);
          x = a.xs;
          y = a.ys;
          z = a.zs;
          ux = stt * cp;
          uy = stt * sp;
          uz = ct;
          rsp = 0.0f;
        }
        W = 1.0f - rsp;
        specular += (double)rsp;
        ev = 1;
        events = 0;
        alive = true;
      }
      pool_next += take;
      need = ballot(!alive && !finished);
    }
    if (ballot(alive) == 0ull) break;
    if (!alive) continue;

    U4 b = philox(pid, ev, 0u, a.key0, a.key1);
    // HOP (kernels.metal:183-187)
    float s = -log_pos(u01nz(b.x)) / a.mut;
    float x0 = x, y0 = y, z0 = z;
    x = fmaf(s, ux, x);
    y = fmaf(s, uy, y);
    z = fmaf(s, uz, z);
    events++;
    bool escaped = false;
    // ESCAPE at z <= 0 (kernels.metal:190-214)
    if (a.bmode == 1 && z <= 0.0f) {
      U4 aux = philox(pid, ev, 1u, a.key0, a.key1);
      addi(&st, TI_SURFACE_HITS, 1);
      float uz1;
      float rf = fresnel(a.n1, a.n2, -uz, &uz1);
      if (u01(aux.x) > rf) {
        float s2 = z0 / -uz;
        float xe = fmaf(s2, ux, x0);
        float ye = fmaf(s2, uy, y0);
        float r = sqrtf(xe * xe + ye * ye);
        int ir = rbin(r / a.dr, a.nr);
        int ia = 0;
        while (ia < a.na - 1 && uz1 < cos_tab[ia]) ia++;
        unsafeAtomicAdd(&t_ra[ir * a.na + ia], (double)W);
        addf(&st, TF_RD, (double)W);
        addi(&st, TI_DETECTIONS, 1);
        addi(&st, TI_SUB_CROSS, 1);
        escaped = true;
        alive = false;
      } else {
        z = -z;
        uz = -uz;
      }
    }
    if (escaped) continue;
    // DROP (kernels.metal:224-234)
    float absorb = W * a.one_m_albedo;
    W = W - absorb;
    if (a.tally_rz) {
      float r = sqrtf(x * x + y * y);
      int ir = rbin(r / a.dr, a.nr);
      int iz = rbin(fabsf(z) / a.dz, a.nz);
      unsafeAtomicAdd(&t_rz[ir * (a.nz + 1) + iz], (double)absorb);
    }
    absorbed += (double)absorb;
    // SPIN (kernels.metal:243-279, HG honoured)
    float ct = hg_cos(a.g, u01(b.y));
    spin(ct, u01(b.z), ux, uy, uz);
    c_int++;
    // ROULETTE (kernels.metal:289-294)
    if (W < a.thr) {
      if (u01(b.w) <= a.chance) {
        float nw = W / a.chance;
        addf(&st, TF_ROULETTE_GAIN, (double)nw - (double)W);
        addi(&st, TI_ROULETTE_SURVIVALS, 1);
        W = nw;
      } else {
        addf(&st, TF_ROULETTE_LOSS, (double)W);
        addi(&st, TI_ROULETTE_KILLS, 1);
        alive = false;
        continue;
      }
    }
    if (events >= a.max_events) {
      addf(&st, TF_CAPPED, (double)W);
      addi(&st, TI_CAPPED, 1);
      alive = false;
      continue;
    }
    ev++;
  }

  addf(&st, TF_ABSORBED, absorbed);
  addf(&st, TF_SPECULAR, specular);
  addi(&st, TI_SUB_SAME, c_int);
  addi(&st, TI_INTERACTIONS, c_int);
  addi(&st, TI_PHOTONS, c_phot);
  __syncthreads();
  if (LDS) {
    double* part = a.part + (int64_t)blockIdx.x * (n_ra + n_rz);
    for (int i = threadIdx.x; i < n_ra + n_rz; i += blockDim.x) part[i] = tal[i];
  }
  if (threadIdx.x < TF_COUNT) unsafeAtomicAdd(&((double*)a.tot)[threadIdx.x], ((double*)&st)[threadIdx.x]);
  if (threadIdx.x < TI_COUNT)
    atomicAdd(&(&a.tot->photons)[threadIdx.x], (unsigned long long)(&st.photons)[threadIdx.x]);
}

// out[j] += sum_b part[b * stride + j], b in ascending order (deterministic).
__global__ void reduce_partials_kernel(const double* __restrict__ part, int grid, int64_t stride,
                                       int64_t len, double* __restrict__ out) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= len) return;
  double s = 0.0;
  for (int b = 0; b < grid; ++b) s += part[(int64_t)b * stride + j];
  out[j] += s;
}

}  // namespace

hipError_t launch_mcsub(const SubArgs& a, int grid, int lds_bytes, hipStream_t s) {
  if (a.lds_tallies) {
    hipFuncSetAttribute((const void*)mcsub_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        lds_bytes);
    hipLaunchKernelGGL(mcsub_kernel<true>, dim3(grid), dim3(kSubBlock), lds_bytes, s, a);
  } else {
    hipLaunchKernelGGL(mcsub_kernel<false>, dim3(grid), dim3(kSubBlock), 0, s, a);
  }
  return hipGetLastError();
}

int mcsub_blocks_per_cu(int lds_bytes) {
  int nb = 0;
  hipError_t e = lds_bytes > 0
      ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mcsub_kernel<true>, kSubBlock, lds_bytes)
      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, mcsub_kernel<false>, kSubBlock, 0);
  return (e == hipSuccess && nb > 0) ? nb : 1;
}

hipError_t launch_reduce_partials(const double* part, int grid, int64_t stride, int64_t len,
                                  double* out, hipStream_t s) {
  if (len <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s,
                     part, grid, stride, len, out);
  return hipGetLastError();
}

}  // namespace nmc
