// mcsub.hip - semi-infinite homogeneous medium with cylindrical tallies
// (reference MCSubKernel, NMC/NMCCode/kernels.metal:59-319), plain and
// batched.
//
// One kernel serves both: a launch transports the photons of `n_items` work
// items (media with their own optical properties, Philox stream and photon
// budget) from one global photon queue.  Queue index q in [0, sum budgets)
// maps to item j by binary search over the budget prefix sums and to photon id
// first + (q - prefix[j]) within that item; a plain run is a batch of one.
// Since an item's photons depend only on (seed, item stream, photon id), item j
// of a batch reproduces, photon for photon, a plain run of that medium.
//
// Every loop iteration is one event of every live lane: hop, surface test
// (Fresnel all-or-none escape / mirror reflection), drop of W(1-albedo) into
// A(r,z), HG spin and roulette.  Lanes are refilled from a wave-private pool of
// queue indices as in mcxyz.hip.  Per-item Rd(r,alpha), A(r,z) and totals are
// privatised per workgroup in LDS (FP64) when they fit, written out once as
// block partials and summed over blocks in a fixed order by a second kernel;
// otherwise they are accumulated with FP64 global atomics.
#include "nmc_internal.h"

namespace nmc {

namespace {

__device__ __forceinline__ int rbin(float q, int n) { return q >= (float)n ? n : (int)q; }

// Largest j in [0, n) with prefix[j] <= q (the work item of queue index q).
__device__ __forceinline__ int find_item(const uint64_t* prefix, int n, uint64_t q) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= q) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ void addf(nmc_totals* t, int k, double v) {
  unsafeAtomicAdd(&((double*)t)[k], v);
}
__device__ __forceinline__ void addi(nmc_totals* t, int k, unsigned long long v) {
  atomicAdd(&(&t->photons)[k], v);
}

template <bool LDS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void mcsub_kernel(SubArgs a) {
  // dynamic LDS: [items] media | [items+1] prefix | [items] totals | tallies (LDS mode)
  extern __shared__ double smem[];
  __shared__ float cos_tab[64];
  __shared__ float4 LT[128];  // log_tab's rows
  __shared__ float2 SC[64];   // sincos2pi_bits' rows
  const int ni = a.n_items;
  const int n_ra = (a.nr + 1) * a.na;
  const int n_rz = a.tally_rz ? (a.nr + 1) * (a.nz + 1) : 0;
  const int per_item = n_ra + n_rz;
  SubMedium* med = (SubMedium*)smem;
  uint64_t* prefix = (uint64_t*)(med + ni);
  nmc_totals* st = (nmc_totals*)(prefix + ni + 1);
  double* tal = (double*)(st + ni);
  double* tally = LDS ? tal : a.tally + (int64_t)(blockIdx.x % (unsigned)a.reps) * ni * per_item;
  nmc_totals* tot = st;
  for (int i = threadIdx.x; i < ni; i += BLOCK) med[i] = a.items[i];
  for (int i = threadIdx.x; i <= ni; i += BLOCK) prefix[i] = a.prefix[i];
  for (int i = threadIdx.x; i < a.na - 1; i += BLOCK) cos_tab[i] = a.cos_tab[i];
  if (threadIdx.x < 128) {
    const float* row = kLogTable[threadIdx.x];
    LT[threadIdx.x] = make_float4(row[0], row[1], row[2], row[3]);
  }
  if (threadIdx.x < 64) SC[threadIdx.x] = make_float2(kSinCosTable[threadIdx.x][0], kSinCosTable[threadIdx.x][1]);
  for (int i = threadIdx.x; i < ni * (TF_COUNT + TI_COUNT); i += BLOCK) ((double*)st)[i] = 0.0;
  if (LDS)
    for (int i = threadIdx.x; i < ni * per_item; i += BLOCK) tal[i] = 0.0;
  __syncthreads();
  const uint64_t first = a.state->first;
  const uint64_t total = prefix[ni];
  // the bins' divisions r / dr and |z| / dz by Markstein's correction on a
  // reciprocal formed once (IEEE division's result, in fewer instructions)
  const bool dr_ok = a.dr >= 0x1p-60f && a.dr <= 0x1p60f;
  const bool dz_ok = a.dz >= 0x1p-60f && a.dz <= 0x1p60f;
  const float y_dr = dr_ok ? rcp_rn_mid(a.dr) : 0.0f;
  const float y_dz = dz_ok ? rcp_rn_mid(a.dz) : 0.0f;
  // bin offsets: with the tallies in LDS every index is below 2^24, so the
  // products take 24-bit multiplies (full rate)
  auto mul_idx = [](int u, int v) -> int {
    return LDS ? (int)__umul24((uint32_t)u, (uint32_t)v) : u * v;
  };

  const int lane = lane_id();
  float x = 0.0f, y = 0.0f, z = 0.0f, ux = 0.0f, uy = 0.0f, uz = 1.0f, W = 0.0f;
  SubMedium m{};
  KeySched ks = key_sched(0u);
  int item = 0;
  uint64_t pid = 0;
  uint32_t ev = 0, events = 0, c_int = 0;
  bool alive = false, finished = false;
  uint64_t pool_next = 0, pool_end = 0;  // wave-uniform
  // wave-uniform: the item holding queue range [pool_lo, pool_hi), found once
  // per pool position so most refills need no per-lane search
  int pool_item = 0;
  uint64_t pool_lo = 0, pool_hi = 0;
  double absorbed = 0.0;

  for (;;) {
    uint64_t need = ballot(!alive && !finished);
    if (need && (uint32_t)__popcll(need) < a.refill_min && ballot(alive) != 0ull) need = 0;
    while (need) {
      if (pool_next >= pool_end) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&a.state->counter, (unsigned long long)a.chunk);
        base = __shfl(base, 0);
        if (base >= total) {
          finished = finished || !alive;
          break;
        }
        pool_next = base;
        pool_end = base + a.chunk < total ? base + a.chunk : total;
      }
      if (pool_next < pool_lo || pool_next >= pool_hi) {
        pool_item = find_item(prefix, ni, pool_next);
        pool_lo = prefix[pool_item];
        pool_hi = prefix[pool_item + 1];
      }
      uint64_t avail = pool_end - pool_next;
      uint64_t cnt = (uint64_t)__popcll(need);
      uint32_t take = (uint32_t)(cnt < avail ? cnt : avail);
      uint32_t rank = (uint32_t)__popcll(need & lanemask_lt());
      if (((need >> lane) & 1ull) && rank < take) {
        uint64_t q = pool_next + rank;
        int lo = q < pool_hi ? pool_item : find_item(prefix, ni, q);
        m = med[lo];
        ks = key_sched(m.key1);
        item = m.slot;
        pid = first + (q - prefix[lo]);
        // LAUNCH (kernels.metal:106-166); a pencil beam (radius 0) uses no draw
        float rsp;
        if (a.mcflag == 0) {
          x = 0.0f;
          if (a.radius != 0.0f) {
            U4 b = philox(pid, 0u, 0u, a.key0, m.key1);
            x = a.radius * sqrtf(u01(b.x));
          }
          y = 0.0f;
          z = 0.0f;
          ux = 0.0f;
          uy = 0.0f;
          uz = 1.0f;
          rsp = m.rsp0;
        } else if (a.mcflag == 1) {
          U4 b = philox(pid, 0u, 0u, a.key0, m.key1);
          x = a.radius * sqrtf(-log_pos(u01nz(b.x)));
          y = 0.0f;
          z = 0.0f;
          float xf = a.waist * sqrtf(-log_pos(u01nz(b.y)));
          float tmp = sqrtf((x - xf) * (x - xf) + a.zfocus * a.zfocus);
          float ct = a.zfocus / tmp;
          float uzn;
          rsp = fresnel(m.n2, m.n1, ct, &uzn);
          uz = uzn;
          ux = sqrtf(1.0f - uz * uz);
          uy = 0.0f;
        } else {
          U4 b = philox(pid, 0u, 0u, a.key0, m.key1);
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
        addf(&tot[item], TF_SPECULAR, (double)rsp);
        addi(&tot[item], TI_PHOTONS, 1);
        ev = 1;
        events = 0;
        c_int = 0;
        absorbed = 0.0;
        alive = true;
      }
      pool_next += take;
      need = ballot(!alive && !finished);
    }
    if (ballot(alive) == 0ull) break;
    if (!alive) continue;

    U4 b = philox(pid, ev, 0u, a.key0, ks);
    // HOP (kernels.metal:183-187)
    float s = -log_tab(u01nz(b.x), LT) * m.inv_mut;
    float x0 = x, y0 = y, z0 = z;
    x = fmaf(s, ux, x);
    y = fmaf(s, uy, y);
    z = fmaf(s, uz, z);
    events++;
    bool ended = false;
    // ESCAPE at z <= 0 (kernels.metal:190-214)
    if (a.bmode == 1 && z <= 0.0f) {
      addi(&tot[item], TI_SURFACE_HITS, 1);
      float uz1;
      float rf = fresnel(m.n1, m.n2, -uz, &uz1);
      if (u01_low(b) > rf) {
        float s2 = z0 / -uz;
        float xe = fmaf(s2, ux, x0);
        float ye = fmaf(s2, uy, y0);
        float r = sqrtf(xe * xe + ye * ye);
        int ir = rbin(div_rn_by(r, a.dr, y_dr, dr_ok), a.nr);
        int ia = 0;
        while (ia < a.na - 1 && uz1 < cos_tab[ia]) ia++;
        unsafeAtomicAdd(&tally[mul_idx(item, per_item) + mul_idx(ir, a.na) + ia], (double)W);
        addf(&tot[item], TF_RD, (double)W);
        addi(&tot[item], TI_DETECTIONS, 1);
        addi(&tot[item], TI_SUB_CROSS, 1);
        ended = true;
      } else {
        z = -z;
        uz = -uz;
      }
    }
    if (!ended) {
      // DROP (kernels.metal:224-234)
      float absorb = W * m.one_m_albedo;
      W = W - absorb;
      if (a.tally_rz) {
        float r = sqrtf(x * x + y * y);
        int ir = rbin(div_rn_by(r, a.dr, y_dr, dr_ok), a.nr);
        int iz = rbin(div_rn_by(fabsf(z), a.dz, y_dz, dz_ok), a.nz);
        unsafeAtomicAdd(&tally[mul_idx(item, per_item) + n_ra + mul_idx(ir, a.nz + 1) + iz],
                        (double)absorb);
      }
      absorbed += (double)absorb;
      // SPIN (kernels.metal:243-279, HG honoured)
      float ct = hg_cos(m.hg, u01(b.y));
      spin_w(ct, b.z, SC, ux, uy, uz);
      c_int++;
      // ROULETTE (kernels.metal:289-294)
      if (W < a.thr) {
        if (u01(b.w) <= a.chance) {
          float nw = W / a.chance;
          addf(&tot[item], TF_ROULETTE_GAIN, (double)nw - (double)W);
          addi(&tot[item], TI_ROULETTE_SURVIVALS, 1);
          W = nw;
        } else {
          addf(&tot[item], TF_ROULETTE_LOSS, (double)W);
          addi(&tot[item], TI_ROULETTE_KILLS, 1);
          ended = true;
        }
      }
      if (!ended && events >= a.max_events) {
        addf(&tot[item], TF_CAPPED, (double)W);
        addi(&tot[item], TI_CAPPED, 1);
        ended = true;
      }
    }
    if (ended) {
      addf(&tot[item], TF_ABSORBED, absorbed);
      addi(&tot[item], TI_SUB_SAME, c_int);
      addi(&tot[item], TI_INTERACTIONS, c_int);
      alive = false;
    } else {
      ev++;
    }
  }

  __syncthreads();
  // block partials: [tallies (LDS mode)] [totals]
  const int lds_tal = LDS ? ni * per_item : 0;
  double* part = a.part + (int64_t)blockIdx.x * (lds_tal + ni * (TF_COUNT + TI_COUNT));
  for (int i = threadIdx.x; i < lds_tal; i += BLOCK) part[i] = tal[i];
  double* pt = part + lds_tal;
  for (int i = threadIdx.x; i < ni * (TF_COUNT + TI_COUNT); i += BLOCK) pt[i] = ((double*)st)[i];
}

// out[j] += sum_b part[b * stride + j] for j < len, blocks in ascending order
// (deterministic).  Integer totals are stored as u64 bit patterns in the
// partials and are summed as integers.
__global__ void reduce_partials_kernel(double* __restrict__ part, int grid, int64_t stride,
                                       int64_t len, double* __restrict__ out, int ni, int per_item,
                                       int clear) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= len) return;
  int64_t tal_len = (int64_t)ni * per_item;
  bool is_int = false;
  if (j >= tal_len) {
    int k = (int)((j - tal_len) % (TF_COUNT + TI_COUNT));
    is_int = k >= TF_COUNT;
  }
  if (is_int) {
    unsigned long long s = 0;
    for (int b = 0; b < grid; ++b) s += __double_as_longlong(part[(int64_t)b * stride + j]);
    unsigned long long* o = (unsigned long long*)out;
    o[j] += s;
  } else {
    double s = 0.0;
    for (int b = 0; b < grid; ++b) s += part[(int64_t)b * stride + j];
    out[j] += s;
  }
  if (clear)
    for (int b = 0; b < grid; ++b) part[(int64_t)b * stride + j] = 0.0;
}

__global__ void state_begin_kernel(RunState* st) { st->counter = 0ull; }
__global__ void state_advance_kernel(RunState* st, unsigned long long adv) { st->first += adv; }

}  // namespace

size_t mcsub_lds_bytes(int n_items, int64_t per_item_in_lds) {
  return (size_t)n_items * sizeof(SubMedium) + (size_t)(n_items + 1) * sizeof(uint64_t) +
         (size_t)n_items * sizeof(nmc_totals) + (size_t)n_items * per_item_in_lds * sizeof(double);
}

namespace {
template <bool L>
const void* sub_fn(int block) {
  switch (block) {
    case 256: return (const void*)mcsub_kernel<L, 256>;
    case 512: return (const void*)mcsub_kernel<L, 512>;
    default: return (const void*)mcsub_kernel<L, 1024>;
  }
}
}  // namespace

hipError_t launch_mcsub(const SubArgs& a, int block, int grid, size_t lds_bytes, bool lds_tallies,
                        hipStream_t s) {
  const void* f = lds_tallies ? sub_fn<true>(block) : sub_fn<false>(block);
  SubArgs arg = a;
  void* args[] = {&arg};
  return hipLaunchKernel(f, dim3(grid), dim3(block), args, lds_bytes, s);
}

hipError_t mcsub_configure(int block, size_t lds_bytes, bool lds_tallies) {
  const void* f = lds_tallies ? sub_fn<true>(block) : sub_fn<false>(block);
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}

int mcsub_blocks_per_cu(int block, size_t lds_bytes, bool lds_tallies) {
  int nb = 0;
  const void* f = lds_tallies ? sub_fn<true>(block) : sub_fn<false>(block);
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, block, lds_bytes);
  return e == hipSuccess ? nb : 0;
}

hipError_t launch_reduce_partials(double* part, int grid, int64_t stride, double* out,
                                  int ni, int per_item, hipStream_t s, bool clear) {
  if (stride <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)((stride + 255) / 256)), dim3(256), 0,
                     s, part, grid, stride, stride, out, ni, per_item, clear ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_state_begin(RunState* st, hipStream_t s) {
  hipLaunchKernelGGL(state_begin_kernel, dim3(1), dim3(1), 0, s, st);
  return hipGetLastError();
}

hipError_t launch_state_advance(RunState* st, uint64_t adv, hipStream_t s) {
  hipLaunchKernelGGL(state_advance_kernel, dim3(1), dim3(1), 0, s, st, (unsigned long long)adv);
  return hipGetLastError();
}

}  // namespace nmc
