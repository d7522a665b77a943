// mcxyz.hip - voxelised 3-D photon transport (reference MCXYZKernel,
// NMC/NMCCode/kernels.metal:321-754, and the two-term variant's sampling,
// kernels.metal:1053-1070).
//
// Launch model: persistent wave64 workgroups.  Each lane carries one photon;
// the loop body advances every live lane by one piece of its current step
// (to the next interaction, voxel face or detector plane) and, for lanes whose
// step is used up, performs the interaction (spin + roulette + next step
// length).  A wave refills its idle lanes only once at least `refill_min` of
// them are idle (or none is live), so the launch code runs for many lanes at
// once; refills draw from a wave-private pool of photon ids replenished
// `chunk` ids at a time from one device counter.  A photon's random numbers
// depend only on its id, never on the lane or the order of execution.
//
// Fluence: deposits are summed per voxel visit in a register (FP32) and added
// when the photon leaves the voxel - into the workgroup's LDS tile (FP64, or
// FP32 with twice the voxels: TileT) if the voxel lies inside the tile box,
// else with an FP64 global atomic, or (kFluRecords) appended as an 8-byte
// record to the workgroup's slice of the voxel's bucket.  The records are
// summed per bucket after the launch: by rec_reduce_kernel, or by the next
// launch's workgroups, each once its own photons are done (its LDS then
// holds a bucket's FP64 sums), which puts that work into the next launch's
// drained tail, where workgroups finish one by one.  The tiles are written
// out as FP64 block partials and summed over blocks in a fixed order.  Rare
// events (detections, escapes, roulette) go to LDS totals, one atomic per
// lane; frequent counters stay in registers until the workgroup ends.  The
// source and detector-plane code is kept out of line (source_photon,
// detector_plane) with its parameters in device memory, the loop-invariant
// values the loop tests or indexes with (tile box, grid size, boundary and
// detector modes, fluence and voxel-map pointers) sit in vector registers
// (in_vgpr), and the fluence mode (none, deposit, visit count, records,
// sensitivities) is a kernel variant: the scalar registers, which the Philox
// key schedule largely fills, then hold no parked values the loop must fetch.
#include "nmc_internal.h"

namespace nmc {

namespace {

struct Photon {
  float x, y, z, ux, uy, uz, rux, ruy, ruz, W, sleft, mua, mus, inv_mus;
  int ix, iy, iz, tis;
  // vidx(ix, iy, iz), kept up to date by the crossings: a step of one voxel
  // adds +-1, +-Ny or +-Nx*Ny, exact in 32-bit modular arithmetic also while
  // the photon is outside the grid, so it is right again when it re-enters
  uint32_t lin;
  bool inside, alive;
  uint64_t pid;
  uint32_t ev, events, b1, b2, b3;
  // the current voxel visit's deposits, summed in FP32 (a visit is one to a
  // few pieces) and added to the tallies in FP64 when it ends
  float vdep;
  // sensitivity variant: the voxel visit's sums (sp = sum D L_k, sq = sum D
  // K_k, sza = sum W s e^{-x}, szs = sum W s h(x)) and the tissue's slot
  float sp, sq, sza, szs;
  int slot;
};

__device__ __forceinline__ int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

// Linear voxel index iz*Ny*Nx + ix*Ny + iy (kernels.metal:482) of an in-grid
// voxel, in 32-bit arithmetic: nmc_prepare rejects grids above 2e9 voxels
// (the reference indexes with a 32-bit int), so every partial sum fits.
__device__ __forceinline__ uint32_t vidx(const XyzArgs& a, int ix, int iy, int iz) {
  return ((uint32_t)iz * (uint32_t)a.Nx + (uint32_t)ix) * (uint32_t)a.Ny + (uint32_t)iy;
}

// The photon's tissue becomes `tis`: its id and the optical row it reads.
template <bool SENS>
__device__ __forceinline__ void take_tissue_t(Photon& p, const XyzTissue* T, int tis) {
  p.tis = tis;
  p.mua = T[tis].mua;
  p.mus = T[tis].mus;
  p.inv_mus = T[tis].inv_mus;
  if constexpr (SENS) p.slot = T[tis].slot;
}

__device__ __forceinline__ void lds_addi(nmc_totals* t, int k, unsigned long long v) {
  atomicAdd(&((unsigned long long*)&t->photons)[k], v);
}
__device__ __forceinline__ void lds_addf(nmc_totals* t, int k, double v) {
  unsafeAtomicAdd(&((double*)t)[k], v);
}

// The LDS tile box and the fluence array, in vector registers (in_vgpr).
struct TileBox {
  uint32_t x0, y0, z0, nx, ny, nz;
  gptr<double> F;
};

// Fluence tally of a kernel variant: none (NMC_FLAG_NO_FLUENCE), the
// deposited weight, or the number of voxel visits (the tile-placement pilot).
// Each is a kernel of its own, so the loop carries no per-visit mode test.
// kFluRecords: deposits, with the visits outside the tile appended as binned
// records (tuning.mode 2) instead of FP64 atomics.
enum : int { kFluNone = 0, kFluDeposit = 1, kFluCount = 2, kFluRecords = 3, kFluSens = 4 };

// The sensitivity variant's per-lane path length L and scatterings K per
// tracked tissue slot, in LDS rows [slot][lane]; these point at this lane's
// column (entry j at [j * stride]).
struct SensRows {
  float* L;
  uint32_t* K;
  uint32_t stride;
};

// A workgroup's slice of one bucket's record region, in LDS: the first
// record index, the capacity and the slots taken so far.
struct RecSlice {
  unsigned long long base;
  uint32_t cap, cnt;
};

// The record stream of a kernel variant (kFluRecords only): the slices in
// LDS and the record array, in vector registers.
struct RecOut {
  RecSlice* slices;
  gptr<unsigned long long> rec;   // as uint2 {voxel index in the bucket, weight bits}
};

// Adds the deposits of the current voxel visit to the tile or the global grid
// (kFluCount: adds 1 per visit instead; kFluRecords: appends a record instead
// of the global atomic while the bucket's slice has room), and to the lane's
// absorbed total.
//
// kFluSens also folds the visit into the per-voxel sensitivities (no
// reference counterpart): with D the visit's deposit in tissue k, L_t and K_t
// the photon's path length and scatterings in tracked tissue t, it adds
// d/dmua_t = -L_t D and d/dmus_t = (K_t/mus_t - L_t) D for t != k, and the
// visit's sums -sp + sza and sq/mus_k - sp - szs for t = k (the layered
// walk's per-segment terms, summed over the segments in the voxel).
// The voxel of a visit: its linear index, its coordinates relative to the
// tile box's origin (below the box they wrap to large values, so one
// unsigned test per axis places them) and whether it lies in the grid.  A
// crossing takes it before it moves the photon's voxel indices on.
struct VisitAt {
  uint32_t lin, lx, ly, lz;
  bool inside;
};

__device__ __forceinline__ VisitAt visit_at(const Photon& p, const TileBox& tb) {
  return VisitAt{p.lin, (uint32_t)p.ix - tb.x0, (uint32_t)p.iy - tb.y0, (uint32_t)p.iz - tb.z0,
                 p.inside};
}

template <int FMODE, typename TileT>
__device__ __forceinline__ void flush_visit(Photon& p, const VisitAt& at, const XyzArgs& a,
                                            const TileBox& tb, TileT* tile, const RecOut& ro,
                                            double& absorbed, const SensRows& sr) {
  constexpr bool COUNT = FMODE == kFluCount;
  if constexpr (FMODE == kFluSens) {
    if (at.inside && (p.vdep != 0.0f || p.sza != 0.0f)) {
      const size_t nvox = (size_t)a.Nx * (size_t)a.Ny * (size_t)a.Nz;
      const gptr<double> out = (gptr<double>)a.sens;
      for (int j = 0; j < a.sens_n; ++j) {
        double ga, gs;
        if (j == p.slot) {
          ga = -(double)p.sp + (double)p.sza;
          gs = (double)p.sq * (double)p.inv_mus - (double)p.sp - (double)p.szs;
        } else {
          const float Lj = sr.L[j * sr.stride];
          const uint32_t Kj = sr.K[j * sr.stride];
          if (p.vdep == 0.0f || (Lj == 0.0f && Kj == 0u)) continue;
          ga = -(double)Lj * (double)p.vdep;
          gs = ((double)Kj * (double)a.sens_inv_mus[j] - (double)Lj) * (double)p.vdep;
        }
        global_add(out + ((size_t)j * nvox + at.lin), ga);
        global_add(out + ((size_t)(a.sens_n + j) * nvox + at.lin), gs);
      }
    }
    p.sp = p.sq = p.sza = p.szs = 0.0f;
  }
  absorbed += (double)p.vdep;
  if constexpr (FMODE == kFluNone) {
    p.vdep = 0.0f;
    return;
  }
  if (at.inside && (COUNT || p.vdep != 0.0f)) {
    double v = COUNT ? 1.0 : (double)p.vdep;
    if (at.lx < tb.nx && at.ly < tb.ny && at.lz < tb.nz) {
      unsafeAtomicAdd(&tile[(at.lz * tb.nx + at.lx) * tb.ny + at.ly], (TileT)v);
    } else {
      const uint32_t k = at.lin;
      if (FMODE == kFluRecords) {
        RecSlice* sl = &ro.slices[k >> kBucketShift];
        // the slice's base and capacity are read before the cursor atomic (a
        // volatile LDS read stays in place), so one wait covers all three
        const unsigned long long sbase =
            *(volatile __attribute__((address_space(3))) unsigned long long*)&sl->base;
        const uint32_t scap = sl->cap;
        const uint32_t slot = atomicAdd(&sl->cnt, 1u);
        if (slot < scap) {
          ro.rec[sbase + slot] = ((unsigned long long)__float_as_uint((float)v) << 32) |
                                 (k & (kBucketVox - 1u));
          p.vdep = 0.0f;
          return;
        }
      }
      global_add(&tb.F[k], v);
    }
  }
  p.vdep = 0.0f;
}

// Event tallies of the transport loop into the workgroup totals in LDS, at
// an address the compiler cannot prove wave-uniform (zv: a zero in a vector
// register).  For a uniform address it would sum the active lanes first, an
// FP64 add per active lane in a loop that the whole wave pays each time any
// lane takes the branch; this way each lane issues its own LDS atomic, one
// instruction for the wave.
struct LoopTotals {
  nmc_totals* st;
  uint32_t zv;
  __device__ __forceinline__ void addf(int k, double v) const {
    unsafeAtomicAdd(&((double*)st)[(uint32_t)k + zv], v);
  }
  __device__ __forceinline__ void addi(int k) const {
    atomicAdd(&((unsigned long long*)&st->photons)[(uint32_t)k + zv], 1ull);
  }
};

// Escape through a grid face under boundaryflag 1 (or the top under 2).
__device__ __forceinline__ void tally_escape(const LoopTotals& t, const XyzArgs& a, int iz, float W) {
  if (iz >= a.Nz) {
    t.addf(TF_TT, (double)W);
    t.addi(TI_ESC_BOTTOM);
  } else if (iz < 0) {
    t.addf(TF_TOP, (double)W);
    t.addi(TI_ESC_TOP);
  } else {
    t.addf(TF_LATERAL, (double)W);
    t.addi(TI_ESC_SIDE);
  }
}

// Start position, direction and weight of photon `pid` for the run's source
// (kernels.metal:366-463), and the specular loss of a focused-Gaussian beam
// entering the tissue (0 otherwise).  Not inlined: refills are rare, and the
// source parameters then stay out of the transport loop's registers.
struct SourceState {
  float x, y, z, ux, uy, uz, W, spec;
};

__device__ __noinline__ SourceState source_photon(uint64_t pid, uint32_t key0, uint32_t key1,
                                                  const XyzAux* __restrict__ aux) {
  const XyzAux& s = *aux;
  float x, y, z, ux, uy, uz, W = 1.0f, spec = 0.0f;
  if (s.launchflag == 1) {
    x = s.xs; y = s.ys; z = s.zs;
    ux = s.ux0; uy = s.uy0; uz = s.uz0;
  } else {
    U4 b = philox(pid, 0u, 0u, key0, key1);
    if (s.mcflag == 0 || s.mcflag == 1) {
      float r;
      if (s.mcflag == 0) r = s.radius * sqrtf(u01nz(b.x));
      else r = s.radius * sqrtf(-log_pos(u01nz(b.x)));
      float sp, cp;
      sincos2pi(u01nz(b.y), &sp, &cp);
      x = s.xs + r * cp;
      y = s.ys + r * sp;
      z = s.zs;
      float rf = s.waist * sqrtf(u01nz(b.z));
      sincos2pi(u01nz(b.w), &sp, &cp);
      float xf = rf * cp, yf = rf * sp;
      float ddx = x - xf, ddy = y - yf;
      float tmp = sqrtf(ddx * ddx + ddy * ddy + s.zfocus * s.zfocus);
      ux = -ddx / tmp;
      uy = -ddy / tmp;
      uz = sqrtf(1.0f - ux * ux - uy * uy);
      if (s.mcflag == 1) {
        float ca2;
        float rsp = fresnel(s.n2, s.n1, fabsf(uz), &ca2);
        spec = W * rsp;
        W = W * (1.0f - rsp);
        snell(s.n2, s.n1, &ux, &uy, &uz);
      }
    } else if (s.mcflag == 2) {
      float ct = 1.0f - 2.0f * u01(b.x);
      float stt = sqrtf(1.0f - ct * ct);
      float sp, cp;
      sincos2pi(u01(b.y), &sp, &cp);
      x = s.xs; y = s.ys; z = s.zs;
      ux = stt * cp; uy = stt * sp; uz = ct;
    } else if (s.mcflag == 3) {
      x = s.radius * (u01nz(b.x) * 2.0f - 1.0f);
      y = s.radius * (u01nz(b.y) * 2.0f - 1.0f);
      z = s.zs;
      ux = 0.0f; uy = 0.0f; uz = 1.0f;
    } else {
      x = 0.0f; y = 0.0f; z = s.zs;
      ux = 0.0f; uy = 0.0f; uz = 1.0f;
    }
  }
  return SourceState{x, y, z, ux, uy, uz, W, spec};
}

// A photon reaching the detector plane z = zsurf (kernels.metal:556-630):
// the Fresnel reflectance `rsp` of the plane and, for the transmitted part
// `wt`, whether it is detected (where = 1: inside the detector annulus and
// numerical aperture; its weight and count go to Ryx/DetPhot2D here) or
// leaves outside it (where = 2); where = 0 when nothing is transmitted.
struct DetectorHit {
  float rsp, wt;
  int where;
};

__device__ __noinline__ DetectorHit detector_plane(float x, float y, float uz, float W, int ix,
                                                   int iy, const XyzAux* __restrict__ aux) {
  const XyzAux& d = *aux;
  float uzt;
  DetectorHit h{fresnel(d.n1, d.n2, fabsf(uz), &uzt), 0.0f, 0};
  if (h.rsp < 1.0f) {
    h.wt = W * (1.0f - h.rsp);
    float ro = sqrtf(x * x + y * y);
    if (ro >= d.rmin && ro <= d.rmax && uzt >= d.cos_na) {
      int64_t k = (int64_t)clampi(ix, d.Nx) * d.Ny + clampi(iy, d.Ny);
      unsafeAtomicAdd(&d.ryx[k], (double)h.wt);
      atomicAdd(&d.det2d[k], 1u);
      h.where = 1;
    } else {
      h.where = 2;
    }
  }
  return h;
}

// LAUNCH (kernels.metal:366-486).  Leaves p.alive false if the photon starts
// outside a grid whose faces absorb it.
__device__ __forceinline__ void launch(Photon& p, const XyzArgs& a, const XyzTissue* T,
                                      const float4* LT, const LoopTotals& lt) {
  const SourceState o = source_photon(p.pid, a.key0, a.key1, a.aux);
  if (o.spec != 0.0f) lt.addf(TF_SPECULAR, (double)o.spec);
  const float x = o.x, y = o.y, z = o.z, ux = o.ux, uy = o.uy, uz = o.uz, W = o.W;
  p.x = x; p.y = y; p.z = z; p.ux = ux; p.uy = uy; p.uz = uz; p.W = W;
  p.ix = (int)floorf(a.hx + x / a.dx);
  p.iy = (int)floorf(a.hy + y / a.dy);
  p.iz = (int)floorf(z / a.dz);
  p.inside = p.ix >= 0 && p.ix < a.Nx && p.iy >= 0 && p.iy < a.Ny && p.iz >= 0 && p.iz < a.Nz;
  p.vdep = 0.0f;
  if (!p.inside && (a.bmode == 1 || (a.bmode == 2 && p.iz < 0))) {
    tally_escape(lt, a, p.iz, W);
    p.alive = false;
    return;
  }
  p.lin = vidx(a, p.ix, p.iy, p.iz);
  p.tis = a.V[p.inside ? p.lin : vidx(a, clampi(p.ix, a.Nx), clampi(p.iy, a.Ny), clampi(p.iz, a.Nz))];
  p.mua = T[p.tis].mua;
  p.mus = T[p.tis].mus;
  p.inv_mus = T[p.tis].inv_mus;
  p.rux = rcp_rn(ux);
  p.ruy = rcp_rn(uy);
  p.ruz = rcp_rn(uz);
  p.ev = 1;
  p.events = 0;
  U4 b = philox(p.pid, 1u, 0u, a.key0, a.key1);
  p.b1 = b.y; p.b2 = b.z; p.b3 = b.w;
  p.sleft = -log_tab(u01nz(b.x), LT);
  p.alive = true;
}

// rec_reduce_kernel's body, shared with the transport kernel's fused
// reduction: F[voxels of bucket b] += the records of workgroup slices
// [w0, w1) of bucket b, summed in FP64 in acc (kBucketVox doubles of LDS),
// then added to F with one FP64 atomic per non-zero voxel.  Every one of the
// NT threads of the workgroup calls it.
template <int NT>
__device__ __forceinline__ void reduce_part(const RecReduceArgs& r, const RecPart part, double* acc) {
  for (uint32_t i = threadIdx.x; i < kBucketVox; i += NT) acc[i] = 0.0;
  __syncthreads();
  const unsigned long long base = r.rec_base[part.bucket];
  const uint32_t cap = r.rec_cap[part.bucket];
  for (int w = part.w0; w < part.w1; ++w) {
    const uint32_t taken = r.rec_count[(int64_t)part.bucket * r.grid + w];
    const uint32_t n = taken < cap ? taken : cap;
    const uint2* src = r.rec + base + (unsigned long long)w * cap;
    // four independent loads in flight per lane, then the remainder
    uint32_t i = threadIdx.x;
    for (; i + 3 * NT < n; i += 4 * NT) {
      const uint2 x0 = src[i], x1 = src[i + NT], x2 = src[i + 2 * NT], x3 = src[i + 3 * NT];
      unsafeAtomicAdd(&acc[x0.x], (double)__uint_as_float(x0.y));
      unsafeAtomicAdd(&acc[x1.x], (double)__uint_as_float(x1.y));
      unsafeAtomicAdd(&acc[x2.x], (double)__uint_as_float(x2.y));
      unsafeAtomicAdd(&acc[x3.x], (double)__uint_as_float(x3.y));
    }
    for (; i < n; i += NT) {
      const uint2 x = src[i];
      unsafeAtomicAdd(&acc[x.x], (double)__uint_as_float(x.y));
    }
  }
  __syncthreads();
  const uint32_t v0 = (uint32_t)part.bucket << kBucketShift;
  const uint32_t nv = r.nvox - v0 < kBucketVox ? r.nvox - v0 : kBucketVox;
  for (uint32_t i = threadIdx.x; i < nv; i += NT)
    if (acc[i] != 0.0) unsafeAtomicAdd(&r.F[v0 + i], acc[i]);
}

template <bool TWO_TERM, int BLOCK, typename TileT, int FMODE>
__global__ __launch_bounds__(BLOCK) void mcxyz_kernel(XyzArgs a) {
  // dynamic LDS: the record slices (kFluRecords), then the tile; kFluSens:
  // the per-lane sensitivity rows instead (no tile)
  extern __shared__ double tile_mem[];
  constexpr bool RECS = FMODE == kFluRecords;
  constexpr bool SENS = FMODE == kFluSens;
  const int n_sens = SENS ? a.sens_n : 0;
  const SensRows sr{(float*)tile_mem + threadIdx.x, (uint32_t*)tile_mem + n_sens * BLOCK + threadIdx.x,
                    (uint32_t)BLOCK};
  const int nb = RECS ? a.n_buckets : 0;
  RecSlice* slices = (RecSlice*)tile_mem;
  TileT* tile = (TileT*)(tile_mem + 2 * nb);  // [tile_nz][tile_nx][tile_ny]
  __shared__ XyzTissue T[NMC_NTISS];
  __shared__ nmc_totals st;
  __shared__ float4 LT[128];  // log_tab's rows
  __shared__ float2 SC[64];   // sincos2pi_bits' rows
  const int tile_n = a.tile_nx * a.tile_ny * a.tile_nz;
  for (int i = threadIdx.x; i < NMC_NTISS; i += BLOCK) T[i] = a.tissue[i];
  if (threadIdx.x < 128) {
    const float* row = kLogTable[threadIdx.x];
    LT[threadIdx.x] = make_float4(row[0], row[1], row[2], row[3]);
  }
  if (threadIdx.x < 64) SC[threadIdx.x] = make_float2(kSinCosTable[threadIdx.x][0], kSinCosTable[threadIdx.x][1]);
  for (int i = threadIdx.x; i < tile_n; i += BLOCK) tile[i] = (TileT)0;
  for (int b = threadIdx.x; b < nb; b += BLOCK) {
    const uint32_t cap = a.rec_cap[b];
    slices[b] = RecSlice{a.rec_base[b] + (unsigned long long)blockIdx.x * cap, cap, 0u};
  }
  if (threadIdx.x < TF_COUNT) ((double*)&st)[threadIdx.x] = 0.0;
  if (threadIdx.x < TI_COUNT) (&st.photons)[threadIdx.x] = 0ull;
  __syncthreads();
  const uint64_t first = a.state->first;
  const TileBox tb{in_vgpr((uint32_t)a.tile_x0), in_vgpr((uint32_t)a.tile_y0),
                   in_vgpr((uint32_t)a.tile_z0), in_vgpr((uint32_t)a.tile_nx),
                   in_vgpr((uint32_t)a.tile_ny), in_vgpr((uint32_t)a.tile_nz),
                   FMODE == kFluNone ? nullptr : gptr_in_vgpr(a.F)};
  const gptr<const uint8_t> vmap = gptr_in_vgpr(a.V);
  const RecOut ro{RECS ? slices : nullptr,
                  RECS ? gptr_in_vgpr((unsigned long long*)a.rec) : nullptr};
  // grid size, boundary and detector modes, likewise in vector registers, and
  // used in forms that leave no loop-invariant lane mask for the compiler to
  // keep in scalar registers: the detector plane's side as a sign (-1: the
  // reflectance plane, crossed going up; +1: the transmittance plane, crossed
  // going down; multiplying by it is exact), and the escape rule of the
  // boundary flag as a per-lane select (escapes through the top face, and
  // through any face)
  const int gnx = (int)in_vgpr((uint32_t)a.Nx), gny = (int)in_vgpr((uint32_t)a.Ny),
            gnz = (int)in_vgpr((uint32_t)a.Nz);
  // the linear index steps of a y-, x- and z-crossing: 1, Ny, Nx*Ny
  const uint32_t ny_v = in_vgpr((uint32_t)a.Ny), nxy_v = in_vgpr((uint32_t)a.Nx * (uint32_t)a.Ny);
  const uint32_t esc_top = in_vgpr(a.bmode == 1 || a.bmode == 2 ? 1u : 0u);
  const uint32_t esc_any = in_vgpr(a.bmode == 1 ? 1u : 0u);
  const float dsign = __uint_as_float(
      in_vgpr(__float_as_uint(a.det_state == NMC_DET_REFLECTANCE ? -1.0f : 1.0f)));
  const float zsurf = __uint_as_float(in_vgpr(__float_as_uint(a.zsurf)));

  const int lane = lane_id();
  Photon p;
  p.alive = false;
  p.x = p.y = p.z = 0.0f;
  p.ux = p.uy = 0.0f;
  p.uz = 1.0f;
  p.W = 0.0f;
  p.vdep = 0.0f;
  p.ix = p.iy = p.iz = 0;
  p.tis = 0;
  p.slot = -1;
  p.lin = 0;
  p.inside = false;
  bool finished = false;
  uint64_t pool_next = 0, pool_end = 0;  // wave-uniform
  double absorbed = 0.0, absorbed_out = 0.0;
  uint32_t c_same = 0, c_cross = 0, c_int = 0, c_phot = 0;
  const LoopTotals lt{&st, in_vgpr(0u)};

  for (;;) {
    // ---- refill idle lanes, in batches of at least refill_min ----
    uint64_t idle = ballot(!p.alive && !finished);
    if (idle && ((uint32_t)__popcll(idle) >= a.refill_min || ballot(p.alive) == 0ull)) {
      while (idle) {
        if (pool_next >= pool_end) {
          unsigned long long base = 0;
          if (lane == 0) base = atomicAdd(&a.state->counter, (unsigned long long)a.chunk);
          base = __shfl(base, 0);
          if (base >= a.n) {
            finished = finished || !p.alive;
            break;
          }
          pool_next = base;
          pool_end = base + a.chunk < a.n ? base + a.chunk : a.n;
        }
        uint64_t avail = pool_end - pool_next;
        uint64_t cnt = (uint64_t)__popcll(idle);
        uint32_t take = (uint32_t)(cnt < avail ? cnt : avail);
        uint32_t rank = (uint32_t)__popcll(idle & lanemask_lt());
        if (((idle >> lane) & 1ull) && rank < take) {
          p.pid = first + pool_next + rank;
          c_phot++;
          launch(p, a, T, LT, lt);
          if constexpr (SENS) {
            for (int j = 0; j < n_sens; ++j) {
              sr.L[j * BLOCK] = 0.0f;
              sr.K[j * BLOCK] = 0u;
            }
            p.sp = p.sq = p.sza = p.szs = 0.0f;
            p.slot = p.alive ? T[p.tis].slot : -1;
          }
        }
        pool_next += take;
        idle = ballot(!p.alive && !finished);
      }
    }
    if (ballot(p.alive) == 0ull) break;
    if (!p.alive) continue;

    // ---- HOP: one piece of the current step (kernels.metal:505-685) ----
    float s_int = p.sleft * p.inv_mus;
    // distance to the face ahead on each axis (index + 1 going forward, index
    // going backward; infinite for a zero direction component), computed on
    // every lane and selected, rather than in branches both of which a wave
    // with lanes going either way would issue
    float tx = (((float)(p.ix + (p.ux > 0.0f ? 1 : 0)) - a.hx) * a.dx - p.x) * p.rux;
    float ty = (((float)(p.iy + (p.uy > 0.0f ? 1 : 0)) - a.hy) * a.dy - p.y) * p.ruy;
    float tz = ((float)(p.iz + (p.uz > 0.0f ? 1 : 0)) * a.dz - p.z) * p.ruz;
    tx = p.ux != 0.0f ? fmaxf(tx, 0.0f) : kInf;
    ty = p.uy != 0.0f ? fmaxf(ty, 0.0f) : kInf;
    tz = p.uz != 0.0f ? fmaxf(tz, 0.0f) : kInf;
    int axis;
    float s_face;
    if (tx <= ty && tx <= tz) {
      axis = 0;
      s_face = tx;
    } else if (ty <= tz) {
      axis = 1;
      s_face = ty;
    } else {
      axis = 2;
      s_face = tz;
    }
    // the detector plane, when the photon moves towards it on its detecting
    // side: (zsurf - z) ruz is then positive (a difference of floats near
    // zsurf is at least 2^-40 or so, and |ruz| >= 1, so it cannot underflow)
    const float t_surf = (zsurf - p.z) * p.ruz;
    const float s_surf = p.uz * dsign > 0.0f && t_surf > 0.0f ? t_surf : kInf;
    int kind;
    float s;
    if (s_int <= s_face && s_int <= s_surf) {
      kind = 0;
      s = s_int;
    } else if (s_surf <= s_face) {
      kind = 1;
      s = s_surf;
    } else {
      kind = 2;
      s = s_face;
    }
    p.x = fmaf(s, p.ux, p.x);
    p.y = fmaf(s, p.uy, p.y);
    p.z = fmaf(s, p.uz, p.z);
    // DROP (kernels.metal:523-533)
    const float xa = p.mua * s;
    float em = expm1_neg(-xa);
    float dep = -(p.W * em);
    if constexpr (SENS) {
      // the segment's sensitivity terms, before L of its tissue grows by s
      if (p.slot >= 0) {
        float* Lr = &sr.L[p.slot * BLOCK];
        const float Lv = *Lr;
        p.sp = p.sp + dep * Lv;
        p.sq = p.sq + dep * (float)sr.K[p.slot * BLOCK];
        const float ws = p.W * s;
        p.sza = p.sza + ws * (1.0f + em);
        p.szs = p.szs + ws * score_growth(xa, em);
        *Lr = Lv + s;
      }
    }
    p.W = p.W - dep;
    // inside the grid the deposit joins the visit's sum, which the voxel's
    // flush also adds to the lane's absorbed total
    if (p.inside) {
      p.vdep = p.vdep + dep;
    } else {
      absorbed_out += (double)dep;
    }
    p.events++;
    bool alive = true;
    // a crossing's next tissue id, taken at the end of the iteration; the
    // flag apart from the loaded value, so that testing it waits for no load
    bool crossed = false;
    int pend = 0;
    if (kind == 0) {
      p.sleft = 0.0f;
      c_same++;
    } else {
      c_cross++;
      p.sleft = p.sleft - s * p.mus;
      if (!(p.sleft > 0.0f)) p.sleft = 0.0f;
      // materialised here, before the crossing: the old mus is then dead
      // once the remaining step is formed, and the new tissue's row (read
      // at the end of the iteration) can go straight into its registers
      p.sleft = __uint_as_float(fresh_vgpr(__float_as_uint(p.sleft)));
      if (kind == 2) {
        // voxel face (kernels.metal:641-679).  The next voxel and its map
        // entry come first: the load is then in flight while the visit just
        // ended is flushed (whose LDS cursor atomic and record store would
        // otherwise sit between the crossing and the load).
        // The indices move on in place (selects, not a branch per axis,
        // which a wave with lanes crossing along different axes would issue
        // in turn); the visit that ends keeps its voxel in `at`.
        VisitAt at = visit_at(p, tb);
        // materialised here, so that the old indices are dead before the new
        // ones are formed and those can take the old ones' registers
        at.lin = fresh_vgpr(at.lin);
        at.lx = fresh_vgpr(at.lx);
        at.ly = fresh_vgpr(at.ly);
        at.lz = fresh_vgpr(at.lz);
        const bool fx = p.ux > 0.0f, fy = p.uy > 0.0f, fz = p.uz > 0.0f;
        p.ix += axis == 0 ? (fx ? 1 : -1) : 0;
        p.iy += axis == 1 ? (fy ? 1 : -1) : 0;
        p.iz += axis == 2 ? (fz ? 1 : -1) : 0;
        p.lin += axis == 0 ? (fx ? ny_v : 0u - ny_v)
                 : axis == 1 ? (fy ? 1u : 0xFFFFFFFFu)
                             : (fz ? nxy_v : 0u - nxy_v);
        const bool nin = (uint32_t)p.ix < (uint32_t)gnx && (uint32_t)p.iy < (uint32_t)gny &&
                         (uint32_t)p.iz < (uint32_t)gnz;
        const bool esc = !nin && (p.iz < 0 ? esc_top : esc_any) != 0u;
        int ntis = 0;
        if (!esc) {
          // outside the grid (open faces) a photon reads the nearest boundary
          // voxel's entry; a branch, since most grids let no photon out alive
          uint32_t vi = p.lin;
          if (!nin) vi = vidx(a, clampi(p.ix, gnx), clampi(p.iy, gny), clampi(p.iz, gnz));
          ntis = vmap[vi];
        }
        flush_visit<FMODE>(p, at, a, tb, tile, ro, absorbed, sr);
        p.inside = nin;
        // the new voxel's tissue (0 for a photon that escapes: it ends here
        // and reads none of its row) is taken at the end of the iteration,
        // after the other lanes' interactions: the map load then has their
        // whole body to arrive in
        pend = ntis;
        crossed = true;
        if (esc) {
          tally_escape(lt, a, p.iz, p.W);
          alive = false;
        }
      }
    }
    if (kind == 1) {
      // detector plane (kernels.metal:556-630), after the crossing's branch
      // rather than beside it: the crossing is then the one path that moves
      // the voxel indices, which lets them move on in their registers
      p.z = zsurf;
      lt.addi(TI_SURFACE_HITS);
      const DetectorHit d = detector_plane(p.x, p.y, p.uz, p.W, p.ix, p.iy, a.aux);
      if (d.where == 1) {
        lt.addf(TF_RD, (double)d.wt);
        lt.addi(TI_DETECTIONS);
      } else if (d.where == 2) {
        lt.addf(TF_RD_OUTSIDE, (double)d.wt);
      }
      p.uz = -p.uz;
      p.ruz = -p.ruz;
      p.W = p.W * d.rsp;
      if (!(p.W > 0.0f)) alive = false;
    }
    if (alive && !(p.W > 0.0f)) alive = false;  // all weight absorbed
    if (alive && p.events >= a.max_events) {
      lt.addf(TF_CAPPED, (double)p.W);
      lt.addi(TI_CAPPED);
      alive = false;
    }
    // a crossing that used up the step interacts in the new voxel's tissue
    if (crossed && alive && !(p.sleft > 0.0f)) {
      take_tissue_t<SENS>(p, T, pend);
      crossed = false;
    }
    if (alive && !(p.sleft > 0.0f)) {
      // the next event's random block and step length (kernels.metal:495-502)
      // do not depend on the spin or the roulette; computed first, they give
      // the tissue-table reads of the phase function time to arrive
      const U4 nb = philox(p.pid, p.ev + 1u, 0u, a.key0, a.key1);
      const float nsl = -log_tab(u01nz(nb.x), LT);
      // SPIN (kernels.metal:695-730); tissue ids are below NMC_NTISS, which
      // lets the row offset take a 24-bit multiply
      __builtin_assume((uint32_t)p.tis < (uint32_t)NMC_NTISS);
      float ct = TWO_TERM ? tt_cos(T[p.tis].tt, u01(p.b1)) : hg_cos(T[p.tis].hg, u01(p.b1));
      spin_w(ct, p.b2, SC, p.ux, p.uy, p.uz);
      rcp_rn3(p.ux, p.uy, p.uz, p.rux, p.ruy, p.ruz);
      c_int++;
      if constexpr (SENS) {
        if (p.slot >= 0) sr.K[p.slot * BLOCK]++;
      }
      // ROULETTE (kernels.metal:740-744)
      if (p.W < a.thr) {
        if (u01(p.b3) <= a.chance) {
          float nw = p.W / a.chance;
          lt.addf(TF_ROULETTE_GAIN, (double)nw - (double)p.W);
          lt.addi(TI_ROULETTE_SURVIVALS);
          p.W = nw;
        } else {
          lt.addf(TF_ROULETTE_LOSS, (double)p.W);
          lt.addi(TI_ROULETTE_KILLS);
          alive = false;
        }
      }
      if (alive) {
        p.ev++;
        p.b1 = nb.y;
        p.b2 = nb.z;
        p.b3 = nb.w;
        p.sleft = nsl;
      }
    }
    if (crossed) take_tissue_t<SENS>(p, T, pend);
    if (!alive) {
      flush_visit<FMODE>(p, visit_at(p, tb), a, tb, tile, ro, absorbed, sr);
      p.alive = false;
    }
  }

  // ---- workgroup totals and tile -> global ----
  lds_addf(&st, TF_ABSORBED, absorbed);
  lds_addf(&st, TF_ABSORBED_OUTSIDE, absorbed_out);
  lds_addi(&st, TI_SUB_SAME, c_same);
  lds_addi(&st, TI_SUB_CROSS, c_cross);
  lds_addi(&st, TI_INTERACTIONS, c_int);
  lds_addi(&st, TI_PHOTONS, c_phot);
  __syncthreads();
  if (tile_n) {
    double* part = a.tile_part + (int64_t)blockIdx.x * tile_n;
    for (int i = threadIdx.x; i < tile_n; i += BLOCK) part[i] = (double)tile[i];
  }
  for (int b = threadIdx.x; b < nb; b += BLOCK)
    a.rec_count[(int64_t)b * gridDim.x + blockIdx.x] = slices[b].cnt;
  if (threadIdx.x < TF_COUNT) unsafeAtomicAdd(&((double*)a.tot)[threadIdx.x], ((double*)&st)[threadIdx.x]);
  if (threadIdx.x < TI_COUNT)
    atomicAdd(&(&a.tot->photons)[threadIdx.x], (unsigned long long)(&st.photons)[threadIdx.x]);

  // ---- the previous launch's records, while the other workgroups drain ----
  if constexpr (RECS) {
    if (a.prev_n_parts > 0) {
      __shared__ uint32_t next_part;
      for (;;) {
        // also orders the tile's copy-out above before its LDS is reused
        __syncthreads();
        if (threadIdx.x == 0) next_part = atomicAdd(a.prev_next, 1u);
        __syncthreads();
        const uint32_t p = next_part;
        if (p >= (uint32_t)a.prev_n_parts) break;
        reduce_part<BLOCK>(a.prev, a.prev.parts[p], tile_mem);
      }
    }
  }
}

// F[tile voxel i] += sum_b part[b][i], blocks in ascending order.
__global__ void tile_reduce_kernel(const double* __restrict__ part, int grid, XyzArgs a) {
  int n = a.tile_nx * a.tile_ny * a.tile_nz;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int b = 0; b < grid; ++b) s += part[(int64_t)b * n + i];
  if (s != 0.0) {
    int ly = i % a.tile_ny, t = i / a.tile_ny;
    int lx = t % a.tile_nx, lz = t / a.tile_nx;
    a.F[vidx(a, a.tile_x0 + lx, a.tile_y0 + ly, a.tile_z0 + lz)] += s;
  }
}

// One RecPart per workgroup.
__global__ __launch_bounds__(1024) void rec_reduce_kernel(RecReduceArgs r) {
  extern __shared__ double acc[];   // [kBucketVox]
  reduce_part<1024>(r, r.parts[blockIdx.x], acc);
}

template <bool TT, typename TileT, int FMODE>
hipError_t launch_block(const XyzArgs& a, int block, int grid, size_t lds, hipStream_t s) {
  switch (block) {
    case 256:
      hipLaunchKernelGGL((mcxyz_kernel<TT, 256, TileT, FMODE>), dim3(grid), dim3(256), lds, s, a);
      break;
    case 512:
      hipLaunchKernelGGL((mcxyz_kernel<TT, 512, TileT, FMODE>), dim3(grid), dim3(512), lds, s, a);
      break;
    default:
      hipLaunchKernelGGL((mcxyz_kernel<TT, 1024, TileT, FMODE>), dim3(grid), dim3(1024), lds, s, a);
      break;
  }
  return hipGetLastError();
}

template <bool TT, typename TileT, int FMODE>
const void* kernel_of(int block) {
  return block == 256   ? (const void*)mcxyz_kernel<TT, 256, TileT, FMODE>
         : block == 512 ? (const void*)mcxyz_kernel<TT, 512, TileT, FMODE>
                        : (const void*)mcxyz_kernel<TT, 1024, TileT, FMODE>;
}

// The record variant exists for the FP32 tile only (tuning.mode 2).
template <int FMODE>
const void* kernel_for(bool two_term, bool tile32, int block) {
  if constexpr (FMODE == kFluRecords) {
    return two_term ? kernel_of<true, float, kFluRecords>(block)
                    : kernel_of<false, float, kFluRecords>(block);
  } else {
    if (two_term)
      return tile32 ? kernel_of<true, float, FMODE>(block) : kernel_of<true, double, FMODE>(block);
    return tile32 ? kernel_of<false, float, FMODE>(block) : kernel_of<false, double, FMODE>(block);
  }
}

template <int FMODE>
hipError_t launch_variant(const XyzArgs& a, bool two_term, bool tile32, int block, int grid,
                          size_t lds_bytes, hipStream_t s) {
  if constexpr (FMODE == kFluRecords) {
    return two_term ? launch_block<true, float, kFluRecords>(a, block, grid, lds_bytes, s)
                    : launch_block<false, float, kFluRecords>(a, block, grid, lds_bytes, s);
  } else {
    if (two_term)
      return tile32 ? launch_block<true, float, FMODE>(a, block, grid, lds_bytes, s)
                    : launch_block<true, double, FMODE>(a, block, grid, lds_bytes, s);
    return tile32 ? launch_block<false, float, FMODE>(a, block, grid, lds_bytes, s)
                  : launch_block<false, double, FMODE>(a, block, grid, lds_bytes, s);
  }
}

const void* kernel_for_mode(int fmode, bool two_term, bool tile32, int block) {
  return fmode == kFluNone      ? kernel_for<kFluNone>(two_term, tile32, block)
         : fmode == kFluCount   ? kernel_for<kFluCount>(two_term, tile32, block)
         : fmode == kFluRecords ? kernel_for<kFluRecords>(two_term, tile32, block)
                                : kernel_for<kFluDeposit>(two_term, tile32, block);
}

// The variants a launch configuration may run: with records (tile32 and the
// record stream) also the record variant.
int last_mode(bool records) { return records ? kFluRecords : kFluCount; }

}  // namespace

hipError_t launch_mcxyz(const XyzArgs& a, bool two_term, bool tile32, int block, int grid,
                        size_t lds_bytes, hipStream_t s) {
  if (!a.F) return launch_variant<kFluNone>(a, two_term, tile32, block, grid, lds_bytes, s);
  if (a.count_visits)
    return launch_variant<kFluCount>(a, two_term, tile32, block, grid, lds_bytes, s);
  if (a.n_buckets > 0 && tile32)
    return launch_variant<kFluRecords>(a, two_term, tile32, block, grid, lds_bytes, s);
  return launch_variant<kFluDeposit>(a, two_term, tile32, block, grid, lds_bytes, s);
}

hipError_t mcxyz_configure(bool two_term, bool tile32, int block, size_t lds_bytes) {
  for (int m = kFluNone; m <= last_mode(tile32); ++m) {
    hipError_t e = hipFuncSetAttribute(kernel_for_mode(m, two_term, tile32, block),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Workgroups per CU that every fluence variant reaches.
int mcxyz_blocks_per_cu(bool two_term, bool tile32, int block, size_t lds_bytes) {
  int best = 0;
  for (int m = kFluNone; m <= last_mode(tile32); ++m) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel_for_mode(m, two_term, tile32, block),
                                                     block, lds_bytes) != hipSuccess)
      return 0;
    best = m == kFluNone || nb < best ? nb : best;
  }
  return best;
}

hipError_t launch_mcxyz_sens(const XyzArgs& a, bool two_term, int grid, size_t lds, hipStream_t s) {
  if (two_term)
    hipLaunchKernelGGL((mcxyz_kernel<true, kSensBlock, double, kFluSens>), dim3(grid),
                       dim3(kSensBlock), lds, s, a);
  else
    hipLaunchKernelGGL((mcxyz_kernel<false, kSensBlock, double, kFluSens>), dim3(grid),
                       dim3(kSensBlock), lds, s, a);
  return hipGetLastError();
}

hipError_t mcxyz_sens_configure(bool two_term, size_t lds) {
  return hipFuncSetAttribute(two_term ? (const void*)mcxyz_kernel<true, kSensBlock, double, kFluSens>
                                      : (const void*)mcxyz_kernel<false, kSensBlock, double, kFluSens>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

int mcxyz_sens_blocks_per_cu(bool two_term, size_t lds) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &nb,
          two_term ? (const void*)mcxyz_kernel<true, kSensBlock, double, kFluSens>
                   : (const void*)mcxyz_kernel<false, kSensBlock, double, kFluSens>,
          kSensBlock, lds) != hipSuccess)
    return 0;
  return nb;
}

hipError_t rec_reduce_configure() {
  return hipFuncSetAttribute((const void*)rec_reduce_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(kBucketVox * sizeof(double)));
}

hipError_t launch_rec_reduce(const RecReduceArgs& r, int n_parts, hipStream_t s) {
  if (n_parts <= 0) return hipSuccess;
  hipLaunchKernelGGL(rec_reduce_kernel, dim3(n_parts), dim3(1024), kBucketVox * sizeof(double), s, r);
  return hipGetLastError();
}

hipError_t launch_tile_reduce(const double* part, int grid, const XyzArgs& a, hipStream_t s) {
  int n = a.tile_nx * a.tile_ny * a.tile_nz;
  if (n <= 0 || !a.F) return hipSuccess;
  hipLaunchKernelGGL(tile_reduce_kernel, dim3((n + 255) / 256), dim3(256), 0, s, part, grid, a);
  return hipGetLastError();
}

}  // namespace nmc
