// layered.hip - layered media with per-layer refractive index, batched over
// stacks, with perturbation-MC path records and the backward kernel.
//
// The walk is the mcxyz loop (kernels.metal:493-750: continuous absorption
// W *= exp(-mua s), HG or two-term spin, roulette) with the voxel faces
// replaced by layer interfaces, where the photon is reflected or refracted
// all-or-none with Fresnel probability (the scheme of the reference's mcsub
// surface, kernels.metal:190-214).  Each photon keeps its path length L_l
// and scattering count K_l per layer; when it leaves through the top its
// (stack, Rd bin, weight, L, K) record is appended to an HBM buffer, from
// which the backward kernel forms dLoss/dmua_l = -sum g w L_l and
// dLoss/dmus_l = sum g w (K_l/mus_l - L_l).
//
// Absorbed-weight sensitivities (kernel variant SENS): the expected weight
// absorbed in layer l is the integral over the path of mua_l W, and the
// density of the path up to a point has the score sum_t (K_t/mus_t - L_t)
// with respect to mus_t (scatterings K_t and length L_t before that point;
// absorption is carried by the weight).  So for a segment of length s in
// layer l, entered with weight W and depositing D = W (1 - e^{-x}), x = mua_l s,
// with L_t, K_t the photon's totals before the segment:
//   dD/dmua_t = -L_t D + [t = l] W s e^{-x}
//   dD/dmus_t = (K_t/mus_t - L_t) D - [t = l] W s h(x),
//   h(x) = (1 - e^{-x} (1 + x)) / x  (the score's growth along the segment).
// Within one visit to layer l the totals of the other layers stay fixed, so
// the photon keeps only the visit's sums of D L_l, D K_l, W s e^{-x} and
// W s h(x), and adds 2 nl terms per visit to its stack's LDS sums.
//
// Launch model as mcxyz.hip: persistent wave64 workgroups, one photon per
// lane, batched lane refills from a device queue spanning every stack's
// budget (queue index -> stack by binary search over budget prefix sums).
// The Rd(r) placement (LDS histogram or global replica) and the path records
// are kernel variants, and the record width and the Rd / record pointers sit
// in vector registers (in_vgpr), so the loop keeps no masks or parked values
// in scalar registers.
#include <algorithm>

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

struct LPhoton {
  float x, y, z, ux, uy, uz, ruz, W, sleft;
  float zt, zb, mua, mus, inv_mus;
  float Lcur;
  int item, l, nl, row0;
  uint32_t Kcur;
  uint64_t pid;
  uint32_t ev, events, hits, b1, b2, b3;
  uint32_t n_same, n_cross, n_int;
  double acur;
  // SENS: the layer's L and K at the start of the visit, and the visit's
  // sums of D L_l, D K_l, W s e^{-x} and W s h(x)
  float Lbase, sp, sq, sza, szs;
  uint32_t Kbase;
  bool alive;
  bool rec_pending;  // left through the top; record written at the next wave-wide point
  int rec_ir;
  KeySched ks;  // this stack's key schedule
  float L[NMC_MAX_LAYERS];
  uint32_t K[NMC_MAX_LAYERS];
};

// Element i of an LDS table, its offset formed with a 24-bit multiply (full
// rate; a plain index compiles to a 32-bit multiply at a quarter of it):
// the tables hold at most 2048 stacks x NMC_MAX_LAYERS rows.
template <typename T>
__device__ __forceinline__ T& lds_at(T* base, int i) {
  return *(T*)((char*)base + __umul24((uint32_t)i, (uint32_t)sizeof(T)));
}

template <bool SENS>
__device__ __forceinline__ void load_layer(LPhoton& p, const LayRow* rows) {
  const LayRow& r = lds_at(rows, p.row0 + p.l);
  p.zt = r.zt;
  p.zb = r.zb;
  p.mua = r.mua;
  p.mus = r.mus;
  p.inv_mus = r.inv_mus;
  if (SENS) {
#pragma unroll
    for (int i = 0; i < NMC_MAX_LAYERS; ++i) {
      if (i == p.l) {
        p.Lbase = p.L[i];
        p.Kbase = p.K[i];
      }
    }
  }
}

// The sensitivity terms of one segment (SENS): D = dep deposited over length
// s from weight w_pre, x = mua s, em = expm1(-x), before Lcur and Kcur grow.
__device__ __forceinline__ void sens_segment(LPhoton& p, float dep, float w_pre, float s, float x,
                                             float em) {
  p.sp = p.sp + dep * (p.Lbase + p.Lcur);
  p.sq = p.sq + dep * (float)(p.Kbase + p.Kcur);
  const float ws = w_pre * s;
  p.sza = p.sza + ws * (1.0f + em);
  p.szs = p.szs + ws * score_growth(x, em);
}

// Folds the current layer's path length, scatter count and absorption into
// the per-layer records / LDS absorption tally; SENS: adds the visit's terms
// to the stack's sensitivity sums sens[item][0 mua | 1 mus][l][t].
template <bool SENS>
__device__ __forceinline__ void fold_layer(LPhoton& p, double* absorbed, double* sens,
                                           const LayRow* rows) {
  if (SENS && (p.acur != 0.0 || p.sza != 0.0f)) {
    double* S = sens + (int64_t)p.item * kSensPerStack + p.l * NMC_MAX_LAYERS;
    const double D = p.acur;
#pragma unroll
    for (int t = 0; t < NMC_MAX_LAYERS; ++t) {
      if (t < p.nl && t != p.l && D != 0.0) {
        const double Lt = (double)p.L[t];
        unsafeAtomicAdd(&S[t], -Lt * D);
        unsafeAtomicAdd(&S[NMC_MAX_LAYERS * NMC_MAX_LAYERS + t],
                        ((double)p.K[t] * (double)rows[p.row0 + t].inv_mus - Lt) * D);
      }
    }
    unsafeAtomicAdd(&S[p.l], -(double)p.sp + (double)p.sza);
    unsafeAtomicAdd(&S[NMC_MAX_LAYERS * NMC_MAX_LAYERS + p.l],
                    (double)p.sq * (double)p.inv_mus - (double)p.sp - (double)p.szs);
    p.sp = p.sq = p.sza = p.szs = 0.0f;
  }
#pragma unroll
  for (int i = 0; i < NMC_MAX_LAYERS; ++i) {
    if (i == p.l) {
      p.L[i] = p.L[i] + p.Lcur;
      p.K[i] += p.Kcur;
    }
  }
  if (p.acur != 0.0) unsafeAtomicAdd(&absorbed[p.item * NMC_MAX_LAYERS + p.l], p.acur);
  p.Lcur = 0.0f;
  p.Kcur = 0u;
  p.acur = 0.0;
}

// Writes the records of the wave's lanes that left through the top since the
// last call (all lanes of the wave call it together).  Slots come from a
// wave-private run [next, end) of kRecChunk slots reserved with one global
// atomic; a flush that overruns the run continues in a fresh one.
// rec_nl and rec are the record width and buffer in vector registers
// (in_vgpr); the per-layer slot tests below then cost a compare each instead
// of a scalar-register mask kept across the loop.
__device__ __forceinline__ void flush_records(LPhoton& p, const LayArgs& a, int lane,
                                              uint64_t& next, uint64_t& end,
                                              uint32_t& written, uint32_t rec_nl_v,
                                              gptr<uint32_t> rec) {
  uint64_t m = ballot(p.rec_pending);
  if (!m) return;
  uint64_t c = (uint64_t)__popcll(m);
  uint64_t rem = end - next;
  unsigned long long fresh = 0;
  if (c > rem) {
    if (lane == 0) fresh = atomicAdd(&a.rec_count[0], (unsigned long long)kRecChunk);
    fresh = __shfl(fresh, 0);
  }
  if (p.rec_pending) {
    uint64_t rank = (uint64_t)__popcll(m & lanemask_lt());
    uint64_t k = rank < rem ? next + rank : fresh + (rank - rem);
    if (k < a.cap) {
      const uint32_t nl = fresh_vgpr(rec_nl_v);
      gptr<uint32_t> r = rec + k * (uint64_t)(3u + 2u * nl);
      r[0] = (uint32_t)p.item;
      r[1] = (uint32_t)p.rec_ir;
      r[2] = __float_as_uint(p.W);
#pragma unroll
      for (uint32_t i = 0; i < NMC_MAX_LAYERS; ++i) {
        if (i < nl) {
          r[3 + i] = __float_as_uint(p.L[i]);
          r[3 + nl + i] = p.K[i];
        }
      }
      written++;
    }
    p.rec_pending = false;
  }
  if (c > rem) {
    next = fresh + (c - rem);
    end = fresh + kRecChunk;
  } else {
    next += c;
  }
}

// RD_LDS: Rd(r) in an LDS histogram (else the block's global replica);
// RECORDS: photons leaving through the top write path records; SENS: the
// absorbed-weight sensitivities.  Each is a kernel variant, so the loop holds
// no mask for them.
template <bool RD_LDS, bool RECORDS, bool SENS>
__global__ __launch_bounds__(kLayBlock) void layered_kernel(LayArgs a) {
  extern __shared__ double smem[];
  __shared__ float4 LT[128];  // log_tab's rows
  __shared__ float2 SC[64];   // sincos2pi_bits' rows
  // [items] absorbed[MAXL] | [items] totals | [items][nr+1] Rd (RD_LDS only) |
  // [items] headers | [items+1] prefix | [rows] | [items] sensitivities (SENS)
  const int ni = a.n_items;
  const int rd_len = RD_LDS ? ni * (a.nr + 1) : 0;
  double* absorbed = smem;
  nmc_totals* st = (nmc_totals*)(absorbed + ni * NMC_MAX_LAYERS);
  double* rd_tile = (double*)(st + ni);
  LayStackDev* hdr = (LayStackDev*)(rd_tile + rd_len);
  uint64_t* prefix = (uint64_t*)(hdr + ni);
  LayRow* rows = (LayRow*)(prefix + ni + 1);
  // after the rows, rounded up to 8 bytes (layered_lds_bytes counts the pad)
  double* sens = SENS ? (double*)(((size_t)(rows + a.n_rows) + 7u) & ~(size_t)7u) : nullptr;
  gptr<double> rd_rep =
      RD_LDS ? nullptr
             : gptr_in_vgpr(a.rd + (int64_t)(blockIdx.x % (unsigned)a.rd_reps) * ni * (a.nr + 1));
  const uint32_t rd_row = in_vgpr((uint32_t)(a.nr + 1));
  const uint32_t rec_nl_v = in_vgpr((uint32_t)a.rec_nl);
  const gptr<uint32_t> rec_v = RECORDS ? gptr_in_vgpr(a.rec) : nullptr;
  for (int i = threadIdx.x; i < ni * (NMC_MAX_LAYERS + TF_COUNT + TI_COUNT) + rd_len; i += kLayBlock)
    smem[i] = 0.0;
  for (int i = threadIdx.x; i < ni; i += kLayBlock) hdr[i] = a.stacks[i];
  for (int i = threadIdx.x; i <= ni; i += kLayBlock) prefix[i] = a.prefix[i];
  for (int i = threadIdx.x; i < a.n_rows; i += kLayBlock) rows[i] = a.rows[i];
  if (threadIdx.x < 128) {
    const float* row = kLogTable[threadIdx.x];
    LT[threadIdx.x] = make_float4(row[0], row[1], row[2], row[3]);
  }
  if (threadIdx.x < 64) SC[threadIdx.x] = make_float2(kSinCosTable[threadIdx.x][0], kSinCosTable[threadIdx.x][1]);
  if (SENS)
    for (int i = threadIdx.x; i < ni * kSensPerStack; i += kLayBlock) sens[i] = 0.0;
  __syncthreads();
  const uint64_t first = a.state->first;
  const uint64_t total = prefix[ni];

  const int lane = lane_id();
  LPhoton p;
  p.alive = false;
  p.ks = key_sched(0u);
  p.item = 0;
  p.l = 0;
  p.acur = 0.0;
  p.Lbase = p.sp = p.sq = p.sza = p.szs = 0.0f;
  p.Kbase = 0u;
  p.rec_pending = false;
  p.rec_ir = 0;
  bool finished = false;
  uint64_t pool_next = 0, pool_end = 0;
  uint64_t rec_next = 0, rec_end = 0;  // wave-uniform record slot run
  uint32_t rec_written = 0;

  for (;;) {
    if (RECORDS) flush_records(p, a, lane, rec_next, rec_end, rec_written, rec_nl_v, rec_v);
    uint64_t idle = ballot(!p.alive && !finished);
    if (idle && ((uint32_t)__popcll(idle) >= a.refill_min || ballot(p.alive) == 0ull)) {
      while (idle) {
        if (pool_next >= pool_end) {
          unsigned long long base = 0;
          if (lane == 0) base = atomicAdd(&a.state->counter, (unsigned long long)a.chunk);
          base = __shfl(base, 0);
          if (base >= total) {
            finished = finished || !p.alive;
            break;
          }
          pool_next = base;
          pool_end = base + a.chunk < total ? base + a.chunk : total;
        }
        uint64_t avail = pool_end - pool_next;
        uint64_t cnt = (uint64_t)__popcll(idle);
        uint32_t take = (uint32_t)(cnt < avail ? cnt : avail);
        uint32_t rank = (uint32_t)__popcll(idle & lanemask_lt());
        if (((idle >> lane) & 1ull) && rank < take) {
          uint64_t q = pool_next + rank;
          int lo = 0, hi = ni - 1;
          while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= q) lo = mid;
            else hi = mid - 1;
          }
          const LayStackDev& H = hdr[lo];
          p.item = lo;
          p.nl = H.nl;
          p.row0 = H.row0;
          p.pid = first + (q - prefix[lo]);
          // LAUNCH: pencil beam at the origin along +z, specular removed
          addi(&st[lo], TI_PHOTONS, 1);
          addf(&st[lo], TF_SPECULAR, (double)H.rsp);
          p.W = 1.0f - H.rsp;
          p.x = p.y = p.z = 0.0f;
          p.ux = p.uy = 0.0f;
          p.uz = 1.0f;
          p.ruz = rcp_rn(p.uz);
          p.l = 0;
#pragma unroll
          for (int i = 0; i < NMC_MAX_LAYERS; ++i) {
            p.L[i] = 0.0f;
            p.K[i] = 0u;
          }
          load_layer<SENS>(p, rows);
          p.hits = 0;
          p.n_same = p.n_cross = p.n_int = 0;
          p.ev = 1;
          p.events = 0;
          p.sp = p.sq = p.sza = p.szs = 0.0f;
          p.Lcur = 0.0f;
          p.Kcur = 0u;
          p.acur = 0.0;
          p.ks = key_sched(H.key1);
          U4 b = philox(p.pid, 1u, 0u, a.key0, p.ks);
          p.b1 = b.y;
          p.b2 = b.z;
          p.b3 = b.w;
          p.sleft = -log_tab(u01nz(b.x), LT);
          p.alive = true;
        }
        pool_next += take;
        idle = ballot(!p.alive && !finished);
      }
    }
    if (ballot(p.alive) == 0ull) break;
    if (!p.alive) continue;

    nmc_totals* T = &lds_at(st, p.item);
    // ---- HOP to the next interaction or interface ----
    float s_int = p.sleft * p.inv_mus;
    // the interface ahead (bottom going down, top going up), as selects
    const float s_b = p.uz != 0.0f ? fmaxf(((p.uz > 0.0f ? p.zb : p.zt) - p.z) * p.ruz, 0.0f)
                                   : kInf;
    bool inter = s_int <= s_b;
    float s = inter ? s_int : s_b;
    p.x = fmaf(s, p.ux, p.x);
    p.y = fmaf(s, p.uy, p.y);
    p.z = fmaf(s, p.uz, p.z);
    const float x = p.mua * s;
    // -x <= 0: the one-path expm1, equal to expm1_f there bit for bit
    float em = expm1_neg(-x);
    float dep = -(p.W * em);
    if (SENS) sens_segment(p, dep, p.W, s, x, em);
    p.W = p.W - dep;
    p.acur += (double)dep;
    p.Lcur = p.Lcur + s;
    p.events++;
    bool alive = true;
    if (inter) {
      p.sleft = 0.0f;
      p.n_same++;
    } else {
      p.n_cross++;
      p.sleft = p.sleft - s * p.mus;
      if (!(p.sleft > 0.0f)) p.sleft = 0.0f;
      bool up = p.uz < 0.0f;
      p.z = up ? p.zt : p.zb;
      const LayStackDev& H = lds_at(hdr, p.item);
      float ni_ = lds_at(rows, p.row0 + p.l).n;
      float nt = up ? (p.l == 0 ? H.n_above : lds_at(rows, p.row0 + p.l - 1).n)
                    : (p.l == p.nl - 1 ? H.n_below : lds_at(rows, p.row0 + p.l + 1).n);
      float ca2;
      float r = fresnel(ni_, nt, fabsf(p.uz), &ca2);
      U4 aux = philox(p.pid, p.hits, 2u, a.key0, p.ks);
      p.hits++;
      if (u01(aux.x) < r) {
        p.uz = -p.uz;
        p.ruz = -p.ruz;
      } else if (up && p.l == 0) {
        // leaves through the top: Rd(r) and the path record
        fold_layer<SENS>(p, absorbed, sens, rows);
        float re = sqrtf(p.x * p.x + p.y * p.y);
        int ir = rbin(re / a.dr, a.nr);
        // the bins near r = 0 take a large share of all exits: an LDS
        // histogram (or a per-XCD replica) keeps them from serialising the
        // whole chip on a few global addresses
        if (RD_LDS)
          unsafeAtomicAdd(&rd_tile[(uint32_t)p.item * rd_row + (uint32_t)ir], (double)p.W);
        else
          global_add(&rd_rep[(uint32_t)p.item * rd_row + (uint32_t)ir], (double)p.W);
        addf(T, TF_RD, (double)p.W);
        addi(T, TI_DETECTIONS, 1);
        if (RECORDS) {
          p.rec_pending = true;
          p.rec_ir = ir;
        }
        alive = false;
      } else if (!up && p.l == p.nl - 1) {
        fold_layer<SENS>(p, absorbed, sens, rows);
        addf(T, TF_TT, (double)p.W);
        addi(T, TI_ESC_BOTTOM, 1);
        alive = false;
      } else {
        fold_layer<SENS>(p, absorbed, sens, rows);
        if (ni_ != nt) {
          p.ux = p.ux * ni_ / nt;
          p.uy = p.uy * ni_ / nt;
          p.uz = up ? -ca2 : ca2;
          p.ruz = rcp_rn(p.uz);
        }
        p.l += up ? -1 : 1;
        load_layer<SENS>(p, rows);
      }
    }
    if (alive && !(p.W > 0.0f)) alive = false;  // all weight absorbed
    if (alive && p.events >= a.max_events) {
      addf(T, TF_CAPPED, (double)p.W);
      addi(T, TI_CAPPED, 1);
      alive = false;
    }
    if (alive && !(p.sleft > 0.0f)) {
      const LayRow& R = lds_at(rows, p.row0 + p.l);
      float ct = R.phase ? tt_cos(R.tt, u01(p.b1)) : hg_cos(R.hg, u01(p.b1));
      spin_w(ct, p.b2, SC, p.ux, p.uy, p.uz);
      p.ruz = rcp_rn(p.uz);
      p.Kcur++;
      p.n_int++;
      if (p.W < a.thr) {
        if (u01(p.b3) <= a.chance) {
          float nw = p.W / a.chance;
          addf(T, TF_ROULETTE_GAIN, (double)nw - (double)p.W);
          addi(T, TI_ROULETTE_SURVIVALS, 1);
          p.W = nw;
        } else {
          addf(T, TF_ROULETTE_LOSS, (double)p.W);
          addi(T, TI_ROULETTE_KILLS, 1);
          alive = false;
        }
      }
      if (alive) {
        p.ev++;
        U4 b = philox(p.pid, p.ev, 0u, a.key0, p.ks);
        p.b1 = b.y;
        p.b2 = b.z;
        p.b3 = b.w;
        p.sleft = -log_tab(u01nz(b.x), LT);
      }
    }
    if (!alive) {
      if (p.acur != 0.0 || (SENS && p.sza != 0.0f)) fold_layer<SENS>(p, absorbed, sens, rows);
      addi(T, TI_SUB_SAME, p.n_same);
      addi(T, TI_SUB_CROSS, p.n_cross);
      addi(T, TI_INTERACTIONS, p.n_int);
      addi(T, TI_SURFACE_HITS, p.hits);
      p.alive = false;
    }
  }

  if (RECORDS) {
    // the unused tail of the wave's slot run is marked as holes; the count of
    // written records goes out with one atomic per wave
    const int rs = 3 + 2 * a.rec_nl;
    for (uint64_t k = rec_next + (uint64_t)lane; k < rec_end; k += 64)
      if (k < a.cap) a.rec[k * (uint64_t)rs] = kRecHole;
    uint32_t w = rec_written;
    for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
    if (lane == 0 && w) atomicAdd(&a.rec_count[1], (unsigned long long)w);
  }

  __syncthreads();
  // absorbed per layer: FP64 LDS -> partials; totals: absorbed sum + the rest
  const int per = ni * (NMC_MAX_LAYERS + TF_COUNT + TI_COUNT);
  for (int j = threadIdx.x; j < ni; j += kLayBlock) {
    double s = 0.0;
    for (int l = 0; l < NMC_MAX_LAYERS; ++l) s += absorbed[j * NMC_MAX_LAYERS + l];
    ((double*)&st[j])[TF_ABSORBED] = s;
  }
  __syncthreads();
  double* part = a.part + (int64_t)blockIdx.x * per;
  for (int i = threadIdx.x; i < per; i += kLayBlock) part[i] = smem[i];
  if (RD_LDS) {
    double* rp = a.rd + (int64_t)blockIdx.x * rd_len;
    for (int i = threadIdx.x; i < rd_len; i += kLayBlock) rp[i] = rd_tile[i];
  }
  if (SENS) {
    double* sp = a.sens_part + (int64_t)blockIdx.x * ni * kSensPerStack;
    for (int i = threadIdx.x; i < ni * kSensPerStack; i += kLayBlock) sp[i] = sens[i];
  }
}

// Gradient accumulation over the records.  Each thread keeps the running sums
// of the stack of its last record in registers and flushes them to LDS when
// the stack changes; blocks then add their LDS sums to the outputs.
__global__ __launch_bounds__(256) void layered_backward_kernel(LayBackArgs b) {
  extern __shared__ double acc[];   // [n_items][2][MAXL]
  const int ni = b.n_items;
  for (int i = threadIdx.x; i < ni * 2 * NMC_MAX_LAYERS; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  uint64_t m = b.n_rec < b.cap ? b.n_rec : b.cap;
  int cur = -1;
  double ga[NMC_MAX_LAYERS], gs[NMC_MAX_LAYERS];
#pragma unroll
  for (int l = 0; l < NMC_MAX_LAYERS; ++l) ga[l] = gs[l] = 0.0;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const int rs = 3 + 2 * b.rec_nl;
    const uint32_t* r = b.rec + k * (uint64_t)rs;
    const uint32_t jr = r[0];
    if (jr >= (uint32_t)ni) continue;   // a hole (kRecHole) or a foreign stack index
    const int j = (int)jr;
    if (j != cur) {
      if (cur >= 0) {
#pragma unroll
        for (int l = 0; l < NMC_MAX_LAYERS; ++l) {
          if (ga[l] != 0.0) unsafeAtomicAdd(&acc[(cur * 2) * NMC_MAX_LAYERS + l], ga[l]);
          if (gs[l] != 0.0) unsafeAtomicAdd(&acc[(cur * 2 + 1) * NMC_MAX_LAYERS + l], gs[l]);
          ga[l] = gs[l] = 0.0;
        }
      }
      cur = j;
    }
    double g = b.grad_rd[(int64_t)j * (b.nr + 1) + r[1]] * (double)__uint_as_float(r[2]) * b.inv_n[j];
    int nl = b.nl[j];
#pragma unroll
    for (int l = 0; l < NMC_MAX_LAYERS; ++l) {
      if (l < nl && l < b.rec_nl) {
        double L = (double)__uint_as_float(r[3 + l]);
        double K = (double)r[3 + b.rec_nl + l];
        ga[l] += -g * L;
        gs[l] += g * (K / (double)b.mus[j * NMC_MAX_LAYERS + l] - L);
      }
    }
  }
  if (cur >= 0) {
#pragma unroll
    for (int l = 0; l < NMC_MAX_LAYERS; ++l) {
      if (ga[l] != 0.0) unsafeAtomicAdd(&acc[(cur * 2) * NMC_MAX_LAYERS + l], ga[l]);
      if (gs[l] != 0.0) unsafeAtomicAdd(&acc[(cur * 2 + 1) * NMC_MAX_LAYERS + l], gs[l]);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ni * 2 * NMC_MAX_LAYERS; i += blockDim.x) {
    double v = acc[i];
    if (v == 0.0) continue;
    int j = i / (2 * NMC_MAX_LAYERS), w = (i / NMC_MAX_LAYERS) & 1, l = i % NMC_MAX_LAYERS;
    unsafeAtomicAdd(&(w ? b.grad_mus : b.grad_mua)[j * NMC_MAX_LAYERS + l], v);
  }
}

}  // namespace

size_t layered_lds_bytes(int n_items, int n_rows, int rd_len, bool sens) {
  return ((size_t)n_items * (NMC_MAX_LAYERS + TF_COUNT + TI_COUNT) + (size_t)rd_len) * sizeof(double) +
         (size_t)n_items * sizeof(LayStackDev) + (size_t)(n_items + 1) * sizeof(uint64_t) +
         (size_t)n_rows * sizeof(LayRow) +
         (sens ? 8 + (size_t)n_items * kSensPerStack * sizeof(double) : 0);
}

namespace {
const void* layered_variant(int v) {
  switch (v) {
    case 0: return (const void*)layered_kernel<false, false, false>;
    case 1: return (const void*)layered_kernel<false, true, false>;
    case 2: return (const void*)layered_kernel<true, false, false>;
    case 3: return (const void*)layered_kernel<true, true, false>;
    case 4: return (const void*)layered_kernel<false, false, true>;
    case 5: return (const void*)layered_kernel<false, true, true>;
    case 6: return (const void*)layered_kernel<true, false, true>;
    default: return (const void*)layered_kernel<true, true, true>;
  }
}
}  // namespace

hipError_t layered_configure(size_t lds_bytes) {
  for (int v = 0; v < 8; ++v) {
    hipError_t e = hipFuncSetAttribute(layered_variant(v), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Workgroups per CU that every variant reaches.
int layered_blocks_per_cu(size_t lds_bytes) {
  int best = 0;
  for (int v = 0; v < 8; ++v) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, layered_variant(v), kLayBlock, lds_bytes) !=
        hipSuccess)
      return 0;
    best = v == 0 || nb < best ? nb : best;
  }
  return best;
}

hipError_t launch_layered(const LayArgs& a, int grid, size_t lds_bytes, hipStream_t s) {
  const int v = (a.sens ? 4 : 0) + (a.rd_lds ? 2 : 0) + (a.records ? 1 : 0);
  LayArgs arg = a;
  void* args[] = {&arg};
  return hipLaunchKernel(layered_variant(v), dim3(grid), dim3(kLayBlock), args, lds_bytes, s);
}

hipError_t launch_layered_backward(const LayBackArgs& b, hipStream_t s) {
  size_t lds = (size_t)b.n_items * 2 * NMC_MAX_LAYERS * sizeof(double);
  uint64_t m = b.n_rec < b.cap ? b.n_rec : b.cap;
  int grid = (int)std::min<uint64_t>(2048, (m + 255) / 256);
  if (grid < 1) grid = 1;
  hipError_t e = hipFuncSetAttribute((const void*)layered_backward_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(layered_backward_kernel, dim3(grid), dim3(256), lds, s, b);
  return hipGetLastError();
}

}  // namespace nmc
