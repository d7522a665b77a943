// nmc_abi.hip - host runtime behind include/nmc.h.
//
// Plays the role of the reference's MetalMC object (NMC/NMCCode/MetalMC.m):
// owns the device buffers, validates and uploads the configuration, launches
// the transport kernels on its own HIP stream and returns the tallies.  A
// run is the sequence [reset queue counter, transport kernel, reduction of
// block partials, advance the first photon id]; all four steps are kernels,
// so the sequence can be captured once as a hipGraph and replayed.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "nmc_internal.h"

using namespace nmc;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int ok() {
  g_err.clear();
  return NMC_OK;
}

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(e_ == hipErrorOutOfMemory ? NMC_ERR_NOMEM : NMC_ERR_DEVICE,         \
                  "%s failed: %s", #expr, hipGetErrorString(e_));                     \
  } while (0)

#define NMC_TRY(expr)              \
  do {                             \
    int s_ = (expr);               \
    if (s_ != NMC_OK) return s_;   \
  } while (0)

bool finite_nonneg(float v) { return std::isfinite(v) && v >= 0.0f; }

void ext_defaults(nmc_ext_config* e) {
  std::memset(e, 0, sizeof(*e));
  e->struct_size = sizeof(*e);
  e->version = NMC_EXT_VERSION;
  e->seed = 1;
  e->threshold = NMC_THRESHOLD;
  e->chance = NMC_CHANCE;
  e->max_events = 100000000u;
  e->nr = 100;
  e->nz = 100;
  e->na = 30;
  e->dr = 0.01f;
  e->dz = 0.01f;
}

// Two-term constants, derived in double from the float parameters
// (q1, q2, q3 of kernels.metal:900-905).
void tt_prepare(float gf, float af, float gb, float ab, float cc, TTCoef* t) {
  std::memset(t, 0, sizeof(*t));
  t->cc = cc;
  t->isof = gf == 0.0f;
  t->isob = gb == 0.0f;
  double g = gf, a = af;
  if (!t->isof) {
    t->c1f = (float)((1.0 + g * g) / (2.0 * g));
    t->c2f = (float)(1.0 / (2.0 * g));
    t->iq2f = (float)(1.0 / std::pow(1.0 - g, 2.0 * a));
    t->iq3f = (float)(1.0 / std::pow(1.0 + g, 2.0 * a));
    t->ef = (float)(-1.0 / a);
  }
  g = gb;
  a = ab;
  if (!t->isob) {
    t->c1b = (float)((1.0 + g * g) / (2.0 * g));
    t->c2b = (float)(1.0 / (2.0 * g));
    t->iq2b = (float)(1.0 / std::pow(1.0 - g, 2.0 * a));
    t->iq3b = (float)(1.0 / std::pow(1.0 + g, 2.0 * a));
    t->eb = (float)(-1.0 / a);
  }
}

bool flt_eq_h(float a, float b) {
  float e = b * 0x1p-23f * 2.0f;
  return (a >= b - e) && (a <= b + e);
}

// Henyey-Greenstein constants of anisotropy g, derived in double.
HGCoef hg_prepare(float g) {
  HGCoef h{};
  h.mode = flt_eq_h(g, 0.0f) ? 1 : (flt_eq_h(g, 1.0f) ? 2 : 0);
  double gd = g;
  h.a = (float)(1.0 + gd * gd);
  h.b = (float)(1.0 - gd * gd);
  h.c = (float)(1.0 - gd);
  h.d = (float)(2.0 * gd);
  h.e = h.mode == 0 ? (float)(1.0 / (2.0 * gd)) : 0.0f;
  return h;
}

// 1/mu rounded once; FLT_MAX for mu = 0 (step lengths then exceed any face).
float inv_pos(float mu) { return mu > 0.0f ? (float)(1.0 / (double)mu) : 3.40282347e+38f; }

// One medium of an mcsub run (mut, albedo, specular as the reference computes
// them, kernels.metal:99-119).
SubMedium make_medium(float mua, float mus, float g, float n1, float n2, uint32_t key1) {
  SubMedium m{};
  m.mut = mua + mus;
  m.inv_mut = inv_pos(m.mut);
  m.one_m_albedo = 1.0f - mus / m.mut;
  m.hg = hg_prepare(g);
  m.n1 = n1;
  m.n2 = n2;
  float tr = n1 / n2;
  tr = (1.0f - tr) / (1.0f + tr);
  m.rsp0 = tr * tr;
  m.key1 = key1;
  return m;
}

constexpr int kMaxBatchItems = 2048;
// mcsub_kernel's static LDS: the exit-angle cosine table (256 B), log_tab's
// 128 rows (2 KB), sincos2pi_bits' 64 rows (512 B), and slack
constexpr size_t kSubStaticLds = 256 + 128 * 16 + 64 * 8 + 256;
// layered_kernel's static LDS: log_tab's 128 rows, sincos2pi_bits' 64 rows,
// and slack
constexpr size_t kLayStaticLds = 128 * 16 + 64 * 8 + 1024;

// A semi-infinite last layer that neither scatters nor absorbs would let a
// photon travel forever.
bool last_layer_semi_infinite_needs_scatter(const nmc_stack& s) {
  const nmc_layer& L = s.layer[s.n_layers - 1];
  bool semi = !(L.thickness > 0.0f && std::isfinite(L.thickness));
  return semi && !(L.mua > 0.0f) && !(L.mus > 0.0f);
}

}  // namespace

struct nmc_handle {
  int sim = 0, device = 0, num_cu = 0;
  size_t lds_max = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipEvent_t k0 = nullptr, k1 = nullptr;  // around the last transport kernel (eager runs)
  bool k_timed = false;
  bool prepared = false, timed = false, batch = false;
  RunParams rp{};
  TissueParams tp{};
  nmc_ext_config ext{};
  nmc_tuning tune_req{};   // as requested by the caller
  nmc_tuning tune{};       // resolved
  uint64_t next_photon = 0;
  // device state
  RunState* dState = nullptr;
  nmc_totals* dTot = nullptr;
  // mcxyz
  uint8_t* dV = nullptr;
  XyzTissue* dTis = nullptr;
  XyzAux* dAux = nullptr;
  double* dF = nullptr;
  double* dTilePart = nullptr;
  XyzRecords recs;             // binned deposit records (tuning.mode 2)
  // per-tissue sensitivities (nmc_mcxyz_sens): [2][sens_n][nvox]
  double* dSens = nullptr;
  int sens_n = 0;
  int32_t sens_ids[NMC_SENS_MAX] = {};
  size_t sens_lds = 0;
  int sens_grid = 0;
  double* dRyx = nullptr;
  uint32_t* dDet = nullptr;
  size_t xyz_lds = 0;
  int xyz_grid = 0;
  // mcsub (plain = batch of one)
  int n_items = 0;
  std::vector<nmc_work_item> items;
  std::vector<int> order;      // queue order of the items (longest photons first)
  std::vector<uint64_t> prefix;
  SubMedium* dItems = nullptr;
  uint64_t* dPrefix = nullptr;
  float* dCos = nullptr;
  double* dOut = nullptr;   // [n_items * per_item tallies | n_items nmc_totals]
  double* dRep = nullptr;   // global-mode tally replicas [sub_reps][n_items * per_item]
  int sub_reps = 1;
  double* dPart = nullptr;
  int64_t n_ra = 0, n_rz = 0;
  bool sub_lds = false;
  size_t sub_lds_bytes = 0;
  int sub_grid = 0;
  // layered media
  nmc_layered_config lcfg{};
  std::vector<nmc_stack> stacks;
  LayStackDev* dHdr = nullptr;
  LayRow* dRows = nullptr;
  uint64_t* dLPrefix = nullptr;
  double* dLRd = nullptr;      // [n][nr+1]
  double* dLOut = nullptr;     // [n][MAXL] absorbed | [n] totals
  double* dLPart = nullptr;
  double* dLRdPart = nullptr;  // Rd partials [lay_rd_count][n][nr+1]
  double* dLSens = nullptr;    // absorbed-weight sensitivities [n][2][MAXL][MAXL]
  double* dLSensPart = nullptr;  // their block partials [grid][n][2][MAXL][MAXL]
  bool sens_valid = false;
  int lay_rd_lds = 0, lay_rd_count = 0;
  unsigned long long* dRecCount = nullptr;
  uint32_t* dRec = nullptr;
  uint64_t rec_cap = 0;
  int rec_nl = 0;
  bool rec_valid = false;      // records belong to the properties now uploaded
  double* dGrad = nullptr;     // grad_rd [n][nr+1] | inv_n [n] | grad_mua, grad_mus [n][MAXL]
  float* dMus = nullptr;
  int* dNl = nullptr;
  int lay_rows = 0, lay_grid = 0;
  size_t lay_lds = 0;
  // graph mode
  bool graph_mode = false;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_n = 0;
  uint64_t graph_seed = 0;     // the Philox key the captured launch arguments hold
  uint32_t graph_stream = 0;
  uint32_t graph_launches = 0; // transport launches the captured graph holds
  uint32_t launches = 0;       // transport launches of the last run
};

namespace {

template <class T>
void free_ptr(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

void drop_graph(nmc_handle* h) {
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  h->graph_exec = nullptr;
  h->graph_n = 0;
}

void free_records(XyzRecords& r) {
  free_ptr(r.base);
  free_ptr(r.cap);
  free_ptr(r.rec);
  free_ptr(r.count);
  free_ptr(r.parts);
  free_ptr(r.rec1);
  free_ptr(r.count1);
  free_ptr(r.next);
  r = XyzRecords{};
}

void free_buffers(nmc_handle* h) {
  drop_graph(h);
  free_records(h->recs);
  free_ptr(h->dV);
  free_ptr(h->dTis);
  free_ptr(h->dAux);
  free_ptr(h->dF);
  free_ptr(h->dTilePart);
  free_ptr(h->dSens);
  h->sens_n = 0;
  free_ptr(h->dRyx);
  free_ptr(h->dDet);
  free_ptr(h->dItems);
  free_ptr(h->dPrefix);
  free_ptr(h->dCos);
  free_ptr(h->dOut);
  free_ptr(h->dRep);
  free_ptr(h->dPart);
  free_ptr(h->dHdr);
  free_ptr(h->dRows);
  free_ptr(h->dLPrefix);
  free_ptr(h->dLRd);
  free_ptr(h->dLOut);
  free_ptr(h->dLPart);
  free_ptr(h->dLRdPart);
  free_ptr(h->dLSens);
  free_ptr(h->dLSensPart);
  free_ptr(h->dRecCount);
  free_ptr(h->dRec);
  free_ptr(h->dGrad);
  free_ptr(h->dMus);
  free_ptr(h->dNl);
  h->prepared = false;
}

int64_t nvox(const RunParams* rp) { return (int64_t)rp->Nx * rp->Ny * rp->Nz; }

int64_t per_item(const nmc_handle* h) { return h->n_ra + h->n_rz; }

// ---- mcxyz launch configuration -------------------------------------------
// Photons per workgroup in one launch with the FP32 LDS tile (tuning.split
// default): 1e8 photons on the default grid of one 1024-lane workgroup per CU
// of a 256-CU device, the launch size the FP32 tile's accuracy was measured at.
constexpr int32_t kFp32TileSplit = 390625;

}  // namespace

namespace nmc {

// Resolves a tuning request for an mcxyz run (no side effects on any handle).
// part_cap >= 0 bounds the tile's block partials (grid x tile voxels doubles):
// the automatic box shrinks to fit it and an explicit box beyond it fails.
int resolve_xyz_launch(const RunParams& rp, int sim, bool has_fluence, const nmc_tuning& r,
                       size_t lds_max, int num_cu, int64_t part_cap, XyzLaunch* out) {
  nmc_tuning t{};
  t.block = r.block ? r.block : 1024;
  if (t.block != 256 && t.block != 512 && t.block != 1024)
    return fail(NMC_ERR_ARG, "tuning.block %d not 256, 512 or 1024", t.block);
  t.refill_min = r.refill_min ? r.refill_min : 16;
  if (t.refill_min < 1 || t.refill_min > 64) return fail(NMC_ERR_ARG, "tuning.refill_min not in 1..64");
  t.chunk = r.chunk;
  if (r.split < 0) return fail(NMC_ERR_ARG, "tuning.split must be >= 0");
  bool tt = sim == NMC_SIM_MCXYZ_TT;
  if (r.mode < 0 || r.mode > 2)
    return fail(NMC_ERR_ARG, "tuning.mode (tile element) not 0, 1 or 2");
  if (r.rec_cap_pct < 0) return fail(NMC_ERR_ARG, "tuning.rec_cap_pct must be >= 0");
  t.mode = r.mode;
  t.rec_cap_pct = r.rec_cap_pct;
  const bool tile32 = t.mode >= 1;
  t.split = tile32 ? (r.split ? r.split : kFp32TileSplit) : 0;
  int n_buckets = 0;
  if (t.mode == 2 && has_fluence) {
    if (part_cap >= 0)
      return fail(NMC_ERR_ARG, "tuning.mode 2 (binned deposit records) needs its record buffer; "
                  "not available without allocation");
    const int64_t nv = (int64_t)rp.Nx * rp.Ny * rp.Nz;
    const int64_t nb = (nv + kBucketVox - 1) / kBucketVox;
    if (nb > kMaxBuckets)
      return fail(NMC_ERR_ARG, "tuning.mode 2: %lld voxels need %lld record buckets (at most %d)",
                  (long long)nv, (long long)nb, kMaxBuckets);
    n_buckets = (int)nb;
  }
  const size_t elem = tile32 ? sizeof(float) : sizeof(double);
  // the kernel's static LDS: tissue rows, totals, log_tab's 128 rows,
  // sincos2pi_bits' 64 rows, slack
  const size_t static_lds =
      sizeof(XyzTissue) * NMC_NTISS + sizeof(nmc_totals) + 128 * 16 + 64 * 8 + 256;
  const size_t rec_lds = (size_t)n_buckets * 16;
  const int64_t lds_budget =
      lds_max > static_lds + rec_lds ? (int64_t)((lds_max - static_lds - rec_lds) / elem) : 0;
  // the kernels' dynamic-LDS limit is set to the whole budget, so no other
  // configuration can fall below what it was resolved for
  HIP_TRY(mcxyz_configure(tt, tile32, t.block, rec_lds + (size_t)lds_budget * elem));
  int64_t budget = lds_budget;
  if (part_cap >= 0) {
    const int occ0 = mcxyz_blocks_per_cu(tt, tile32, t.block, 0);
    const int64_t most_blocks = (int64_t)num_cu * std::max(1, occ0);
    budget = std::min<int64_t>(budget, part_cap / most_blocks);
  }
  if (!has_fluence || r.tile[3] < 0 || (r.tile[3] == 0 && budget < 1)) {
    for (int i = 0; i < 6; ++i) t.tile[i] = 0;
  } else if (r.tile[3] > 0) {
    for (int i = 0; i < 6; ++i) t.tile[i] = r.tile[i];
  } else {
    // automatic box: centred on the source in x,y, from the source depth down
    // (centred on it for an interior point source), about 2/3 as deep as wide
    int nz = (int)std::lround(0.66 * std::cbrt((double)budget));
    nz = std::max(1, std::min(nz, rp.Nz));
    int nxy = (int)std::floor(std::sqrt((double)budget / nz));
    int nx = std::max(1, std::min(nxy, rp.Nx)), ny = std::max(1, std::min(nxy, rp.Ny));
    float sx = rp.launchflag == 1 || rp.mcflag == 2 ? rp.xs : (rp.mcflag == 4 ? 0.0f : rp.xs);
    float sy = rp.launchflag == 1 || rp.mcflag == 2 ? rp.ys : (rp.mcflag == 4 ? 0.0f : rp.ys);
    int cx = (int)std::floor(0.5f * rp.Nx + sx / rp.dx), cy = (int)std::floor(0.5f * rp.Ny + sy / rp.dy);
    int cz = (int)std::floor(rp.zs / rp.dz);
    int z0 = (rp.mcflag == 2 || rp.launchflag == 1) ? cz - nz / 2 : cz;
    t.tile[0] = std::max(0, std::min(cx - nx / 2, rp.Nx - nx));
    t.tile[1] = std::max(0, std::min(cy - ny / 2, rp.Ny - ny));
    t.tile[2] = std::max(0, std::min(z0, rp.Nz - nz));
    t.tile[3] = nx;
    t.tile[4] = ny;
    t.tile[5] = nz;
  }
  int64_t tile_n = (int64_t)t.tile[3] * t.tile[4] * t.tile[5];
  if (tile_n > 0) {
    if (t.tile[0] < 0 || t.tile[1] < 0 || t.tile[2] < 0 || t.tile[0] + t.tile[3] > rp.Nx ||
        t.tile[1] + t.tile[4] > rp.Ny || t.tile[2] + t.tile[5] > rp.Nz)
      return fail(NMC_ERR_ARG, "tuning.tile box outside the grid");
    if (tile_n > lds_budget) return fail(NMC_ERR_ARG, "tuning.tile needs %lld voxels, LDS holds %lld",
                                         (long long)tile_n, (long long)lds_budget);
  }
  size_t lds = rec_lds + (size_t)tile_n * elem;
  // fused record reduction: the transport kernel's LDS must hold a bucket's
  // FP64 sums (automatic: only when the tile already takes that much)
  if (r.rec_fuse < -1 || r.rec_fuse > 1) return fail(NMC_ERR_ARG, "tuning.rec_fuse not -1, 0 or 1");
  constexpr size_t kFuseLds = kBucketVox * sizeof(double);
  t.rec_fuse = -1;
  if (n_buckets > 0 && (r.rec_fuse == 1 || (r.rec_fuse == 0 && lds >= kFuseLds))) {
    if ((size_t)lds_budget * elem + rec_lds < kFuseLds)
      return fail(NMC_ERR_ARG, "tuning.rec_fuse: the LDS cannot hold a bucket's sums");
    lds = std::max(lds, kFuseLds);
    t.rec_fuse = 1;
  }
  int occ = mcxyz_blocks_per_cu(tt, tile32, t.block, lds);
  if (occ <= 0) return fail(NMC_ERR_ARG, "kernel does not fit on a CU with block %d, %zu B LDS", t.block, lds);
  t.blocks_per_cu = r.blocks_per_cu > 0 ? std::min(r.blocks_per_cu, occ) : occ;
  const int grid = num_cu * t.blocks_per_cu;
  if (part_cap >= 0 && (int64_t)grid * tile_n > part_cap)
    return fail(NMC_ERR_ARG, "tuning.tile: %d workgroups x %lld voxels of FP64 block partials exceed "
                "the %lld doubles the scratch array holds", grid, (long long)tile_n, (long long)part_cap);
  out->t = t;
  out->lds = lds;
  out->grid = grid;
  out->tile_n = tile_n;
  out->n_buckets = n_buckets;
  return ok();
}

uint32_t pick_chunk(uint64_t n, int waves, int req) {
  if (req > 0) return (uint32_t)req;
  uint64_t c = n / ((uint64_t)waves * 8ull);
  return (uint32_t)std::max<uint64_t>(64, std::min<uint64_t>(1024, c));
}

// Launch arguments of an mcxyz run that follow from the configuration alone
// (the device pointers, photon count and queue chunk are set by the caller).
XyzArgs xyz_args(const RunParams& rp, const nmc_ext_config& e, int sim, uint32_t key0, uint32_t key1) {
  (void)sim;
  XyzArgs a{};
  a.Nx = rp.Nx;
  a.Ny = rp.Ny;
  a.Nz = rp.Nz;
  a.dx = rp.dx;
  a.dy = rp.dy;
  a.dz = rp.dz;
  a.hx = 0.5f * (float)rp.Nx;
  a.hy = 0.5f * (float)rp.Ny;
  a.bmode = rp.boundaryflag;
  a.det_state = rp.det_state;
  a.zsurf = rp.zsurf;
  a.thr = e.threshold;
  a.chance = e.chance;
  a.max_events = e.max_events;
  a.key0 = key0;
  a.key1 = key1;
  a.count_visits = (e.flags & NMC_FLAG_COUNT_VISITS) != 0;
  return a;
}

// The reduction arguments of record slot `slot` of R.
RecReduceArgs rec_reduce_args(const XyzRecords& R, int slot, double* F, int grid, uint32_t nvox) {
  RecReduceArgs rr{};
  rr.parts = R.parts;
  rr.rec_base = R.base;
  rr.rec_cap = R.cap;
  rr.rec = slot ? R.rec1 : R.rec;
  rr.rec_count = slot ? R.count1 : R.count;
  rr.F = F;
  rr.grid = grid;
  rr.nvox = nvox;
  return rr;
}

// Enqueues the reduction of the records still pending in R (fused mode: the
// last launch's), on stream s.
int flush_records(XyzRecords& R, double* F, int grid, uint32_t nvox, hipStream_t s) {
  if (R.pending < 0) return ok();
  const RecReduceArgs rr = rec_reduce_args(R, R.pending, F, grid, nvox);
  R.pending = -1;
  HIP_TRY(launch_rec_reduce(rr, R.n_parts, s));
  return ok();
}

// Enqueues photons [first, first + n) of an mcxyz run on stream s, `first`
// being a.state->first: [reset queue counter, transport kernel, tile
// reduction, record reduction, advance first] per launch.  With the FP32
// tile each workgroup sums its share of a launch's deposits in FP32; runs
// are split into launches of at most tuning.split photons per workgroup
// (each launch flushes its tile into the FP64 array), which bounds those
// sums to the size the FP32 tile's accuracy was measured at, whatever the
// grid.  The launches take consecutive photon ids, so the photons are the
// same.  With the fused record reduction (R->rec1 set, `fuse`) a launch
// writes the record slot the previous one did not and reduces the previous
// one's records itself; the last launch's stay pending in R for the next
// launch or flush_records.
int enqueue_xyz(XyzArgs a, bool two_term, const XyzLaunch& L, uint64_t n, hipStream_t s,
                uint32_t* launches, XyzRecords* R, hipEvent_t k0, hipEvent_t k1, bool fuse) {
  const bool recs = R && R->n_buckets > 0 && a.F && !a.count_visits;
  fuse = fuse && recs && R->rec1 != nullptr;
  const uint32_t nvox = (uint32_t)((uint64_t)a.Nx * a.Ny * a.Nz);
  a.n_buckets = recs ? R->n_buckets : 0;
  a.rec_base = recs ? R->base : nullptr;
  a.rec_cap = recs ? R->cap : nullptr;
  a.prev_n_parts = 0;
  a.prev_next = nullptr;
  if (recs && !fuse) NMC_TRY(flush_records(*R, a.F, L.grid, nvox, s));
  a.tile_x0 = L.t.tile[0];
  a.tile_y0 = L.t.tile[1];
  a.tile_z0 = L.t.tile[2];
  a.tile_nx = L.t.tile[3];
  a.tile_ny = L.t.tile[4];
  a.tile_nz = L.t.tile[5];
  a.refill_min = (uint32_t)L.t.refill_min;
  const uint64_t cap = L.t.mode >= 1 ? (uint64_t)L.t.split * (uint64_t)L.grid : 0;
  const uint64_t step = cap > 0 ? cap : n;
  for (uint64_t done = 0; done < n; done += step) {
    const uint64_t part = std::min<uint64_t>(step, n - done);
    a.n = part;
    a.chunk = pick_chunk(part, L.grid * (L.t.block / 64), L.t.chunk);
    const int slot = fuse ? R->slot : 0;
    if (recs) {
      a.rec = slot ? R->rec1 : R->rec;
      a.rec_count = slot ? R->count1 : R->count;
      if (fuse && R->pending >= 0) {
        a.prev = rec_reduce_args(*R, R->pending, a.F, L.grid, nvox);
        a.prev_n_parts = R->n_parts;
        a.prev_next = R->next;
        HIP_TRY(hipMemsetAsync(R->next, 0, sizeof(uint32_t), s));
      } else {
        a.prev_n_parts = 0;
      }
    }
    HIP_TRY(launch_state_begin(a.state, s));
    if (k0) HIP_TRY(hipEventRecord(k0, s));
    HIP_TRY(launch_mcxyz(a, two_term, L.t.mode >= 1, L.t.block, L.grid, L.lds, s));
    if (k1) HIP_TRY(hipEventRecord(k1, s));
    HIP_TRY(launch_tile_reduce(a.tile_part, L.grid, a, s));
    if (recs) {
      if (fuse) {
        R->pending = slot;
        R->slot = slot ^ 1;
      } else {
        HIP_TRY(launch_rec_reduce(rec_reduce_args(*R, 0, a.F, L.grid, nvox), R->n_parts, s));
      }
    }
    HIP_TRY(launch_state_advance(a.state, part, s));
    ++*launches;
  }
  return ok();
}

// Device tables of an mcxyz configuration: the tissue rows (index 0..18) and
// the source / detector block (its ryx and det2d pointers left null).
void xyz_tables(const RunParams& rp, const TissueParams& tp, int sim, XyzTissue* tis, XyzAux* aux) {
  for (int t = 0; t < NMC_NTISS; ++t) {
    std::memset(&tis[t], 0, sizeof(XyzTissue));
    tis[t].mua = tp.muav[t];
    tis[t].mus = tp.musv[t];
    tis[t].inv_mus = inv_pos(tp.musv[t]);
    tis[t].g = tp.gv[t];
    tis[t].hg = hg_prepare(tp.gv[t]);
    if (sim == NMC_SIM_MCXYZ_TT) tt_prepare(tp.gf[t], tp.af[t], tp.gb[t], tp.ab[t], tp.CC[t], &tis[t].tt);
    tis[t].slot = -1;
  }
  std::memset(aux, 0, sizeof(*aux));
  aux->mcflag = rp.mcflag;
  aux->launchflag = rp.launchflag;
  aux->xs = rp.xs;
  aux->ys = rp.ys;
  aux->zs = rp.zs;
  aux->ux0 = rp.ux0;
  aux->uy0 = rp.uy0;
  aux->uz0 = rp.uz0;
  aux->radius = rp.radius;
  aux->waist = rp.waist;
  aux->zfocus = rp.zfocus;
  aux->n1 = rp.n1;
  aux->n2 = rp.n2;
  if (rp.xd == 0.0f) {
    aux->rmin = 0.0f;
    aux->rmax = rp.det_radius;
  } else {
    aux->rmin = rp.xd - rp.det_radius;
    aux->rmax = aux->rmin + 2.0f * rp.det_radius;
  }
  aux->cos_na = rp.na >= 90.0f ? -1.0f : (float)std::cos((double)rp.na * 3.14159265358979323846 / 180.0);
  aux->Nx = rp.Nx;
  aux->Ny = rp.Ny;
}

}  // namespace nmc

namespace {

uint32_t key0_of(const nmc_handle* h) { return (uint32_t)h->ext.seed; }
uint32_t key1_of(const nmc_handle* h, uint32_t stream) { return (uint32_t)(h->ext.seed >> 32) + stream; }

// Visits outside the tile box per photon, per record bucket: a visit-count
// launch of kRecPilot photons (no tile, FP64 counts) on scratch buffers, so
// the handle's tallies and photon state are untouched.
constexpr uint64_t kRecPilot = 200000;

int record_rates(nmc_handle* h, const XyzLaunch& L, std::vector<double>* rate) {
  const RunParams& rp = h->rp;
  const int64_t nv = nvox(&rp), n2 = (int64_t)rp.Nx * rp.Ny;
  double* cnt = nullptr;
  double* ryx = nullptr;
  uint32_t* det = nullptr;
  nmc_totals* tot = nullptr;
  RunState* st = nullptr;
  XyzAux* aux = nullptr;
  auto release = [&] {
    free_ptr(cnt);
    free_ptr(ryx);
    free_ptr(det);
    free_ptr(tot);
    free_ptr(st);
    free_ptr(aux);
  };
  hipError_t e = hipMalloc(&cnt, nv * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&ryx, n2 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&det, n2 * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&tot, sizeof(nmc_totals));
  if (e == hipSuccess) e = hipMalloc(&st, sizeof(RunState));
  if (e == hipSuccess) e = hipMalloc(&aux, sizeof(XyzAux));
  if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, nv * sizeof(double), h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(ryx, 0, n2 * sizeof(double), h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(det, 0, n2 * sizeof(uint32_t), h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(tot, 0, sizeof(nmc_totals), h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(st, 0, sizeof(RunState), h->stream);
  if (e == hipSuccess) {
    std::vector<XyzTissue> tis(NMC_NTISS);
    XyzAux ax{};
    xyz_tables(rp, h->tp, h->sim, tis.data(), &ax);
    ax.ryx = ryx;
    ax.det2d = det;
    e = hipMemcpyAsync(aux, &ax, sizeof ax, hipMemcpyHostToDevice, h->stream);
  }
  if (e != hipSuccess) {
    release();
    return fail(e == hipErrorOutOfMemory ? NMC_ERR_NOMEM : NMC_ERR_DEVICE, "record pilot: %s",
                hipGetErrorString(e));
  }
  XyzArgs a = xyz_args(rp, h->ext, h->sim, key0_of(h), key1_of(h, h->ext.stream_id));
  a.count_visits = 1;
  a.aux = aux;
  a.V = h->dV;
  a.tissue = h->dTis;
  a.F = cnt;
  a.tot = tot;
  a.state = st;
  XyzLaunch P = L;
  for (int i = 0; i < 6; ++i) P.t.tile[i] = 0;
  P.t.mode = 0;
  P.t.split = 0;
  P.tile_n = 0;
  P.lds = 0;
  P.n_buckets = 0;
  uint32_t launches = 0;
  int rc = enqueue_xyz(a, h->sim == NMC_SIM_MCXYZ_TT, P, kRecPilot, h->stream, &launches, nullptr,
                       nullptr, nullptr, false);
  std::vector<double> host;
  if (rc == NMC_OK) {
    host.resize(nv);
    e = hipMemcpyAsync(host.data(), cnt, nv * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess)
      rc = fail(NMC_ERR_DEVICE, "record pilot: %s", hipGetErrorString(e));
  }
  release();
  if (rc != NMC_OK) return rc;
  rate->assign(L.n_buckets, 0.0);
  const int* b = L.t.tile;
  for (int64_t v = 0; v < nv; ++v) {
    if (host[v] == 0.0) continue;
    const int64_t iy = v % rp.Ny, ix = (v / rp.Ny) % rp.Nx, iz = v / ((int64_t)rp.Ny * rp.Nx);
    const bool in_tile = ix >= b[0] && ix < b[0] + b[3] && iy >= b[1] && iy < b[1] + b[4] &&
                         iz >= b[2] && iz < b[2] + b[5];
    if (!in_tile) (*rate)[v >> kBucketShift] += host[v];
  }
  for (double& r : *rate) r /= (double)kRecPilot;
  return ok();
}

// Record buffers for launch shape L: each bucket's workgroup slices hold the
// pilot's expected out-of-tile visits of tuning.split photons (plus 15 % for
// uneven shares of the queue, 10 % and 6 sigma of Poisson slack, 32 records),
// scaled by tuning.rec_cap_pct percent when set; the reduction takes about
// 4e6 records per workgroup.  With the fused reduction (tuning.rec_fuse) a
// second record buffer and slot counts of the same shape.
int make_records(nmc_handle* h, const XyzLaunch& L, XyzRecords* R) {
  std::vector<double> rate;
  NMC_TRY(record_rates(h, L, &rate));
  const int nb = L.n_buckets, grid = L.grid;
  const double per_wg = 1.15 * (double)L.t.split;
  const double pct = L.t.rec_cap_pct > 0 ? L.t.rec_cap_pct / 100.0 : 1.0;
  std::vector<unsigned long long> base(nb);
  std::vector<uint32_t> cap(nb);
  std::vector<RecPart> parts;
  unsigned long long total = 0;
  for (int b = 0; b < nb; ++b) {
    const double e = rate[b] * per_wg;
    const double c = std::ceil((1.1 * e + 6.0 * std::sqrt(e) + 32.0) * pct);
    if (c > 4.0e9) return fail(NMC_ERR_ARG, "tuning.split: record slices above 2^32 records");
    cap[b] = (uint32_t)c;
    base[b] = total;
    total += (unsigned long long)cap[b] * grid;
    const double expect = rate[b] * (double)L.t.split * grid;
    int np = (int)std::min<double>(grid, std::max(1.0, std::ceil(expect / 4.0e6)));
    for (int k = 0; k < np; ++k)
      parts.push_back(RecPart{b, (int)((int64_t)grid * k / np), (int)((int64_t)grid * (k + 1) / np), 0});
  }
  R->n_buckets = nb;
  R->n_parts = (int)parts.size();
  R->n_rec = total;
  HIP_TRY(rec_reduce_configure());
  HIP_TRY(hipMalloc(&R->base, nb * sizeof(unsigned long long)));
  HIP_TRY(hipMalloc(&R->cap, nb * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&R->count, (size_t)nb * grid * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&R->parts, parts.size() * sizeof(RecPart)));
  HIP_TRY(hipMalloc(&R->rec, std::max<unsigned long long>(total, 1) * sizeof(uint2)));
  if (L.t.rec_fuse == 1) {
    HIP_TRY(hipMalloc(&R->count1, (size_t)nb * grid * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&R->rec1, std::max<unsigned long long>(total, 1) * sizeof(uint2)));
    HIP_TRY(hipMalloc(&R->next, sizeof(uint32_t)));
  }
  HIP_TRY(hipMemcpy(R->base, base.data(), nb * sizeof(unsigned long long), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(R->cap, cap.data(), nb * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(R->parts, parts.data(), parts.size() * sizeof(RecPart), hipMemcpyHostToDevice));
  return ok();
}

// Resolves h->tune_req for an mcxyz handle.  Transactional: every check runs
// and every new buffer is allocated before the handle changes, so a rejected
// request leaves the previous configuration (and its buffers) in force.
int resolve_xyz_tuning(nmc_handle* h) {
  XyzLaunch L{};
  NMC_TRY(resolve_xyz_launch(h->rp, h->sim, h->dF != nullptr, h->tune_req, h->lds_max, h->num_cu, -1, &L));
  XyzRecords R;
  if (L.n_buckets > 0) {
    int st = make_records(h, L, &R);
    if (st != NMC_OK) {
      free_records(R);
      return st;
    }
  }
  double* part = nullptr;
  if (L.tile_n > 0) {
    hipError_t e = hipMalloc(&part, (size_t)L.grid * L.tile_n * sizeof(double));
    if (e != hipSuccess) {
      free_records(R);
      return fail(e == hipErrorOutOfMemory ? NMC_ERR_NOMEM : NMC_ERR_DEVICE, "tile partials: %s",
                  hipGetErrorString(e));
    }
  }
  // commit: the cached graph holds the old buffers, so it goes first
  drop_graph(h);
  free_ptr(h->dTilePart);
  free_records(h->recs);
  h->dTilePart = part;
  h->recs = R;
  h->xyz_lds = L.lds;
  h->xyz_grid = L.grid;
  h->tune = L.t;
  return ok();
}

// ---- mcsub launch configuration -------------------------------------------
// Transactional like resolve_xyz_tuning.
int resolve_sub_tuning(nmc_handle* h) {
  const nmc_tuning& r = h->tune_req;
  nmc_tuning t{};
  // tally placement: tuning.mode = 1 LDS, 2 global atomics, 0 automatic
  // (LDS whenever the per-item tallies fit)
  int mode = r.mode;
  if (mode < 0 || mode > 2) return fail(NMC_ERR_ARG, "tuning.mode (tally placement) not in 0..2");
  bool want_lds = mode != 2;
  // kSubStaticLds: the kernel's static LDS (cosine table, log_tab's rows)
  size_t lds_full = mcsub_lds_bytes(h->n_items, per_item(h)) + kSubStaticLds;
  const bool sub_lds = want_lds && lds_full <= h->lds_max;
  const size_t lds = mcsub_lds_bytes(h->n_items, sub_lds ? per_item(h) : 0);
  if (lds + kSubStaticLds > h->lds_max)
    return fail(NMC_ERR_ARG, "%d work items need %zu B of LDS", h->n_items, lds);
  t.block = r.block ? r.block : 1024;
  if (t.block != 256 && t.block != 512 && t.block != 1024)
    return fail(NMC_ERR_ARG, "mcsub block must be 256, 512 or 1024");
  t.refill_min = r.refill_min ? r.refill_min : 1;  // mcsub's launch is cheap: refill at once
  if (t.refill_min < 1 || t.refill_min > 64) return fail(NMC_ERR_ARG, "tuning.refill_min not in 1..64");
  t.chunk = r.chunk;
  t.mode = sub_lds ? 1 : 2;
  HIP_TRY(mcsub_configure(t.block, h->lds_max - kSubStaticLds, sub_lds));
  int occ = mcsub_blocks_per_cu(t.block, lds, sub_lds);
  if (occ <= 0) return fail(NMC_ERR_ARG, "mcsub kernel does not fit on a CU");
  t.blocks_per_cu = r.blocks_per_cu > 0 ? std::min(r.blocks_per_cu, occ) : occ;
  const int grid = h->num_cu * t.blocks_per_cu;
  size_t stride = (size_t)h->n_items * ((sub_lds ? per_item(h) : 0) + TF_COUNT + TI_COUNT);
  double* part = nullptr;
  double* rep = nullptr;
  int reps = 1;
  HIP_TRY(hipMalloc(&part, (size_t)grid * stride * sizeof(double)));
  // global-mode tallies go to per-XCD replicas (the bins near the beam axis
  // take a large share of all adds); zeroed here, then by each reduction
  if (!sub_lds) {
    reps = replica_count((int64_t)h->n_items * per_item(h), grid);
    size_t rep_n = (size_t)reps * h->n_items * per_item(h);
    hipError_t e = hipMalloc(&rep, rep_n * sizeof(double));
    if (e == hipSuccess) e = hipMemsetAsync(rep, 0, rep_n * sizeof(double), h->stream);
    if (e != hipSuccess) {
      free_ptr(part);
      free_ptr(rep);
      return fail(e == hipErrorOutOfMemory ? NMC_ERR_NOMEM : NMC_ERR_DEVICE,
                  "mcsub tally replicas: %s", hipGetErrorString(e));
    }
  }
  // commit: the cached graph holds the old buffers, so it goes first
  drop_graph(h);
  free_ptr(h->dPart);
  free_ptr(h->dRep);
  h->dPart = part;
  h->dRep = rep;
  h->sub_reps = reps;
  h->sub_lds = sub_lds;
  h->sub_lds_bytes = lds;
  h->sub_grid = grid;
  h->tune = t;
  return ok();
}


// Enqueues one run (no events, no capture handling).
int enqueue_xyz_sens(nmc_handle* h, XyzArgs a, uint64_t n, hipStream_t s, hipEvent_t k0, hipEvent_t k1);

int enqueue_run(nmc_handle* h, uint64_t n) {
  hipStream_t s = h->stream;
  if (h->sim == NMC_SIM_MCSUB) {
    HIP_TRY(launch_state_begin(h->dState, s));
    const RunParams& rp = h->rp;
    const nmc_ext_config& e = h->ext;
    SubArgs a{};
    a.mcflag = rp.mcflag;
    a.bmode = rp.boundaryflag;
    a.nr = e.nr;
    a.nz = e.nz;
    a.na = e.na;
    a.dr = e.dr;
    a.dz = e.dz;
    a.radius = rp.radius;
    a.waist = rp.waist;
    a.zfocus = rp.zfocus;
    a.xs = rp.xs;
    a.ys = rp.ys;
    a.zs = rp.zs;
    a.thr = e.threshold;
    a.chance = e.chance;
    a.max_events = e.max_events;
    a.key0 = key0_of(h);
    uint64_t total = h->prefix[h->n_items];
    a.chunk = pick_chunk(total, h->sub_grid * (h->tune.block / 64), h->tune.chunk);
    a.refill_min = (uint32_t)h->tune.refill_min;
    a.n_items = h->n_items;
    a.items = h->dItems;
    a.prefix = h->dPrefix;
    a.cos_tab = h->dCos;
    a.tally_rz = h->n_rz > 0;
    a.tally = h->sub_lds ? h->dOut : h->dRep;
    a.reps = h->sub_lds ? 1 : h->sub_reps;
    a.part = h->dPart;
    a.state = h->dState;
    HIP_TRY(launch_mcsub(a, h->tune.block, h->sub_grid, h->sub_lds_bytes, h->sub_lds, s));
    h->launches++;
    {
      int64_t lds_per = h->sub_lds ? per_item(h) : 0;
      int64_t stride = (int64_t)h->n_items * (lds_per + TF_COUNT + TI_COUNT);
      double* out = h->dOut + (h->sub_lds ? 0 : (int64_t)h->n_items * per_item(h));
      HIP_TRY(launch_reduce_partials(h->dPart, h->sub_grid, stride, out, h->n_items, (int)lds_per, s));
      if (!h->sub_lds)
        HIP_TRY(launch_reduce_partials(h->dRep, h->sub_reps, (int64_t)h->n_items * per_item(h), h->dOut,
                                       h->n_items, (int)per_item(h), s, true));
    }
    uint64_t adv = 0;
    for (int j = 0; j < h->n_items; ++j) adv = std::max<uint64_t>(adv, h->prefix[j + 1] - h->prefix[j]);
    HIP_TRY(launch_state_advance(h->dState, adv, s));
  } else {
    XyzArgs a = xyz_args(h->rp, h->ext, h->sim, key0_of(h), key1_of(h, h->ext.stream_id));
    a.aux = h->dAux;
    a.V = h->dV;
    a.tissue = h->dTis;
    a.F = h->dF;
    a.tile_part = h->dTilePart;
    a.tot = h->dTot;
    a.state = h->dState;
    XyzLaunch L{h->tune, h->xyz_lds, h->xyz_grid, 0, 0};
    uint32_t launches = 0;
    L.n_buckets = h->recs.n_buckets;
    // transport-kernel events only outside graph capture
    hipEvent_t k0 = h->graph_mode ? nullptr : h->k0, k1 = h->graph_mode ? nullptr : h->k1;
    if (h->sens_n > 0) {
      NMC_TRY(enqueue_xyz_sens(h, a, n, s, k0, k1));
      launches = 1;
    } else {
      // graph mode: every launch reduces its own records (a captured graph
      // holds no state carried between replays)
      NMC_TRY(enqueue_xyz(a, h->sim == NMC_SIM_MCXYZ_TT, L, n, s, &launches, &h->recs, k0, k1,
                          !h->graph_mode));
    }
    h->k_timed = k0 != nullptr;
    h->launches += launches;
  }
  return ok();
}

// A run of the sensitivity variant: one launch of n photons, every deposit
// to the FP64 arrays (no tile, no records).
int enqueue_xyz_sens(nmc_handle* h, XyzArgs a, uint64_t n, hipStream_t s, hipEvent_t k0, hipEvent_t k1) {
  a.tile_x0 = a.tile_y0 = a.tile_z0 = 0;
  a.tile_nx = a.tile_ny = a.tile_nz = 0;
  a.n_buckets = 0;
  a.refill_min = (uint32_t)h->tune.refill_min;
  a.sens = h->dSens;
  a.sens_n = h->sens_n;
  for (int j = 0; j < h->sens_n; ++j) a.sens_inv_mus[j] = inv_pos(h->tp.musv[h->sens_ids[j]]);
  a.n = n;
  a.chunk = pick_chunk(n, h->sens_grid * (kSensBlock / 64), h->tune.chunk);
  HIP_TRY(launch_state_begin(a.state, s));
  if (k0) HIP_TRY(hipEventRecord(k0, s));
  HIP_TRY(launch_mcxyz_sens(a, h->sim == NMC_SIM_MCXYZ_TT, h->sens_grid, h->sens_lds, s));
  if (k1) HIP_TRY(hipEventRecord(k1, s));
  HIP_TRY(launch_state_advance(a.state, n, s));
  return ok();
}

// Enqueues the reduction of the records the last fused mcxyz launch left
// pending (nothing for other handles), on the handle's stream: every entry
// point that reads or changes the fluence, or runs without fusing, calls it
// first.
int flush_pending(nmc_handle* h) {
  if (h->recs.pending < 0) return ok();
  return flush_records(h->recs, h->dF, h->xyz_grid, (uint32_t)nvox(&h->rp), h->stream);
}

int write_first(nmc_handle* h) {
  RunState st{0ull, (unsigned long long)h->next_photon};
  HIP_TRY(hipMemcpyAsync(h->dState, &st, sizeof st, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return ok();
}

// Runs (n photons for mcxyz; every item's budget for mcsub), via the graph in
// graph mode.
int run_async(nmc_handle* h, uint64_t n) {
  HIP_TRY(hipSetDevice(h->device));
  // graph and sensitivity runs reduce their own records; pending ones from
  // earlier eager launches go first, outside any capture
  if (h->graph_mode || h->sens_n > 0) NMC_TRY(flush_pending(h));
  HIP_TRY(hipEventRecord(h->e0, h->stream));
  h->launches = 0;
  if (h->graph_mode) {
    if (!h->graph_exec || h->graph_n != n || h->graph_seed != h->ext.seed ||
        h->graph_stream != h->ext.stream_id) {
      drop_graph(h);
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      int st = enqueue_run(h, n);
      hipError_t ce = hipStreamEndCapture(h->stream, &g);
      if (st != NMC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      HIP_TRY(ce);
      hipError_t ie = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIP_TRY(ie);
      h->graph_n = n;
      h->graph_seed = h->ext.seed;
      h->graph_stream = h->ext.stream_id;
      h->graph_launches = h->launches;
    }
    HIP_TRY(hipGraphLaunch(h->graph_exec, h->stream));
    h->launches = h->graph_launches;
  } else {
    NMC_TRY(enqueue_run(h, n));
  }
  HIP_TRY(hipEventRecord(h->e1, h->stream));
  h->timed = true;
  return ok();
}

// Queue order: items whose photons live longest (highest albedo) first, so
// the end of a launch is made of short-lived photons.
void order_items(nmc_handle* h) {
  int n = h->n_items;
  h->order.resize(n);
  for (int j = 0; j < n; ++j) h->order[j] = j;
  std::stable_sort(h->order.begin(), h->order.end(), [&](int a, int b) {
    const nmc_work_item &x = h->items[a], &y = h->items[b];
    double ax = x.mus / std::max(1e-30, (double)x.mua + x.mus);
    double ay = y.mus / std::max(1e-30, (double)y.mua + y.mus);
    return ax > ay;
  });
}

int upload_media(nmc_handle* h);

int alloc_sub_outputs(nmc_handle* h) {
  free_ptr(h->dOut);
  size_t n = (size_t)h->n_items * (per_item(h) + TF_COUNT + TI_COUNT);
  HIP_TRY(hipMalloc(&h->dOut, n * sizeof(double)));
  HIP_TRY(hipMalloc(&h->dItems, h->n_items * sizeof(SubMedium)));
  HIP_TRY(hipMalloc(&h->dPrefix, (h->n_items + 1) * sizeof(uint64_t)));
  order_items(h);
  NMC_TRY(upload_media(h));
  std::vector<float> cos_tab(64, 0.0f);
  for (int j = 0; j < h->ext.na; ++j)
    cos_tab[j] = (float)std::cos((double)(j + 1) * (3.14159265358979323846 / 2.0) / (double)h->ext.na);
  HIP_TRY(hipMalloc(&h->dCos, 64 * sizeof(float)));
  HIP_TRY(hipMemcpy(h->dCos, cos_tab.data(), 64 * sizeof(float), hipMemcpyHostToDevice));
  return ok();
}

// Uploads the media in queue order with their budget prefix sums (the Philox
// key of a medium depends on the seed).
int upload_media(nmc_handle* h) {
  int n = h->n_items;
  std::vector<SubMedium> med(n);
  h->prefix.assign(n + 1, 0);
  for (int q = 0; q < n; ++q) {
    const nmc_work_item& it = h->items[h->order[q]];
    med[q] = make_medium(it.mua, it.mus, it.g, it.n1, it.n2, key1_of(h, it.stream_id));
    med[q].slot = h->order[q];
    h->prefix[q + 1] = h->prefix[q] + it.n_photons;
  }
  HIP_TRY(hipMemcpy(h->dItems, med.data(), med.size() * sizeof(SubMedium), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->dPrefix, h->prefix.data(), h->prefix.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  return ok();
}

int check_tissues(int sim, const TissueParams* tp, const bool* used);

// The checks of nmc_validate; scan = false skips the voxel map (the
// allocation-free entry point checks a device-resident map separately).
int validate_impl(int sim, const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                  const nmc_ext_config* ext, bool scan) {
  if (!rp || !tp) return fail(NMC_ERR_ARG, "RunParams and TissueParams are required");
  if (sim != NMC_SIM_MCSUB && sim != NMC_SIM_MCXYZ && sim != NMC_SIM_MCXYZ_TT)
    return fail(NMC_ERR_UNSUPPORTED, "unsupported simulation kind %d", sim);
  nmc_ext_config e;
  ext_defaults(&e);
  if (ext) {
    if (ext->struct_size != sizeof(nmc_ext_config) || ext->version != NMC_EXT_VERSION)
      return fail(NMC_ERR_ARG, "nmc_ext_config size/version mismatch (%u/%u)", ext->struct_size,
                  ext->version);
    e = *ext;
  }
  if (!(e.threshold >= 0.0f) || !(e.chance > 0.0f && e.chance <= 1.0f))
    return fail(NMC_ERR_ARG, "roulette threshold/chance out of range");
  if (e.max_events == 0) return fail(NMC_ERR_ARG, "max_events must be positive");
  if (!(rp->n1 > 0.0f) || !(rp->n2 > 0.0f) || !std::isfinite(rp->n1) || !std::isfinite(rp->n2))
    return fail(NMC_ERR_ARG, "refractive indices must be positive and finite");
  if (rp->boundaryflag < 0 || rp->boundaryflag > 2)
    return fail(NMC_ERR_ARG, "boundaryflag %d not in 0..2", rp->boundaryflag);
  if (sim == NMC_SIM_MCSUB) {
    if (rp->mcflag < 0 || rp->mcflag > 2) return fail(NMC_ERR_ARG, "mcsub mcflag %d not in 0..2", rp->mcflag);
    // the medium is the half-space z > 0 with its surface at z = 0 for every
    // source; zs places only the isotropic point source (the reference uses
    // its compile-time zs_mcsub = 0, definitions.h:47, both as the launch
    // depth and as the escape plane, kernels.metal:113,131,152,190)
    if ((rp->mcflag == 0 || rp->mcflag == 1) && rp->zs != 0.0f)
      return fail(NMC_ERR_ARG, "mcsub beams (mcflag 0, 1) enter at the surface z = 0: zs must be 0");
    if (rp->mcflag == 2 && !(rp->zs >= 0.0f && std::isfinite(rp->zs) && std::isfinite(rp->xs) &&
                             std::isfinite(rp->ys)))
      return fail(NMC_ERR_ARG, "mcsub point source: zs must be finite and >= 0 (inside the medium)");
    if (e.nr <= 0 || e.nz <= 0 || e.na <= 0 || e.na > 64 || !(e.dr > 0.0f) || !(e.dz > 0.0f))
      return fail(NMC_ERR_ARG, "mcsub bins invalid (nr,nz >= 1, 1 <= na <= 64, dr,dz > 0)");
    float mua = tp->muav[1], mus = tp->musv[1], g = tp->gv[1];
    if (!finite_nonneg(mua) || !finite_nonneg(mus) || !(mua + mus > 0.0f))
      return fail(NMC_ERR_ARG, "tissue 1: mua, mus must be finite, >= 0 and not both 0");
    if (!(g >= -1.0f && g <= 1.0f)) return fail(NMC_ERR_ARG, "tissue 1: g not in [-1,1]");
    return ok();
  }
  if (rp->Nx <= 0 || rp->Ny <= 0 || rp->Nz <= 0)
    return fail(NMC_ERR_ARG, "grid size %dx%dx%d invalid", rp->Nx, rp->Ny, rp->Nz);
  if ((double)rp->Nx * rp->Ny * rp->Nz > 2.0e9) return fail(NMC_ERR_ARG, "grid larger than 2e9 voxels");
  if (!(rp->dx > 0.0f) || !(rp->dy > 0.0f) || !(rp->dz > 0.0f) || !std::isfinite(rp->dx) ||
      !std::isfinite(rp->dy) || !std::isfinite(rp->dz))
    return fail(NMC_ERR_ARG, "voxel sizes must be positive and finite");
  if (rp->launchflag != 1 && (rp->mcflag < 0 || rp->mcflag > 4))
    return fail(NMC_ERR_ARG, "mcflag %d not in 0..4", rp->mcflag);
  if (rp->det_state != NMC_DET_REFLECTANCE && rp->det_state != NMC_DET_TRANSMITTANCE)
    return fail(NMC_ERR_ARG, "det_state %d not 0 or 1", rp->det_state);
  if (rp->launchflag == 1) {
    float n2 = rp->ux0 * rp->ux0 + rp->uy0 * rp->uy0 + rp->uz0 * rp->uz0;
    if (!(std::fabs(n2 - 1.0f) < 1e-3f)) return fail(NMC_ERR_ARG, "launch direction (ux0,uy0,uz0) is not a unit vector");
  }
  if (!scan) return ok();
  if (!v) return fail(NMC_ERR_ARG, "voxel map is required for mcxyz");
  bool used[256] = {false};
  int64_t n = nvox(rp);
  for (int64_t i = 0; i < n; ++i) used[v[i]] = true;
  return check_tissues(sim, tp, used);
}

// The tissue rows a voxel map uses (used[id]) must exist and be valid.
int check_tissues(int sim, const TissueParams* tp, const bool* used) {
  for (int t = 0; t < 256; ++t) {
    if (!used[t]) continue;
    if (t >= NMC_NTISS) return fail(NMC_ERR_ARG, "voxel tissue id %d >= %d", t, NMC_NTISS);
    if (!finite_nonneg(tp->muav[t]) || !finite_nonneg(tp->musv[t]))
      return fail(NMC_ERR_ARG, "tissue %d: mua/mus must be finite and >= 0", t);
    if (sim == NMC_SIM_MCXYZ_TT) {
      if (!(tp->CC[t] >= 0.0f && tp->CC[t] <= 1.0f)) return fail(NMC_ERR_ARG, "tissue %d: CC not in [0,1]", t);
      if (!(tp->gf[t] >= 0.0f && tp->gf[t] < 1.0f) || !(tp->gb[t] >= 0.0f && tp->gb[t] < 1.0f))
        return fail(NMC_ERR_ARG, "tissue %d: gf/gb not in [0,1)", t);
      if ((tp->gf[t] != 0.0f && !(tp->af[t] > 0.0f)) || (tp->gb[t] != 0.0f && !(tp->ab[t] > 0.0f)))
        return fail(NMC_ERR_ARG, "tissue %d: af/ab must be > 0", t);
    } else if (!(tp->gv[t] >= -1.0f && tp->gv[t] <= 1.0f)) {
      return fail(NMC_ERR_ARG, "tissue %d: g not in [-1,1]", t);
    }
  }
  return ok();
}

}  // namespace

extern "C" {

void nmc_ext_config_init(nmc_ext_config* cfg) {
  if (cfg) ext_defaults(cfg);
}

const char* nmc_last_error(void) { return g_err.c_str(); }

const char* nmc_version(void) {
  return "nmc_hip 2 gfx950 -O3 -ffp-contract=off wave64 philox4x32-10";
}

int nmc_validate(int sim, const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                 const nmc_ext_config* ext) {
  return validate_impl(sim, rp, tp, v, ext, true);
}

int nmc_create(int sim_type, int device, nmc_handle** out) {
  if (!out) return fail(NMC_ERR_ARG, "out is null");
  *out = nullptr;
  if (sim_type != NMC_SIM_MCSUB && sim_type != NMC_SIM_MCXYZ && sim_type != NMC_SIM_MCXYZ_TT &&
      sim_type != NMC_SIM_LAYERED)
    return fail(NMC_ERR_UNSUPPORTED, "unsupported simulation kind %d", sim_type);
  int count = 0;
  HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(NMC_ERR_ARG, "device %d not present (%d devices)", device, count);
  HIP_TRY(hipSetDevice(device));
  nmc_handle* h = new nmc_handle();
  h->sim = sim_type;
  h->device = device;
  int lds = 0;
  hipError_t e = hipDeviceGetAttribute(&h->num_cu, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  h->lds_max = (size_t)lds;
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&h->e0);
  if (e == hipSuccess) e = hipEventCreate(&h->e1);
  if (e == hipSuccess) e = hipEventCreate(&h->k0);
  if (e == hipSuccess) e = hipEventCreate(&h->k1);
  if (e == hipSuccess) e = hipMalloc(&h->dTot, sizeof(nmc_totals));
  if (e == hipSuccess) e = hipMalloc(&h->dState, sizeof(RunState));
  if (e != hipSuccess) {
    nmc_destroy(h);
    return fail(NMC_ERR_DEVICE, "device setup failed: %s", hipGetErrorString(e));
  }
  *out = h;
  return ok();
}

static int prepare_sub(nmc_handle* h, const RunParams* rp, const nmc_ext_config* ext,
                       const nmc_work_item* items, int n_items, bool batch) {
  h->rp = *rp;
  ext_defaults(&h->ext);
  if (ext) h->ext = *ext;
  h->batch = batch;
  h->n_items = n_items;
  h->items.assign(items, items + n_items);
  h->n_ra = (int64_t)(h->ext.nr + 1) * h->ext.na;
  h->n_rz = (h->ext.flags & NMC_FLAG_NO_FLUENCE) ? 0 : (int64_t)(h->ext.nr + 1) * (h->ext.nz + 1);
  NMC_TRY(alloc_sub_outputs(h));
  NMC_TRY(resolve_sub_tuning(h));
  h->prepared = true;
  return nmc_reset_tallies(h);
}

int nmc_prepare(nmc_handle* h, const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                const nmc_ext_config* ext) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  NMC_TRY(nmc_validate(h->sim, rp, tp, v, ext));
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  free_buffers(h);
  h->tp = *tp;
  h->next_photon = 0;
  h->timed = false;
  if (h->sim == NMC_SIM_MCSUB) {
    nmc_ext_config e;
    ext_defaults(&e);
    if (ext) e = *ext;
    nmc_work_item it{};
    it.mua = tp->muav[1];
    it.mus = tp->musv[1];
    it.g = tp->gv[1];
    it.n1 = rp->n1;
    it.n2 = rp->n2;
    it.stream_id = e.stream_id;
    it.n_photons = 0;
    NMC_TRY(prepare_sub(h, rp, ext, &it, 1, false));
    return write_first(h);
  }
  h->rp = *rp;
  ext_defaults(&h->ext);
  if (ext) h->ext = *ext;
  int64_t n = nvox(rp), n2 = (int64_t)rp->Nx * rp->Ny;
  HIP_TRY(hipMalloc(&h->dV, n));
  HIP_TRY(hipMemcpy(h->dV, v, n, hipMemcpyHostToDevice));
  if (!(h->ext.flags & NMC_FLAG_NO_FLUENCE)) HIP_TRY(hipMalloc(&h->dF, n * sizeof(double)));
  HIP_TRY(hipMalloc(&h->dRyx, n2 * sizeof(double)));
  HIP_TRY(hipMalloc(&h->dDet, n2 * sizeof(uint32_t)));
  std::vector<XyzTissue> tis(NMC_NTISS);
  XyzAux aux{};
  xyz_tables(*rp, *tp, h->sim, tis.data(), &aux);
  HIP_TRY(hipMalloc(&h->dTis, NMC_NTISS * sizeof(XyzTissue)));
  HIP_TRY(hipMemcpy(h->dTis, tis.data(), NMC_NTISS * sizeof(XyzTissue), hipMemcpyHostToDevice));
  aux.ryx = h->dRyx;
  aux.det2d = h->dDet;
  HIP_TRY(hipMalloc(&h->dAux, sizeof(XyzAux)));
  HIP_TRY(hipMemcpy(h->dAux, &aux, sizeof(XyzAux), hipMemcpyHostToDevice));
  NMC_TRY(resolve_xyz_tuning(h));
  h->prepared = true;
  NMC_TRY(nmc_reset_tallies(h));
  return write_first(h);
}

int nmc_prepare_batch(nmc_handle* h, const RunParams* rp, const nmc_ext_config* ext,
                      const nmc_work_item* items, int n_items) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (h->sim != NMC_SIM_MCSUB) return fail(NMC_ERR_UNSUPPORTED, "batched runs need an NMC_SIM_MCSUB handle");
  if (!items || n_items <= 0 || n_items > kMaxBatchItems)
    return fail(NMC_ERR_ARG, "n_items must be in 1..%d", kMaxBatchItems);
  TissueParams tp{};
  for (int j = 0; j < n_items; ++j) {
    const nmc_work_item& it = items[j];
    tp.muav[1] = it.mua;
    tp.musv[1] = it.mus;
    tp.gv[1] = it.g;
    RunParams r = *rp;
    r.n1 = it.n1;
    r.n2 = it.n2;
    int st = nmc_validate(NMC_SIM_MCSUB, &r, &tp, nullptr, ext);
    if (st != NMC_OK) return fail(st, "work item %d: %s", j, g_err.c_str());
  }
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  free_buffers(h);
  h->next_photon = 0;
  h->timed = false;
  NMC_TRY(prepare_sub(h, rp, ext, items, n_items, true));
  return write_first(h);
}

int nmc_reset_tallies(nmc_handle* h) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (!h->prepared) return fail(NMC_ERR_STATE, "nmc_prepare has not succeeded");
  HIP_TRY(hipSetDevice(h->device));
  // records still awaiting their reduction belong to the tallies cleared here
  h->recs.pending = -1;
  HIP_TRY(hipMemsetAsync(h->dTot, 0, sizeof(nmc_totals), h->stream));
  if (h->sim == NMC_SIM_MCSUB) {
    size_t n = (size_t)h->n_items * (per_item(h) + TF_COUNT + TI_COUNT);
    HIP_TRY(hipMemsetAsync(h->dOut, 0, n * sizeof(double), h->stream));
  } else {
    int64_t n = nvox(&h->rp), n2 = (int64_t)h->rp.Nx * h->rp.Ny;
    if (h->dF) HIP_TRY(hipMemsetAsync(h->dF, 0, n * sizeof(double), h->stream));
    if (h->dSens)
      HIP_TRY(hipMemsetAsync(h->dSens, 0, 2 * (size_t)h->sens_n * n * sizeof(double), h->stream));
    HIP_TRY(hipMemsetAsync(h->dRyx, 0, n2 * sizeof(double), h->stream));
    HIP_TRY(hipMemsetAsync(h->dDet, 0, n2 * sizeof(uint32_t), h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return ok();
}

int nmc_set_seed(nmc_handle* h, uint64_t seed, uint64_t first_photon) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (!h->prepared) return fail(NMC_ERR_STATE, "nmc_prepare has not succeeded");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->ext.seed = seed;
  h->next_photon = first_photon;
  if (h->sim == NMC_SIM_MCSUB) NMC_TRY(upload_media(h));
  return write_first(h);
}

int nmc_set_tuning(nmc_handle* h, const nmc_tuning* t) {
  if (!h || !t) return fail(NMC_ERR_ARG, "null argument");
  if (t->refill_min < 0 || t->refill_min > 64) return fail(NMC_ERR_ARG, "tuning.refill_min not in 1..64");
  const nmc_tuning prev = h->tune_req;
  h->tune_req = *t;
  if (!h->prepared) return ok();
  HIP_TRY(hipSetDevice(h->device));
  NMC_TRY(flush_pending(h));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->sim == NMC_SIM_LAYERED) {
    // the grid, and the buffers sized by it, are fixed at nmc_layered_prepare;
    // the refill batch and the queue chunk can change between runs
    h->tune.refill_min = t->refill_min ? t->refill_min : 16;
    h->tune.chunk = t->chunk;
    return ok();
  }
  // a rejected request changes nothing: the resolution is transactional and
  // the request it replaced stays the one a later nmc_prepare resolves
  int st = h->sim == NMC_SIM_MCSUB ? resolve_sub_tuning(h) : resolve_xyz_tuning(h);
  if (st != NMC_OK) h->tune_req = prev;
  return st;
}

int nmc_last_transport_ms(nmc_handle* h, float* ms) {
  if (!h || !ms) return fail(NMC_ERR_ARG, "null argument");
  if (h->sim != NMC_SIM_MCXYZ && h->sim != NMC_SIM_MCXYZ_TT)
    return fail(NMC_ERR_UNSUPPORTED, "transport-kernel timing is kept for mcxyz handles");
  if (!h->timed || !h->k_timed)
    return fail(NMC_ERR_STATE, "no eager mcxyz run has been timed (graph replays keep no kernel events)");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipEventSynchronize(h->k1));
  HIP_TRY(hipEventElapsedTime(ms, h->k0, h->k1));
  return ok();
}

int nmc_mcxyz_sens(nmc_handle* h, const int32_t* tissue_ids, int n_tissues) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (h->sim != NMC_SIM_MCXYZ && h->sim != NMC_SIM_MCXYZ_TT)
    return fail(NMC_ERR_UNSUPPORTED, "sensitivities are computed for mcxyz handles");
  if (!h->prepared) return fail(NMC_ERR_STATE, "nmc_prepare has not succeeded");
  if (!h->dF) return fail(NMC_ERR_STATE, "sensitivities need the fluence tally (no NMC_FLAG_NO_FLUENCE)");
  if (n_tissues < 0 || n_tissues > NMC_SENS_MAX)
    return fail(NMC_ERR_ARG, "n_tissues %d not in 0..%d", n_tissues, NMC_SENS_MAX);
  if (n_tissues > 0 && !tissue_ids) return fail(NMC_ERR_ARG, "null tissue_ids");
  for (int j = 0; j < n_tissues; ++j) {
    if (tissue_ids[j] < 0 || tissue_ids[j] >= NMC_NTISS)
      return fail(NMC_ERR_ARG, "tissue id %d not in 0..%d", tissue_ids[j], NMC_NTISS - 1);
    for (int i = 0; i < j; ++i)
      if (tissue_ids[i] == tissue_ids[j]) return fail(NMC_ERR_ARG, "tissue id %d given twice", tissue_ids[j]);
  }
  HIP_TRY(hipSetDevice(h->device));
  NMC_TRY(flush_pending(h));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const bool tt = h->sim == NMC_SIM_MCXYZ_TT;
  const size_t lds = (size_t)2 * n_tissues * kSensBlock * sizeof(uint32_t);
  int grid = 0;
  double* buf = nullptr;
  if (n_tissues > 0) {
    HIP_TRY(mcxyz_sens_configure(tt, lds));
    const int occ = mcxyz_sens_blocks_per_cu(tt, lds);
    if (occ <= 0) return fail(NMC_ERR_DEVICE, "the sensitivity kernel does not fit on a CU");
    grid = h->num_cu * occ;
    const size_t bytes = 2 * (size_t)n_tissues * (size_t)nvox(&h->rp) * sizeof(double);
    hipError_t e = hipMalloc(&buf, bytes);
    if (e != hipSuccess)
      return fail(e == hipErrorOutOfMemory ? NMC_ERR_NOMEM : NMC_ERR_DEVICE, "sensitivity arrays: %s",
                  hipGetErrorString(e));
    e = hipMemset(buf, 0, bytes);
    if (e != hipSuccess) {
      (void)hipFree(buf);  // the error being reported is the memset's
      return fail(NMC_ERR_DEVICE, "sensitivity arrays: %s", hipGetErrorString(e));
    }
  }
  // the tissue table's slots
  std::vector<XyzTissue> tis(NMC_NTISS);
  XyzAux aux{};
  xyz_tables(h->rp, h->tp, h->sim, tis.data(), &aux);
  for (int j = 0; j < n_tissues; ++j) tis[tissue_ids[j]].slot = j;
  hipError_t e = hipMemcpy(h->dTis, tis.data(), NMC_NTISS * sizeof(XyzTissue), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (buf) (void)hipFree(buf);  // the error being reported is the copy's
    return fail(NMC_ERR_DEVICE, "tissue table: %s", hipGetErrorString(e));
  }
  // commit: the cached graph runs the previous variant
  drop_graph(h);
  free_ptr(h->dSens);
  h->dSens = buf;
  h->sens_n = n_tissues;
  for (int j = 0; j < n_tissues; ++j) h->sens_ids[j] = tissue_ids[j];
  h->sens_lds = lds;
  h->sens_grid = grid;
  return ok();
}

int nmc_mcxyz_sens_results(nmc_handle* h, double* d_mua, double* d_mus) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (!h->prepared || !h->dSens) return fail(NMC_ERR_STATE, "no tissues selected by nmc_mcxyz_sens");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const size_t per = (size_t)h->sens_n * (size_t)nvox(&h->rp);
  if (d_mua) HIP_TRY(hipMemcpy(d_mua, h->dSens, per * sizeof(double), hipMemcpyDeviceToHost));
  if (d_mus) HIP_TRY(hipMemcpy(d_mus, h->dSens + per, per * sizeof(double), hipMemcpyDeviceToHost));
  return ok();
}

int nmc_last_launch_count(nmc_handle* h, uint32_t* n) {
  if (!h || !n) return fail(NMC_ERR_ARG, "null argument");
  if (!h->timed) return fail(NMC_ERR_STATE, "no run has been timed");
  *n = h->launches;
  return ok();
}

int nmc_get_tuning(nmc_handle* h, nmc_tuning* t) {
  if (!h || !t) return fail(NMC_ERR_ARG, "null argument");
  if (!h->prepared) return fail(NMC_ERR_STATE, "nmc_prepare has not succeeded");
  *t = h->tune;
  return ok();
}

int nmc_set_graph_mode(nmc_handle* h, int enable) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->device));
  NMC_TRY(flush_pending(h));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->graph_mode = enable != 0;
  drop_graph(h);
  return ok();
}

int nmc_run_async(nmc_handle* h, uint64_t n_photons) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (!h->prepared) return fail(NMC_ERR_STATE, "nmc_prepare has not succeeded");
  if (n_photons == 0) return ok();
  if (h->sim == NMC_SIM_MCSUB) {
    if (h->batch) return fail(NMC_ERR_STATE, "handle holds a batch; use nmc_run_batch");
    if (h->items[0].n_photons != n_photons) {
      // a run still in flight reads the old budget
      HIP_TRY(hipSetDevice(h->device));
      HIP_TRY(hipStreamSynchronize(h->stream));
      h->items[0].n_photons = n_photons;
      h->prefix[1] = n_photons;
      HIP_TRY(hipMemcpy(h->dPrefix, h->prefix.data(), 2 * sizeof(uint64_t), hipMemcpyHostToDevice));
    }
  }
  NMC_TRY(run_async(h, n_photons));
  h->next_photon += n_photons;
  return ok();
}

int nmc_synchronize(nmc_handle* h) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->device));
  NMC_TRY(flush_pending(h));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return ok();
}

int nmc_run(nmc_handle* h, uint64_t n_photons) {
  NMC_TRY(nmc_run_async(h, n_photons));
  return nmc_synchronize(h);
}

int nmc_run_batch_async(nmc_handle* h) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (!h->prepared || !h->batch) return fail(NMC_ERR_STATE, "nmc_prepare_batch has not succeeded");
  uint64_t adv = 0;
  for (int j = 0; j < h->n_items; ++j) adv = std::max<uint64_t>(adv, h->items[j].n_photons);
  NMC_TRY(run_async(h, h->prefix[h->n_items]));
  h->next_photon += adv;
  return ok();
}

int nmc_run_batch(nmc_handle* h) {
  NMC_TRY(nmc_run_batch_async(h));
  return nmc_synchronize(h);
}

int nmc_last_kernel_ms(nmc_handle* h, float* ms) {
  if (!h || !ms) return fail(NMC_ERR_ARG, "null argument");
  if (!h->timed) return fail(NMC_ERR_STATE, "no run has been timed");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipEventSynchronize(h->e1));
  HIP_TRY(hipEventElapsedTime(ms, h->e0, h->e1));
  return ok();
}

int nmc_get_results(nmc_handle* h, nmc_results* out) {
  if (!h || !out) return fail(NMC_ERR_ARG, "null argument");
  if (!h->prepared) return fail(NMC_ERR_STATE, "nmc_prepare has not succeeded");
  HIP_TRY(hipSetDevice(h->device));
  NMC_TRY(flush_pending(h));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->sim == NMC_SIM_MCSUB) {
    if (h->batch) return fail(NMC_ERR_STATE, "handle holds a batch; use nmc_get_batch_results");
    if (out->rd_ra) HIP_TRY(hipMemcpy(out->rd_ra, h->dOut, h->n_ra * sizeof(double), hipMemcpyDeviceToHost));
    if (out->a_rz) {
      if (!h->n_rz) return fail(NMC_ERR_STATE, "A(r,z) tally disabled by NMC_FLAG_NO_FLUENCE");
      HIP_TRY(hipMemcpy(out->a_rz, h->dOut + h->n_ra, h->n_rz * sizeof(double), hipMemcpyDeviceToHost));
    }
    HIP_TRY(hipMemcpy(&out->totals, h->dOut + per_item(h), sizeof(nmc_totals), hipMemcpyDeviceToHost));
    return ok();
  }
  int64_t n = nvox(&h->rp), n2 = (int64_t)h->rp.Nx * h->rp.Ny;
  if (out->fluence) {
    if (!h->dF) return fail(NMC_ERR_STATE, "fluence tally disabled by NMC_FLAG_NO_FLUENCE");
    HIP_TRY(hipMemcpy(out->fluence, h->dF, n * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (out->ryx) HIP_TRY(hipMemcpy(out->ryx, h->dRyx, n2 * sizeof(double), hipMemcpyDeviceToHost));
  if (out->det2d) HIP_TRY(hipMemcpy(out->det2d, h->dDet, n2 * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&out->totals, h->dTot, sizeof(nmc_totals), hipMemcpyDeviceToHost));
  return ok();
}

int nmc_add_results(nmc_handle* h, nmc_results* out) {
  if (!h || !out) return fail(NMC_ERR_ARG, "null argument");
  if (!h->prepared) return fail(NMC_ERR_STATE, "nmc_prepare has not succeeded");
  if (h->sim == NMC_SIM_MCSUB && h->batch)
    return fail(NMC_ERR_STATE, "handle holds a batch; use nmc_get_batch_results");
  int64_t na = 0, nb = 0, nc = 0;   // element counts of the three array slots
  if (h->sim == NMC_SIM_MCSUB) {
    na = h->n_ra;
    nb = h->n_rz;
  } else {
    na = nvox(&h->rp);
    nb = nc = (int64_t)h->rp.Nx * h->rp.Ny;
  }
  const bool xyz = h->sim != NMC_SIM_MCSUB;
  double* a_out = xyz ? out->fluence : out->rd_ra;
  double* b_out = xyz ? out->ryx : out->a_rz;
  std::vector<double> a(a_out ? na : 0), b(b_out ? nb : 0);
  std::vector<uint32_t> c(xyz && out->det2d ? nc : 0);
  nmc_results tmp{};
  if (xyz) {
    tmp.fluence = a_out ? a.data() : nullptr;
    tmp.ryx = b_out ? b.data() : nullptr;
    tmp.det2d = out->det2d ? c.data() : nullptr;
  } else {
    tmp.rd_ra = a_out ? a.data() : nullptr;
    tmp.a_rz = b_out ? b.data() : nullptr;
  }
  NMC_TRY(nmc_get_results(h, &tmp));
  for (size_t i = 0; i < a.size(); ++i) a_out[i] += a[i];
  for (size_t i = 0; i < b.size(); ++i) b_out[i] += b[i];
  for (size_t i = 0; i < c.size(); ++i) out->det2d[i] += c[i];
  double* tf = (double*)&out->totals;
  const double* sf = (const double*)&tmp.totals;
  for (int k = 0; k < TF_COUNT; ++k) tf[k] += sf[k];
  uint64_t* ti = &out->totals.photons;
  const uint64_t* si = &tmp.totals.photons;
  for (int k = 0; k < TI_COUNT; ++k) ti[k] += si[k];
  return ok();
}

int nmc_get_batch_results(nmc_handle* h, double* tally, nmc_totals* totals) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (!h->prepared || !h->batch) return fail(NMC_ERR_STATE, "nmc_prepare_batch has not succeeded");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  int64_t nt = (int64_t)h->n_items * per_item(h);
  if (tally) HIP_TRY(hipMemcpy(tally, h->dOut, nt * sizeof(double), hipMemcpyDeviceToHost));
  if (totals)
    HIP_TRY(hipMemcpy(totals, h->dOut + nt, h->n_items * sizeof(nmc_totals), hipMemcpyDeviceToHost));
  return ok();
}

// ---------------------------------------------------------------------------
// Layered media
// ---------------------------------------------------------------------------
static int validate_stacks(const nmc_layered_config* cfg, const nmc_stack* st, int n) {
  if (!cfg || !st) return fail(NMC_ERR_ARG, "config and stacks are required");
  if (cfg->struct_size != sizeof(nmc_layered_config) || cfg->version != NMC_EXT_VERSION)
    return fail(NMC_ERR_ARG, "nmc_layered_config size/version mismatch");
  if (n <= 0 || n > kMaxBatchItems) return fail(NMC_ERR_ARG, "n_stacks must be in 1..%d", kMaxBatchItems);
  if (cfg->nr <= 0 || !(cfg->dr > 0.0f)) return fail(NMC_ERR_ARG, "Rd bins invalid (nr >= 1, dr > 0)");
  if (!(cfg->threshold >= 0.0f) || !(cfg->chance > 0.0f && cfg->chance <= 1.0f) || cfg->max_events == 0)
    return fail(NMC_ERR_ARG, "roulette / max_events settings invalid");
  for (int j = 0; j < n; ++j) {
    const nmc_stack& s = st[j];
    if (s.n_layers < 1 || s.n_layers > NMC_MAX_LAYERS)
      return fail(NMC_ERR_ARG, "stack %d: n_layers %d not in 1..%d", j, s.n_layers, NMC_MAX_LAYERS);
    if (!(s.n_above > 0.0f) || !(s.n_below > 0.0f)) return fail(NMC_ERR_ARG, "stack %d: ambient index must be > 0", j);
    for (int l = 0; l < s.n_layers; ++l) {
      const nmc_layer& L = s.layer[l];
      bool last = l == s.n_layers - 1;
      if (!last && !(L.thickness > 0.0f && std::isfinite(L.thickness)))
        return fail(NMC_ERR_ARG, "stack %d layer %d: thickness must be > 0 (only the last may be semi-infinite)", j, l);
      if (!finite_nonneg(L.mua) || !finite_nonneg(L.mus) || !(L.n > 0.0f) || !std::isfinite(L.n))
        return fail(NMC_ERR_ARG, "stack %d layer %d: mua, mus >= 0 and n > 0 required", j, l);
      if (L.phase == 0) {
        if (!(L.g >= -1.0f && L.g <= 1.0f)) return fail(NMC_ERR_ARG, "stack %d layer %d: g not in [-1,1]", j, l);
      } else if (L.phase == 1) {
        if (!(L.cc >= 0.0f && L.cc <= 1.0f) || !(L.gf >= 0.0f && L.gf < 1.0f) || !(L.gb >= 0.0f && L.gb < 1.0f) ||
            (L.gf != 0.0f && !(L.af > 0.0f)) || (L.gb != 0.0f && !(L.ab > 0.0f)))
          return fail(NMC_ERR_ARG, "stack %d layer %d: two-term parameters out of range", j, l);
      } else {
        return fail(NMC_ERR_ARG, "stack %d layer %d: phase %d not 0 or 1", j, l, L.phase);
      }
    }
    if (last_layer_semi_infinite_needs_scatter(s))
      return fail(NMC_ERR_ARG, "stack %d: a semi-infinite last layer needs mua > 0 or mus > 0", j);
  }
  return ok();
}

// Per-stack header and per-layer rows (interface depths summed in double,
// rounded once to float).
static void build_tables(nmc_handle* h, std::vector<LayStackDev>& hdr, std::vector<LayRow>& rows,
                         std::vector<float>& mus, std::vector<int>& nl) {
  int n = (int)h->stacks.size();
  hdr.assign(n, LayStackDev{});
  rows.clear();
  mus.assign((size_t)n * NMC_MAX_LAYERS, 0.0f);
  nl.assign(n, 0);
  for (int j = 0; j < n; ++j) {
    const nmc_stack& s = h->stacks[j];
    LayStackDev& H = hdr[j];
    H.nl = s.n_layers;
    H.row0 = (int)rows.size();
    H.n_above = s.n_above;
    H.n_below = s.n_below;
    H.key1 = key1_of(h, s.stream_id);
    // normal-incidence specular reflection at the top (fresnel() ca1 = 1 branch)
    float n0 = s.layer[0].n;
    float d = (s.n_above + n0), r = (n0 - s.n_above) / d;
    float e = n0 * 0x1p-23f * 2.0f;
    H.rsp = (s.n_above >= n0 - e && s.n_above <= n0 + e) ? 0.0f : r * r;
    double z = 0.0;
    for (int l = 0; l < s.n_layers; ++l) {
      const nmc_layer& L = s.layer[l];
      LayRow R{};
      R.zt = (float)z;
      bool last = l == s.n_layers - 1;
      if (last && !(L.thickness > 0.0f && std::isfinite(L.thickness))) {
        R.zb = std::numeric_limits<float>::infinity();
      } else {
        z += (double)L.thickness;
        R.zb = (float)z;
      }
      R.mua = L.mua;
      R.mus = L.mus;
      R.inv_mus = inv_pos(L.mus);
      R.n = L.n;
      R.hg = hg_prepare(L.g);
      R.phase = L.phase;
      if (L.phase) tt_prepare(L.gf, L.af, L.gb, L.ab, L.cc, &R.tt);
      rows.push_back(R);
      mus[(size_t)j * NMC_MAX_LAYERS + l] = L.mus;
    }
    nl[j] = s.n_layers;
  }
}

static int upload_stacks(nmc_handle* h) {
  std::vector<LayStackDev> hdr;
  std::vector<LayRow> rows;
  std::vector<float> mus;
  std::vector<int> nl;
  build_tables(h, hdr, rows, mus, nl);
  int n = (int)h->stacks.size();
  std::vector<uint64_t> prefix(n + 1, 0);
  std::vector<double> inv_n(n);
  for (int j = 0; j < n; ++j) {
    prefix[j + 1] = prefix[j] + h->stacks[j].n_photons;
    inv_n[j] = h->stacks[j].n_photons ? 1.0 / (double)h->stacks[j].n_photons : 0.0;
  }
  h->prefix = prefix;
  HIP_TRY(hipMemcpy(h->dHdr, hdr.data(), n * sizeof(LayStackDev), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->dRows, rows.data(), rows.size() * sizeof(LayRow), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->dLPrefix, prefix.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->dMus, mus.data(), mus.size() * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->dNl, nl.data(), n * sizeof(int), hipMemcpyHostToDevice));
  double* inv = h->dGrad + (size_t)n * (h->lcfg.nr + 1);
  HIP_TRY(hipMemcpy(inv, inv_n.data(), n * sizeof(double), hipMemcpyHostToDevice));
  return ok();
}

int nmc_layered_prepare(nmc_handle* h, const nmc_layered_config* cfg, const nmc_stack* stacks,
                        int n_stacks) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (h->sim != NMC_SIM_LAYERED) return fail(NMC_ERR_UNSUPPORTED, "handle is not NMC_SIM_LAYERED");
  NMC_TRY(validate_stacks(cfg, stacks, n_stacks));
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  free_buffers(h);
  h->lcfg = *cfg;
  h->stacks.assign(stacks, stacks + n_stacks);
  h->ext.seed = cfg->seed;
  h->next_photon = 0;
  h->timed = false;
  int n = n_stacks;
  h->lay_rows = 0;
  h->rec_nl = 0;
  uint64_t total = 0;
  for (int j = 0; j < n; ++j) {
    h->lay_rows += stacks[j].n_layers;
    h->rec_nl = std::max(h->rec_nl, (int)stacks[j].n_layers);
    total += stacks[j].n_photons;
  }
  HIP_TRY(hipMalloc(&h->dHdr, n * sizeof(LayStackDev)));
  HIP_TRY(hipMalloc(&h->dRows, h->lay_rows * sizeof(LayRow)));
  HIP_TRY(hipMalloc(&h->dLPrefix, (n + 1) * sizeof(uint64_t)));
  HIP_TRY(hipMalloc(&h->dLRd, (size_t)n * (cfg->nr + 1) * sizeof(double)));
  HIP_TRY(hipMalloc(&h->dLOut, (size_t)n * (NMC_MAX_LAYERS + TF_COUNT + TI_COUNT) * sizeof(double)));
  HIP_TRY(hipMalloc(&h->dMus, (size_t)n * NMC_MAX_LAYERS * sizeof(float)));
  HIP_TRY(hipMalloc(&h->dNl, n * sizeof(int)));
  HIP_TRY(hipMalloc(&h->dGrad, ((size_t)n * (cfg->nr + 1) + n + 2 * (size_t)n * NMC_MAX_LAYERS) * sizeof(double)));
  HIP_TRY(hipMalloc(&h->dRecCount, 2 * sizeof(unsigned long long)));
  h->rec_cap = 0;
  h->rec_valid = false;
  const bool sens = cfg->absorbed_grad != 0;
  h->sens_valid = false;
  h->lay_lds = layered_lds_bytes(n, h->lay_rows, 0, sens);
  if (h->lay_lds + kLayStaticLds > h->lds_max)
    return fail(NMC_ERR_ARG, "%d stacks / %d layers need %zu B of LDS (%zu available)", n, h->lay_rows,
                h->lay_lds, h->lds_max);
  // Rd(r) goes to an LDS histogram when it fits without costing occupancy
  // (tuning.mode 0 or 1), else, or with tuning.mode 2, to global replicas
  if (h->tune_req.mode < 0 || h->tune_req.mode > 2)
    return fail(NMC_ERR_ARG, "tuning.mode (Rd placement) not in 0..2");
  const int rd_len = n * (cfg->nr + 1);
  const size_t lds_rd = layered_lds_bytes(n, h->lay_rows, rd_len, sens);
  const bool rd_fits = lds_rd + kLayStaticLds <= h->lds_max;
  HIP_TRY(layered_configure(rd_fits ? lds_rd : h->lay_lds));
  int occ = layered_blocks_per_cu(h->lay_lds);
  if (occ <= 0) return fail(NMC_ERR_ARG, "layered kernel does not fit on a CU");
  h->lay_rd_lds = h->tune_req.mode != 2 && rd_fits && layered_blocks_per_cu(lds_rd) >= occ;
  if (h->lay_rd_lds) h->lay_lds = lds_rd;
  if (h->tune_req.blocks_per_cu > 0) occ = std::min(occ, h->tune_req.blocks_per_cu);
  h->lay_grid = h->num_cu * occ;
  h->tune = nmc_tuning{};
  h->tune.block = kLayBlock;
  h->tune.blocks_per_cu = occ;
  h->tune.refill_min = h->tune_req.refill_min ? h->tune_req.refill_min : 16;
  h->tune.chunk = h->tune_req.chunk;
  h->tune.mode = h->lay_rd_lds ? 1 : 2;
  HIP_TRY(hipMalloc(&h->dLPart, (size_t)h->lay_grid * n * (NMC_MAX_LAYERS + TF_COUNT + TI_COUNT) * sizeof(double)));
  h->lay_rd_count = h->lay_rd_lds ? h->lay_grid : replica_count(rd_len, h->lay_grid);
  HIP_TRY(hipMalloc(&h->dLRdPart, (size_t)h->lay_rd_count * rd_len * sizeof(double)));
  if (sens) {
    HIP_TRY(hipMalloc(&h->dLSens, (size_t)n * kSensPerStack * sizeof(double)));
    HIP_TRY(hipMalloc(&h->dLSensPart, (size_t)h->lay_grid * n * kSensPerStack * sizeof(double)));
  }
  HIP_TRY(hipMemsetAsync(h->dLRdPart, 0, (size_t)h->lay_rd_count * rd_len * sizeof(double), h->stream));
  if (cfg->records) {
    // every photon leaves at most one record; each wave may leave the unused
    // tail of its last slot run (< kRecChunk slots) as holes
    uint64_t waves = (uint64_t)h->lay_grid * (kLayBlock / 64);
    h->rec_cap = cfg->record_capacity ? cfg->record_capacity : total + waves * kRecChunk;
    if (h->rec_cap == 0) h->rec_cap = 1;
    HIP_TRY(hipMalloc(&h->dRec, h->rec_cap * (3 + 2 * h->rec_nl) * sizeof(uint32_t)));
  }
  NMC_TRY(upload_stacks(h));
  h->prepared = true;
  return write_first(h);
}

int nmc_layered_update(nmc_handle* h, const nmc_stack* stacks, int n_stacks) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (h->sim != NMC_SIM_LAYERED || !h->prepared) return fail(NMC_ERR_STATE, "nmc_layered_prepare has not succeeded");
  if (n_stacks != (int)h->stacks.size()) return fail(NMC_ERR_ARG, "update must keep the number of stacks");
  NMC_TRY(validate_stacks(&h->lcfg, stacks, n_stacks));
  uint64_t total = 0;
  for (int j = 0; j < n_stacks; ++j) {
    if (stacks[j].n_layers != h->stacks[j].n_layers) return fail(NMC_ERR_ARG, "update must keep each stack's layer count");
    total += stacks[j].n_photons;
  }
  const uint64_t holes = (uint64_t)h->lay_grid * (kLayBlock / 64) * kRecChunk;
  if (h->lcfg.records && !h->lcfg.record_capacity && total + holes > h->rec_cap)
    return fail(NMC_ERR_ARG, "budgets exceed the record capacity set at prepare");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->stacks.assign(stacks, stacks + n_stacks);
  h->rec_valid = false;
  h->sens_valid = false;
  return upload_stacks(h);
}

int nmc_layered_run(nmc_handle* h) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (h->sim != NMC_SIM_LAYERED || !h->prepared) return fail(NMC_ERR_STATE, "nmc_layered_prepare has not succeeded");
  HIP_TRY(hipSetDevice(h->device));
  h->rec_valid = false;
  int n = (int)h->stacks.size();
  hipStream_t s = h->stream;
  const int rd_len = n * (h->lcfg.nr + 1);
  HIP_TRY(hipMemsetAsync(h->dLRd, 0, (size_t)rd_len * sizeof(double), s));
  HIP_TRY(hipMemsetAsync(h->dLOut, 0, (size_t)n * (NMC_MAX_LAYERS + TF_COUNT + TI_COUNT) * sizeof(double), s));
  HIP_TRY(hipMemsetAsync(h->dRecCount, 0, 2 * sizeof(unsigned long long), s));
  h->sens_valid = false;
  if (h->dLSens) HIP_TRY(hipMemsetAsync(h->dLSens, 0, (size_t)n * kSensPerStack * sizeof(double), s));
  HIP_TRY(hipEventRecord(h->e0, s));
  HIP_TRY(launch_state_begin(h->dState, s));
  LayArgs a{};
  a.n_items = n;
  a.n_rows = h->lay_rows;
  a.nr = h->lcfg.nr;
  a.dr = h->lcfg.dr;
  a.thr = h->lcfg.threshold;
  a.chance = h->lcfg.chance;
  a.max_events = h->lcfg.max_events;
  a.key0 = key0_of(h);
  uint64_t total = h->prefix[n];
  a.chunk = pick_chunk(total, h->lay_grid * (kLayBlock / 64), h->tune.chunk);
  a.refill_min = (uint32_t)h->tune.refill_min;
  a.stacks = h->dHdr;
  a.rows = h->dRows;
  a.prefix = h->dLPrefix;
  a.rd = h->dLRdPart;
  a.rd_lds = h->lay_rd_lds;
  a.rd_reps = h->lay_rd_lds ? 1 : h->lay_rd_count;
  a.part = h->dLPart;
  a.records = h->lcfg.records != 0;
  a.rec_nl = h->rec_nl;
  a.cap = h->rec_cap;
  a.rec_count = h->dRecCount;
  a.rec = h->dRec;
  a.state = h->dState;
  a.sens = h->dLSens != nullptr;
  a.sens_part = h->dLSensPart;
  HIP_TRY(launch_layered(a, h->lay_grid, h->lay_lds, s));
  h->launches = 1;
  HIP_TRY(launch_reduce_partials(h->dLPart, h->lay_grid, (int64_t)n * (NMC_MAX_LAYERS + TF_COUNT + TI_COUNT),
                                 h->dLOut, n, NMC_MAX_LAYERS, s));
  HIP_TRY(launch_reduce_partials(h->dLRdPart, h->lay_rd_count, rd_len, h->dLRd, n, h->lcfg.nr + 1, s,
                                 !h->lay_rd_lds));
  if (h->dLSens)
    HIP_TRY(launch_reduce_partials(h->dLSensPart, h->lay_grid, (int64_t)n * kSensPerStack, h->dLSens, n,
                                   kSensPerStack, s));
  HIP_TRY(hipEventRecord(h->e1, s));
  h->timed = true;
  HIP_TRY(hipStreamSynchronize(s));
  h->rec_valid = true;
  h->sens_valid = h->dLSens != nullptr;
  return ok();
}

int nmc_layered_absorbed_grad(nmc_handle* h, double* d_mua, double* d_mus) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (h->sim != NMC_SIM_LAYERED || !h->prepared) return fail(NMC_ERR_STATE, "nmc_layered_prepare has not succeeded");
  if (!h->dLSens) return fail(NMC_ERR_STATE, "absorbed_grad was not enabled at prepare");
  if (!h->sens_valid) return fail(NMC_ERR_STATE, "no nmc_layered_run since the last change of properties");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  const size_t n = h->stacks.size(), per = (size_t)NMC_MAX_LAYERS * NMC_MAX_LAYERS;
  std::vector<double> host(n * kSensPerStack);
  HIP_TRY(hipMemcpy(host.data(), h->dLSens, host.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t j = 0; j < n; ++j) {
    if (d_mua) std::copy(&host[j * kSensPerStack], &host[j * kSensPerStack] + per, d_mua + j * per);
    if (d_mus) std::copy(&host[j * kSensPerStack + per], &host[j * kSensPerStack] + 2 * per, d_mus + j * per);
  }
  return ok();
}

int nmc_layered_results(nmc_handle* h, double* rd, double* absorbed, nmc_totals* totals,
                        uint64_t* n_records) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (h->sim != NMC_SIM_LAYERED || !h->prepared) return fail(NMC_ERR_STATE, "nmc_layered_prepare has not succeeded");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  size_t n = h->stacks.size();
  if (rd) HIP_TRY(hipMemcpy(rd, h->dLRd, n * (h->lcfg.nr + 1) * sizeof(double), hipMemcpyDeviceToHost));
  if (absorbed) HIP_TRY(hipMemcpy(absorbed, h->dLOut, n * NMC_MAX_LAYERS * sizeof(double), hipMemcpyDeviceToHost));
  if (totals)
    HIP_TRY(hipMemcpy(totals, h->dLOut + n * NMC_MAX_LAYERS, n * sizeof(nmc_totals), hipMemcpyDeviceToHost));
  if (n_records) {
    unsigned long long c[2] = {0, 0};
    HIP_TRY(hipMemcpy(c, h->dRecCount, sizeof c, hipMemcpyDeviceToHost));
    *n_records = c[1];
  }
  return ok();
}

int nmc_layered_records(nmc_handle* h, uint32_t* out, uint64_t max_records, uint64_t* n_out,
                        int* rec_nl_out) {
  if (!h || !n_out) return fail(NMC_ERR_ARG, "null argument");
  if (h->sim != NMC_SIM_LAYERED || !h->prepared) return fail(NMC_ERR_STATE, "nmc_layered_prepare has not succeeded");
  if (!h->lcfg.records) return fail(NMC_ERR_STATE, "records were not enabled at prepare");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  unsigned long long cnt[2] = {0, 0};
  HIP_TRY(hipMemcpy(cnt, h->dRecCount, sizeof cnt, hipMemcpyDeviceToHost));
  const uint64_t slots = std::min<uint64_t>(cnt[0], h->rec_cap);
  const size_t rs = 3 + 2 * (size_t)h->rec_nl;
  if (rec_nl_out) *rec_nl_out = h->rec_nl;
  *n_out = 0;
  if (!out || max_records == 0 || slots == 0) return ok();
  std::vector<uint32_t> host(slots * rs);
  HIP_TRY(hipMemcpy(host.data(), h->dRec, host.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  uint64_t n = 0;
  for (uint64_t k = 0; k < slots && n < max_records; ++k) {
    const uint32_t* r = &host[k * rs];
    if (r[0] == kRecHole) continue;
    std::copy(r, r + rs, out + n * rs);
    ++n;
  }
  *n_out = n;
  return ok();
}

int nmc_layered_backward(nmc_handle* h, const double* grad_rd, double* grad_mua, double* grad_mus) {
  if (!h || !grad_rd) return fail(NMC_ERR_ARG, "null argument");
  if (h->sim != NMC_SIM_LAYERED || !h->prepared) return fail(NMC_ERR_STATE, "nmc_layered_prepare has not succeeded");
  if (!h->lcfg.records) return fail(NMC_ERR_STATE, "records were not enabled at prepare");
  if (!h->rec_valid)
    return fail(NMC_ERR_STATE, "no path records for the current properties: nmc_layered_run first");
  HIP_TRY(hipSetDevice(h->device));
  hipStream_t s = h->stream;
  size_t n = h->stacks.size(), nb = n * (h->lcfg.nr + 1);
  unsigned long long cnt[2] = {0, 0};   // slots handed out, records written
  HIP_TRY(hipMemcpy(cnt, h->dRecCount, sizeof cnt, hipMemcpyDeviceToHost));
  const unsigned long long nrec = cnt[0];
  if (nrec > h->rec_cap)
    return fail(NMC_ERR_NOMEM, "%llu record slots exceed the capacity %llu; raise record_capacity",
                (unsigned long long)nrec, (unsigned long long)h->rec_cap);
  double* d_grad_rd = h->dGrad;
  double* d_inv = h->dGrad + nb;
  double* d_gmua = d_inv + n;
  double* d_gmus = d_gmua + n * NMC_MAX_LAYERS;
  HIP_TRY(hipMemcpyAsync(d_grad_rd, grad_rd, nb * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemsetAsync(d_gmua, 0, 2 * n * NMC_MAX_LAYERS * sizeof(double), s));
  LayBackArgs b{};
  b.n_items = (int)n;
  b.nr = h->lcfg.nr;
  b.n_rec = nrec;
  b.cap = h->rec_cap;
  b.rec_nl = h->rec_nl;
  b.rec = h->dRec;
  b.grad_rd = d_grad_rd;
  b.mus = h->dMus;
  b.nl = h->dNl;
  b.inv_n = d_inv;
  b.grad_mua = d_gmua;
  b.grad_mus = d_gmus;
  HIP_TRY(hipEventRecord(h->e0, s));
  HIP_TRY(launch_layered_backward(b, s));
  HIP_TRY(hipEventRecord(h->e1, s));
  if (grad_mua) HIP_TRY(hipMemcpyAsync(grad_mua, d_gmua, n * NMC_MAX_LAYERS * sizeof(double), hipMemcpyDeviceToHost, s));
  if (grad_mus) HIP_TRY(hipMemcpyAsync(grad_mus, d_gmus, n * NMC_MAX_LAYERS * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return ok();
}

void nmc_destroy(nmc_handle* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_buffers(h);
  free_ptr(h->dTot);
  free_ptr(h->dState);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->k0) (void)hipEventDestroy(h->k0);
  if (h->k1) (void)hipEventDestroy(h->k1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free mcxyz run (nmc_mcxyz_run_noalloc).  Its scratch lives in a
// static device array of fixed size, one copy per GPU, held until the process
// ends:
//   [0, 16)        RunState (queue counter, first photon id)
//   [64, 96)       bitmask of the tissue ids the voxel map uses (256 bits)
//   [256, 1024)    XyzAux (source and detector block)
//   [1024, 4096)   XyzTissue[NMC_NTISS]
//   [4096, end)    FP64 block partials of the LDS tile (grid x tile voxels)
// In host code `scratch` names a host-side stand-in, never the device memory:
// its device address comes from hipGetSymbolAddress, once per device.
// ---------------------------------------------------------------------------
constexpr size_t kNoallocHead = 4096;
constexpr size_t kNoallocStateOff = 0, kNoallocMaskOff = 64, kNoallocAuxOff = 256,
                 kNoallocTisOff = 1024;
static_assert(NMC_NOALLOC_SCRATCH_BYTES > kNoallocHead, "scratch holds the head block");
static_assert(sizeof(RunState) <= kNoallocMaskOff, "RunState fits its slot");
static_assert(kNoallocAuxOff + sizeof(XyzAux) <= kNoallocTisOff, "XyzAux fits its slot");
static_assert(kNoallocTisOff + NMC_NTISS * sizeof(XyzTissue) <= kNoallocHead, "tissue table fits");
static_assert(NMC_NOALLOC_MAX_PARTIALS * sizeof(double) == NMC_NOALLOC_SCRATCH_BYTES - kNoallocHead,
              "partials fill the rest");

static __device__ __align__(256) unsigned char scratch[NMC_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
unsigned char* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

// Device address of `scratch` on device `dev`, which must be current: looked
// up on the first call for that device and kept.
int scratch_on(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(NMC_ERR_DEVICE, "device %d out of range", dev);
  std::call_once(g_scratch_once[dev], [dev] {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? (unsigned char*)p : nullptr;
  });
  if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
    return fail(NMC_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                hipGetErrorString(g_scratch_err[dev]));
  *out = g_scratch_ptr[dev];
  return ok();
}

__global__ void noalloc_mask_clear(uint32_t* mask) {
  if (threadIdx.x < 8) mask[threadIdx.x] = 0u;
}

// mask |= bit v[i] for every voxel (per-workgroup LDS mask first; a lane adds
// a bit only when its tissue id differs from the one it added last).
__global__ void __launch_bounds__(256) noalloc_mask_scan(const uint8_t* __restrict__ v, uint64_t n,
                                                         uint32_t* mask) {
  __shared__ uint32_t m[8];
  if (threadIdx.x < 8) m[threadIdx.x] = 0u;
  __syncthreads();
  uint32_t last = 256u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t t = v[i];
    if (t != last) {
      atomicOr(&m[t >> 5], 1u << (t & 31u));
      last = t;
    }
  }
  __syncthreads();
  if (threadIdx.x < 8 && m[threadIdx.x]) atomicOr(&mask[threadIdx.x], m[threadIdx.x]);
}

struct NoallocTables {
  XyzTissue tis[NMC_NTISS];
  XyzAux aux;
  RunState state;
};

// Writes the run's tables (passed by value) into the scratch array.
__global__ void noalloc_init(NoallocTables in, XyzTissue* tis, XyzAux* aux, RunState* st) {
  if (threadIdx.x < NMC_NTISS) tis[threadIdx.x] = in.tis[threadIdx.x];
  if (threadIdx.x == 0) {
    *aux = in.aux;
    *st = in.state;
  }
}

}  // namespace

extern "C" int nmc_mcxyz_run_noalloc(int sim_type, const RunParams* rp, const TissueParams* tp,
                                     const nmc_ext_config* ext, const nmc_tuning* tuning,
                                     const uint8_t* d_v, uint64_t first_photon, uint64_t n_photons,
                                     double* d_fluence, double* d_ryx, uint32_t* d_det2d,
                                     nmc_totals* d_totals, void* stream) {
  if (sim_type != NMC_SIM_MCXYZ && sim_type != NMC_SIM_MCXYZ_TT)
    return fail(NMC_ERR_UNSUPPORTED, "nmc_mcxyz_run_noalloc runs NMC_SIM_MCXYZ or NMC_SIM_MCXYZ_TT");
  NMC_TRY(validate_impl(sim_type, rp, tp, nullptr, ext, false));
  if (!d_v || !d_ryx || !d_det2d || !d_totals)
    return fail(NMC_ERR_ARG, "d_v, d_ryx, d_det2d and d_totals are required");
  nmc_ext_config e;
  ext_defaults(&e);
  if (ext) e = *ext;
  const bool fluence = !(e.flags & NMC_FLAG_NO_FLUENCE) && d_fluence != nullptr;
  int dev = 0, num_cu = 0, lds = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  unsigned char* base = nullptr;
  NMC_TRY(scratch_on(dev, &base));
  nmc_tuning req{};
  if (tuning) req = *tuning;
  XyzLaunch L{};
  NMC_TRY(resolve_xyz_launch(*rp, sim_type, fluence, req, (size_t)lds, num_cu,
                             (int64_t)NMC_NOALLOC_MAX_PARTIALS, &L));
  hipStream_t s = (hipStream_t)stream;
  // the tissue ids the device-resident voxel map uses, checked on the host
  uint32_t* dmask = (uint32_t*)(base + kNoallocMaskOff);
  const uint64_t nv = (uint64_t)nvox(rp);
  hipLaunchKernelGGL(noalloc_mask_clear, dim3(1), dim3(64), 0, s, dmask);
  HIP_TRY(hipGetLastError());
  const uint64_t scan_blocks = std::min<uint64_t>((nv + 255) / 256, 4096);
  hipLaunchKernelGGL(noalloc_mask_scan, dim3((unsigned)scan_blocks), dim3(256), 0, s, d_v, nv, dmask);
  HIP_TRY(hipGetLastError());
  uint32_t mask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(mask, dmask, sizeof mask, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  bool used[256];
  for (int t = 0; t < 256; ++t) used[t] = ((mask[t >> 5] >> (t & 31)) & 1u) != 0u;
  NMC_TRY(check_tissues(sim_type, tp, used));
  if (n_photons == 0) return ok();
  NoallocTables tab;
  std::memset(&tab, 0, sizeof tab);
  xyz_tables(*rp, *tp, sim_type, tab.tis, &tab.aux);
  tab.aux.ryx = d_ryx;
  tab.aux.det2d = d_det2d;
  tab.state.counter = 0ull;
  tab.state.first = (unsigned long long)first_photon;
  XyzTissue* dtis = (XyzTissue*)(base + kNoallocTisOff);
  XyzAux* daux = (XyzAux*)(base + kNoallocAuxOff);
  RunState* dstate = (RunState*)(base + kNoallocStateOff);
  hipLaunchKernelGGL(noalloc_init, dim3(1), dim3(64), 0, s, tab, dtis, daux, dstate);
  HIP_TRY(hipGetLastError());
  XyzArgs a = xyz_args(*rp, e, sim_type, (uint32_t)e.seed, (uint32_t)(e.seed >> 32) + e.stream_id);
  a.aux = daux;
  a.V = d_v;
  a.tissue = dtis;
  a.F = fluence ? d_fluence : nullptr;
  a.tile_part = (double*)(base + kNoallocHead);
  a.tot = d_totals;
  a.state = dstate;
  uint32_t launches = 0;
  NMC_TRY(enqueue_xyz(a, sim_type == NMC_SIM_MCXYZ_TT, L, n_photons, s, &launches, nullptr,
                      nullptr, nullptr, false));
  return ok();
}
