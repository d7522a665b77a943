// nmc_abi.hip - host runtime behind include/nmc.h.
//
// Plays the role of the reference's MetalMC object (NMC/NMCCode/MetalMC.m):
// owns the device buffers, validates and uploads the configuration, launches
// the transport kernels on its own HIP stream and returns the tallies.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
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

}  // namespace

struct nmc_handle {
  int sim = 0, device = 0, num_cu = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool prepared = false, timed = false;
  RunParams rp{};
  TissueParams tp{};
  nmc_ext_config ext{};
  uint64_t next_photon = 0;
  // device buffers
  uint8_t* dV = nullptr;
  XyzTissue* dTis = nullptr;
  double* dF = nullptr;
  double* dRyx = nullptr;
  uint32_t* dDet = nullptr;
  double* dRa = nullptr;
  double* dRz = nullptr;
  double* dPart = nullptr;
  float* dCos = nullptr;
  nmc_totals* dTot = nullptr;
  unsigned long long* dCounter = nullptr;
  int sub_grid = 0, lds_bytes = 0;
  int64_t n_ra = 0, n_rz = 0;
};

namespace {

void free_buffers(nmc_handle* h) {
  void* bufs[] = {h->dV, h->dTis, h->dF, h->dRyx, h->dDet, h->dRa, h->dRz, h->dPart, h->dCos,
                  h->dTot, h->dCounter};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  h->dV = nullptr;
  h->dTis = nullptr;
  h->dF = h->dRyx = h->dRa = h->dRz = h->dPart = nullptr;
  h->dDet = nullptr;
  h->dCos = nullptr;
  h->dTot = nullptr;
  h->dCounter = nullptr;
  h->prepared = false;
}

int64_t nvox(const RunParams* rp) { return (int64_t)rp->Nx * rp->Ny * rp->Nz; }

}  // namespace

extern "C" {

void nmc_ext_config_init(nmc_ext_config* cfg) {
  if (cfg) ext_defaults(cfg);
}

const char* nmc_last_error(void) { return g_err.c_str(); }

const char* nmc_version(void) {
  return "nmc_hip 1 gfx950 -O3 -ffp-contract=off wave64 philox4x32-10";
}

int nmc_validate(int sim, const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                 const nmc_ext_config* ext) {
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
  if (!v) return fail(NMC_ERR_ARG, "voxel map is required for mcxyz");
  bool used[256] = {false};
  int64_t n = nvox(rp);
  for (int64_t i = 0; i < n; ++i) used[v[i]] = true;
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

int nmc_create(int sim_type, int device, nmc_handle** out) {
  if (!out) return fail(NMC_ERR_ARG, "out is null");
  *out = nullptr;
  if (sim_type != NMC_SIM_MCSUB && sim_type != NMC_SIM_MCXYZ && sim_type != NMC_SIM_MCXYZ_TT)
    return fail(NMC_ERR_UNSUPPORTED, "unsupported simulation kind %d", sim_type);
  int count = 0;
  HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(NMC_ERR_ARG, "device %d not present (%d devices)", device, count);
  HIP_TRY(hipSetDevice(device));
  nmc_handle* h = new nmc_handle();
  h->sim = sim_type;
  h->device = device;
  hipError_t e = hipDeviceGetAttribute(&h->num_cu, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&h->e0);
  if (e == hipSuccess) e = hipEventCreate(&h->e1);
  if (e == hipSuccess) e = hipMalloc(&h->dTot, sizeof(nmc_totals));
  if (e == hipSuccess) e = hipMalloc(&h->dCounter, sizeof(unsigned long long));
  if (e != hipSuccess) {
    nmc_destroy(h);
    return fail(NMC_ERR_DEVICE, "device setup failed: %s", hipGetErrorString(e));
  }
  *out = h;
  return ok();
}

int nmc_prepare(nmc_handle* h, const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                const nmc_ext_config* ext) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  int st = nmc_validate(h->sim, rp, tp, v, ext);
  if (st != NMC_OK) return st;
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  // keep dTot / dCounter across prepares
  nmc_totals* tot = h->dTot;
  unsigned long long* ctr = h->dCounter;
  h->dTot = nullptr;
  h->dCounter = nullptr;
  free_buffers(h);
  h->dTot = tot;
  h->dCounter = ctr;
  h->rp = *rp;
  h->tp = *tp;
  ext_defaults(&h->ext);
  if (ext) h->ext = *ext;
  h->next_photon = 0;
  h->timed = false;

  if (h->sim == NMC_SIM_MCSUB) {
    const nmc_ext_config& e = h->ext;
    h->n_ra = (int64_t)(e.nr + 1) * e.na;
    h->n_rz = (e.flags & NMC_FLAG_NO_FLUENCE) ? 0 : (int64_t)(e.nr + 1) * (e.nz + 1);
    HIP_TRY(hipMalloc(&h->dRa, h->n_ra * sizeof(double)));
    if (h->n_rz) HIP_TRY(hipMalloc(&h->dRz, h->n_rz * sizeof(double)));
    std::vector<float> cos_tab(64, 0.0f);
    for (int j = 0; j < e.na; ++j)
      cos_tab[j] = (float)std::cos((double)(j + 1) * (3.14159265358979323846 / 2.0) / (double)e.na);
    HIP_TRY(hipMalloc(&h->dCos, 64 * sizeof(float)));
    HIP_TRY(hipMemcpy(h->dCos, cos_tab.data(), 64 * sizeof(float), hipMemcpyHostToDevice));
    int lds_max = 0;
    HIP_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
    int64_t bytes = (h->n_ra + h->n_rz) * (int64_t)sizeof(double);
    // keep room for the static LDS (totals, cosine table)
    h->lds_bytes = bytes + 2048 <= lds_max ? (int)bytes : 0;
    if (h->lds_bytes) {
      h->sub_grid = h->num_cu * mcsub_blocks_per_cu(h->lds_bytes);
      HIP_TRY(hipMalloc(&h->dPart, (size_t)h->sub_grid * (h->n_ra + h->n_rz) * sizeof(double)));
    } else {
      h->sub_grid = h->num_cu * mcsub_blocks_per_cu(0);
    }
  } else {
    int64_t n = nvox(rp), n2 = (int64_t)rp->Nx * rp->Ny;
    HIP_TRY(hipMalloc(&h->dV, n));
    HIP_TRY(hipMemcpy(h->dV, v, n, hipMemcpyHostToDevice));
    if (!(h->ext.flags & NMC_FLAG_NO_FLUENCE)) HIP_TRY(hipMalloc(&h->dF, n * sizeof(double)));
    HIP_TRY(hipMalloc(&h->dRyx, n2 * sizeof(double)));
    HIP_TRY(hipMalloc(&h->dDet, n2 * sizeof(uint32_t)));
    std::vector<XyzTissue> tis(NMC_NTISS);
    for (int t = 0; t < NMC_NTISS; ++t) {
      std::memset(&tis[t], 0, sizeof(XyzTissue));
      tis[t].mua = tp->muav[t];
      tis[t].mus = tp->musv[t];
      tis[t].g = tp->gv[t];
      if (h->sim == NMC_SIM_MCXYZ_TT) tt_prepare(tp->gf[t], tp->af[t], tp->gb[t], tp->ab[t], tp->CC[t], &tis[t].tt);
    }
    HIP_TRY(hipMalloc(&h->dTis, NMC_NTISS * sizeof(XyzTissue)));
    HIP_TRY(hipMemcpy(h->dTis, tis.data(), NMC_NTISS * sizeof(XyzTissue), hipMemcpyHostToDevice));
  }
  h->prepared = true;
  return nmc_reset_tallies(h);
}

int nmc_reset_tallies(nmc_handle* h) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (!h->prepared) return fail(NMC_ERR_STATE, "nmc_prepare has not succeeded");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipMemsetAsync(h->dTot, 0, sizeof(nmc_totals), h->stream));
  if (h->sim == NMC_SIM_MCSUB) {
    HIP_TRY(hipMemsetAsync(h->dRa, 0, h->n_ra * sizeof(double), h->stream));
    if (h->dRz) HIP_TRY(hipMemsetAsync(h->dRz, 0, h->n_rz * sizeof(double), h->stream));
  } else {
    int64_t n = nvox(&h->rp), n2 = (int64_t)h->rp.Nx * h->rp.Ny;
    if (h->dF) HIP_TRY(hipMemsetAsync(h->dF, 0, n * sizeof(double), h->stream));
    HIP_TRY(hipMemsetAsync(h->dRyx, 0, n2 * sizeof(double), h->stream));
    HIP_TRY(hipMemsetAsync(h->dDet, 0, n2 * sizeof(uint32_t), h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return ok();
}

int nmc_set_seed(nmc_handle* h, uint64_t seed, uint64_t first_photon) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (!h->prepared) return fail(NMC_ERR_STATE, "nmc_prepare has not succeeded");
  h->ext.seed = seed;
  h->next_photon = first_photon;
  return ok();
}

static uint32_t pick_chunk(uint64_t n, int waves) {
  uint64_t c = n / ((uint64_t)waves * 8ull);
  if (c < 64) c = 64;
  if (c > 1024) c = 1024;
  return (uint32_t)c;
}

int nmc_run_async(nmc_handle* h, uint64_t n_photons) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  if (!h->prepared) return fail(NMC_ERR_STATE, "nmc_prepare has not succeeded");
  if (n_photons == 0) return ok();
  HIP_TRY(hipSetDevice(h->device));
  const nmc_ext_config& e = h->ext;
  uint32_t key0 = (uint32_t)e.seed, key1 = (uint32_t)(e.seed >> 32) + e.stream_id;
  HIP_TRY(launch_zero_u64(h->dCounter, 1, h->stream));
  HIP_TRY(hipEventRecord(h->e0, h->stream));
  if (h->sim == NMC_SIM_MCSUB) {
    const RunParams& rp = h->rp;
    SubArgs a{};
    float mua = h->tp.muav[1], mus = h->tp.musv[1];
    a.mut = mua + mus;
    a.one_m_albedo = 1.0f - mus / a.mut;
    a.g = h->tp.gv[1];
    a.n1 = rp.n1;
    a.n2 = rp.n2;
    float tr = rp.n1 / rp.n2;
    tr = (1.0f - tr) / (1.0f + tr);
    a.rsp0 = tr * tr;
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
    a.key0 = key0;
    a.key1 = key1;
    a.first = h->next_photon;
    a.n = n_photons;
    a.chunk = pick_chunk(n_photons, h->sub_grid * (kSubBlock / 64));
    a.cos_tab = h->dCos;
    a.tally_rz = h->dRz != nullptr;
    a.lds_tallies = h->lds_bytes > 0;
    a.part = h->dPart;
    a.rd_ra = h->dRa;
    a.a_rz = h->dRz;
    a.tot = h->dTot;
    a.counter = h->dCounter;
    HIP_TRY(launch_mcsub(a, h->sub_grid, h->lds_bytes, h->stream));
    if (a.lds_tallies) {
      int64_t stride = h->n_ra + h->n_rz;  // block partial = [Rd(r,a) | A(r,z)]
      HIP_TRY(launch_reduce_partials(h->dPart, h->sub_grid, stride, h->n_ra, h->dRa, h->stream));
      if (h->n_rz)
        HIP_TRY(launch_reduce_partials(h->dPart + h->n_ra, h->sub_grid, stride, h->n_rz, h->dRz,
                                       h->stream));
    }
  } else {
    const RunParams& rp = h->rp;
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
    a.mcflag = rp.mcflag;
    a.launchflag = rp.launchflag;
    a.zsurf = rp.zsurf;
    a.n1 = rp.n1;
    a.n2 = rp.n2;
    if (rp.xd == 0.0f) {
      a.rmin = 0.0f;
      a.rmax = rp.det_radius;
    } else {
      a.rmin = rp.xd - rp.det_radius;
      a.rmax = a.rmin + 2.0f * rp.det_radius;
    }
    a.cos_na = rp.na >= 90.0f ? -1.0f : (float)std::cos((double)rp.na * 3.14159265358979323846 / 180.0);
    a.xs = rp.xs;
    a.ys = rp.ys;
    a.zs = rp.zs;
    a.ux0 = rp.ux0;
    a.uy0 = rp.uy0;
    a.uz0 = rp.uz0;
    a.radius = rp.radius;
    a.waist = rp.waist;
    a.zfocus = rp.zfocus;
    a.thr = e.threshold;
    a.chance = e.chance;
    a.max_events = e.max_events;
    a.key0 = key0;
    a.key1 = key1;
    a.first = h->next_photon;
    a.n = n_photons;
    bool tt = h->sim == NMC_SIM_MCXYZ_TT;
    int grid = h->num_cu * mcxyz_blocks_per_cu(tt);
    a.chunk = pick_chunk(n_photons, grid * (kXyzBlock / 64));
    a.V = h->dV;
    a.tissue = h->dTis;
    a.F = h->dF;
    a.ryx = h->dRyx;
    a.det2d = h->dDet;
    a.tot = h->dTot;
    a.counter = h->dCounter;
    HIP_TRY(launch_mcxyz(a, tt, grid, h->stream));
  }
  HIP_TRY(hipEventRecord(h->e1, h->stream));
  h->timed = true;
  h->next_photon += n_photons;
  return ok();
}

int nmc_synchronize(nmc_handle* h) {
  if (!h) return fail(NMC_ERR_ARG, "null handle");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return ok();
}

int nmc_run(nmc_handle* h, uint64_t n_photons) {
  int st = nmc_run_async(h, n_photons);
  if (st != NMC_OK) return st;
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
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->sim == NMC_SIM_MCSUB) {
    if (out->rd_ra) HIP_TRY(hipMemcpy(out->rd_ra, h->dRa, h->n_ra * sizeof(double), hipMemcpyDeviceToHost));
    if (out->a_rz) {
      if (!h->dRz) return fail(NMC_ERR_STATE, "A(r,z) tally disabled by NMC_FLAG_NO_FLUENCE");
      HIP_TRY(hipMemcpy(out->a_rz, h->dRz, h->n_rz * sizeof(double), hipMemcpyDeviceToHost));
    }
  } else {
    int64_t n = nvox(&h->rp), n2 = (int64_t)h->rp.Nx * h->rp.Ny;
    if (out->fluence) {
      if (!h->dF) return fail(NMC_ERR_STATE, "fluence tally disabled by NMC_FLAG_NO_FLUENCE");
      HIP_TRY(hipMemcpy(out->fluence, h->dF, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (out->ryx) HIP_TRY(hipMemcpy(out->ryx, h->dRyx, n2 * sizeof(double), hipMemcpyDeviceToHost));
    if (out->det2d) HIP_TRY(hipMemcpy(out->det2d, h->dDet, n2 * sizeof(uint32_t), hipMemcpyDeviceToHost));
  }
  HIP_TRY(hipMemcpy(&out->totals, h->dTot, sizeof(nmc_totals), hipMemcpyDeviceToHost));
  return ok();
}

void nmc_destroy(nmc_handle* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_buffers(h);
  if (h->dTot) (void)hipFree(h->dTot);
  if (h->dCounter) (void)hipFree(h->dCounter);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

}  // extern "C"
