// nmc_internal.h - argument blocks shared by the host runtime (nmc_abi.hip)
// and the transport kernels.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nmc.h"
#include "nmc_device.h"

namespace nmc {

// Device totals: the exact layout of nmc_totals (11 doubles, then 12 u64).
enum TotF {
  TF_SPECULAR, TF_RD, TF_RD_OUTSIDE, TF_ABSORBED, TF_ABSORBED_OUTSIDE, TF_TT, TF_LATERAL,
  TF_TOP, TF_ROULETTE_GAIN, TF_ROULETTE_LOSS, TF_CAPPED, TF_COUNT
};
enum TotI {
  TI_PHOTONS, TI_DETECTIONS, TI_SURFACE_HITS, TI_ESC_BOTTOM, TI_ESC_SIDE, TI_ESC_TOP,
  TI_ROULETTE_KILLS, TI_ROULETTE_SURVIVALS, TI_CAPPED, TI_SUB_SAME, TI_SUB_CROSS,
  TI_INTERACTIONS, TI_COUNT
};
static_assert(sizeof(nmc_totals) == TF_COUNT * 8 + TI_COUNT * 8, "nmc_totals layout");

// Per-tissue row of the mcxyz material table (held in LDS by the kernel).
struct XyzTissue {
  float mua, mus, g, pad;
  TTCoef tt;
};

struct XyzArgs {
  int Nx, Ny, Nz;
  float dx, dy, dz, hx, hy;
  int bmode, det_state, mcflag, launchflag;
  float zsurf, n1, n2, rmin, rmax, cos_na;
  float xs, ys, zs, ux0, uy0, uz0, radius, waist, zfocus;
  float thr, chance;
  uint32_t max_events;
  uint32_t key0, key1;
  uint64_t first, n;
  uint32_t chunk;
  const uint8_t* V;
  const XyzTissue* tissue;
  double* F;                        // may be null (NMC_FLAG_NO_FLUENCE)
  double* ryx;
  uint32_t* det2d;
  nmc_totals* tot;
  unsigned long long* counter;      // photon-id dispenser, zeroed before launch
};

struct SubArgs {
  float mut, one_m_albedo, g, n1, n2, rsp0;
  int mcflag, bmode, nr, nz, na;
  float dr, dz, radius, waist, zfocus, xs, ys, zs;
  float thr, chance;
  uint32_t max_events;
  uint32_t key0, key1;
  uint64_t first, n;
  uint32_t chunk;
  const float* cos_tab;             // na-1 upper bin-edge cosines
  int tally_rz;                     // 0: skip the A(r,z) tally
  int lds_tallies;                  // 1: tallies privatised in LDS, flushed to `part`
  double* part;                     // [grid][(nr+1)*na + (nr+1)*(nz+1)] block partials
  double* rd_ra;                    // direct global tallies when lds_tallies == 0
  double* a_rz;
  nmc_totals* tot;
  unsigned long long* counter;
};

constexpr int kXyzBlock = 256;
constexpr int kSubBlock = 1024;

hipError_t launch_mcxyz(const XyzArgs& a, bool two_term, int grid, hipStream_t s);
hipError_t launch_mcsub(const SubArgs& a, int grid, int lds_bytes, hipStream_t s);
hipError_t launch_reduce_partials(const double* part, int grid, int64_t stride, int64_t len,
                                  double* out, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s);
int mcxyz_blocks_per_cu(bool two_term);
int mcsub_blocks_per_cu(int lds_bytes);

}  // namespace nmc
