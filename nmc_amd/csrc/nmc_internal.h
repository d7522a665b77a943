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
  float mua, mus, inv_mus, g;       // inv_mus = 1/mus (FLT_MAX for mus = 0)
  HGCoef hg;
  TTCoef tt;
  int32_t slot;                     // sensitivity slot (nmc_mcxyz_sens), -1: none
};

// Photon-queue state read by the kernels from device memory, so that a
// captured graph can advance `first` between replays.
struct RunState {
  unsigned long long counter;       // next queue index (zeroed before each launch)
  unsigned long long first;         // photon id of queue index 0 (per item)
};

// Parameters of the rare paths of an mcxyz run: the source
// (kernels.metal:366-463 launch modes) and the detector plane
// (kernels.metal:556-630).  They live in device memory, not among the kernel
// arguments, and are read by non-inlined functions, so the transport loop
// does not hold them in scalar registers (with them there, the kernel spilled
// 75 SGPRs into VGPR lanes and reloaded them inside the loop).
struct XyzAux {
  int mcflag, launchflag;
  float xs, ys, zs, ux0, uy0, uz0, radius, waist, zfocus;
  float n1, n2, rmin, rmax, cos_na;
  int Nx, Ny;
  double* ryx;                      // [Nx][Ny] detected weight
  uint32_t* det2d;                  // [Nx][Ny] detected photons
};

// Buckets of the binned deposit records: 2^14 voxels, whose FP64 sums fill
// 128 KB of LDS in the reduction.
constexpr int kBucketShift = 14;
constexpr uint32_t kBucketVox = 1u << kBucketShift;
// At most this many buckets (their slices take 16 B of the transport
// kernel's LDS each): grids up to 2048 x 16384 = 33.5 M voxels.
constexpr int kMaxBuckets = 2048;

// One workgroup of the record reduction: bucket b, workgroup slices [w0, w1).
struct RecPart {
  int bucket, w0, w1, pad;
};

struct RecReduceArgs {
  const RecPart* parts;
  const unsigned long long* rec_base;
  const uint32_t* rec_cap;
  const uint2* rec;
  const uint32_t* rec_count;
  double* F;
  int grid;                         // transport workgroups (slices per bucket)
  uint32_t nvox;
};

struct XyzArgs {
  int Nx, Ny, Nz;
  float dx, dy, dz, hx, hy;
  int bmode, det_state;
  float zsurf;
  const XyzAux* aux;
  float thr, chance;
  uint32_t max_events;
  uint32_t key0, key1;
  uint64_t n;                       // photons in this launch
  uint32_t chunk;                   // queue indices taken per wave refill
  uint32_t refill_min;              // refill once this many lanes of a wave are idle
  int count_visits;                 // 1: F counts voxel visits (tile placement pilot)
  int tile_x0, tile_y0, tile_z0;    // LDS fluence tile origin (voxel indices)
  int tile_nx, tile_ny, tile_nz;    // tile extent; 0 = no tile
  const uint8_t* V;
  const XyzTissue* tissue;
  double* F;                        // may be null (NMC_FLAG_NO_FLUENCE)
  double* tile_part;                // [grid][tile voxels] block partials of the tile
  nmc_totals* tot;
  RunState* state;
  // binned deposit records (tuning.mode 2): a visit outside the tile to a
  // voxel of bucket b (voxels [b kBucketVox, (b + 1) kBucketVox)) is appended
  // to the workgroup's slice rec[rec_base[b] + blockIdx.x rec_cap[b] + slot]
  // of that bucket, slot from a counter in LDS; beyond the slice's capacity it
  // falls back to the FP64 atomic.  n_buckets = 0: no records.
  int n_buckets;
  const unsigned long long* rec_base;   // [n_buckets] first record of the bucket's region
  const uint32_t* rec_cap;              // [n_buckets] records per workgroup slice
  uint2* rec;                           // {voxel index within the bucket, weight bits (f32)}
  uint32_t* rec_count;                  // [n_buckets][grid] slots taken (may exceed rec_cap)
  // per-tissue sensitivities (nmc_mcxyz_sens): sens[0 or 1][slot][voxel]
  // (d/dmua, d/dmus) for sens_n tracked tissues, whose 1/mus are given; null:
  // none.  Runs with them use their own kernel variant (no tile, no records).
  double* sens;
  int sens_n;
  float sens_inv_mus[NMC_SENS_MAX];
  // the previous launch's records (tuning.mode 2, fused reduction): a
  // workgroup whose own photons are done takes parts of them one at a time
  // from *prev_next, sums each in its LDS (which the tile no longer needs)
  // and adds the sums to F, so the reduction of one launch runs in the
  // drained tail of the next.  prev_n_parts = 0: nothing to reduce.
  RecReduceArgs prev;
  int prev_n_parts;
  uint32_t* prev_next;
};

// One medium of a (batched) mcsub launch.
struct SubMedium {
  float mut, inv_mut, one_m_albedo, n1, n2, rsp0;
  uint32_t key1;
  int slot;                         // output row (queue order may differ)
  HGCoef hg;
};

struct SubArgs {
  int mcflag, bmode, nr, nz, na;
  float dr, dz, radius, waist, zfocus, xs, ys, zs;
  float thr, chance;
  uint32_t max_events;
  uint32_t key0;
  uint32_t chunk;
  uint32_t refill_min;              // refill once this many lanes of a wave are idle
  int n_items;
  const SubMedium* items;           // [n_items] in queue order
  const uint64_t* prefix;           // [n_items+1] queue index where item j starts
  const float* cos_tab;             // na-1 upper bin-edge cosines
  int tally_rz;                     // 0: skip the A(r,z) tally
  // global mode: [reps][n_items][(nr+1)*na + (nr+1)*(nz+1)], FP64 atomics into
  // replica blockIdx % reps, summed into the output after the launch
  double* tally;
  int reps;
  double* part;                     // block partials: [LDS tallies] [totals]
  RunState* state;
};

// Layered media: one row per (stack, layer), one header per stack.
struct LayRow {
  float zt, zb, mua, mus, inv_mus, n;
  int phase, pad;
  HGCoef hg;
  TTCoef tt;
};
struct LayStackDev {
  int nl, row0;
  float n_above, n_below, rsp;
  uint32_t key1;
  int pad[2];
};

struct LayArgs {
  int n_items, n_rows, nr;
  float dr, thr, chance;
  uint32_t max_events, key0, chunk, refill_min;
  const LayStackDev* stacks;        // [n_items]
  const LayRow* rows;               // [n_rows]
  const uint64_t* prefix;           // [n_items+1]
  // Rd(r) partials, reduced over their first index after the launch:
  // rd_lds = 1: [grid][n_items][nr+1], each block's LDS histogram written out;
  // rd_lds = 0: [rd_reps][n_items][nr+1], FP64 global atomics into replica
  // blockIdx % rd_reps (blocks are dealt to the XCDs round-robin, so any
  // multiple of 8 replicas keeps each replica on one XCD).
  double* rd;
  int rd_lds, rd_reps;
  double* part;                     // block partials of [absorbed | totals]
  int records;
  int rec_nl;                       // layers stored per record (deepest stack)
  uint64_t cap;
  unsigned long long* rec_count;    // [0] slots handed out, [1] records written
  uint32_t* rec;                    // [cap][3 + 2 rec_nl]: stack, bin, w, L[rec_nl], K[rec_nl];
                                    // stack 0xFFFFFFFF marks an unused slot
  RunState* state;
  // absorbed-weight sensitivities (sens = 1): block partials
  // [grid][n_items][2][MAXL][MAXL] of d(absorbed in layer l)/d(mua_t | mus_t)
  int sens;
  double* sens_part;
};

// Doubles of the absorbed-weight sensitivities per stack: [2][MAXL][MAXL].
constexpr int kSensPerStack = 2 * NMC_MAX_LAYERS * NMC_MAX_LAYERS;

// Record slots a wave reserves at once (one global atomic per kRecChunk records).
constexpr int kRecChunk = 256;
constexpr uint32_t kRecHole = 0xFFFFFFFFu;

struct LayBackArgs {
  int n_items, nr;
  int rec_nl;
  uint64_t n_rec, cap;
  const uint32_t* rec;              // as LayArgs::rec
  const double* grad_rd;            // [n_items][nr+1]
  const float* mus;                 // [n_items][NMC_MAX_LAYERS]
  const int* nl;                    // [n_items]
  const double* inv_n;              // [n_items] 1 / budget
  double* grad_mua;                 // [n_items][NMC_MAX_LAYERS] (accumulated)
  double* grad_mus;
};

constexpr int kLayBlock = 1024;
size_t layered_lds_bytes(int n_items, int n_rows, int rd_len, bool sens = false);
// Global tally replicas (Rd(r) of the layered kernel, Rd/A of batched mcsub)
// when the tallies do not fit in LDS: as many as kRepBytes allows, at least
// kMinReps (one per XCD), at most one per block.
constexpr int kMinReps = 8;
constexpr size_t kRepBytes = (size_t)256 << 20;
inline int replica_count(int64_t len, int grid) {
  int64_t r = (int64_t)(kRepBytes / (sizeof(double) * (size_t)(len > 0 ? len : 1)));
  if (r > grid) r = grid;
  if (r < kMinReps) r = kMinReps;
  return (int)r;
}
hipError_t layered_configure(size_t lds_bytes);
int layered_blocks_per_cu(size_t lds_bytes);
hipError_t launch_layered(const LayArgs& a, int grid, size_t lds_bytes, hipStream_t s);
hipError_t launch_layered_backward(const LayBackArgs& b, hipStream_t s);


// A resolved mcxyz launch shape: tuning (tile box, block, ...), dynamic LDS
// bytes, workgroups, and tile voxels (0: no tile).
struct XyzLaunch {
  nmc_tuning t;
  size_t lds;
  int grid;
  int64_t tile_n;
  int n_buckets;                    // > 0: binned deposit records (tuning.mode 2)
};

// Device buffers of the binned deposit records of an mcxyz handle.
struct XyzRecords {
  int n_buckets = 0;
  unsigned long long* base = nullptr;   // [n_buckets]
  uint32_t* cap = nullptr;              // [n_buckets]
  uint2* rec = nullptr;                 // sum over b of cap[b] * grid records
  uint32_t* count = nullptr;            // [n_buckets][grid]
  RecPart* parts = nullptr;             // reduction workgroups
  int n_parts = 0;
  uint64_t n_rec = 0;                   // records the buffer holds
  // fused reduction (when the transport kernel's LDS holds a bucket's FP64
  // sums): a second record buffer, so that consecutive launches alternate
  // between slot 0 (rec, count) and slot 1 (rec1, count1) and each reduces
  // the other slot's records of the launch before it; null: one slot, every
  // launch followed by rec_reduce_kernel
  uint2* rec1 = nullptr;
  uint32_t* count1 = nullptr;
  uint32_t* next = nullptr;             // part counter of the fused reduction
  // host-side state: the slot the next launch writes and the slot whose
  // records still await their reduction (-1: none)
  int slot = 0;
  int pending = -1;
};

// The sensitivity variant: 1024-lane workgroups, lds = its per-lane rows.
hipError_t launch_mcxyz_sens(const XyzArgs& a, bool two_term, int grid, size_t lds, hipStream_t s);
hipError_t mcxyz_sens_configure(bool two_term, size_t lds);
int mcxyz_sens_blocks_per_cu(bool two_term, size_t lds);
constexpr int kSensBlock = 1024;
hipError_t launch_mcxyz(const XyzArgs& a, bool two_term, bool tile32, int block, int grid,
                        size_t lds_bytes, hipStream_t s);
hipError_t mcxyz_configure(bool two_term, bool tile32, int block, size_t lds_bytes);
int mcxyz_blocks_per_cu(bool two_term, bool tile32, int block, size_t lds_bytes);
hipError_t launch_tile_reduce(const double* part, int grid, const XyzArgs& a, hipStream_t s);
hipError_t launch_rec_reduce(const RecReduceArgs& r, int n_parts, hipStream_t s);
hipError_t rec_reduce_configure();
hipError_t launch_mcsub(const SubArgs& a, int block, int grid, size_t lds_bytes, bool lds_tallies,
                        hipStream_t s);
hipError_t mcsub_configure(int block, size_t lds_bytes, bool lds_tallies);
int mcsub_blocks_per_cu(int block, size_t lds_bytes, bool lds_tallies);
size_t mcsub_lds_bytes(int n_items, int64_t per_item_in_lds);
// out[j] += sum over b < grid of part[b * stride + j] (integer sums for the
// nmc_totals counters); with clear, each part[] element is zeroed once read,
// so replicas start the next run at zero without a memset.
hipError_t launch_reduce_partials(double* part, int grid, int64_t stride, double* out,
                                  int ni, int per_item, hipStream_t s, bool clear = false);
hipError_t launch_state_begin(RunState* st, hipStream_t s);
hipError_t launch_state_advance(RunState* st, uint64_t adv, hipStream_t s);

}  // namespace nmc
