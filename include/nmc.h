/*
 * nmc.h - C ABI of the MI355X photon-transport library (libnmc_hip.so).
 *
 * The library replaces the Metal host runtime of NMC (`MetalMC`,
 * NMC/NMCCode/MetalMC.h:25-31) and the kernels it dispatches
 * (NMC/NMCCode/kernels.metal:59-754).  Signatures use plain C types and
 * pointers only; every entry point returns an `int` status (NMC_OK == 0) and
 * records a message readable through nmc_last_error().  Nothing in the library
 * calls exit().
 *
 * Geometry/material inputs are the reference's own structs, byte-compatible
 * with NMC/NMCCode/definitions.h:186-218 (checked by static asserts below), so
 * a caller that already fills `RunParams`/`TissueParams` for NMC can pass them
 * unchanged.  Everything NMC fixed at compile time (PHOTON_TOTAL, BINS, dr/dz,
 * THRESHOLD, CHANCE, ...; definitions.h:16-82) travels in `nmc_ext_config`.
 */
#ifndef NMC_H
#define NMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants shared with the reference (definitions.h:57-58,66-67) ---- */
#define NMC_NTISS 19          /* tissue table length, index 0 unused       */
#define NMC_STRLEN 32
#define NMC_THRESHOLD 0.0001f /* roulette threshold                         */
#define NMC_CHANCE 0.1f       /* roulette survival chance                   */

/* ---- status codes ---- */
#define NMC_OK 0
#define NMC_ERR_ARG -1        /* invalid argument / configuration           */
#define NMC_ERR_STATE -2      /* call out of order (e.g. run before prepare) */
#define NMC_ERR_DEVICE -3     /* HIP runtime error                          */
#define NMC_ERR_NOMEM -4      /* allocation failed / larger than device     */
#define NMC_ERR_UNSUPPORTED -5

/* ---- simulation kinds (definitions.h:151-164, MC_SIM_TYPE) ---- */
#define NMC_SIM_MCSUB 0       /* semi-infinite medium, cylindrical tallies  */
#define NMC_SIM_MCXYZ 1       /* voxelised 3-D medium                       */
#define NMC_SIM_MCXYZ_TT 2    /* mcxyz with the two-term phase function     */

/* ---- detection side (definitions.h:167-171) ---- */
#define NMC_DET_REFLECTANCE 0
#define NMC_DET_TRANSMITTANCE 1

/*
 * RunParams, byte-compatible with definitions.h:186-209 (172 bytes).
 * The three enums of the reference are stored as int32.
 */
typedef struct RunParams {
    char myname[NMC_STRLEN];
    int32_t mcflag, launchflag, boundaryflag, speckleflag, semiflag;
    float xfocus, yfocus, zfocus;
    float ux0, uy0, uz0;
    float radius, det_radius, xd;
    float waist;
    float zsurf;
    float n1, n2;
    float na, lambda, lc;
    int32_t det_state;
    int32_t ph_quant;
    float dx, dy, dz;
    int32_t Nx, Ny, Nz, Nt;
    float xs, ys, zs;
    float time_min;
    int32_t mckernelflag;
} RunParams;

/* TissueParams, byte-compatible with definitions.h:211-218 (608 bytes). */
typedef struct TissueParams {
    float muav[NMC_NTISS];
    float musv[NMC_NTISS];
    float gv[NMC_NTISS];
    float gf[NMC_NTISS], af[NMC_NTISS], gb[NMC_NTISS], ab[NMC_NTISS], CC[NMC_NTISS];
} TissueParams;

#define NMC_EXT_VERSION 1

/* flags in nmc_ext_config.flags */
#define NMC_FLAG_NO_FLUENCE 1u   /* skip the 3-D / (r,z) absorption tally   */
#define NMC_FLAG_COUNT_VISITS 2u /* mcxyz: the fluence array counts voxel
                                    visits instead of deposited weight (used
                                    by the autotuner to place the LDS tile)  */

/*
 * Run-time settings that NMC fixed at compile time (definitions.h:16-82) plus
 * the Philox stream selection.  No counterpart struct exists in the reference.
 */
typedef struct nmc_ext_config {
    uint32_t struct_size;   /* sizeof(nmc_ext_config)                       */
    uint32_t version;       /* NMC_EXT_VERSION                              */
    uint64_t seed;          /* Philox key = (seed lo, seed hi + stream_id)  */
    uint32_t stream_id;
    uint32_t flags;         /* NMC_FLAG_*                                   */
    float threshold;        /* roulette threshold (default NMC_THRESHOLD)   */
    float chance;           /* roulette survival  (default NMC_CHANCE)      */
    uint32_t max_events;    /* per-photon substep+interaction watchdog      */
    /* mcsub cylindrical binning (replaces BINS, dr_mcsub, dz_mcsub)        */
    int32_t nr, nz, na;     /* real bins; one overflow bin is appended to r and z */
    float dr, dz;
    int32_t reserved[8];
} nmc_ext_config;

/*
 * Scalar totals of one tally set, all weights in photon-weight units (not yet
 * divided by the photon count).  The integer fields are exact event counts.
 */
typedef struct nmc_totals {
    double specular;          /* weight reflected at launch (Rsp)            */
    double rd;                /* weight escaping through the detector window */
    double rd_outside;        /* weight escaping the surface outside it      */
    double absorbed;          /* weight deposited inside the grid / medium   */
    double absorbed_outside;  /* weight deposited outside the grid           */
    double tt;                /* weight leaving through the bottom (z >= Nz*dz) */
    double lateral;           /* weight leaving through the x/y sides        */
    double top;               /* weight leaving through the grid top (z < 0) */
    double roulette_gain;     /* weight created by roulette survivals        */
    double roulette_loss;     /* weight removed by roulette kills            */
    double capped;            /* weight of photons stopped by max_events     */
    uint64_t photons;
    uint64_t detections;      /* escapes through the detector window         */
    uint64_t surface_hits;    /* arrivals at the zsurf plane / mcsub surface */
    uint64_t escapes_bottom, escapes_side, escapes_top;
    uint64_t roulette_kills, roulette_survivals, capped_photons;
    uint64_t substeps_same;   /* hop pieces that end in an interaction       */
    uint64_t substeps_cross;  /* hop pieces that end on a voxel face / plane */
    uint64_t interactions;    /* scattering events (spin + roulette)         */
} nmc_totals;

/*
 * Host-side result pointers; any may be NULL to skip that array.  Sizes:
 *  mcxyz: fluence Nz*Nx*Ny (index iz*Ny*Nx + ix*Ny + iy, kernels.metal:482),
 *         ryx and det2d Nx*Ny (index ix*Ny + iy).
 *  mcsub: rd_ra (nr+1)*na   (index ir*na + ia, ir == nr is the overflow bin),
 *         a_rz  (nr+1)*(nz+1) (index ir*(nz+1) + iz; F[ir][iz] of kernels.metal:233).
 * Values are raw weights in double precision, accumulated over every
 * nmc_run since the last nmc_reset_tallies.
 */
typedef struct nmc_results {
    double* fluence;
    double* ryx;
    uint32_t* det2d;
    double* rd_ra;
    double* a_rz;
    nmc_totals totals;
} nmc_results;

typedef struct nmc_handle nmc_handle;

/* Defaults for every field of nmc_ext_config. */
void nmc_ext_config_init(nmc_ext_config* cfg);

/* Checks a configuration without touching a device (the checks nmc_prepare
 * runs first): NMC_OK or an error code with nmc_last_error() set.  The
 * reference performs no validation (mcio.h:16-189 reads unchecked). */
int nmc_validate(int sim_type, const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                 const nmc_ext_config* ext);

/* Replaces -[MetalMC initWithDevice::] (MetalMC.h:25, MetalMC.m:96-177):
 * selects the device and the simulation kind. */
int nmc_create(int sim_type, int device, nmc_handle** out);

/* Replaces -[MetalMC prepareData:::] (MetalMC.h:26, MetalMC.m:179-383):
 * validates and copies parameters and the voxel map (NULL for mcsub) to the
 * device and allocates the device tallies (zeroed).  `ext` may be NULL. */
int nmc_prepare(nmc_handle* h, const RunParams* rp, const TissueParams* tp,
                const uint8_t* v, const nmc_ext_config* ext);

/* Replaces -[MetalMC setNumRunsSeed::] (MetalMC.h:30, MetalMC.m:498-505):
 * the 64-bit Philox seed and the id of the next photon to launch. */
int nmc_set_seed(nmc_handle* h, uint64_t seed, uint64_t first_photon);

/* Replaces -[MetalMC sendComputeCommand] (MetalMC.h:27, MetalMC.m:385-407):
 * transports photons [first_photon, first_photon + n) and blocks until done;
 * afterwards first_photon advances by n. */
int nmc_run(nmc_handle* h, uint64_t n_photons);

/* Same as nmc_run but returns after enqueueing on the handle's stream. */
int nmc_run_async(nmc_handle* h, uint64_t n_photons);
/* Blocks until every enqueued run is done, its tallies complete (with the
 * fused record reduction, tuning.rec_fuse, this first enqueues the reduction
 * of the last launch's records). */
int nmc_synchronize(nmc_handle* h);

/* Milliseconds between HIP events recorded around the transport kernel(s) of
 * the last nmc_run / nmc_run_async, on the stream they were launched on. */
int nmc_last_kernel_ms(nmc_handle* h, float* ms);

/* Replaces -[MetalMC getComputeResults:...] (MetalMC.h:28, MetalMC.m:508-824):
 * copies the accumulated tallies into the caller's arrays. */
int nmc_get_results(nmc_handle* h, nmc_results* out);
/* getComputeResults' own semantics (MetalMC.m:521,787,809): ADDS the
 * accumulated tallies into the caller's arrays and totals (+=), for callers
 * that sum several handles or runs into one set of arrays. */
int nmc_add_results(nmc_handle* h, nmc_results* out);

/* Zeroes the device tallies (the reference re-creates buffers instead). */
int nmc_reset_tallies(nmc_handle* h);

void nmc_destroy(nmc_handle* h);

/*
 * Launch tuning (no reference counterpart: NMC dispatches a fixed 10 000-thread
 * grid, MetalMC.m:483-495).  Zero fields mean "library default".
 */
typedef struct nmc_tuning {
    int32_t block;          /* threads per workgroup: 256, 512 or 1024          */
    int32_t blocks_per_cu;  /* workgroups per CU (<= occupancy limit)           */
    int32_t refill_min;     /* idle lanes per wave before a refill (1..64)      */
    int32_t chunk;          /* photon ids a wave takes from the queue at once   */
    int32_t tile[6];        /* mcxyz LDS fluence tile x0,y0,z0,nx,ny,nz;
                               nx == 0: automatic box, nx < 0: no tile          */
    int32_t mode;           /* mcsub: tally placement, 0 automatic (LDS when the
                               per-item tallies fit), 1 LDS, 2 global atomics;
                               mcxyz: LDS tile element, 0 FP64, 1 FP32 (twice
                               the voxels; sums within a workgroup in FP32,
                               across workgroups in FP64; runs are split into
                               launches of tuning.split photons per
                               workgroup, the same photons), 2 FP32 tile with
                               binned deposit records: a voxel visit outside
                               the tile appends an 8-byte record (voxel,
                               weight rounded to FP32) to the workgroup's
                               slice of the voxel's 16384-voxel bucket
                               instead of an FP64 atomic (slice capacities
                               from a visit-count pilot at set-up; a full
                               slice falls back to the atomic), and a
                               reduction sums each bucket's records in FP64
                               into the fluence; grids up to 2048 buckets;
                               layered: Rd(r) placement, read at
                               nmc_layered_prepare, 0 or 1 LDS histogram when
                               it fits without lowering occupancy, 2 global
                               replicas (effective value reported as 1 / 2)   */
    int32_t split;          /* mcxyz with the FP32 tile: photons per workgroup
                               in one launch (0: 390625, i.e. 1e8 photons on
                               one 1024-lane workgroup per CU of 256 CUs); a
                               longer run becomes consecutive launches of the
                               same photons, each flushing its tile in FP64  */
    int32_t rec_cap_pct;    /* mode 2: record slice capacity in per cent of
                               the pilot's estimate (0: 100)                 */
    int32_t rec_fuse;       /* mode 2: 0 automatic, on when the tile's LDS
                               holds a bucket's FP64 sums (128 KB); -1 off.
                               On: two record buffers; consecutive launches
                               alternate between them, and the workgroups of
                               each launch, once their own photons are done,
                               reduce the previous launch's records into the
                               fluence.  The last launch's records are
                               reduced by the next launch or, at the latest,
                               by nmc_synchronize / nmc_run / nmc_get_results
                               (effective value reported as 1 / -1)          */
    int32_t reserved[2];
} nmc_tuning;

/* A rejected request (NMC_ERR_ARG, or an allocation failure) leaves the
 * handle's previous configuration and buffers unchanged. */
int nmc_set_tuning(nmc_handle* h, const nmc_tuning* t);
/* Effective values after defaults and clamping (tile resolved to a box). */
int nmc_get_tuning(nmc_handle* h, nmc_tuning* t);
/* mcxyz: milliseconds of the last transport-kernel launch of the last eager
 * run (HIP events on the handle's stream around that kernel alone, without
 * the tile and record reductions that nmc_last_kernel_ms includes). */
int nmc_last_transport_ms(nmc_handle* h, float* ms);
/* Transport-kernel launches of the last run (no reference counterpart; more
 * than one when an FP32-tile mcxyz run is split by tuning.split). */
int nmc_last_launch_count(nmc_handle* h, uint32_t* n);

/*
 * Per-tissue fluence sensitivities of the voxel walk (no reference
 * counterpart; the reference's absorbed-weight tally is F3D,
 * kernels.metal:530,550, normalised to fluence in main.m:481-490).
 * nmc_mcxyz_sens selects up to NMC_SENS_MAX distinct tissue ids (n_tissues
 * = 0: none) of a prepared mcxyz handle with the fluence tally; later runs
 * then also accumulate, for every voxel v and selected tissue t, the
 * perturbation-MC derivatives of the raw absorbed weight F[v] (the
 * `fluence` array of nmc_get_results) with respect to mua_t and mus_t: each
 * deposit weighted by the photon's path length and scatterings in every
 * tissue before it.  Those runs use their own kernel variant: every deposit
 * goes to the FP64 array directly (no LDS tile, no records; tuning.mode and
 * tile are ignored), the photons and F are the same.  The sensitivities are
 * zeroed here and by nmc_reset_tallies.  For the fluence phi = F / (N vol
 * mua_k) of a voxel of tissue k: dphi/dmua_t = dF/dmua_t / (N vol mua_k)
 * - [t = k] phi / mua_k, dphi/dmus_t = dF/dmus_t / (N vol mua_k).
 */
#define NMC_SENS_MAX 8
int nmc_mcxyz_sens(nmc_handle* h, const int32_t* tissue_ids, int n_tissues);
/* d_mua, d_mus: n_tissues x Nz*Nx*Ny doubles each (voxel order of the
 * fluence array), the sums of all runs since the selection or reset. */
int nmc_mcxyz_sens_results(nmc_handle* h, double* d_mua, double* d_mus);

/*
 * Batched mcsub: many media (e.g. wavelengths), each with its own optical
 * properties, Philox stream and photon budget, transported by ONE persistent
 * launch from a device photon queue.  No counterpart in the reference, whose
 * notebooks run one `nmc` process per configuration (example3.ipynb:862-902).
 * Item j's photons are those of a plain run with the same medium and
 * ext->stream_id = items[j].stream_id.
 */
typedef struct nmc_work_item {
    float mua, mus, g;      /* medium (tissue type 1 of a plain mcsub run)     */
    float n1, n2;           /* medium / outside refractive index               */
    uint32_t stream_id;
    uint64_t n_photons;     /* budget per nmc_run_batch                        */
} nmc_work_item;

/* Beam and bins from rp/ext (as nmc_prepare on an NMC_SIM_MCSUB handle);
 * per-item tallies are (nr+1)*na Rd(r,alpha) bins plus, unless
 * NMC_FLAG_NO_FLUENCE, (nr+1)*(nz+1) A(r,z) bins. */
int nmc_prepare_batch(nmc_handle* h, const RunParams* rp, const nmc_ext_config* ext,
                      const nmc_work_item* items, int n_items);
/* Transports photons [first, first + n_photons_j) of every item j in one
 * launch; first then advances by max_j n_photons_j. */
int nmc_run_batch(nmc_handle* h);
int nmc_run_batch_async(nmc_handle* h);
/* tally: n_items x per-item bins (NULL to skip); totals: n_items entries. */
int nmc_get_batch_results(nmc_handle* h, double* tally, nmc_totals* totals);

/* enable != 0: the launch sequence of nmc_run / nmc_run_batch (and their
 * async forms) is captured once as a hipGraph and replayed on later calls
 * with the same photon count and seed. */
int nmc_set_graph_mode(nmc_handle* h, int enable);

/*
 * Layered media with per-layer refractive index and gradients (no reference
 * counterpart: NMC has a single global n1/n2 used only at zsurf,
 * definitions.h:196, kernels.metal:579).  Semantics follow the mcxyz loop
 * (continuous absorption, W *= exp(-mua s)) with all-or-none Fresnel
 * reflection / Snell refraction at every interface.  Light enters at the
 * origin along +z (pencil); Rd(r) counts weight leaving through the top.
 */
#define NMC_SIM_LAYERED 3
#define NMC_MAX_LAYERS 8

typedef struct nmc_layer {
    float thickness;              /* cm; last layer: <= 0 means semi-infinite */
    float mua, mus, n;            /* /cm, /cm, refractive index               */
    float g;                      /* HG anisotropy (phase == 0)               */
    float gf, af, gb, ab, cc;     /* two-term parameters (phase == 1), as the
                                     gf/af/gb/ab/CC of TissueParams           */
    int32_t phase;                /* 0 HG, 1 two-term                          */
    int32_t reserved;
} nmc_layer;

typedef struct nmc_stack {
    int32_t n_layers;             /* 1..NMC_MAX_LAYERS                         */
    float n_above, n_below;       /* ambient indices above / below the stack  */
    uint32_t stream_id;           /* Philox stream of this stack              */
    uint64_t n_photons;           /* budget per nmc_layered_run               */
    nmc_layer layer[NMC_MAX_LAYERS];
} nmc_stack;

typedef struct nmc_layered_config {
    uint32_t struct_size;         /* sizeof(nmc_layered_config)               */
    uint32_t version;             /* NMC_EXT_VERSION                          */
    uint64_t seed;
    int32_t nr;                   /* Rd(r) bins; bin nr is the overflow bin   */
    float dr;
    float threshold, chance;      /* roulette                                 */
    uint32_t max_events;
    int32_t records;              /* 1: keep per-detected-photon path records */
    uint64_t record_capacity;     /* 0: sum of the budgets                    */
    int32_t absorbed_grad;        /* 1: the forward run also accumulates the
                                     absorbed-weight sensitivities
                                     (nmc_layered_absorbed_grad)              */
    int32_t reserved[7];
} nmc_layered_config;

/* Allocates tallies (and the record buffer) for n_stacks stacks and uploads
 * them on an NMC_SIM_LAYERED handle. */
int nmc_layered_prepare(nmc_handle* h, const nmc_layered_config* cfg, const nmc_stack* stacks,
                        int n_stacks);
/* New optical properties / budgets for the same stack shapes (optimisers). */
int nmc_layered_update(nmc_handle* h, const nmc_stack* stacks, int n_stacks);
/* Zeroes tallies and records, then transports photons [first, first + n_j)
 * of every stack j in one launch (first from nmc_set_seed, default 0; it is
 * not advanced, so repeated runs reuse the same photons). */
int nmc_layered_run(nmc_handle* h);
/* rd: n_stacks x (nr+1) raw escaping weight; absorbed: n_stacks x
 * NMC_MAX_LAYERS; totals: n_stacks; n_records: records stored (any may be NULL). */
int nmc_layered_results(nmc_handle* h, double* rd, double* absorbed, nmc_totals* totals,
                        uint64_t* n_records);
/* Copies the path records of the last run to the caller, max_records at most,
 * each 3 + 2 * rec_nl u32 words: stack, Rd bin, weight (f32 bits), L[rec_nl]
 * (f32 bits), K[rec_nl], where rec_nl is the deepest stack's layer count; the
 * order is unspecified.  *n_out = records copied; *rec_nl_out = rec_nl. */
int nmc_layered_records(nmc_handle* h, uint32_t* out, uint64_t max_records, uint64_t* n_out,
                        int* rec_nl_out);
/* Perturbation-MC backward pass over the stored records: given
 * grad_rd[j][ir] = dLoss/d(rd[j][ir] / n_photons_j), returns
 * grad_mua[j][l] = sum over records of grad_rd * w * (-L_l) / n_j and
 * grad_mus[j][l] = sum of grad_rd * w * (K_l / mus_l - L_l) / n_j, where L_l and
 * K_l are the photon's path length and scattering count in layer l.  The
 * records must come from the current properties: after nmc_layered_update the
 * call fails with NMC_ERR_STATE until nmc_layered_run has run again. */
int nmc_layered_backward(nmc_handle* h, const double* grad_rd, double* grad_mua,
                         double* grad_mus);

/*
 * Allocation-free mcxyz run.  No counterpart in the reference: for callers
 * that cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc, code that must not allocate between captures).  Transports
 * photons [first_photon, first_photon + n_photons) - the photons that
 * nmc_prepare + nmc_set_seed(h, ext->seed, first_photon) + nmc_run(h, n) give
 * on a handle of the same kind - and ADDS the tallies into the caller's
 * device arrays, as getComputeResults adds into caller-zeroed arrays
 * (MetalMC.m:521): d_fluence Nz*Nx*Ny doubles (NULL or NMC_FLAG_NO_FLUENCE:
 * no fluence tally), d_ryx Nx*Ny doubles, d_det2d Nx*Ny u32, d_totals one
 * nmc_totals.  rp, tp, ext (may be NULL) and tuning (may be NULL) are host
 * structs; d_v is the voxel map in device memory (Nx*Ny*Nz bytes).  Runs on
 * the current device, on `stream` (a hipStream_t; NULL: the null stream).
 * The call checks the voxel map's tissue ids on the device and synchronises
 * `stream` once for that, then enqueues the run and returns; it calls no
 * hipMalloc or hipFree.  Scratch (tissue table, source block, queue state,
 * FP64 block partials of the LDS tile) lives in one static device array of
 * NMC_NOALLOC_SCRATCH_BYTES per GPU, kept until the process ends: a tile whose
 * block partials (workgroups x tile voxels) exceed NMC_NOALLOC_MAX_PARTIALS
 * doubles is rejected with NMC_ERR_ARG (the automatic tile shrinks to fit).
 * Every call on a GPU shares that array: calls that overlap on one GPU (on
 * different streams, unsynchronised) race.
 */
#define NMC_NOALLOC_MAX_PARTIALS (2u * 1024u * 1024u)
#define NMC_NOALLOC_SCRATCH_BYTES (4096u + 8u * NMC_NOALLOC_MAX_PARTIALS)
int nmc_mcxyz_run_noalloc(int sim_type, const RunParams* rp, const TissueParams* tp,
                          const nmc_ext_config* ext, const nmc_tuning* tuning, const uint8_t* d_v,
                          uint64_t first_photon, uint64_t n_photons, double* d_fluence,
                          double* d_ryx, uint32_t* d_det2d, nmc_totals* d_totals, void* stream);

/* Per-layer absorbed-weight sensitivities of the last nmc_layered_run (the
 * handle prepared with absorbed_grad = 1): d_mua[j][l][t] and d_mus[j][l][t]
 * (each n_stacks x NMC_MAX_LAYERS x NMC_MAX_LAYERS) are the derivatives of
 * the raw weight absorbed in layer l of stack j (nmc_layered_results'
 * `absorbed`) with respect to mua and mus of layer t: the perturbation-MC
 * estimate over every deposit, each weighted by the photon's per-layer path
 * lengths and scattering counts before it (no reference counterpart; the
 * reference's absorbed tallies are F3D / A(r,z), kernels.metal:530,550,
 * 224-234). */
int nmc_layered_absorbed_grad(nmc_handle* h, double* d_mua, double* d_mus);

/* Message of the last failing call on this host thread ("" if none). */
const char* nmc_last_error(void);

/* Library build string (offload target, compile flags that affect results). */
const char* nmc_version(void);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#ifdef __cplusplus
static_assert(sizeof(RunParams) == 172, "RunParams must match definitions.h:186-209");
static_assert(sizeof(TissueParams) == 608, "TissueParams must match definitions.h:211-218");
static_assert(offsetof(RunParams, n1) == 96 && offsetof(RunParams, det_state) == 116 &&
                  offsetof(RunParams, Nx) == 136 && offsetof(RunParams, mckernelflag) == 168,
              "RunParams field offsets");
#else
_Static_assert(sizeof(RunParams) == 172, "RunParams must match definitions.h:186-209");
_Static_assert(sizeof(TissueParams) == 608, "TissueParams must match definitions.h:211-218");
#endif

#endif /* NMC_H */
