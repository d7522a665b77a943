/*
 * nmc_oracle.h - CPU oracle for the photon-transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: tests/, the smoke entry
 * point and the CPU-baseline leg of the benchmark may load it; the product
 * library (nmc_amd/) never does.  See nmc_oracle.c for the algorithm and the
 * reference lines each function restates.
 */
#ifndef NMC_ORACLE_H
#define NMC_ORACLE_H

#include <stdint.h>
#include "../include/nmc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* primitives (exported for the unit tests) */
void nmco_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float nmco_logf(float x);
float nmco_logf_tab(float x);
float nmco_expf(float x);
float nmco_expm1f(float x);
void nmco_sincos2pi(float u, float* s, float* c);
float nmco_fresnel(float n1, float n2, float ca1, float* ca2);
float nmco_hg_cos(float g, float u);
float nmco_tt_cos(float gf, float af, float gb, float ab, float cc, float u);
void nmco_spin(float cost, float u_psi, float* ux, float* uy, float* uz);
void nmco_spin_w(float cost, uint32_t w_psi, float* ux, float* uy, float* uz);
void nmco_sincos2pi_bits(uint32_t x, float* s, float* c);
void nmco_sincos2pi_bits_n(const uint32_t* x, float* s, float* c, int64_t n);

/* vectorised helpers for the tests (n elements each) */
void nmco_logf_n(const float* x, float* y, int64_t n);
void nmco_logf_tab_n(const float* x, float* y, int64_t n);
void nmco_expf_n(const float* x, float* y, int64_t n);
void nmco_expm1f_n(const float* x, float* y, int64_t n);
void nmco_sincos2pi_n(const float* u, float* s, float* c, int64_t n);

/*
 * Transport photons [first, first + n) of the configuration and accumulate
 * into `out` (arrays as documented for nmc_results in include/nmc.h; arrays
 * must be zeroed by the caller before the first call).  n_threads <= 0 uses
 * OpenMP's default.  Returns NMC_OK or NMC_ERR_ARG.
 */
int nmco_mcxyz(const RunParams* rp, const TissueParams* tp, const uint8_t* v,
               const nmc_ext_config* ext, uint64_t first, uint64_t n, int n_threads,
               nmc_results* out);
/* nmco_mcxyz plus the per-voxel sensitivities of the absorbed weight (the
 * fluence array) to mua and mus of tissues ids[0..n_ids) (n_ids <=
 * NMC_SENS_MAX, distinct ids): sens[0][j][voxel] += d/dmua, sens[1][j][voxel]
 * += d/dmus of tissue ids[j] (2 n_ids nvox doubles, zeroed by the caller). */
int nmco_mcxyz_sens(const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                    const nmc_ext_config* ext, uint64_t first, uint64_t n, int n_threads,
                    nmc_results* out, const int32_t* ids, int n_ids, double* sens);
int nmco_mcsub(const RunParams* rp, const TissueParams* tp, const nmc_ext_config* ext,
               uint64_t first, uint64_t n, int n_threads, nmc_results* out);

/* sens (may be NULL): [n_stacks][2][MAXL][MAXL] += d(absorbed in layer l) /
 * d(mua_t | mus_t), the perturbation-MC sums over every deposit. */
int nmco_layered(const nmc_layered_config* cfg, const nmc_stack* stacks, int n_stacks,
                 uint64_t first, int n_threads, double* rd, double* absorbed, nmc_totals* totals,
                 uint64_t* n_rec, uint64_t cap, uint32_t* rec_item, uint32_t* rec_bin,
                 float* rec_w, float* rec_L, uint32_t* rec_K, double* sens);
void nmco_layered_backward(const nmc_layered_config* cfg, const nmc_stack* stacks, int n_stacks,
                           uint64_t n_rec, uint64_t cap, const uint32_t* rec_item,
                           const uint32_t* rec_bin, const float* rec_w, const float* rec_L,
                           const uint32_t* rec_K, const double* grad_rd, double* grad_mua,
                           double* grad_mus);

/* Launch states of photons [first, first + n) (test entry points for the
 * source models): mcxyz 8 floats per photon (x, y, z, ux, uy, uz, W,
 * specular loss), mcsub 7 (x, y, z, ux, uy, uz, specular reflectance). */
int nmco_xyz_source(const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                    const nmc_ext_config* ext, uint64_t first, uint64_t n, float* out);
int nmco_sub_source(const RunParams* rp, const TissueParams* tp, const nmc_ext_config* ext,
                    uint64_t first, uint64_t n, float* out);

/* Seconds of the last nmco_mcxyz / nmco_mcsub transport loop (CLOCK_MONOTONIC). */
double nmco_last_seconds(void);
int nmco_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
