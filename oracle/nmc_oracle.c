/*
 * nmc_oracle.c - CPU restatement of NMC's photon-transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never linked into
 * it.  Written photon-by-photon in the shape of the reference kernels so that
 * it reads against them line by line; the GPU library reorganises the same
 * arithmetic into a lane-refilling state machine.
 *
 * Algorithm sources (paths relative to the reference root):
 *   mcxyz  NMC/NMCCode/kernels.metal:321-754   (MCXYZKernel)
 *   mcsub  NMC/NMCCode/kernels.metal:59-319    (MCSubKernel)
 *   two-term phase function kernels.metal:894-905,1053-1070 (MCXYZKernelTT)
 *   Fresnel / Snell NMC/NMCCode/helpers.metal:243-314, FltEq helpers.metal:21-25
 *   constants NMC/NMCCode/definitions.h:16-82
 *
 * Deliberate departures from the reference, shared with the GPU library:
 *   - RNG: Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as
 *     easy as 1, 2, 3") keyed by (seed, stream) with counter (photon id,
 *     event index, tag) instead of RandomGen (helpers.metal:319-371).
 *   - log/exp/expm1/sin/cos are this file's own float routines (below) so
 *     that CPU and GPU produce identical trajectories (the voxel walk's step
 *     log by the table of ../nmc_amd/csrc/nmc_tables.h, whose rows the
 *     tests recompute from their definition); +,-,*,/,sqrt, fma are
 *     IEEE-754 binary32 with round-to-nearest, no contraction.
 *   - Voxel stepping tracks integer voxel indices (a DDA walk) instead of the
 *     `ls` face nudge of kernels.metal:538 and SameVoxel (helpers.metal:27-39);
 *     the photon still deposits into the voxel the piece of step lies in.
 *   - absorb = -W*expm1(-mua*s) instead of W*(1-exp(-mua*s)) (same quantity,
 *     no cancellation for small mua*s).
 *   - Reference defects are not reproduced: mcsub HG overwrite (kernels.metal
 *     :252), bflag never restored (:487), zsurf sleft reassignment (:553-571),
 *     unreflected photons outside the detector aperture (:580-623), hard
 *     kills W<1e-6 / num_scatt>10000 (:747-748), RAY_DEPTH truncation
 *     (:503,683).  A per-photon event watchdog (max_events) replaces the kills
 *     and its weight is tallied separately.
 */
#include "nmc_oracle.h"
#include "../nmc_amd/csrc/nmc_tables.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define INF_F __builtin_inff()

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw 2011).                          */
/* ------------------------------------------------------------------------ */
/* Planted mistakes for the oracle's own tests (tests/test_oracle_pins.py
 * builds copies with -DNMCO_MUTANT=k and checks that each pin fails on them;
 * the oracle proper is built with NMCO_MUTANT = 0):
 *   1  flat beam radius R*xi instead of R*sqrt(xi)
 *   2  the remaining step keeps its physical length across a change of
 *      tissue or layer (no rescale of the optical depth by mus)
 *   3  mcsub drops its weight at the position before the hop
 *   4  mcsub bins counted from 1 (the first bin left empty)
 *   5  the absorbed-weight sensitivity to mus drops the score's growth
 *      within a segment (the W s h(x) term)
 *   6  the voxel sensitivity to mua drops the own-tissue term W s e^{-x}
 *   7  the voxel sensitivity to mus drops the own-tissue W s h(x) term  */
#ifndef NMCO_MUTANT
#define NMCO_MUTANT 0
#endif

void nmco_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Stream of one photon: counter (id lo, id hi, event index, tag). */
static inline void block(const uint32_t key[2], uint64_t pid, uint32_t ev, uint32_t tag,
                         uint32_t out[4])
{
    uint32_t ctr[4] = {(uint32_t)pid, (uint32_t)(pid >> 32), ev, tag};
    nmco_philox4x32_10(ctr, key, out);
}

/* uniform in [0,1): top 24 bits, the reference's (u32>>8)*2^-24 mapping */
static inline float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }
/* uniform in (0,1]: used where the reference loops `while ((rnd=rand())<=0)` */
static inline float u01nz(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }
/* a fifth uniform in [0,1) from the low bytes of words 0-2, which the four
 * top-24-bit uniforms of the block leave unused (mcsub's escape test) */
static inline float u01_low(const uint32_t b[4])
{
    uint32_t v = ((b[0] & 0xffu) << 16) | ((b[1] & 0xffu) << 8) | (b[2] & 0xffu);
    return (float)v * 0x1p-24f;
}

/* ------------------------------------------------------------------------ */
/* Transcendentals.                                                          */
/* ------------------------------------------------------------------------ */
static inline float bits_f(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline uint32_t f_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/*
 * Natural log of a positive normal float.  x = 2^k * m with m in
 * [sqrt(2)/2, sqrt(2)); log(m) = log(1+f) = 2 atanh(s), s = f/(2+f), written as
 * f - (hfsq - s*(hfsq + R)) with R the atanh series 2z/3 + 2z^2/5 + 2z^3/7 +
 * 2z^4/9, z = s^2 (the reduction of Cody & Waite / fdlibm's e_logf, with the
 * plain Taylor coefficients).  k*ln2 is split hi+lo, hi with 7 trailing zero
 * bits so k*hi is exact.
 */
float nmco_logf(float x)
{
    if (!(x > 0.0f)) return x == 0.0f ? -INF_F : __builtin_nanf("");
    uint32_t b = f_bits(x);
    int k = (int)(b >> 23) - 127;
    float m = bits_f((b & 0x007fffffu) | 0x3f800000u); /* [1,2) */
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; k += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = z * fmaf(z, fmaf(z, fmaf(z, 0x1.c71c72p-3f, 0x1.24924ap-2f), 0x1.99999ap-2f),
                       0x1.555556p-1f);
    float hfsq = 0.5f * f * f;
    float fk = (float)k;
    return fk * 0x1.62e4p-1f - ((hfsq - (s * (hfsq + R) + fk * 0x1.7f7d1cp-20f)) - f);
}

/*
 * The voxel walk's step-length log, by table (Tang's reduction): x = 2^k m'
 * with m' in [3/4, 3/2) (the mantissa, halved with k + 1 when >= 3/2); the top
 * seven mantissa bits pick row i = (rc, L_hi, L_lo) of NMC_LOG_TABLE (built
 * as gen_tables.py states: rc = 1/midpoint with 8 significant bits, so
 * r = m' rc - 1 is exact; -log(rc) split so that k ln2_hi + L_hi is exact);
 * log(x) = (k ln2_hi + L_hi) + ((k ln2_lo + L_lo) + log1p(r)),
 * log1p(r) = r + r^2 (-1/2 + r (1/3 - r/4)).  Positive normal x only.
 */
static const float log_table[128][4] = NMC_LOG_TABLE;

float nmco_logf_tab(float x)
{
    uint32_t b = f_bits(x);
    uint32_t i = (b >> 16) & 0x7fu;
    int up = i >= 64u;
    float m = bits_f((b & 0x007fffffu) | (up ? 0x3f000000u : 0x3f800000u));
    float fk = (float)((int)(b >> 23) - (up ? 126 : 127));
    const float* t = log_table[i];
    float r = fmaf(m, t[0], -1.0f);
    float p = fmaf(r, fmaf(r, -0x1p-2f, 0x1.555556p-2f), -0.5f);
    float q = fmaf(r * r, p, r);
    float hi = fmaf(fk, 0x1.62e4p-1f, t[1]);
    float lo = fmaf(fk, 0x1.7f7d1cp-20f, t[2]);
    return hi + (lo + q);
}

/* expm1(r) for |r| <= ln2/2: r + r^2 p(r) with p of degree 5 in Horner
 * form; p's coefficients approximate the minimax fit of (expm1(r) - r) / r^2
 * on that interval in relative error (2.4e-10), rounded to float. */
static inline float em1_core(float r)
{
    float p = fmaf(r, 0x1.a032bcp-13f, 0x1.6d71e6p-10f);
    p = fmaf(r, p, 0x1.11118ap-7f);
    p = fmaf(r, p, 0x1.5554b2p-5f);
    p = fmaf(r, p, 0x1.555554p-3f);
    p = fmaf(r, p, 0.5f);
    return fmaf(r * r, p, r);
}

/*
 * exp(x) = 2^k exp(r), k = rint(x/ln2), r = x - k ln2 (Cody-Waite, two fma
 * steps).  Results below 2^-126 (x < -87) are flushed to 0, above FLT_MAX
 * give +inf.
 */
float nmco_expf(float x)
{
    if (x < -87.0f) return 0.0f;
    if (x > 88.5f) return INF_F;
    float k = rintf(x * 0x1.715476p+0f);
    float r = fmaf(-k, 0x1.62e4p-1f, x);
    r = fmaf(-k, 0x1.7f7d1cp-20f, r);
    float e = 1.0f + em1_core(r);
    int ki = (int)k;
    if (ki > 127) { e = e * 2.0f; ki -= 1; }
    return e * bits_f((uint32_t)(ki + 127) << 23);
}

float nmco_expm1f(float x)
{
    if (fabsf(x) <= 0x1.62e430p-2f) return em1_core(x);
    return nmco_expf(x) - 1.0f;
}

/*
 * sin(2 pi u), cos(2 pi u) for u in [0, 1]: exact quarter-turn reduction
 * u = q/4 + f, |f| <= 1/8, then Taylor series of sin(2 pi f) to f^9 and of
 * cos(2 pi f) to f^10 (coefficients (2pi)^n/n! rounded to float).
 */
void nmco_sincos2pi(float u, float* s, float* c)
{
    float q = rintf(u * 4.0f);
    float f = u - q * 0.25f;
    float t = f * f;
    float ps = fmaf(t, 0x1.507834p+5f, -0x1.32d2ccp+6f);
    ps = fmaf(t, ps, 0x1.466bc6p+6f);
    ps = fmaf(t, ps, -0x1.4abbcep+5f);
    ps = fmaf(t, ps, 0x1.921fb6p+2f);
    ps = ps * f;
    float pc = fmaf(t, -0x1.a6d1f2p+4f, 0x1.e1f506p+5f);
    pc = fmaf(t, pc, -0x1.55d3c8p+6f);
    pc = fmaf(t, pc, 0x1.03c1f0p+6f);
    pc = fmaf(t, pc, -0x1.3bd3ccp+4f);
    pc = fmaf(t, pc, 1.0f);
    switch (((int)q) & 3) {
    case 0: *s = ps; *c = pc; break;
    case 1: *s = pc; *c = -ps; break;
    case 2: *s = -ps; *c = -pc; break;
    default: *s = -pc; *c = ps; break;
    }
}

void nmco_logf_n(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = nmco_logf(x[i]); }
void nmco_logf_tab_n(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = nmco_logf_tab(x[i]); }
void nmco_expf_n(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = nmco_expf(x[i]); }
void nmco_expm1f_n(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = nmco_expm1f(x[i]); }
void nmco_sincos2pi_n(const float* u, float* s, float* c, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) nmco_sincos2pi(u[i], s + i, c + i);
}

/* ------------------------------------------------------------------------ */
/* Optics (helpers.metal:21-25, 243-314).                                    */
/* ------------------------------------------------------------------------ */
static inline int flt_eq(float a, float b)
{
    float e = b * 0x1p-23f * 2.0f; /* FLT_EPSILON * (1 + error), error = 1 */
    return (a >= b - e) && (a <= b + e);
}

/* Unpolarised Fresnel reflectance n1 -> n2 at incidence cosine ca1; writes
 * the transmission cosine (helpers.metal:243-291). */
float nmco_fresnel(float n1, float n2, float ca1, float* ca2)
{
    if (flt_eq(n1, n2)) { *ca2 = ca1; return 0.0f; }
    if (ca1 > 1.0f - 1.0e-6f) {
        *ca2 = ca1;
        float r = (n2 - n1) / (n2 + n1);
        return r * r;
    }
    if (ca1 < 1.0e-6f) { *ca2 = 0.0f; return 1.0f; }
    float sa1 = sqrtf(1.0f - ca1 * ca1);
    float sa2 = n1 * sa1 / n2;
    if (sa2 >= 1.0f) { *ca2 = 0.0f; return 1.0f; }
    float c2 = sqrtf(1.0f - sa2 * sa2);
    *ca2 = c2;
    float cap = ca1 * c2 - sa1 * sa2;
    float cam = ca1 * c2 + sa1 * sa2;
    float sap = sa1 * c2 + ca1 * sa2;
    float sam = sa1 * c2 - ca1 * sa2;
    return 0.5f * sam * sam * (cam * cam + cap * cap) / (sap * sap * cam * cam);
}

/* Snell refraction of a direction from index n1 into n2 (helpers.metal:293-314). */
static void snell(float n1, float n2, float* ux, float* uy, float* uz)
{
    float c = 1.0f - (1.0f - *uz * *uz) * n1 * n1 / n2 / n2;
    if (c < 0.0f) {
        if (fabsf(*uz) > 1.0f) *uz = *uz < 0.0f ? -1.0f : 1.0f;
        return;
    }
    *ux = *ux * n1 / n2;
    *uy = *uy * n1 / n2;
    *uz = sqrtf(c);
}

/* ------------------------------------------------------------------------ */
/* Scattering (kernels.metal:695-730, 1053-1070).                            */
/* ------------------------------------------------------------------------ */
/* HG sampling (kernels.metal:695-703) from constants derived in double:
 * a = 1+g^2, b = 1-g^2, c = 1-g, d = 2g, e = 1/(2g); cos = (a - t^2) e with
 * t = b / (c + d u); mode 1: g == 0 (isotropic), 2: g == 1 (forward). */
typedef struct { float a, b, c, d, e; int mode; } hg_coef;

static void hg_prepare(float g, hg_coef* h)
{
    h->mode = flt_eq(g, 0.0f) ? 1 : (flt_eq(g, 1.0f) ? 2 : 0);
    double gd = g;
    h->a = (float)(1.0 + gd * gd);
    h->b = (float)(1.0 - gd * gd);
    h->c = (float)(1.0 - gd);
    h->d = (float)(2.0 * gd);
    h->e = h->mode == 0 ? (float)(1.0 / (2.0 * gd)) : 0.0f;
}

static inline float hg_sample(const hg_coef* h, float u)
{
    float ct;
    if (h->mode == 1) ct = 2.0f * u - 1.0f;
    else if (h->mode == 2) ct = 1.0f;
    else {
        float t = h->b / (h->c + h->d * u);
        ct = (h->a - t * t) * h->e;
    }
    return fminf(fmaxf(ct, -1.0f), 1.0f);
}

float nmco_hg_cos(float g, float u)
{
    hg_coef h;
    hg_prepare(g, &h);
    return hg_sample(&h, u);
}

/* 1/mu rounded once from double; FLT_MAX for mu = 0. */
static inline float inv_pos(float mu) { return mu > 0.0f ? (float)(1.0 / (double)mu) : 3.40282347e+38f; }

/* Per-tissue constants of the two-term function, derived in double on the
 * host (q1, q2, q3 of kernels.metal:900-905). */
typedef struct { float cc, c1f, c2f, iq2f, iq3f, ef, c1b, c2b, iq2b, iq3b, eb; int isof, isob; } tt_coef;

static void tt_prepare(float gf, float af, float gb, float ab, float cc, tt_coef* t)
{
    t->cc = cc;
    t->isof = gf == 0.0f;
    t->isob = gb == 0.0f;
    double g = gf, a = af;
    if (!t->isof) {
        t->c1f = (float)((1.0 + g * g) / (2.0 * g));
        t->c2f = (float)(1.0 / (2.0 * g));
        t->iq2f = (float)(1.0 / pow(1.0 - g, 2.0 * a));
        t->iq3f = (float)(1.0 / pow(1.0 + g, 2.0 * a));
        t->ef = (float)(-1.0 / a);
    }
    g = gb; a = ab;
    if (!t->isob) {
        t->c1b = (float)((1.0 + g * g) / (2.0 * g));
        t->c2b = (float)(1.0 / (2.0 * g));
        t->iq2b = (float)(1.0 / pow(1.0 - g, 2.0 * a));
        t->iq3b = (float)(1.0 / pow(1.0 + g, 2.0 * a));
        t->eb = (float)(-1.0 / a);
    }
}

static inline float pow_neg(float b, float e)
{
    if (e == -2.0f) return 1.0f / (b * b);
    return nmco_expf(e * nmco_logf(b));
}

/* One uniform picks the branch (forward with probability CC) and, rescaled
 * to [0,1) within the branch, samples it. */
static float tt_cos(const tt_coef* t, float u)
{
    float ct;
    if (u < t->cc) {
        float v = u / t->cc;
        if (t->isof) ct = 2.0f * v - 1.0f;
        else ct = t->c1f - t->c2f * pow_neg(v * t->iq2f + (1.0f - v) * t->iq3f, t->ef);
    } else {
        float v = (u - t->cc) / (1.0f - t->cc);
        if (t->isob) ct = 2.0f * v - 1.0f;
        else ct = -(t->c1b - t->c2b * pow_neg(v * t->iq2b + (1.0f - v) * t->iq3b, t->eb));
    }
    return fminf(fmaxf(ct, -1.0f), 1.0f);
}

float nmco_tt_cos(float gf, float af, float gb, float ab, float cc, float u)
{
    tt_coef t;
    tt_prepare(gf, af, gb, ab, cc, &t);
    return tt_cos(&t, u);
}

/*
 * sin(2 pi u), cos(2 pi u) of the uniform u = (x >> 8) 2^-24 of random word x
 * (the walks' azimuth): u = j/64 + f, j = x >> 26, f = ((x >> 8) mod 2^18)
 * 2^-24 (exact), a = 2 pi f < 0.0982; with row j = (sin, cos)(2 pi j/64) of
 * NMC_SINCOS_TABLE (double values rounded to float, quarter turns exact):
 * s = s_j cos a + c_j sin a, c = c_j cos a - s_j sin a, sin a = a + a^3 (-1/6
 * + a^2/120), cos a = 1 + a^2 (-1/2 + a^2/24).
 */
static const float sincos_table[64][2] = NMC_SINCOS_TABLE;

void nmco_sincos2pi_bits(uint32_t x, float* s, float* c)
{
    const float* t = sincos_table[x >> 26];
    float a = (float)((x >> 8) & 0x3ffffu) * 0x1.921fb6p-22f;
    float a2 = a * a;
    float sa = fmaf(a2 * a, fmaf(a2, 0x1.111112p-7f, -0x1.555556p-3f), a);
    float ca = fmaf(a2, fmaf(a2, 0x1.555556p-5f, -0.5f), 1.0f);
    *s = fmaf(t[0], ca, t[1] * sa);
    *c = fmaf(t[1], ca, -(t[0] * sa));
}

void nmco_sincos2pi_bits_n(const uint32_t* x, float* s, float* c, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) nmco_sincos2pi_bits(x[i], s + i, c + i);
}

/* New direction after deflection cos(theta) = cost and azimuth psi given by
 * sp = sin(psi), cp = cos(psi) (kernels.metal:705-730; sin(psi) is computed
 * directly).  Pole formula where the reference's float test 1 - |uz| <= 1e-12
 * (kernels.metal:713) holds, i.e. |uz| >= 1, or where ux^2 + uy^2 < 2^-100 (a
 * direction slightly off unit length), which bounds the general formula's
 * quotient.  nmco_spin: psi = 2 pi u_psi; nmco_spin_w (the walks): psi from
 * random word w_psi by nmco_sincos2pi_bits. */
static void spin_sc(float cost, float sp, float cp, float* ux, float* uy, float* uz)
{
    float x = *ux, y = *uy, z = *uz;
    float r2 = x * x + y * y;
    if (fabsf(z) >= 1.0f || !(r2 >= 0x1p-100f)) {
        float sint = sqrtf(1.0f - cost * cost);
        *ux = sint * cp;
        *uy = sint * sp;
        *uz = z >= 0.0f ? cost : -cost;
    } else {
        /* st = sin(theta) / sqrt(ux^2 + uy^2) as one square root; the uz term
         * sin(theta) * sqrt(ux^2 + uy^2) is st * (ux^2 + uy^2) */
        float st = sqrtf((1.0f - cost * cost) / r2);
        *ux = st * (x * z * cp - y * sp) + x * cost;
        *uy = st * (y * z * cp + x * sp) + y * cost;
        *uz = -(st * cp * r2) + z * cost;
    }
}

void nmco_spin(float cost, float u_psi, float* ux, float* uy, float* uz)
{
    float sp, cp;
    nmco_sincos2pi(u_psi, &sp, &cp);
    spin_sc(cost, sp, cp, ux, uy, uz);
}

void nmco_spin_w(float cost, uint32_t w_psi, float* ux, float* uy, float* uz)
{
    float sp, cp;
    nmco_sincos2pi_bits(w_psi, &sp, &cp);
    spin_sc(cost, sp, cp, ux, uy, uz);
}

/* ------------------------------------------------------------------------ */
/* Shared run bookkeeping.                                                   */
/* ------------------------------------------------------------------------ */
static double g_last_seconds;
double nmco_last_seconds(void) { return g_last_seconds; }
int nmco_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void ext_defaults(const nmc_ext_config* in, nmc_ext_config* e)
{
    memset(e, 0, sizeof(*e));
    e->struct_size = sizeof(*e);
    e->version = NMC_EXT_VERSION;
    e->seed = 1;
    e->threshold = NMC_THRESHOLD;
    e->chance = NMC_CHANCE;
    e->max_events = 100000000u;
    e->nr = 100; e->nz = 100; e->na = 30;
    e->dr = 0.01f; e->dz = 0.01f;
    if (in) *e = *in;
}

static void totals_add(nmc_totals* a, const nmc_totals* b)
{
    a->specular += b->specular; a->rd += b->rd; a->rd_outside += b->rd_outside;
    a->absorbed += b->absorbed; a->absorbed_outside += b->absorbed_outside;
    a->tt += b->tt; a->lateral += b->lateral; a->top += b->top;
    a->roulette_gain += b->roulette_gain; a->roulette_loss += b->roulette_loss;
    a->capped += b->capped;
    a->photons += b->photons; a->detections += b->detections; a->surface_hits += b->surface_hits;
    a->escapes_bottom += b->escapes_bottom; a->escapes_side += b->escapes_side;
    a->escapes_top += b->escapes_top; a->roulette_kills += b->roulette_kills;
    a->roulette_survivals += b->roulette_survivals; a->capped_photons += b->capped_photons;
    a->substeps_same += b->substeps_same; a->substeps_cross += b->substeps_cross;
    a->interactions += b->interactions;
}

/* Roulette (kernels.metal:740-744).  Returns 0 if the photon is killed. */
static inline int roulette(float* W, float u, float thr, float chance, nmc_totals* t)
{
    if (*W < thr) {
        if (u <= chance) {
            float nw = *W / chance;
            t->roulette_gain += (double)nw - (double)*W;
            t->roulette_survivals++;
            *W = nw;
        } else {
            t->roulette_loss += (double)*W;
            t->roulette_kills++;
            *W = 0.0f;
            return 0;
        }
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* mcxyz (kernels.metal:321-754).                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    int Nx, Ny, Nz;
    float dx, dy, dz, hx, hy;
    int bmode, det_state, mcflag, launchflag, two_term, count_visits;
    float zsurf, n1, n2, rmin, rmax, cos_na;
    float xs, ys, zs, ux0, uy0, uz0, radius, waist, zfocus;
    float thr, chance;
    uint32_t max_events;
    uint32_t key[2];
    const uint8_t* V;
    float mua[NMC_NTISS], mus[NMC_NTISS], inv_mus[NMC_NTISS];
    hg_coef hg[NMC_NTISS];
    tt_coef tt[NMC_NTISS];
    /* sensitivities: tracked tissues ids[0..n_sens), slot[id] (-1: untracked) */
    int n_sens, sens_ids[NMC_SENS_MAX], slot[NMC_NTISS];
} xyz_ctx;

typedef struct { double* F; double* ryx; uint32_t* det; nmc_totals t; double* sens; } xyz_acc;

static float score_growth(float x, float em);

/* Per-voxel absorbed-weight sensitivities (no reference counterpart; the
 * reference's absorbed tally is F3D, kernels.metal:530,550).  The weight a
 * voxel visit in tissue k deposits is the path integral of mua_k W along its
 * segments; as for the layered walk (lay_visit / sens_fold below), with L_t,
 * K_t the photon's path length and scatterings in tissue t before a point, a
 * segment of length s entered with weight W adds
 *   dD/dmua_t = -L_t D + [t = k] W s e^{-x},
 *   dD/dmus_t = (K_t/mus_t - L_t) D - [t = k] W s h(x),  x = mua_k s,
 * and during the visit only L_k and K_k change, so the visit keeps the sums
 * sp = sum D L_k, sq = sum D K_k, sza = sum W s e^{-x}, szs = sum W s h(x)
 * and the deposit total D.  sens[0][j][v], sens[1][j][v]: d/dmua and d/dmus
 * of tissue ids[j] for voxel v. */
typedef struct { double D; float sp, sq, sza, szs; } xyz_visit;

static void xyz_sens_fold(const xyz_ctx* c, int k, const float* Ls, const uint32_t* Ks,
                          const xyz_visit* v, int64_t vox, int64_t nv, double* sens)
{
    if (v->D == 0.0 && v->sza == 0.0f) return;
    const int sk = c->slot[k];
    for (int j = 0; j < c->n_sens; ++j) {
        double ga, gs;
        if (j == sk) {
            ga = -(double)v->sp + (NMCO_MUTANT == 6 ? 0.0 : (double)v->sza);
            gs = (double)v->sq * (double)c->inv_mus[k] - (double)v->sp -
                 (NMCO_MUTANT == 7 ? 0.0 : (double)v->szs);
        } else {
            if (v->D == 0.0 || (Ls[j] == 0.0f && Ks[j] == 0u)) continue;
            ga = -(double)Ls[j] * v->D;
            gs = ((double)Ks[j] * (double)c->inv_mus[c->sens_ids[j]] - (double)Ls[j]) * v->D;
        }
        sens[(int64_t)j * nv + vox] += ga;
        sens[((int64_t)c->n_sens + j) * nv + vox] += gs;
    }
}

static inline float xface(int i, float h, float d) { return ((float)i - h) * d; }

/* LAUNCH (kernels.metal:366-463): start position, direction, weight and the
 * specular loss (focused Gaussian beam entering the tissue, else 0) of photon
 * pid.  o = {x, y, z, ux, uy, uz, W, spec}. */
static void xyz_source(const xyz_ctx* c, uint64_t pid, float o[8])
{
    uint32_t b[4];
    float x, y, z, ux, uy, uz, W = 1.0f, spec = 0.0f;
    if (c->launchflag == 1) {
        x = c->xs; y = c->ys; z = c->zs;
        ux = c->ux0; uy = c->uy0; uz = c->uz0;
    } else {
        block(c->key, pid, 0u, 0u, b);
        if (c->mcflag == 0 || c->mcflag == 1) {
            float r;
            if (c->mcflag == 0) r = NMCO_MUTANT == 1 ? c->radius * u01nz(b[0]) : c->radius * sqrtf(u01nz(b[0]));
            else r = c->radius * sqrtf(-nmco_logf(u01nz(b[0])));
            float sp, cp;
            nmco_sincos2pi(u01nz(b[1]), &sp, &cp);
            x = c->xs + r * cp;
            y = c->ys + r * sp;
            z = c->zs;
            float rf = c->waist * sqrtf(u01nz(b[2]));
            nmco_sincos2pi(u01nz(b[3]), &sp, &cp);
            float xf = rf * cp, yf = rf * sp;
            float ddx = x - xf, ddy = y - yf;
            float tmp = sqrtf(ddx * ddx + ddy * ddy + c->zfocus * c->zfocus);
            ux = -ddx / tmp;
            uy = -ddy / tmp;
            uz = sqrtf(1.0f - ux * ux - uy * uy);
            if (c->mcflag == 1) {
                float ca2;
                float rsp = nmco_fresnel(c->n2, c->n1, fabsf(uz), &ca2);
                spec = W * rsp;
                W = W * (1.0f - rsp);
                snell(c->n2, c->n1, &ux, &uy, &uz);
            }
        } else if (c->mcflag == 2) {
            float ct = 1.0f - 2.0f * u01(b[0]);
            float st = sqrtf(1.0f - ct * ct);
            float sp, cp;
            nmco_sincos2pi(u01(b[1]), &sp, &cp);
            x = c->xs; y = c->ys; z = c->zs;
            ux = st * cp; uy = st * sp; uz = ct;
        } else if (c->mcflag == 3) {
            x = c->radius * (u01nz(b[0]) * 2.0f - 1.0f);
            y = c->radius * (u01nz(b[1]) * 2.0f - 1.0f);
            z = c->zs;
            ux = 0.0f; uy = 0.0f; uz = 1.0f;
        } else {
            x = 0.0f; y = 0.0f; z = c->zs;
            ux = 0.0f; uy = 0.0f; uz = 1.0f;
        }
    }
    o[0] = x; o[1] = y; o[2] = z; o[3] = ux; o[4] = uy; o[5] = uz; o[6] = W; o[7] = spec;
}

static void xyz_photon(const xyz_ctx* c, uint64_t pid, xyz_acc* a)
{
    nmc_totals* t = &a->t;
    uint32_t b[4];
    t->photons++;

    float o[8];
    xyz_source(c, pid, o);
    float x = o[0], y = o[1], z = o[2], ux = o[3], uy = o[4], uz = o[5], W = o[6];
    t->specular += (double)o[7];

    int ix = (int)floorf(c->hx + x / c->dx);
    int iy = (int)floorf(c->hy + y / c->dy);
    int iz = (int)floorf(z / c->dz);
    int inside = ix >= 0 && ix < c->Nx && iy >= 0 && iy < c->Ny && iz >= 0 && iz < c->Nz;
    if (!inside && (c->bmode == 1 || (c->bmode == 2 && iz < 0))) {
        if (iz >= c->Nz) { t->tt += (double)W; t->escapes_bottom++; }
        else if (iz < 0) { t->top += (double)W; t->escapes_top++; }
        else { t->lateral += (double)W; t->escapes_side++; }
        return;
    }
#define CLAMPI(v, n) ((v) < 0 ? 0 : ((v) >= (n) ? (n) - 1 : (v)))
#define VIDX(i, j, k) ((int64_t)(k) * c->Ny * c->Nx + (int64_t)(i) * c->Ny + (j))
    int tis = c->V[VIDX(CLAMPI(ix, c->Nx), CLAMPI(iy, c->Ny), CLAMPI(iz, c->Nz))];
    float rux = 1.0f / ux, ruy = 1.0f / uy, ruz = 1.0f / uz;

    uint32_t ev = 1;
    block(c->key, pid, ev, 0u, b);
    float sleft = -nmco_logf_tab(u01nz(b[0]));
    uint32_t events = 0;
    /* Deposits of one voxel visit are summed in float (a visit is one to a
     * few pieces) and added, in double, to F and to the absorbed total when
     * the photon leaves the voxel or ends. */
    float vdep = 0.0f;
    /* sensitivities: path length and scatterings per tracked tissue, and the
     * current visit's sums */
    float Ls[NMC_SENS_MAX] = {0};
    uint32_t Ks[NMC_SENS_MAX] = {0};
    xyz_visit vs = {0.0, 0.0f, 0.0f, 0.0f, 0.0f};
    const int64_t nvox = (int64_t)c->Nx * c->Ny * c->Nz;
#define FLUSH()                                                                        \
    do {                                                                               \
        t->absorbed += (double)vdep;                                                   \
        if (inside && a->F && (c->count_visits || vdep != 0.0f))                       \
            a->F[VIDX(ix, iy, iz)] += c->count_visits ? 1.0 : (double)vdep;             \
        if (a->sens && inside) {                                                       \
            vs.D = (double)vdep;                                                       \
            xyz_sens_fold(c, tis, Ls, Ks, &vs, VIDX(ix, iy, iz), nvox, a->sens);        \
        }                                                                              \
        vdep = 0.0f;                                                                   \
        vs.sp = vs.sq = vs.sza = vs.szs = 0.0f;                                        \
    } while (0)

    for (;;) {
        /* HOP: one piece of the step (kernels.metal:505-685) */
        float mua = c->mua[tis], mus = c->mus[tis];
        float s_int = sleft * c->inv_mus[tis];
        float tx = INF_F, ty = INF_F, tz = INF_F;
        if (ux > 0.0f) tx = (xface(ix + 1, c->hx, c->dx) - x) * rux;
        else if (ux < 0.0f) tx = (xface(ix, c->hx, c->dx) - x) * rux;
        if (uy > 0.0f) ty = (xface(iy + 1, c->hy, c->dy) - y) * ruy;
        else if (uy < 0.0f) ty = (xface(iy, c->hy, c->dy) - y) * ruy;
        if (uz > 0.0f) tz = (xface(iz + 1, 0.0f, c->dz) - z) * ruz;
        else if (uz < 0.0f) tz = (xface(iz, 0.0f, c->dz) - z) * ruz;
        tx = fmaxf(tx, 0.0f); ty = fmaxf(ty, 0.0f); tz = fmaxf(tz, 0.0f);
        int axis;
        float s_face;
        if (tx <= ty && tx <= tz) { axis = 0; s_face = tx; }
        else if (ty <= tz) { axis = 1; s_face = ty; }
        else { axis = 2; s_face = tz; }
        float s_surf = INF_F;
        if (c->det_state == NMC_DET_REFLECTANCE) {
            if (uz < 0.0f && z > c->zsurf) s_surf = fmaxf((c->zsurf - z) * ruz, 0.0f);
        } else {
            if (uz > 0.0f && z < c->zsurf) s_surf = fmaxf((c->zsurf - z) * ruz, 0.0f);
        }
        int kind; /* 0 interaction, 1 surface, 2 face */
        float s;
        if (s_int <= s_face && s_int <= s_surf) { kind = 0; s = s_int; }
        else if (s_surf <= s_face) { kind = 1; s = s_surf; }
        else { kind = 2; s = s_face; }

        x = fmaf(s, ux, x);
        y = fmaf(s, uy, y);
        z = fmaf(s, uz, z);
        /* DROP (kernels.metal:523-533) */
        const float xa = mua * s;
        float em = nmco_expm1f(-xa);
        float dep = -(W * em);
        if (a->sens && c->slot[tis] >= 0) {
            const int sl = c->slot[tis];
            vs.sp = vs.sp + dep * Ls[sl];
            vs.sq = vs.sq + dep * (float)Ks[sl];
            const float ws = W * s;
            vs.sza = vs.sza + ws * (1.0f + em);
            vs.szs = vs.szs + ws * score_growth(xa, em);
            Ls[sl] = Ls[sl] + s;
        }
        W = W - dep;
        if (inside) {
            vdep = vdep + dep;
        } else {
            t->absorbed_outside += (double)dep;
        }
        events++;
        int alive = 1;
        if (kind == 0) {
            sleft = 0.0f;
            t->substeps_same++;
        } else {
            t->substeps_cross++;
            sleft = sleft - s * mus;
            if (!(sleft > 0.0f)) sleft = 0.0f;
            if (kind == 1) {
                /* detector plane (kernels.metal:556-630) */
                z = c->zsurf;
                t->surface_hits++;
                float uzt;
                float rsp = nmco_fresnel(c->n1, c->n2, fabsf(uz), &uzt);
                if (rsp < 1.0f) {
                    float wt = W * (1.0f - rsp);
                    float ro = sqrtf(x * x + y * y);
                    if (ro >= c->rmin && ro <= c->rmax && uzt >= c->cos_na) {
                        int px = CLAMPI(ix, c->Nx), py = CLAMPI(iy, c->Ny);
                        if (a->ryx) a->ryx[(int64_t)px * c->Ny + py] += (double)wt;
                        if (a->det) a->det[(int64_t)px * c->Ny + py] += 1u;
                        t->rd += (double)wt;
                        t->detections++;
                    } else {
                        t->rd_outside += (double)wt;
                    }
                }
                uz = -uz;
                ruz = -ruz;
                W = W * rsp;
                if (!(W > 0.0f)) alive = 0;
            } else {
                /* voxel face (kernels.metal:641-679) */
                FLUSH();
                if (axis == 0) ix += ux > 0.0f ? 1 : -1;
                else if (axis == 1) iy += uy > 0.0f ? 1 : -1;
                else iz += uz > 0.0f ? 1 : -1;
                inside = ix >= 0 && ix < c->Nx && iy >= 0 && iy < c->Ny && iz >= 0 && iz < c->Nz;
                if (!inside && (c->bmode == 1 || (c->bmode == 2 && iz < 0))) {
                    if (iz >= c->Nz) { t->tt += (double)W; t->escapes_bottom++; }
                    else if (iz < 0) { t->top += (double)W; t->escapes_top++; }
                    else { t->lateral += (double)W; t->escapes_side++; }
                    alive = 0;
                } else {
                    int nt = c->V[VIDX(CLAMPI(ix, c->Nx), CLAMPI(iy, c->Ny), CLAMPI(iz, c->Nz))];
                    if (NMCO_MUTANT == 2 && nt != tis && c->mus[tis] > 0.0f)
                        sleft = sleft * (c->mus[nt] / c->mus[tis]);
                    tis = nt;
                }
            }
        }
        if (alive && !(W > 0.0f)) alive = 0; /* all weight absorbed */
        if (!alive) { FLUSH(); return; }
        if (events >= c->max_events) {
            FLUSH();
            t->capped += (double)W;
            t->capped_photons++;
            return;
        }
        if (sleft > 0.0f) continue;

        /* SPIN (kernels.metal:695-730) */
        float ct = c->two_term ? tt_cos(&c->tt[tis], u01(b[1])) : hg_sample(&c->hg[tis], u01(b[1]));
        nmco_spin_w(ct, b[2], &ux, &uy, &uz);
        rux = 1.0f / ux; ruy = 1.0f / uy; ruz = 1.0f / uz;
        t->interactions++;
        if (a->sens && c->slot[tis] >= 0) Ks[c->slot[tis]]++;
        /* ROULETTE (kernels.metal:740-744) */
        if (!roulette(&W, u01(b[3]), c->thr, c->chance, t)) { FLUSH(); return; }
        ev++;
        block(c->key, pid, ev, 0u, b);
        sleft = -nmco_logf_tab(u01nz(b[0]));
    }
#undef FLUSH
#undef VIDX
}

static int xyz_setup(const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                     const nmc_ext_config* e, xyz_ctx* c)
{
    memset(c, 0, sizeof(*c));
    if (!rp || !tp || !v) return NMC_ERR_ARG;
    if (rp->Nx <= 0 || rp->Ny <= 0 || rp->Nz <= 0 || !(rp->dx > 0) || !(rp->dy > 0) || !(rp->dz > 0))
        return NMC_ERR_ARG;
    c->Nx = rp->Nx; c->Ny = rp->Ny; c->Nz = rp->Nz;
    c->dx = rp->dx; c->dy = rp->dy; c->dz = rp->dz;
    c->hx = 0.5f * (float)rp->Nx; c->hy = 0.5f * (float)rp->Ny;
    c->bmode = rp->boundaryflag; c->det_state = rp->det_state;
    c->mcflag = rp->mcflag; c->launchflag = rp->launchflag;
    c->two_term = rp->mckernelflag == NMC_SIM_MCXYZ_TT;
    c->zsurf = rp->zsurf; c->n1 = rp->n1; c->n2 = rp->n2;
    if (rp->xd == 0.0f) { c->rmin = 0.0f; c->rmax = rp->det_radius; }
    else { c->rmin = rp->xd - rp->det_radius; c->rmax = c->rmin + 2.0f * rp->det_radius; }
    c->cos_na = rp->na >= 90.0f ? -1.0f : (float)cos((double)rp->na * 3.14159265358979323846 / 180.0);
    c->xs = rp->xs; c->ys = rp->ys; c->zs = rp->zs;
    c->ux0 = rp->ux0; c->uy0 = rp->uy0; c->uz0 = rp->uz0;
    c->radius = rp->radius; c->waist = rp->waist; c->zfocus = rp->zfocus;
    c->thr = e->threshold; c->chance = e->chance; c->max_events = e->max_events;
    c->count_visits = (e->flags & NMC_FLAG_COUNT_VISITS) != 0;
    c->key[0] = (uint32_t)e->seed;
    c->key[1] = (uint32_t)(e->seed >> 32) + e->stream_id;
    c->V = v;
    for (int i = 0; i < NMC_NTISS; ++i) {
        c->slot[i] = -1;
        c->mua[i] = tp->muav[i]; c->mus[i] = tp->musv[i]; c->inv_mus[i] = inv_pos(tp->musv[i]);
        hg_prepare(tp->gv[i], &c->hg[i]);
        if (c->two_term) tt_prepare(tp->gf[i], tp->af[i], tp->gb[i], tp->ab[i], tp->CC[i], &c->tt[i]);
    }
    int64_t nv = (int64_t)rp->Nx * rp->Ny * rp->Nz;
    for (int64_t i = 0; i < nv; ++i)
        if (v[i] >= NMC_NTISS) return NMC_ERR_ARG;
    return NMC_OK;
}

int nmco_mcxyz(const RunParams* rp, const TissueParams* tp, const uint8_t* v,
               const nmc_ext_config* ext, uint64_t first, uint64_t n, int n_threads,
               nmc_results* out)
{
    return nmco_mcxyz_sens(rp, tp, v, ext, first, n, n_threads, out, NULL, 0, NULL);
}

int nmco_mcxyz_sens(const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                    const nmc_ext_config* ext, uint64_t first, uint64_t n, int n_threads,
                    nmc_results* out, const int32_t* ids, int n_ids, double* sens)
{
    nmc_ext_config e;
    ext_defaults(ext, &e);
    if (n_ids < 0 || n_ids > NMC_SENS_MAX || (n_ids > 0 && (!ids || !sens))) return NMC_ERR_ARG;
    xyz_ctx* c = (xyz_ctx*)malloc(sizeof(xyz_ctx));
    if (!c) return NMC_ERR_NOMEM;
    int st = xyz_setup(rp, tp, v, &e, c);
    if (st != NMC_OK) { free(c); return st; }
    c->n_sens = n_ids;
    for (int j = 0; j < n_ids; ++j) {
        if (ids[j] < 0 || ids[j] >= NMC_NTISS || c->slot[ids[j]] >= 0) { free(c); return NMC_ERR_ARG; }
        c->sens_ids[j] = ids[j];
        c->slot[ids[j]] = j;
    }
    int64_t nv = (int64_t)c->Nx * c->Ny * c->Nz, n2d = (int64_t)c->Nx * c->Ny;
    const int64_t ns = 2 * (int64_t)n_ids * nv;
    int nt = n_threads > 0 ? n_threads : nmco_max_threads();
    if ((uint64_t)nt > n) nt = n > 0 ? (int)n : 1;
    xyz_acc* acc = (xyz_acc*)calloc((size_t)nt, sizeof(xyz_acc));
    int fail = 0;
    for (int i = 0; i < nt; ++i) {
        if (n_ids > 0) {
            acc[i].sens = i == 0 ? sens : (double*)calloc((size_t)ns, sizeof(double));
            fail |= acc[i].sens == NULL;
        }
        if (out->fluence && !(e.flags & NMC_FLAG_NO_FLUENCE)) {
            acc[i].F = i == 0 ? out->fluence : (double*)calloc((size_t)nv, sizeof(double));
            fail |= acc[i].F == NULL;
        }
        if (out->ryx) { acc[i].ryx = i == 0 ? out->ryx : (double*)calloc((size_t)n2d, sizeof(double)); fail |= acc[i].ryx == NULL; }
        if (out->det2d) { acc[i].det = i == 0 ? out->det2d : (uint32_t*)calloc((size_t)n2d, sizeof(uint32_t)); fail |= acc[i].det == NULL; }
    }
    double t0 = now_s();
    if (!fail) {
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int i = 0; i < nt; ++i) {
            uint64_t lo = first + n * (uint64_t)i / (uint64_t)nt;
            uint64_t hi = first + n * (uint64_t)(i + 1) / (uint64_t)nt;
            for (uint64_t p = lo; p < hi; ++p) xyz_photon(c, p, &acc[i]);
        }
    }
    g_last_seconds = now_s() - t0;
    for (int i = 0; i < nt; ++i) {
        if (i > 0) {
            if (acc[i].F) { for (int64_t j = 0; j < nv; ++j) out->fluence[j] += acc[i].F[j]; free(acc[i].F); }
            if (acc[i].ryx) { for (int64_t j = 0; j < n2d; ++j) out->ryx[j] += acc[i].ryx[j]; free(acc[i].ryx); }
            if (acc[i].det) { for (int64_t j = 0; j < n2d; ++j) out->det2d[j] += acc[i].det[j]; free(acc[i].det); }
            if (acc[i].sens) { for (int64_t j = 0; j < ns; ++j) sens[j] += acc[i].sens[j]; free(acc[i].sens); }
        }
        totals_add(&out->totals, &acc[i].t);
    }
    free(acc);
    free(c);
    return fail ? NMC_ERR_NOMEM : NMC_OK;
}

/* Launch states of photons [first, first + n) of an mcxyz configuration
 * (test entry point for the source models): out[8 i .. 8 i + 7] = x, y, z,
 * ux, uy, uz, W, specular loss of photon first + i. */
int nmco_xyz_source(const RunParams* rp, const TissueParams* tp, const uint8_t* v,
                    const nmc_ext_config* ext, uint64_t first, uint64_t n, float* out)
{
    nmc_ext_config e;
    ext_defaults(ext, &e);
    xyz_ctx* c = (xyz_ctx*)malloc(sizeof(xyz_ctx));
    if (!c) return NMC_ERR_NOMEM;
    int st = xyz_setup(rp, tp, v, &e, c);
    if (st == NMC_OK)
        for (uint64_t i = 0; i < n; ++i) xyz_source(c, first + i, out + 8 * i);
    free(c);
    return st;
}

/* ------------------------------------------------------------------------ */
/* mcsub (kernels.metal:59-319).                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    float mua, mus, g, n1, n2, mut, inv_mut, one_m_albedo, rsp0;
    hg_coef hg;
    int mcflag, bmode, nr, nz, na;
    float dr, dz, radius, waist, zfocus, xs, ys, zs;
    float thr, chance;
    uint32_t max_events;
    uint32_t key[2];
    float cos_tab[64]; /* cos of the upper edge of exit-angle bin j */
} sub_ctx;

typedef struct { double* rd_ra; double* a_rz; nmc_totals t; } sub_acc;

static inline int rbin(float q, int n) { return q >= (float)n ? n : (int)q; }

/* LAUNCH (kernels.metal:106-166): o = {x, y, z, ux, uy, uz, rsp}. */
static void sub_source(const sub_ctx* c, uint64_t pid, float o[7])
{
    uint32_t b[4];
    float x, y, z, ux, uy, uz, rsp;
    block(c->key, pid, 0u, 0u, b);
    if (c->mcflag == 0) {
        /* uniform collimated beam at the surface (kernels.metal:106-120) */
        x = NMCO_MUTANT == 1 ? c->radius * u01(b[0]) : c->radius * sqrtf(u01(b[0]));
        y = 0.0f; z = 0.0f;
        ux = 0.0f; uy = 0.0f; uz = 1.0f;
        rsp = c->rsp0;
    } else if (c->mcflag == 1) {
        /* Gaussian beam (kernels.metal:121-145) */
        x = c->radius * sqrtf(-nmco_logf(u01nz(b[0]))); y = 0.0f; z = 0.0f;
        float xf = c->waist * sqrtf(-nmco_logf(u01nz(b[1])));
        float tmp = sqrtf((x - xf) * (x - xf) + c->zfocus * c->zfocus);
        float st = -(x - xf) / tmp, ct = c->zfocus / tmp;
        float uzn;
        rsp = nmco_fresnel(c->n2, c->n1, ct, &uzn);
        (void)st;
        uz = uzn;
        ux = sqrtf(1.0f - uz * uz);
        uy = 0.0f;
    } else {
        /* isotropic point source (kernels.metal:146-160) */
        float ct = 1.0f - 2.0f * u01(b[0]);
        float st = sqrtf(1.0f - ct * ct);
        float sp, cp;
        nmco_sincos2pi(u01(b[1]), &sp, &cp);
        x = c->xs; y = c->ys; z = c->zs;
        ux = st * cp; uy = st * sp; uz = ct;
        rsp = 0.0f;
    }
    o[0] = x; o[1] = y; o[2] = z; o[3] = ux; o[4] = uy; o[5] = uz; o[6] = rsp;
}

static void sub_photon(const sub_ctx* c, uint64_t pid, sub_acc* a)
{
    nmc_totals* t = &a->t;
    uint32_t b[4];
    t->photons++;
    float o[7];
    sub_source(c, pid, o);
    float x = o[0], y = o[1], z = o[2], ux = o[3], uy = o[4], uz = o[5], rsp = o[6];
    float W = 1.0f - rsp;
    t->specular += (double)rsp;

    uint32_t ev = 1, events = 0;
    for (;;) {
        block(c->key, pid, ev, 0u, b);
        /* HOP (kernels.metal:183-187) */
        float s = -nmco_logf_tab(u01nz(b[0])) * c->inv_mut;
        float x0 = x, y0 = y, z0 = z;
        x = fmaf(s, ux, x);
        y = fmaf(s, uy, y);
        z = fmaf(s, uz, z);
        events++;
        /* ESCAPE at z <= 0 (kernels.metal:190-214) */
        if (c->bmode == 1 && z <= 0.0f) {
            t->surface_hits++;
            float uz1;
            float rf = nmco_fresnel(c->n1, c->n2, -uz, &uz1);
            if (u01_low(b) > rf) {
                float s2 = z0 / -uz;
                x = fmaf(s2, ux, x0);
                y = fmaf(s2, uy, y0);
                float r = sqrtf(x * x + y * y);
                int ir = rbin(r / c->dr, c->nr);
                int ia = 0;
                while (ia < c->na - 1 && uz1 < c->cos_tab[ia]) ia++;
                if (NMCO_MUTANT == 4) {
                    ir = ir < c->nr ? ir + 1 : c->nr;
                    ia = ia < c->na - 1 ? ia + 1 : c->na - 1;
                }
                if (a->rd_ra) a->rd_ra[(int64_t)ir * c->na + ia] += (double)W;
                t->rd += (double)W;
                t->detections++;
                t->substeps_cross++;
                return;
            }
            z = -z;
            uz = -uz;
        }
        t->substeps_same++;
        /* DROP (kernels.metal:224-234) */
        float absorb = W * c->one_m_albedo;
        W = W - absorb;
        const float xa = NMCO_MUTANT == 3 ? x0 : x, ya = NMCO_MUTANT == 3 ? y0 : y;
        const float za = NMCO_MUTANT == 3 ? z0 : z;
        float r = sqrtf(xa * xa + ya * ya);
        int ir = rbin(r / c->dr, c->nr);
        int iz = rbin(fabsf(za) / c->dz, c->nz);
        if (NMCO_MUTANT == 4) {
            ir = ir < c->nr ? ir + 1 : c->nr;
            iz = iz < c->nz ? iz + 1 : c->nz;
        }
        if (a->a_rz) a->a_rz[(int64_t)ir * (c->nz + 1) + iz] += (double)absorb;
        t->absorbed += (double)absorb;
        /* SPIN (kernels.metal:243-279, with g honoured) */
        float ct = hg_sample(&c->hg, u01(b[1]));
        nmco_spin_w(ct, b[2], &ux, &uy, &uz);
        t->interactions++;
        if (!roulette(&W, u01(b[3]), c->thr, c->chance, t)) return;
        if (events >= c->max_events) {
            t->capped += (double)W;
            t->capped_photons++;
            return;
        }
        ev++;
    }
}

static int sub_setup(const RunParams* rp, const TissueParams* tp, const nmc_ext_config* ext,
                     nmc_ext_config* e_out, sub_ctx* cp)
{
    nmc_ext_config e;
    ext_defaults(ext, &e);
    *e_out = e;
    if (!rp || !tp || e.nr <= 0 || e.nz <= 0 || e.na <= 0 || e.na > 64 || !(e.dr > 0) || !(e.dz > 0))
        return NMC_ERR_ARG;
    sub_ctx c;
    memset(&c, 0, sizeof(c));
    c.mua = tp->muav[1]; c.mus = tp->musv[1]; c.g = tp->gv[1];
    c.n1 = rp->n1; c.n2 = rp->n2;
    if (!(c.mua >= 0) || !(c.mus >= 0) || !(c.mua + c.mus > 0)) return NMC_ERR_ARG;
    c.mut = c.mua + c.mus;
    c.inv_mut = inv_pos(c.mut);
    hg_prepare(c.g, &c.hg);
    c.one_m_albedo = 1.0f - c.mus / c.mut;
    float tr = c.n1 / c.n2;
    tr = (1.0f - tr) / (1.0f + tr);
    c.rsp0 = tr * tr;
    c.mcflag = rp->mcflag; c.bmode = rp->boundaryflag;
    c.nr = e.nr; c.nz = e.nz; c.na = e.na; c.dr = e.dr; c.dz = e.dz;
    c.radius = rp->radius; c.waist = rp->waist; c.zfocus = rp->zfocus;
    c.xs = rp->xs; c.ys = rp->ys; c.zs = rp->zs;
    c.thr = e.threshold; c.chance = e.chance; c.max_events = e.max_events;
    c.key[0] = (uint32_t)e.seed;
    c.key[1] = (uint32_t)(e.seed >> 32) + e.stream_id;
    for (int j = 0; j < c.na; ++j)
        c.cos_tab[j] = (float)cos((double)(j + 1) * (3.14159265358979323846 / 2.0) / (double)c.na);
    *cp = c;
    return NMC_OK;
}

/* Launch states of photons [first, first + n) of an mcsub configuration (test
 * entry point for the source models): out[7 i .. 7 i + 6] = x, y, z, ux, uy,
 * uz, specular reflectance of photon first + i. */
int nmco_sub_source(const RunParams* rp, const TissueParams* tp, const nmc_ext_config* ext,
                    uint64_t first, uint64_t n, float* out)
{
    nmc_ext_config e;
    sub_ctx c;
    int st = sub_setup(rp, tp, ext, &e, &c);
    if (st == NMC_OK)
        for (uint64_t i = 0; i < n; ++i) sub_source(&c, first + i, out + 7 * i);
    return st;
}

int nmco_mcsub(const RunParams* rp, const TissueParams* tp, const nmc_ext_config* ext,
               uint64_t first, uint64_t n, int n_threads, nmc_results* out)
{
    nmc_ext_config e;
    sub_ctx c;
    int st = sub_setup(rp, tp, ext, &e, &c);
    if (st != NMC_OK) return st;

    int64_t nra = (int64_t)(c.nr + 1) * c.na, nrz = (int64_t)(c.nr + 1) * (c.nz + 1);
    int nt = n_threads > 0 ? n_threads : nmco_max_threads();
    if ((uint64_t)nt > n) nt = n > 0 ? (int)n : 1;
    sub_acc* acc = (sub_acc*)calloc((size_t)nt, sizeof(sub_acc));
    int fail = 0;
    for (int i = 0; i < nt; ++i) {
        if (out->rd_ra) { acc[i].rd_ra = i == 0 ? out->rd_ra : (double*)calloc((size_t)nra, sizeof(double)); fail |= !acc[i].rd_ra; }
        if (out->a_rz && !(e.flags & NMC_FLAG_NO_FLUENCE)) { acc[i].a_rz = i == 0 ? out->a_rz : (double*)calloc((size_t)nrz, sizeof(double)); fail |= !acc[i].a_rz; }
    }
    double t0 = now_s();
    if (!fail) {
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int i = 0; i < nt; ++i) {
            uint64_t lo = first + n * (uint64_t)i / (uint64_t)nt;
            uint64_t hi = first + n * (uint64_t)(i + 1) / (uint64_t)nt;
            for (uint64_t p = lo; p < hi; ++p) sub_photon(&c, p, &acc[i]);
        }
    }
    g_last_seconds = now_s() - t0;
    for (int i = 0; i < nt; ++i) {
        if (i > 0) {
            if (acc[i].rd_ra) { for (int64_t j = 0; j < nra; ++j) out->rd_ra[j] += acc[i].rd_ra[j]; free(acc[i].rd_ra); }
            if (acc[i].a_rz) { for (int64_t j = 0; j < nrz; ++j) out->a_rz[j] += acc[i].a_rz[j]; free(acc[i].a_rz); }
        }
        totals_add(&out->totals, &acc[i].t);
    }
    free(acc);
    return fail ? NMC_ERR_NOMEM : NMC_OK;
}

/* ------------------------------------------------------------------------ */
/* Layered media with per-layer index and perturbation-MC records.          */
/* No reference counterpart (NMC has one global n1/n2, definitions.h:196);  */
/* the step/drop/spin/roulette follow mcxyz above (kernels.metal:493-750),  */
/* interfaces use Fresnel all-or-none reflection and Snell refraction as    */
/* mcsub's surface does (kernels.metal:190-214, helpers.metal:243-314).     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nl;
    float n_above, n_below;
    float zt[NMC_MAX_LAYERS], zb[NMC_MAX_LAYERS];
    float mua[NMC_MAX_LAYERS], mus[NMC_MAX_LAYERS], inv_mus[NMC_MAX_LAYERS], n[NMC_MAX_LAYERS];
    hg_coef hg[NMC_MAX_LAYERS];
    int phase[NMC_MAX_LAYERS];
    tt_coef tt[NMC_MAX_LAYERS];
    float rsp;
    uint32_t key[2];
} lay_stack;

typedef struct {
    int nr;
    float dr, thr, chance;
    uint32_t max_events;
} lay_cfg;

typedef struct {
    double* rd;          /* [nr+1] */
    double* absorbed;    /* [NMC_MAX_LAYERS] */
    nmc_totals t;
    /* records (may be NULL) */
    uint64_t* n_rec;
    uint64_t cap;
    uint32_t* rec_item;
    uint32_t* rec_bin;
    float* rec_w;
    float* rec_L;        /* [NMC_MAX_LAYERS][cap] */
    uint32_t* rec_K;     /* [NMC_MAX_LAYERS][cap] */
    uint32_t item;
    double* sens;        /* [2][NMC_MAX_LAYERS][NMC_MAX_LAYERS] or NULL */
} lay_acc;

/* Absorbed-weight sensitivities of one layer visit (no reference
 * counterpart).  The weight absorbed in layer l is the path integral of
 * mua_l W; the density of the path up to a point has the score
 * sum_t (K_t/mus_t - L_t) for mus_t, with K_t, L_t the scatterings and length
 * before that point.  A segment of length s entered with weight W deposits
 * D = W (1 - e^{-x}), x = mua_l s, and contributes
 *   dD/dmua_t = -L_t D + [t = l] W s e^{-x},
 *   dD/dmus_t = (K_t/mus_t - L_t) D - [t = l] W s h(x),
 * h(x) = (1 - e^{-x} (1 + x)) / x.  During a visit to l the other layers'
 * totals are fixed, so a visit keeps the sums of D L_l, D K_l, W s e^{-x}
 * and W s h(x) (pinned by central finite differences with common random
 * numbers and by the closed form of a non-scattering slab,
 * tests/test_oracle_layered.py). */
typedef struct { double D; float sp, sq, sza, szs; } lay_visit;

static float score_growth(float x, float em)
{
    if (x < 0.05f) return x * (0.5f - x * (0.33333334f - x * (0.125f - x * 0.033333335f)));
    return (-em - x * (1.0f + em)) / x;
}

static void sens_fold(const lay_stack* S, int l, const float* L, const uint32_t* K,
                      lay_visit* v, double* sens)
{
    if (sens && (v->D != 0.0 || v->sza != 0.0f)) {
        double* S0 = sens + l * NMC_MAX_LAYERS;
        double* S1 = sens + NMC_MAX_LAYERS * NMC_MAX_LAYERS + l * NMC_MAX_LAYERS;
        for (int t = 0; t < S->nl; ++t) {
            if (t == l || v->D == 0.0) continue;
            S0[t] += -(double)L[t] * v->D;
            S1[t] += ((double)K[t] * (double)S->inv_mus[t] - (double)L[t]) * v->D;
        }
        S0[l] += -(double)v->sp + (double)v->sza;
        S1[l] += (double)v->sq * (double)S->inv_mus[l] - (double)v->sp -
                 (NMCO_MUTANT == 5 ? 0.0 : (double)v->szs);
    }
    v->D = 0.0;
    v->sp = v->sq = v->sza = v->szs = 0.0f;
}

void nmco_layered_prepare_stack(const nmc_stack* s, uint64_t seed, lay_stack* o);

static void lay_photon(const lay_stack* S, const lay_cfg* c, uint64_t pid, lay_acc* a)
{
    nmc_totals* t = &a->t;
    uint32_t b[4], aux[4];
    t->photons++;
    float ca2;
    float rsp = nmco_fresnel(S->n_above, S->n[0], 1.0f, &ca2);
    t->specular += (double)rsp;
    float W = 1.0f - rsp;
    float x = 0.0f, y = 0.0f, z = 0.0f, ux = 0.0f, uy = 0.0f, uz = 1.0f;
    float ruz = 1.0f / uz;
    int l = 0;
    uint32_t hits = 0, ev = 1, events = 0;
    /* path length is summed per layer visit (Lcur) and added to L[l] when
     * the photon leaves the layer */
    float L[NMC_MAX_LAYERS] = {0}, Lcur = 0.0f;
    uint32_t K[NMC_MAX_LAYERS] = {0};
    lay_visit v = {0.0, 0.0f, 0.0f, 0.0f, 0.0f};
#define SENS_FOLD() sens_fold(S, l, L, K, &v, a->sens)
    block(S->key, pid, ev, 0u, b);
    float sleft = -nmco_logf_tab(u01nz(b[0]));
    for (;;) {
        float mua = S->mua[l], mus = S->mus[l];
        float s_int = sleft * S->inv_mus[l];
        float s_b = INF_F;
        if (uz > 0.0f) s_b = fmaxf((S->zb[l] - z) * ruz, 0.0f);
        else if (uz < 0.0f) s_b = fmaxf((S->zt[l] - z) * ruz, 0.0f);
        int inter = s_int <= s_b;
        float s = inter ? s_int : s_b;
        x = fmaf(s, ux, x);
        y = fmaf(s, uy, y);
        z = fmaf(s, uz, z);
        const float xa = mua * s;
        float em = nmco_expm1f(-xa);
        float dep = -(W * em);
        if (a->sens) {
            v.sp = v.sp + dep * (L[l] + Lcur);
            v.sq = v.sq + dep * (float)K[l];
            const float ws = W * s;
            v.sza = v.sza + ws * (1.0f + em);
            v.szs = v.szs + ws * score_growth(xa, em);
        }
        W = W - dep;
        a->absorbed[l] += (double)dep;
        t->absorbed += (double)dep;
        v.D += (double)dep;
        Lcur = Lcur + s;
        events++;
        if (inter) {
            sleft = 0.0f;
            t->substeps_same++;
        } else {
            t->substeps_cross++;
            sleft = sleft - s * mus;
            if (!(sleft > 0.0f)) sleft = 0.0f;
            int up = uz < 0.0f;
            z = up ? S->zt[l] : S->zb[l];
            float ni = S->n[l];
            float nt = up ? (l == 0 ? S->n_above : S->n[l - 1]) : (l == S->nl - 1 ? S->n_below : S->n[l + 1]);
            float r = nmco_fresnel(ni, nt, fabsf(uz), &ca2);
            t->surface_hits++;
            block(S->key, pid, hits, 2u, aux);
            hits++;
            if (u01(aux[0]) < r) {
                uz = -uz;
                ruz = -ruz;
            } else if (up && l == 0) {
                SENS_FOLD();
                L[l] = L[l] + Lcur;
                float re = sqrtf(x * x + y * y);
                int ir = rbin(re / c->dr, c->nr);
                a->rd[ir] += (double)W;
                t->rd += (double)W;
                t->detections++;
                if (a->rec_w) {
                    uint64_t k = (*a->n_rec)++;
                    if (k < a->cap) {
                        a->rec_item[k] = a->item;
                        a->rec_bin[k] = (uint32_t)ir;
                        a->rec_w[k] = W;
                        for (int i = 0; i < NMC_MAX_LAYERS; ++i) {
                            a->rec_L[(uint64_t)i * a->cap + k] = L[i];
                            a->rec_K[(uint64_t)i * a->cap + k] = K[i];
                        }
                    }
                }
                return;
            } else if (!up && l == S->nl - 1) {
                SENS_FOLD();
                t->tt += (double)W;
                t->escapes_bottom++;
                return;
            } else {
                SENS_FOLD();
                L[l] = L[l] + Lcur;
                Lcur = 0.0f;
                if (ni != nt) {
                    ux = ux * ni / nt;
                    uy = uy * ni / nt;
                    uz = up ? -ca2 : ca2;
                    ruz = 1.0f / uz;
                }
                const int lnew = up ? l - 1 : l + 1;
                if (NMCO_MUTANT == 2 && S->mus[l] > 0.0f) sleft = sleft * (S->mus[lnew] / S->mus[l]);
                l = lnew;
            }
        }
        if (!(W > 0.0f)) { SENS_FOLD(); return; } /* all weight absorbed */
        if (events >= c->max_events) {
            SENS_FOLD();
            t->capped += (double)W;
            t->capped_photons++;
            return;
        }
        if (sleft > 0.0f) continue;
        float ct = S->phase[l] ? tt_cos(&S->tt[l], u01(b[1])) : hg_sample(&S->hg[l], u01(b[1]));
        nmco_spin_w(ct, b[2], &ux, &uy, &uz);
        ruz = 1.0f / uz;
        K[l]++;
        t->interactions++;
        if (!roulette(&W, u01(b[3]), c->thr, c->chance, t)) { SENS_FOLD(); return; }
        ev++;
        block(S->key, pid, ev, 0u, b);
        sleft = -nmco_logf_tab(u01nz(b[0]));
    }
#undef SENS_FOLD
}

/* Interface depths summed in double, rounded once to float. */
void nmco_layered_prepare_stack(const nmc_stack* s, uint64_t seed, lay_stack* o)
{
    memset(o, 0, sizeof(*o));
    o->nl = s->n_layers;
    o->n_above = s->n_above;
    o->n_below = s->n_below;
    double z = 0.0;
    for (int l = 0; l < s->n_layers; ++l) {
        const nmc_layer* L = &s->layer[l];
        o->zt[l] = (float)z;
        int last = l == s->n_layers - 1;
        if (last && !(L->thickness > 0.0f && L->thickness < INF_F)) o->zb[l] = INF_F;
        else { z += (double)L->thickness; o->zb[l] = (float)z; }
        o->mua[l] = L->mua; o->mus[l] = L->mus; o->inv_mus[l] = inv_pos(L->mus); o->n[l] = L->n;
        hg_prepare(L->g, &o->hg[l]);
        o->phase[l] = L->phase;
        if (L->phase) tt_prepare(L->gf, L->af, L->gb, L->ab, L->cc, &o->tt[l]);
    }
    o->key[0] = (uint32_t)seed;
    o->key[1] = (uint32_t)(seed >> 32) + s->stream_id;
}

/*
 * Runs photons [first, first + n_j) of each stack j.  rd: [n_stacks][nr+1],
 * absorbed: [n_stacks][NMC_MAX_LAYERS], totals: [n_stacks]; all accumulated
 * (+=).  Records (optional, rec_w != NULL) are appended stack by stack in
 * photon order; *n_rec is the count (records beyond cap are counted, not kept).
 */
int nmco_layered(const nmc_layered_config* cfg, const nmc_stack* stacks, int n_stacks,
                 uint64_t first, int n_threads, double* rd, double* absorbed, nmc_totals* totals,
                 uint64_t* n_rec, uint64_t cap, uint32_t* rec_item, uint32_t* rec_bin,
                 float* rec_w, float* rec_L, uint32_t* rec_K, double* sens)
{
    if (!cfg || !stacks || n_stacks <= 0 || cfg->nr <= 0 || !(cfg->dr > 0.0f)) return NMC_ERR_ARG;
    const int per_sens = 2 * NMC_MAX_LAYERS * NMC_MAX_LAYERS;
    lay_cfg c = {cfg->nr, cfg->dr, cfg->threshold, cfg->chance, cfg->max_events};
    int nt = n_threads > 0 ? n_threads : nmco_max_threads();
    double t0 = now_s();
    uint64_t nrec = 0;
    for (int j = 0; j < n_stacks; ++j) {
        lay_stack S;
        nmco_layered_prepare_stack(&stacks[j], cfg->seed, &S);
        uint64_t n = stacks[j].n_photons;
        int ntj = (uint64_t)nt > n ? (n > 0 ? (int)n : 1) : nt;
        lay_acc* acc = (lay_acc*)calloc((size_t)ntj, sizeof(lay_acc));
        for (int i = 0; i < ntj; ++i) {
            acc[i].rd = (double*)calloc((size_t)(cfg->nr + 1), sizeof(double));
            acc[i].absorbed = (double*)calloc(NMC_MAX_LAYERS, sizeof(double));
            if (sens) acc[i].sens = (double*)calloc((size_t)per_sens, sizeof(double));
        }
        if (rec_w) {
            /* records are written in photon order: run the threads' ranges
             * into private slots, then copy them in order */
            for (int i = 0; i < ntj; ++i) {
                uint64_t lo = n * (uint64_t)i / (uint64_t)ntj, hi = n * (uint64_t)(i + 1) / (uint64_t)ntj;
                uint64_t m = hi - lo;
                acc[i].cap = m;
                acc[i].n_rec = (uint64_t*)calloc(1, sizeof(uint64_t));
                acc[i].rec_item = (uint32_t*)calloc((size_t)m + 1, sizeof(uint32_t));
                acc[i].rec_bin = (uint32_t*)calloc((size_t)m + 1, sizeof(uint32_t));
                acc[i].rec_w = (float*)calloc((size_t)m + 1, sizeof(float));
                acc[i].rec_L = (float*)calloc((size_t)(m + 1) * NMC_MAX_LAYERS, sizeof(float));
                acc[i].rec_K = (uint32_t*)calloc((size_t)(m + 1) * NMC_MAX_LAYERS, sizeof(uint32_t));
                acc[i].item = (uint32_t)j;
            }
        }
#pragma omp parallel for num_threads(ntj) schedule(static)
        for (int i = 0; i < ntj; ++i) {
            uint64_t lo = first + n * (uint64_t)i / (uint64_t)ntj;
            uint64_t hi = first + n * (uint64_t)(i + 1) / (uint64_t)ntj;
            for (uint64_t p = lo; p < hi; ++p) lay_photon(&S, &c, p, &acc[i]);
        }
        for (int i = 0; i < ntj; ++i) {
            for (int b = 0; b <= cfg->nr; ++b) rd[(int64_t)j * (cfg->nr + 1) + b] += acc[i].rd[b];
            for (int b = 0; b < NMC_MAX_LAYERS; ++b) absorbed[(int64_t)j * NMC_MAX_LAYERS + b] += acc[i].absorbed[b];
            if (sens) {
                for (int b = 0; b < per_sens; ++b) sens[(int64_t)j * per_sens + b] += acc[i].sens[b];
                free(acc[i].sens);
            }
            totals_add(&totals[j], &acc[i].t);
            if (rec_w) {
                uint64_t m = *acc[i].n_rec;
                for (uint64_t k = 0; k < m; ++k, ++nrec) {
                    if (nrec >= cap) continue;
                    rec_item[nrec] = acc[i].rec_item[k];
                    rec_bin[nrec] = acc[i].rec_bin[k];
                    rec_w[nrec] = acc[i].rec_w[k];
                    for (int l = 0; l < NMC_MAX_LAYERS; ++l) {
                        rec_L[(uint64_t)l * cap + nrec] = acc[i].rec_L[(uint64_t)l * acc[i].cap + k];
                        rec_K[(uint64_t)l * cap + nrec] = acc[i].rec_K[(uint64_t)l * acc[i].cap + k];
                    }
                }
                free(acc[i].n_rec); free(acc[i].rec_item); free(acc[i].rec_bin);
                free(acc[i].rec_w); free(acc[i].rec_L); free(acc[i].rec_K);
            }
            free(acc[i].rd);
            free(acc[i].absorbed);
        }
        free(acc);
    }
    if (n_rec) *n_rec = nrec;
    g_last_seconds = now_s() - t0;
    return NMC_OK;
}

/*
 * The perturbation-MC estimator over records (the same quantity the GPU
 * backward kernel computes): grad_mua[j][l] = sum_k g_k (-L_kl),
 * grad_mus[j][l] = sum_k g_k (K_kl / mus_jl - L_kl), g_k = grad_rd[j][bin_k]
 * * w_k / n_j for the records k of stack j.
 */
void nmco_layered_backward(const nmc_layered_config* cfg, const nmc_stack* stacks, int n_stacks,
                           uint64_t n_rec, uint64_t cap, const uint32_t* rec_item,
                           const uint32_t* rec_bin, const float* rec_w, const float* rec_L,
                           const uint32_t* rec_K, const double* grad_rd, double* grad_mua,
                           double* grad_mus)
{
    uint64_t m = n_rec < cap ? n_rec : cap;
    for (uint64_t k = 0; k < m; ++k) {
        uint32_t j = rec_item[k];
        if ((int)j >= n_stacks) continue;
        double g = grad_rd[(uint64_t)j * (cfg->nr + 1) + rec_bin[k]] * (double)rec_w[k] /
                   (double)stacks[j].n_photons;
        for (int l = 0; l < stacks[j].n_layers; ++l) {
            double L = (double)rec_L[(uint64_t)l * cap + k];
            double K = (double)rec_K[(uint64_t)l * cap + k];
            grad_mua[(uint64_t)j * NMC_MAX_LAYERS + l] += -g * L;
            grad_mus[(uint64_t)j * NMC_MAX_LAYERS + l] += g * (K / (double)stacks[j].layer[l].mus - L);
        }
    }
}
