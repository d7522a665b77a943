// nmc_device.h - device building blocks shared by the transport kernels.
//
// Every routine here is written so that, compiled with -ffp-contract=off, it
// performs exactly the IEEE binary32 operations listed in its body (explicit
// fmaf only); the CPU oracle performs the same operations, which makes photon
// trajectories identical on both sides for the same Philox stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nmc_tables.h"

namespace nmc {

constexpr float kInf = __builtin_inff();

// 1/x rounded to nearest, bit for bit what IEEE division gives: the hardware
// reciprocal (v_rcp_f32, within 1 ulp) refined by one Newton-Raphson step
// with fused multiply-adds (3 instructions instead of the 11 of the division
// sequence).  Magnitudes outside [2^-125, 2^125] (zero, subnormals, huge
// values, inf, NaN) take the division.  Equality with 1.0f / x for all 2^32
// inputs is checked on the GPU (nmcp_rcp_check, tests/test_gpu_math.py).
__device__ __forceinline__ float rcp_rn_mid(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  const float e = fmaf(-x, r, 1.0f);
  return fmaf(e, r, r);
}

__device__ __forceinline__ float rcp_rn(float x) {
  const float ax = fabsf(x);
  if (!(ax >= 0x1p-125f && ax <= 0x1p125f)) return 1.0f / x;
  return rcp_rn_mid(x);
}

// rcp_rn of the three components of a direction, with one range test for
// all three (the smallest and largest magnitude) instead of one each.
__device__ __forceinline__ void rcp_rn3(float x, float y, float z, float& rx, float& ry,
                                        float& rz) {
  const float lo = fminf(fminf(fabsf(x), fabsf(y)), fabsf(z));
  const float hi = fmaxf(fmaxf(fabsf(x), fabsf(y)), fabsf(z));
  if (lo >= 0x1p-125f && hi <= 0x1p125f) {
    rx = rcp_rn_mid(x);
    ry = rcp_rn_mid(y);
    rz = rcp_rn_mid(z);
  } else {
    rx = rcp_rn(x);
    ry = rcp_rn(y);
    rz = rcp_rn(z);
  }
}

// a / b rounded to nearest, bit for bit what IEEE division gives, by
// Markstein's correction (P. Markstein, IBM J. Res. Dev. 34(1):111-119, 1990):
// with y = RN(1/b) (rcp_rn_mid) and q = RN(a y) within one ulp of a/b, the
// residual r = a - b q is exact and RN(q + r y) = RN(a/b).  6 instructions
// instead of the 11 of the division sequence, but no range handling: only for
// operands with |b| in [2^-125, 2^125], and a = 0 or a, a/b and the residual
// normal.  Each call site states why its operands qualify.  Checked against
// the division on the GPU (nmcp_div_check, tests/test_gpu_math.py):
// exhaustively on log_pos's operands, on every Henyey-Greenstein divisor of
// 256 anisotropies, and on 2^32 pairs each of spin's operands and of general
// operands in [2^-60, 2^60].
__device__ __forceinline__ float div_rn(float a, float b) {
  const float y = rcp_rn_mid(b);
  const float q = a * y;
  const float r = fmaf(-b, q, a);
  return fmaf(r, y, q);
}

// div_rn(a, b) with y = rcp_rn_mid(b) formed once for many dividends a:
// taken for a = 0 or |a| in [2^-60, 2^60] when b_ok says |b| is in
// [2^-60, 2^60] (the operand domain of div_rn's check on 2^32 pairs); any
// other dividend takes the division.
__device__ __forceinline__ float div_rn_by(float a, float b, float y, bool b_ok) {
  const float aa = fabsf(a);
  if (b_ok && (a == 0.0f || (aa >= 0x1p-60f && aa <= 0x1p60f))) {
    const float q = a * y;
    return fmaf(fmaf(-b, q, a), y, q);
  }
  return a / b;
}

// sqrt(x) rounded to nearest, bit for bit sqrtf, for x = 0 or x in
// [2^-100, 2^126]: the hardware square root (within 1 ulp) corrected to the
// nearest by the signs of the residuals x - s' s of its two neighbours s'
// (9 instructions instead of 15: no rescaling of tiny inputs, no special
// values).  Checked against sqrtf on the GPU for every input of that domain
// (nmcp_sqrt_check, tests/test_gpu_math.py).
__device__ __forceinline__ float sqrt_rn(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  const float t = fmaf(-sm, s, x) <= 0.0f ? sm : s;
  return fmaf(-sp, s, x) > 0.0f ? sp : t;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter = (photon id lo,
// photon id hi, event index, tag); key = (seed lo, seed hi + stream).  The key
// schedule is wave-uniform, so it stays in scalar registers.
// ---------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

// a ^ b ^ c in one instruction: CDNA4's three-input bitwise operation
// (v_bitop3_b32) with the truth table of XOR over its operand patterns
// 0xF0, 0xCC, 0xAA, i.e. 0x96; the compiler does not fuse the two XORs itself.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ U4 philox(uint64_t pid, uint32_t ev, uint32_t tag, uint32_t k0,
                                     uint32_t k1) {
  uint32_t c0 = (uint32_t)pid, c1 = (uint32_t)(pid >> 32), c2 = ev, c3 = tag;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
    uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return U4{c0, c1, c2, c3};
}

// The same generator with the per-lane half of the key schedule precomputed
// (k1s[r] = k1 + r * 0xBB67AE85), for loops whose key changes rarely.
struct KeySched {
  uint32_t k1[10];
};

__device__ __forceinline__ KeySched key_sched(uint32_t k1) {
  KeySched s;
#pragma unroll
  for (int r = 0; r < 10; ++r) s.k1[r] = k1 + (uint32_t)r * 0xBB67AE85u;
  return s;
}

__device__ __forceinline__ U4 philox(uint64_t pid, uint32_t ev, uint32_t tag, uint32_t k0,
                                     const KeySched& ks) {
  uint32_t c0 = (uint32_t)pid, c1 = (uint32_t)(pid >> 32), c2 = ev, c3 = tag;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
    uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, ks.k1[r]);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
  }
  return U4{c0, c1, c2, c3};
}

// [0,1): top 24 bits scaled by 2^-24 (the reference's rand() mapping,
// helpers.metal:319-371); (0,1]: the same shifted by one ulp of 2^-24, used
// where the reference rejects zero.
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }
__device__ __forceinline__ float u01nz(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }
// A fifth uniform in [0,1) from the low bytes of words x, y, z, which the
// four top-24-bit uniforms of a block leave unused.
__device__ __forceinline__ float u01_low(const U4& b) {
  uint32_t v = ((b.x & 0xffu) << 16) | ((b.y & 0xffu) << 8) | (b.z & 0xffu);
  return (float)v * 0x1p-24f;
}

// ---------------------------------------------------------------------------
// Float transcendentals (argument domains as used by the transport loop).
// ---------------------------------------------------------------------------

// log of a positive normal float: x = 2^k m, m in [sqrt2/2, sqrt2), s = f/(2+f),
// log(1+f) = f - (hfsq - s (hfsq + R)), R = atanh series 2z/3 + ... + 2z^4/9.
__device__ __forceinline__ float log_pos(float x) {
  uint32_t b = __float_as_uint(x);
  int k = (int)(b >> 23) - 127;
  float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
  if (m > 0x1.6a09e6p+0f) {
    m = m * 0.5f;
    k += 1;
  }
  float f = m - 1.0f;
  // f = 0 or |f| in [2^-24, 0.42], 2 + f in [1.7, 2.5]: div_rn's domain
  float s = div_rn(f, 2.0f + f);
  float z = s * s;
  float R = z * fmaf(z, fmaf(z, fmaf(z, 0x1.c71c72p-3f, 0x1.24924ap-2f), 0x1.99999ap-2f),
                     0x1.555556p-1f);
  float hfsq = 0.5f * f * f;
  float fk = (float)k;
  return fk * 0x1.62e4p-1f - ((hfsq - (s * (hfsq + R) + fk * 0x1.7f7d1cp-20f)) - f);
}

// The 128 {rc, L_hi, L_lo, 0} rows of log_tab (definition in
// gen_tables.py); kernels copy it into LDS.
static __constant__ float kLogTable[128][4] = NMC_LOG_TABLE;

// log of a positive normal float by table (Tang's reduction): x = 2^k m' with
// m' in [3/4, 3/2) (the mantissa, halved with k + 1 when it is >= 3/2); the
// top seven mantissa bits pick row i, r = m' rc_i - 1 is exact (rc_i has 8
// significant bits, |r| < 2^-7) and log(x) = (k ln2_hi + L_hi) + ((k ln2_lo +
// L_lo) + log1p(r)), the first sum exact, log1p(r) = r + r^2 (-1/2 + r (1/3 -
// r/4)).  Within 0.70 ulp of log on every uniform draw k 2^-24 and over
// [e^-30, e^30], monotone on the draws; no division, one LDS row read: 19
// instructions where log_pos takes 35.
__device__ __forceinline__ float log_tab(float x, const float4* __restrict__ tab) {
  const uint32_t b = __float_as_uint(x);
  const uint32_t i = (b >> 16) & 0x7fu;
  const bool up = i >= 64u;
  const float m = __uint_as_float((b & 0x007fffffu) | (up ? 0x3f000000u : 0x3f800000u));
  const float fk = (float)((int)(b >> 23) - (up ? 126 : 127));
  const float4 t = tab[i];
  const float r = fmaf(m, t.x, -1.0f);
  const float p = fmaf(r, fmaf(r, -0x1p-2f, 0x1.555556p-2f), -0.5f);
  const float q = fmaf(r * r, p, r);
  const float hi = fmaf(fk, 0x1.62e4p-1f, t.y);
  const float lo = fmaf(fk, 0x1.7f7d1cp-20f, t.z);
  return hi + (lo + q);
}

// expm1(r), |r| <= ln2/2: r + r^2 p(r) with p of degree 5, Horner with fma;
// p's coefficients are a least-squares approximation to the minimax fit of
// (expm1(r) - r) / r^2 on that interval in relative error (2.4e-10), rounded
// to float: within 0.71 ulp of expm1 over the interval.
__device__ __forceinline__ float em1_core(float r) {
  float p = fmaf(r, 0x1.a032bcp-13f, 0x1.6d71e6p-10f);
  p = fmaf(r, p, 0x1.11118ap-7f);
  p = fmaf(r, p, 0x1.5554b2p-5f);
  p = fmaf(r, p, 0x1.555554p-3f);
  p = fmaf(r, p, 0.5f);
  return fmaf(r * r, p, r);
}

// exp(x) by Cody-Waite reduction k = rint(x/ln2); flushed to 0 below x = -87.
__device__ __forceinline__ float exp_f(float x) {
  if (x < -87.0f) return 0.0f;
  if (x > 88.5f) return kInf;
  float k = rintf(x * 0x1.715476p+0f);
  float r = fmaf(-k, 0x1.62e4p-1f, x);
  r = fmaf(-k, 0x1.7f7d1cp-20f, r);
  float e = 1.0f + em1_core(r);
  int ki = (int)k;
  if (ki > 127) {
    e = e * 2.0f;
    ki -= 1;
  }
  return e * __uint_as_float((uint32_t)(ki + 127) << 23);
}

__device__ __forceinline__ float expm1_f(float x) {
  if (fabsf(x) <= 0x1.62e430p-2f) return em1_core(x);
  return exp_f(x) - 1.0f;
}

// expm1_f(x) for x <= 0, bit for bit, as one path for every lane: |x| <=
// ln2/2 takes k = 0, so r = x exactly (fma(-0, c, x) = x) and the result is
// em1_core(x); otherwise (exp_f(x) - 1) with k in [-126, -1], whose scale
// 2^k is a normal float and needs no overflow step.
__device__ __forceinline__ float expm1_neg(float x) {
  const bool small = fabsf(x) <= 0x1.62e430p-2f;
  const float k = small ? 0.0f : rintf(x * 0x1.715476p+0f);
  float r = fmaf(-k, 0x1.62e4p-1f, x);
  r = fmaf(-k, 0x1.7f7d1cp-20f, r);
  const float p = em1_core(r);
  const float big = (1.0f + p) * __uint_as_float((uint32_t)((int)k + 127) << 23) - 1.0f;
  return small ? p : (x < -87.0f ? -1.0f : big);
}

// h(x) = (1 - e^{-x} (1 + x)) / x for x = mua s >= 0, with em = expm1(-x):
// a series below x = 0.05 (where the direct form cancels), else direct.  The
// growth of a segment's mus score in the perturbation-MC sensitivities.
__device__ __forceinline__ float score_growth(float x, float em) {
  if (x < 0.05f) return x * (0.5f - x * (0.33333334f - x * (0.125f - x * 0.033333335f)));
  return (-em - x * (1.0f + em)) / x;
}

// sin(2 pi u), cos(2 pi u), u in [0,1]: exact quarter-turn reduction, Taylor
// polynomials of sin (to f^9) and cos (to f^10) of 2 pi f, |f| <= 1/8.
__device__ __forceinline__ void sincos2pi(float u, float* s, float* c) {
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
  int qi = ((int)q) & 3;
  float ss = (qi & 1) ? pc : ps;
  float cc = (qi & 1) ? ps : pc;
  *s = (qi & 2) ? -ss : ss;
  *c = (qi == 1 || qi == 2) ? -cc : cc;
}

// The 64 {sin, cos}(2 pi j/64) rows of sincos2pi_bits (definition in
// gen_tables.py); kernels copy them into LDS.
static __constant__ float kSinCosTable[64][2] = NMC_SINCOS_TABLE;

// sin(2 pi u), cos(2 pi u) of the uniform u = (x >> 8) 2^-24 (u01's mapping)
// straight from the random word x: u = j/64 + f with j the top six bits of x
// and f = ((x >> 8) mod 2^18) 2^-24 (exact), a = 2 pi f < 0.0982, and the
// angle sum with row j: s = s_j cos a + c_j sin a, c = c_j cos a - s_j sin a,
// sin a = a + a^3 (-1/6 + a^2/120), cos a = 1 + a^2 (-1/2 + a^2/24).  Within
// 1.2e-7 of sin and cos, exact at the quarter turns; 16 instructions and one
// LDS read where sincos2pi(u01(x)) takes 28.
__device__ __forceinline__ void sincos2pi_bits(uint32_t x, const float2* __restrict__ sc, float& s,
                                               float& c) {
  const float2 t = sc[x >> 26];
  const float a = (float)((x >> 8) & 0x3ffffu) * 0x1.921fb6p-22f;
  const float a2 = a * a;
  const float sa = fmaf(a2 * a, fmaf(a2, 0x1.111112p-7f, -0x1.555556p-3f), a);
  const float ca = fmaf(a2, fmaf(a2, 0x1.555556p-5f, -0.5f), 1.0f);
  s = fmaf(t.x, ca, t.y * sa);
  c = fmaf(t.y, ca, -(t.x * sa));
}

// ---------------------------------------------------------------------------
// Optics (reference helpers.metal:21-25, 243-314).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool flt_eq(float a, float b) {
  float e = b * 0x1p-23f * 2.0f;
  return (a >= b - e) && (a <= b + e);
}

// Unpolarised Fresnel reflectance from n1 into n2 at incidence cosine ca1;
// writes the transmission cosine.
__device__ __forceinline__ float fresnel(float n1, float n2, float ca1, float* ca2) {
  if (flt_eq(n1, n2)) {
    *ca2 = ca1;
    return 0.0f;
  }
  if (ca1 > 1.0f - 1.0e-6f) {
    *ca2 = ca1;
    float r = (n2 - n1) / (n2 + n1);
    return r * r;
  }
  if (ca1 < 1.0e-6f) {
    *ca2 = 0.0f;
    return 1.0f;
  }
  float sa1 = sqrtf(1.0f - ca1 * ca1);
  float sa2 = n1 * sa1 / n2;
  if (sa2 >= 1.0f) {
    *ca2 = 0.0f;
    return 1.0f;
  }
  float c2 = sqrtf(1.0f - sa2 * sa2);
  *ca2 = c2;
  float cap = ca1 * c2 - sa1 * sa2;
  float cam = ca1 * c2 + sa1 * sa2;
  float sap = sa1 * c2 + ca1 * sa2;
  float sam = sa1 * c2 - ca1 * sa2;
  return 0.5f * sam * sam * (cam * cam + cap * cap) / (sap * sap * cam * cam);
}

__device__ __forceinline__ void snell(float n1, float n2, float* ux, float* uy, float* uz) {
  float c = 1.0f - (1.0f - *uz * *uz) * n1 * n1 / n2 / n2;
  if (c < 0.0f) {
    if (fabsf(*uz) > 1.0f) *uz = *uz < 0.0f ? -1.0f : 1.0f;
    return;
  }
  *ux = *ux * n1 / n2;
  *uy = *uy * n1 / n2;
  *uz = sqrtf(c);
}

// ---------------------------------------------------------------------------
// Scattering (reference kernels.metal:695-730, 1053-1070).
// ---------------------------------------------------------------------------
// Henyey-Greenstein sampling (kernels.metal:695-703) with per-tissue
// constants derived on the host in double: a = 1+g^2, b = 1-g^2, c = 1-g,
// d = 2g, e = 1/(2g); mode 1 = isotropic (g == 0), 2 = forward (g == 1).
struct HGCoef {
  float a, b, c, d, e;
  int mode;
};

__device__ __forceinline__ float hg_cos(const HGCoef& h, float u) {
  float ct;
  if (h.mode == 1) {
    ct = 2.0f * u - 1.0f;
  } else if (h.mode == 2) {
    ct = 1.0f;
  } else {
    // b = 1 - g^2 in [0, 1], c + d u in [1 - |g|, 1 + |g|] (|g| < 1 here: at
    // least 2^-24 as floats), quotient in [1 - |g|, 1 + |g|]: div_rn's domain
    float t = div_rn(h.b, h.c + h.d * u);
    ct = (h.a - t * t) * h.e;
  }
  // clamp to [-1, 1] with one median-of-three (ct is never NaN here, where
  // it equals fminf(fmaxf(ct, -1), 1))
  return __builtin_amdgcn_fmed3f(ct, -1.0f, 1.0f);
}

// Two-term phase-function constants per tissue, derived on the host in double
// from (gf, af, gb, ab, CC): c1 = (1+g^2)/(2g), c2 = 1/(2g), iq2 = (1-g)^(-2a),
// iq3 = (1+g)^(-2a), e = -1/a; iso* marks g == 0 (isotropic branch).
struct TTCoef {
  float cc, c1f, c2f, iq2f, iq3f, ef, c1b, c2b, iq2b, iq3b, eb;
  int isof, isob;
};

__device__ __forceinline__ float pow_neg(float b, float e) {
  if (e == -2.0f) return 1.0f / (b * b);
  return exp_f(e * log_pos(b));
}

// The branch's constants are selected first and the inverse CDF evaluated
// once, so lanes of a wave that took different branches share one
// evaluation (the forward branch's ct, or the backward branch's negated).
__device__ __forceinline__ float tt_cos(const TTCoef& t, float u) {
  const bool fwd = u < t.cc;
  float v = (fwd ? u : u - t.cc) / (fwd ? t.cc : 1.0f - t.cc);
  float ct;
  if (fwd ? t.isof : t.isob) {
    ct = 2.0f * v - 1.0f;
  } else {
    float iq2 = fwd ? t.iq2f : t.iq2b, iq3 = fwd ? t.iq3f : t.iq3b;
    float c1 = fwd ? t.c1f : t.c1b, c2 = fwd ? t.c2f : t.c2b;
    ct = c1 - c2 * pow_neg(v * iq2 + (1.0f - v) * iq3, fwd ? t.ef : t.eb);
    if (!fwd) ct = -ct;
  }
  return fminf(fmaxf(ct, -1.0f), 1.0f);
}

// Rotate direction (ux,uy,uz) by polar cosine `cost` and azimuth psi given
// by sp = sin(psi), cp = cos(psi) (spin: psi = 2 pi u_psi; spin_w: psi from
// the random word as sincos2pi_bits takes it).  The pole formula is taken where the reference takes it: its test
// 1 - |uz| <= 1e-12 in float arithmetic (kernels.metal:713) holds exactly
// when |uz| >= 1.  It is also taken when ux^2 + uy^2 < 2^-100 (only a
// direction that is not quite unit length gets there with |uz| < 1), which
// keeps the general formula's quotient below 2^100.
__device__ __forceinline__ void spin_sc(float cost, float sp, float cp, float& ux, float& uy,
                                        float& uz) {
  float x = ux, y = uy, z = uz;
  float r2 = x * x + y * y;
  if (fabsf(z) >= 1.0f || !(r2 >= 0x1p-100f)) {
    // 1 - cos^2 is 0 or at least 2^-24 (cos^2 <= 1 is a float): sqrt_rn's domain
    float sint = sqrt_rn(1.0f - cost * cost);
    ux = sint * cp;
    uy = sint * sp;
    uz = z >= 0.0f ? cost : -cost;
  } else {
    // st = sin(theta) / sqrt(ux^2 + uy^2) as one square root; the uz term
    // sin(theta) * sqrt(ux^2 + uy^2) is st * (ux^2 + uy^2)
    // r2 = 1 - uz^2 up to rounding (about 1.2e-7 or more for a unit vector
    // with |uz| < 1), at least 2^-100 here; 1 - cos^2 is 0 or in [2^-24, 1]:
    // the residual of the quotient stays normal (div_rn's domain), and the
    // quotient (0 or at least 2^-24) is in sqrt_rn's
    float st = sqrt_rn(div_rn(1.0f - cost * cost, r2));
    ux = st * (x * z * cp - y * sp) + x * cost;
    uy = st * (y * z * cp + x * sp) + y * cost;
    uz = -(st * cp * r2) + z * cost;
  }
}

__device__ __forceinline__ void spin(float cost, float u_psi, float& ux, float& uy, float& uz) {
  float sp, cp;
  sincos2pi(u_psi, &sp, &cp);
  spin_sc(cost, sp, cp, ux, uy, uz);
}

__device__ __forceinline__ void spin_w(float cost, uint32_t w_psi, const float2* __restrict__ sc,
                                       float& ux, float& uy, float& uz) {
  float sp, cp;
  sincos2pi_bits(w_psi, sc, sp, cp);
  spin_sc(cost, sp, cp, ux, uy, uz);
}

// A wave-uniform value kept in a vector register.  The transport loops use
// every scalar register the hardware gives a wave, so the compiler would
// otherwise park loop-invariant values in lanes of a vector register and
// fetch each back with a v_readlane (a VALU instruction) wherever the loop
// uses it; a copy per lane costs one v_mov at kernel start and a vector
// register, of which these kernels have spare ones at their occupancy.
__device__ __forceinline__ uint32_t in_vgpr(uint32_t s) {
  uint32_t v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
  return v;
}

// A wave-uniform pointer in a pair of vector registers (see in_vgpr); the
// atomics and loads that use it take a 64-bit vector address anyway.
template <typename T>
__device__ __forceinline__ T* ptr_in_vgpr(T* p) {
  const uint64_t u = (uint64_t)p;
  return (T*)(((uint64_t)in_vgpr((uint32_t)(u >> 32)) << 32) | in_vgpr((uint32_t)u));
}

// A pointer into device (global) memory, typed as such: loads, stores and
// atomics through it compile to global_* instructions.  Through a generic
// pointer they become flat_* instructions, which count on the LDS wait
// counter as well, so every later wait for an LDS access would also wait for
// them (a fluence atomic or record store stalling the next tissue-table read).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

// ptr_in_vgpr for a pointer known to address device memory.
template <typename T>
__device__ __forceinline__ gptr<T> gptr_in_vgpr(T* p) {
  const uint64_t u = (uint64_t)p;
  return (gptr<T>)(((uint64_t)in_vgpr((uint32_t)(u >> 32)) << 32) | in_vgpr((uint32_t)u));
}

// FP64 add to device memory without a return value (global_atomic_add_f64).
__device__ __forceinline__ void global_add(gptr<double> p, double v) {
  (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The same value, redefined for the compiler at this point: comparisons with
// it are evaluated where they are used instead of being hoisted out of the
// loop as lane masks, which would occupy (or spill) scalar registers.
__device__ __forceinline__ uint32_t fresh_vgpr(uint32_t v) {
  asm volatile("" : "+v"(v));
  return v;
}

// ---------------------------------------------------------------------------
// Wave helpers (wave64).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
  int l = __lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

}  // namespace nmc
