/*
 * gsbn_core.h -- state layout and the per-element update rules of the lazy
 * BCPNN procedure, shared by the HIP kernels and the host code of this library.
 *
 * Per-synapse state.  The reference keeps six row-major arrays
 * pij/eij/zi2/zj2/wij/tij indexed (h*C + k)*M + j (Proj.cpp:99-104).  Row
 * updates walk j for fixed (h,k) and column updates walk k for fixed (h,j), so
 * in separate arrays one of the two is always a stride-M walk that touches five
 * different cache lines per synapse.  Here the six values of one synapse form
 * one 32-byte record, records are stored in the same (h*C + k)*M + j order, and
 * a column update reads and writes one aligned 32-byte sector per synapse
 * while a row update streams M consecutive records.
 *
 * Per-row state (pi, ei, zi, ti; Proj.cpp:91-94) is one 16-byte record.
 *
 * Closed-form catch-up.  Over an elapsed gap of D whole steps the traces obey
 * the linear cascade dz/dt = -kz z, de/dt = ke (z - e), dp/dt = kp (e - p)
 * (Proj.cpp:241-247 for the row traces, Proj.cpp:269-277 for the synapses, with
 * kz = kzi for rows and kz = kzi + kzj for the product zi2*zj2).  Its solution
 * is linear in the state:
 *     z' = Ez z
 *     e' = Ee e + Cez z
 *     p' = Ep p + Cpe e + Cpz z
 * with Ee = exp(-ke D), Ep = exp(-kp D), Cez = ke phi(kz,ke), Cpe = kp phi(ke,kp),
 * Cpz = kp ke psi(kz,ke,kp), phi and psi being the first and second divided
 * differences of x -> exp(-x D) (sign chosen so that both are positive).  The
 * coefficients are tabulated per D in double precision on the host (coeff.cpp)
 * and rounded to float; each update is then a fixed sequence of float
 * multiplies and adds evaluated in the order written below.
 */
#ifndef GSBN_CORE_H
#define GSBN_CORE_H

#include <stdint.h>
#include "detmath.h"

typedef struct {
  float pij, eij, zi2, zj2, wij;
  int32_t tij;
  int32_t pad0, pad1;
} gsbn_syn_t;

typedef struct {
  float pi, ei, zi;
  int32_t ti;
} gsbn_rowst_t;

/* coefficient table: GSBN_COEF_STRIDE floats per elapsed gap D */
#define GSBN_COEF_STRIDE 12
enum {
  GSBN_C_EZI = 0,   /* exp(-kzi D) */
  GSBN_C_EZJ = 1,   /* exp(-kzj D) */
  GSBN_C_EE = 2,    /* exp(-ke D) */
  GSBN_C_EP = 3,    /* exp(-kp D) */
  GSBN_C_EZ_IJ = 4, /* ke phi(kzi+kzj, ke) */
  GSBN_C_PE = 5,    /* kp phi(ke, kp) */
  GSBN_C_PZ_IJ = 6, /* kp ke psi(kzi+kzj, ke, kp) */
  GSBN_C_EZ_I = 7,  /* ke phi(kzi, ke) */
  GSBN_C_PZ_I = 8   /* kp ke psi(kzi, ke, kp) */
};

/* value written into the support-contribution buffer for a synapse whose row
 * update took the D<=0 branch (Proj.cpp:428-431 adds nothing to epsc) */
#define GSBN_NO_CONTRIB_BITS 0x7fbadbadu

/* the seven coefficients a synapse update uses */
typedef struct {
  float ezi, ezj, ee, ep, cez, cpe, cpz;
} gsbn_syn_coef_t;

GSBN_HD GSBN_INLINE gsbn_syn_coef_t gsbn_syn_coef(const float* c) {
  gsbn_syn_coef_t k;
  k.ezi = c[GSBN_C_EZI];
  k.ezj = c[GSBN_C_EZJ];
  k.ee = c[GSBN_C_EE];
  k.ep = c[GSBN_C_EP];
  k.cez = c[GSBN_C_EZ_IJ];
  k.cpe = c[GSBN_C_PE];
  k.cpz = c[GSBN_C_PZ_IJ];
  return k;
}

/* Synapse catch-up over a gap.  Row form adds kfti to zi2, column form adds
 * kftj to zj2 (by the caller). */
GSBN_HD GSBN_INLINE void gsbn_syn_catchup_k(const gsbn_syn_coef_t k, float* pij, float* eij,
                                            float* zi2, float* zj2) {
  float z0 = (*zi2) * (*zj2);
  float p = (k.ep * (*pij) + k.cpe * (*eij)) + k.cpz * z0;
  float e = k.ee * (*eij) + k.cez * z0;
  *pij = p;
  *eij = e;
  *zi2 = k.ezi * (*zi2);
  *zj2 = k.ezj * (*zj2);
}

/* the same with coefficients c (pointer to one table entry) */
GSBN_HD GSBN_INLINE void gsbn_syn_catchup(const float* c, float* pij, float* eij,
                                          float* zi2, float* zj2) {
  gsbn_syn_catchup_k(gsbn_syn_coef(c), pij, eij, zi2, zj2);
}

GSBN_HD GSBN_INLINE void gsbn_row_catchup(const float* c, float* pi, float* ei, float* zi) {
  float p = (c[GSBN_C_EP] * (*pi) + c[GSBN_C_PE] * (*ei)) + c[GSBN_C_PZ_I] * (*zi);
  float e = c[GSBN_C_EE] * (*ei) + c[GSBN_C_EZ_I] * (*zi);
  *pi = p;
  *ei = e;
  *zi = c[GSBN_C_EZI] * (*zi);
}

/* Wij = wgain * log((pij + eps^2) / ((pi + eps) (pj + eps)))  (Proj.cpp:461) */
GSBN_HD GSBN_INLINE float gsbn_wij(float pij, float pi, float pj, float wgain, float eps,
                                   float eps2) {
  return wgain * gsbn_logf((pij + eps2) / ((pi + eps) * (pj + eps)));
}

/* ---- reduced-precision trace storage ----
 * float -> bfloat16 / binary16 with round-to-nearest-even (subnormals kept,
 * overflow to infinity, NaN stays NaN), written with integer operations so
 * that host and device round identically. */
GSBN_HD GSBN_INLINE uint16_t gsbn_f32_to_bf16(float x) {
  uint32_t u = gsbn_fbits(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

GSBN_HD GSBN_INLINE float gsbn_bf16_to_f32(uint16_t h) { return gsbn_bitsf((uint32_t)h << 16); }

GSBN_HD GSBN_INLINE uint16_t gsbn_f32_to_f16(float x) {
  const uint32_t u = gsbn_fbits(x);
  const uint32_t sign = (u >> 16) & 0x8000u;
  const uint32_t a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);          /* NaN */
  if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);         /* >= 65520: inf */
  if (a < 0x38800000u) {                                            /* below 2^-14: subnormal */
    if (a < 0x33000000u) return (uint16_t)sign;                     /* < 2^-25: zero */
    const uint32_t e = a >> 23;
    const uint32_t m = (a & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - e;                                /* 14..24 */
    uint32_t q = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return (uint16_t)(sign | q);
  }
  uint32_t r = a - 0x38000000u;                                     /* rebias exponent */
  r += 0xfffu + ((r >> 13) & 1u);
  return (uint16_t)(sign | (r >> 13));
}

GSBN_HD GSBN_INLINE float gsbn_f16_to_f32(uint16_t h) {
  const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  const uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
  const uint32_t m = (uint32_t)h & 0x3ffu;
  if (e == 0) {
    if (m == 0) return gsbn_bitsf(sign);
    /* subnormal: m * 2^-24, exact in float */
    const float v = (float)m * 5.9604644775390625e-08f;
    return sign ? -v : v;
  }
  if (e == 31) return gsbn_bitsf(sign | 0x7f800000u | (m << 13));
  return gsbn_bitsf(sign | ((e + 112u) << 23) | (m << 13));
}

/* 16-byte synapse record of the reduced-precision forms; wij is kept in a
 * separate float array because only row updates touch it */
typedef struct {
  uint16_t pij, eij;
  float zi2, zj2;
  int32_t tij;
} gsbn_syn16_t;

#ifndef GSBN_PREC_F32
#define GSBN_PREC_F32 0
#define GSBN_PREC_BF16 1
#define GSBN_PREC_F16 2
#endif

/* stored value of a P or E trace: bf16(x), or binary16(x * scale) with a
 * power-of-two scale (so that x * scale and the decode's product with
 * 1 / scale are exact) */
GSBN_HD GSBN_INLINE uint16_t gsbn_pack_trace(int prec, float x, float scale) {
  return prec == GSBN_PREC_BF16 ? gsbn_f32_to_bf16(x) : gsbn_f32_to_f16(x * scale);
}
GSBN_HD GSBN_INLINE float gsbn_unpack_trace(int prec, uint16_t h, float inv_scale) {
  return prec == GSBN_PREC_BF16 ? gsbn_bf16_to_f32(h) : gsbn_f16_to_f32(h) * inv_scale;
}

/* table lookup: gaps beyond the table use the last entry, which the builder
 * only allows when that entry is the exact saturated value */
GSBN_HD GSBN_INLINE const float* gsbn_coef(const float* table, int len, int gap) {
  int d = gap < len ? gap : len - 1;
  return table + (int64_t)d * GSBN_COEF_STRIDE;
}

#endif /* GSBN_CORE_H */
