/*
 * philox.h -- counter-based random numbers shared by host and device.
 *
 * The reference draws its per-step uniforms and normals from a
 * std::default_random_engine seeded by std::random_device on the CPU and from
 * a cuRAND XORWOW generator on the GPU (Random.cpp:5-14, 27-63).  Neither is
 * reproducible nor shared between the two paths.  Here every draw is a pure
 * function of (seed, population, simulation step, unit):
 *
 *   Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 *   easy as 1, 2, 3", SC'11) with key = (seed low word, seed high word) and
 *   counter = (unit, step, population, 0).
 *
 * Word 0 gives the uniform U[0,1) as (w0 >> 8) * 2^-24 (exactly representable
 * floats, as std::uniform_real_distribution<float> produces).  Words 1 and 2
 * give one standard normal by the Box-Muller transform
 * z = sqrt(-2 ln u1) cos(2 pi u2) with u1 = ((w1 >> 8) + 1) 2^-24 in (0,1] and
 * u2 = (w2 >> 8) 2^-24 in [0,1), evaluated in double with the deterministic
 * routines of detmath.h and rounded once to float.
 */
#ifndef GSBN_PHILOX_H
#define GSBN_PHILOX_H

#include <stdint.h>
#include "detmath.h"

typedef struct { uint32_t v[4]; } gsbn_u32x4;

GSBN_HD GSBN_INLINE gsbn_u32x4 gsbn_philox4x32_10(gsbn_u32x4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * (uint64_t)ctr.v[0];
    uint64_t p1 = (uint64_t)M1 * (uint64_t)ctr.v[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    gsbn_u32x4 n;
    n.v[0] = hi1 ^ ctr.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ ctr.v[3] ^ k1;
    n.v[3] = lo0;
    ctr = n;
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}

/* One draw per (seed, pop, step, unit): a uniform in [0,1) and a standard
 * normal. */
GSBN_HD GSBN_INLINE void gsbn_draw_unit(uint64_t seed, uint32_t pop, uint32_t step,
                                        uint32_t unit, float* u01, float* z) {
  gsbn_u32x4 c;
  c.v[0] = unit;
  c.v[1] = step;
  c.v[2] = pop;
  c.v[3] = 0u;
  gsbn_u32x4 w = gsbn_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const float inv24f = 5.9604644775390625e-08f; /* 2^-24 */
  *u01 = (float)(w.v[0] >> 8) * inv24f;
  const double inv24 = 5.9604644775390625e-08;
  double u1 = (double)((w.v[1] >> 8) + 1u) * inv24;
  double u2 = (double)(w.v[2] >> 8) * inv24;
  double r = gsbn_sqrt(-2.0 * gsbn_log(u1));
  *z = (float)(r * gsbn_cos2pi(u2));
}

#endif /* GSBN_PHILOX_H */
