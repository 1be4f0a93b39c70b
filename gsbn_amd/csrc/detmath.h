/*
 * detmath.h -- deterministic transcendental functions shared by host and device.
 *
 * The BCPNN update needs exp (WTA soft-max, Pop.cpp:280,286), log (Wij and bj,
 * Proj.cpp:461, Proj.cpp:346; lginp in ProcExtGen.cpp:137-139) and, for the
 * support noise, a Gaussian generator (sqrt, log, cos for Box-Muller).  The
 * reference calls libm on the CPU and libdevice on the GPU, so its two paths
 * never agree bit for bit.  These routines are written once and compiled for
 * both sides; they use only IEEE-754 double +, -, *, / and integer bit
 * manipulation, which round identically on x86-64 SSE2 and on gfx950, so the
 * host and the device produce identical bits as long as both are compiled with
 * floating-point contraction disabled (-ffp-contract=off) and without
 * fast-math.
 *
 * Methods (all evaluated in double, single-precision entry points round once):
 *   exp:  Cody-Waite reduction x = k ln2 + r with the fdlibm split of ln2
 *         (ln2_hi has 11 trailing zero bits so k*ln2_hi is exact), Taylor
 *         polynomial of degree 13 on |r| <= ln2/2, scaling by 2^k.
 *   log:  x = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
 *         log m = 2 atanh(s) by its odd power series to s^23.
 *   sqrt: reciprocal square root seeded from the exponent/mantissa bits and
 *         refined by four Newton steps, then one Newton correction of x*y.
 *   cos(2 pi u): reduction by the symmetries of cos to an angle in [0, pi/4]
 *         and Taylor series of cos / sin of degree 18 / 17.
 * Accuracy of the double results is a few units in the last place of double,
 * so the float entry points are correctly rounded except for arguments whose
 * exact result lies within ~1e-15 relative of a float rounding boundary.
 */
#ifndef GSBN_DETMATH_H
#define GSBN_DETMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define GSBN_HD __host__ __device__
#else
#define GSBN_HD
#endif

#define GSBN_INLINE static inline

GSBN_HD GSBN_INLINE uint64_t gsbn_dbits(double x) {
  union { double d; uint64_t u; } v;
  v.d = x;
  return v.u;
}

GSBN_HD GSBN_INLINE double gsbn_bitsd(uint64_t u) {
  union { double d; uint64_t u; } v;
  v.u = u;
  return v.d;
}

GSBN_HD GSBN_INLINE uint32_t gsbn_fbits(float x) {
  union { float f; uint32_t u; } v;
  v.f = x;
  return v.u;
}

GSBN_HD GSBN_INLINE float gsbn_bitsf(uint32_t u) {
  union { float f; uint32_t u; } v;
  v.u = u;
  return v.f;
}

/* Round to nearest integer, ties to even, for |x| < 2^51 (the only range used).
 * Adding and subtracting 1.5*2^52 performs the rounding in the FPU's default
 * mode on both targets. */
GSBN_HD GSBN_INLINE double gsbn_rint_small(double x) {
  const double magic = 6755399441055744.0; /* 1.5 * 2^52 */
  double y = x + magic;
  return y - magic;
}

/* e^x in double. Valid for all x; returns +inf above 709.78 and 0 below -745. */
GSBN_HD GSBN_INLINE double gsbn_exp(double x) {
  if (x != x) return x;
  if (x > 709.78) return gsbn_bitsd(0x7ff0000000000000ULL);
  if (x < -745.2) return 0.0;
  const double inv_ln2 = 1.44269504088896338700e+00;
  const double ln2_hi = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000 */
  const double ln2_lo = 1.90821492927058770002e-10; /* 0x3dea39ef35793c76 */
  double kd = gsbn_rint_small(x * inv_ln2);
  int k = (int)kd;
  double r = (x - kd * ln2_hi) - kd * ln2_lo;
  /* Taylor series of e^r, Horner form, degree 13; |r| <= 0.3466 gives a
   * truncation error below 3e-18. */
  double p = 1.0 / 6227020800.0;                 /* 1/13! */
  p = p * r + 1.0 / 479001600.0;                 /* 1/12! */
  p = p * r + 1.0 / 39916800.0;                  /* 1/11! */
  p = p * r + 1.0 / 3628800.0;                   /* 1/10! */
  p = p * r + 1.0 / 362880.0;                    /* 1/9!  */
  p = p * r + 1.0 / 40320.0;                     /* 1/8!  */
  p = p * r + 1.0 / 5040.0;                      /* 1/7!  */
  p = p * r + 1.0 / 720.0;                       /* 1/6!  */
  p = p * r + 1.0 / 120.0;                       /* 1/5!  */
  p = p * r + 1.0 / 24.0;                        /* 1/4!  */
  p = p * r + 1.0 / 6.0;                         /* 1/3!  */
  p = p * r + 0.5;                               /* 1/2!  */
  p = p * r + 1.0;
  p = p * r + 1.0;
  /* scale by 2^k; split in two factors so that subnormal results are exact
   * products of normal numbers */
  int k1 = k / 2;
  int k2 = k - k1;
  double s1 = gsbn_bitsd((uint64_t)(int64_t)(k1 + 1023) << 52);
  double s2 = gsbn_bitsd((uint64_t)(int64_t)(k2 + 1023) << 52);
  return (p * s1) * s2;
}

/* natural logarithm in double for x > 0 (x <= 0, NaN handled). */
GSBN_HD GSBN_INLINE double gsbn_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return gsbn_bitsd(0x7ff8000000000000ULL);
  if (x == 0.0) return gsbn_bitsd(0xfff0000000000000ULL);
  uint64_t b = gsbn_dbits(x);
  if ((b >> 52) == 0x7ff) return x; /* +inf */
  int e = 0;
  if ((b >> 52) == 0) { /* subnormal: scale up by 2^64 */
    x = x * 18446744073709551616.0;
    b = gsbn_dbits(x);
    e = -64;
  }
  e += (int)((b >> 52) & 0x7ff) - 1023;
  double m = gsbn_bitsd((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > 1.41421356237309504880) { /* bring m into [sqrt(1/2), sqrt(2)) */
    m = m * 0.5;
    e += 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  /* 2 atanh(s) = 2 s (1 + z/3 + z^2/5 + ... + z^11/23); z <= 0.0295 */
  double q = 1.0 / 23.0;
  q = q * z + 1.0 / 21.0;
  q = q * z + 1.0 / 19.0;
  q = q * z + 1.0 / 17.0;
  q = q * z + 1.0 / 15.0;
  q = q * z + 1.0 / 13.0;
  q = q * z + 1.0 / 11.0;
  q = q * z + 1.0 / 9.0;
  q = q * z + 1.0 / 7.0;
  q = q * z + 1.0 / 5.0;
  q = q * z + 1.0 / 3.0;
  q = q * z + 1.0;
  double logm = 2.0 * s * q;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  double ed = (double)e;
  return ed * ln2_hi + (ed * ln2_lo + logm);
}

/* square root in double for x >= 0 */
GSBN_HD GSBN_INLINE double gsbn_sqrt(double x) {
  if (x != x || x < 0.0) return gsbn_bitsd(0x7ff8000000000000ULL);
  if (x == 0.0) return x;
  uint64_t b = gsbn_dbits(x);
  if ((b >> 52) == 0x7ff) return x;
  double scale = 1.0;
  if ((b >> 52) == 0) { /* subnormal: x*2^128, result * 2^-64 */
    x = x * 340282366920938463463374607431768211456.0;
    scale = 1.0 / 18446744073709551616.0;
    b = gsbn_dbits(x);
  }
  /* seed for 1/sqrt(x): halve the biased exponent, linear mantissa guess */
  double y = gsbn_bitsd(0x5fe6eb50c7b537a9ULL - (b >> 1));
  y = y * (1.5 - 0.5 * x * y * y);
  y = y * (1.5 - 0.5 * x * y * y);
  y = y * (1.5 - 0.5 * x * y * y);
  y = y * (1.5 - 0.5 * x * y * y);
  double s = x * y;
  s = s + 0.5 * y * (x - s * s);
  return s * scale;
}

/* cos(2 pi u) in double for finite u */
GSBN_HD GSBN_INLINE double gsbn_cos2pi(double u) {
  double v = u - gsbn_rint_small(u);          /* v in [-1/2, 1/2] */
  double a = v < 0.0 ? -v : v;                 /* cos is even */
  double sign = 1.0;
  if (a > 0.25) { a = 0.5 - a; sign = -1.0; }  /* cos(pi - t) = -cos t */
  const double two_pi = 6.28318530717958647692;
  double r;
  if (a > 0.125) {                             /* cos t = sin(pi/2 - t) */
    double t = two_pi * (0.25 - a);
    double t2 = t * t;
    double p = -1.0 / 355687428096000.0;       /* -1/17! */
    p = p * t2 + 1.0 / 1307674368000.0;        /*  1/15! */
    p = p * t2 - 1.0 / 6227020800.0;           /* -1/13! */
    p = p * t2 + 1.0 / 39916800.0;             /*  1/11! */
    p = p * t2 - 1.0 / 362880.0;               /* -1/9!  */
    p = p * t2 + 1.0 / 5040.0;                 /*  1/7!  */
    p = p * t2 - 1.0 / 120.0;                  /* -1/5!  */
    p = p * t2 + 1.0 / 6.0;                    /*  1/3!  */
    r = t - t * t2 * p;
  } else {
    double t = two_pi * a;
    double t2 = t * t;
    double p = 1.0 / 6402373705728000.0;       /*  1/18! */
    p = p * t2 - 1.0 / 20922789888000.0;       /* -1/16! */
    p = p * t2 + 1.0 / 87178291200.0;          /*  1/14! */
    p = p * t2 - 1.0 / 479001600.0;            /* -1/12! */
    p = p * t2 + 1.0 / 3628800.0;              /*  1/10! */
    p = p * t2 - 1.0 / 40320.0;                /* -1/8!  */
    p = p * t2 + 1.0 / 720.0;                  /*  1/6!  */
    p = p * t2 - 1.0 / 24.0;                   /* -1/4!  */
    p = p * t2 + 0.5;                          /*  1/2!  */
    r = 1.0 - t2 * p;
  }
  return sign * r;
}

GSBN_HD GSBN_INLINE float gsbn_expf(float x) { return (float)gsbn_exp((double)x); }
GSBN_HD GSBN_INLINE float gsbn_logf(float x) { return (float)gsbn_log((double)x); }

#endif /* GSBN_DETMATH_H */
