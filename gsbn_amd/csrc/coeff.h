/*
 * coeff.h -- double-precision construction of the closed-form catch-up
 * coefficients (see gsbn_core.h for their meaning).
 */
#ifndef GSBN_COEFF_H
#define GSBN_COEFF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Rates per step, taken from the single-precision values the update uses
 * (dt/tauzi, dt/tauzj, dt/taue, dt/taup * prn; Proj.cpp:38-41). */
typedef struct {
  double kzi, kzj, ke, kp;
} gsbn_rates_t;

/* Divided differences of F(x) = exp(-D x): returns F[x0,x1] in *f01 and
 * F[x0,x1,x2] in *f012. */
void gsbn_dd_exp(double x0, double x1, double x2, int64_t gap, double* f01, double* f012);

/* One table entry (GSBN_COEF_STRIDE floats) for an elapsed gap >= 0. */
void gsbn_coeff_entry(const gsbn_rates_t* r, int64_t gap, float* out);

/* First gap from which every coefficient keeps its limit value forever, or -1
 * if the table never saturates (negative learning rate). */
int64_t gsbn_coeff_saturation(const gsbn_rates_t* r, int64_t max_len);

#ifdef __cplusplus
}
#endif

#endif
