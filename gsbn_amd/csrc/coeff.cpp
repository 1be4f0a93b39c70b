/*
 * coeff.cpp -- FP64 decay-coefficient table.
 *
 * The reference evaluates the closed-form catch-up directly in float, dividing
 * by rate differences such as (ke - kzi) and (kzi - kp + kzj)
 * (Proj.cpp:241-247, 269-277).  With close time constants (the shipped
 * data/10x10_rnd.prototxt uses tauzi=5 ms, tauzj=5.001 ms, taue=5.0002 ms) the
 * differences are ~1e-5 of the rates and the float expression loses most of its
 * digits.  The coefficients here are divided differences of x -> exp(-D x)
 * computed as entries of the matrix exponential of a 3x3 upper bidiagonal
 * matrix (the Opitz representation), evaluated by scaling and squaring as in
 * McCurdy, Ng and Parlett, "Accurate computation of divided differences of the
 * exponential function", Math. Comp. 43 (1984).  All products in the squaring
 * phase have operands of equal sign, so no cancellation occurs; the result is
 * accurate to ~1e-11 relative even when the rates coincide.
 */
#include "coeff.h"
#include "gsbn_core.h"

#include <math.h>

namespace {

struct UT3 {
  double d0, d1, d2, u01, u12, u02;
};

UT3 ut_mul(const UT3& a, const UT3& b) {
  UT3 r;
  r.d0 = a.d0 * b.d0;
  r.d1 = a.d1 * b.d1;
  r.d2 = a.d2 * b.d2;
  r.u01 = a.d0 * b.u01 + a.u01 * b.d1;
  r.u12 = a.d1 * b.u12 + a.u12 * b.d2;
  r.u02 = (a.d0 * b.u02 + a.u01 * b.u12) + a.u02 * b.d2;
  return r;
}

}  // namespace

extern "C" void gsbn_dd_exp(double x0, double x1, double x2, int64_t gap, double* f01,
                            double* f012) {
  const double D = (double)gap;
  double m = x0 < x1 ? x0 : x1;
  m = m < x2 ? m : x2;
  const double y0 = x0 - m, y1 = x1 - m, y2 = x2 - m;
  double ymax = y0 > y1 ? y0 : y1;
  ymax = ymax > y2 ? ymax : y2;
  /* B = -D J' with J' bidiagonal, diagonal y, superdiagonal 1/D */
  double norm = D * ymax;
  if (norm < 1.0) norm = 1.0;
  int s = 0;
  double scale = 1.0;
  while (norm > 0.5) {
    norm *= 0.5;
    scale *= 0.5;
    ++s;
  }
  UT3 c;
  c.d0 = -D * y0 * scale;
  c.d1 = -D * y1 * scale;
  c.d2 = -D * y2 * scale;
  c.u01 = -scale;
  c.u12 = -scale;
  c.u02 = 0.0;
  /* Taylor series of exp(C), degree 18 */
  UT3 e = {1.0, 1.0, 1.0, 0.0, 0.0, 0.0};
  UT3 t = e;
  for (int n = 1; n <= 18; ++n) {
    t = ut_mul(t, c);
    const double inv = (double)n;
    t.d0 /= inv;
    t.d1 /= inv;
    t.d2 /= inv;
    t.u01 /= inv;
    t.u12 /= inv;
    t.u02 /= inv;
    e.d0 += t.d0;
    e.d1 += t.d1;
    e.d2 += t.d2;
    e.u01 += t.u01;
    e.u12 += t.u12;
    e.u02 += t.u02;
  }
  for (int i = 0; i < s; ++i) e = ut_mul(e, e);
  const double em = gsbn_exp(-D * m);
  /* exp(B)_{01} = G[y0,y1]/D and exp(B)_{02} = G[y0,y1,y2]/D^2 */
  *f01 = (e.u01 * D) * em;
  *f012 = (e.u02 * (D * D)) * em;
}

extern "C" void gsbn_coeff_entry(const gsbn_rates_t* r, int64_t gap, float* out) {
  for (int i = 0; i < GSBN_COEF_STRIDE; ++i) out[i] = 0.0f;
  if (gap <= 0) {
    out[GSBN_C_EZI] = 1.0f;
    out[GSBN_C_EZJ] = 1.0f;
    out[GSBN_C_EE] = 1.0f;
    out[GSBN_C_EP] = 1.0f;
    return;
  }
  const double D = (double)gap;
  const double a = r->kzi + r->kzj;
  double f01, f012, g01, g012, h01, h012;
  gsbn_dd_exp(a, r->ke, r->kp, gap, &f01, &f012);
  gsbn_dd_exp(r->ke, r->kp, r->kp, gap, &g01, &g012);
  gsbn_dd_exp(r->kzi, r->ke, r->kp, gap, &h01, &h012);
  out[GSBN_C_EZI] = (float)gsbn_exp(-r->kzi * D);
  out[GSBN_C_EZJ] = (float)gsbn_exp(-r->kzj * D);
  out[GSBN_C_EE] = (float)gsbn_exp(-r->ke * D);
  out[GSBN_C_EP] = (float)gsbn_exp(-r->kp * D);
  out[GSBN_C_EZ_IJ] = (float)(r->ke * -f01);
  out[GSBN_C_PE] = (float)(r->kp * -g01);
  out[GSBN_C_PZ_IJ] = (float)((r->kp * r->ke) * f012);
  out[GSBN_C_EZ_I] = (float)(r->ke * -h01);
  out[GSBN_C_PZ_I] = (float)((r->kp * r->ke) * h012);
}

static int at_limit(const gsbn_rates_t* r, const float* c) {
  const float ep_lim = r->kp > 0.0 ? 0.0f : 1.0f;
  return c[GSBN_C_EZI] == 0.0f && c[GSBN_C_EZJ] == 0.0f && c[GSBN_C_EE] == 0.0f &&
         c[GSBN_C_EP] == ep_lim && c[GSBN_C_EZ_IJ] == 0.0f && c[GSBN_C_PE] == 0.0f &&
         c[GSBN_C_PZ_IJ] == 0.0f && c[GSBN_C_EZ_I] == 0.0f && c[GSBN_C_PZ_I] == 0.0f;
}

/* Every coefficient is a sum of decaying exponentials with non-negative
 * weights that is unimodal in D (a convolution of at most three exponential
 * densities); past 8/kmin all of them are non-increasing, so the first gap at
 * or after that point where all sit at their limit value starts saturation. */
extern "C" int64_t gsbn_coeff_saturation(const gsbn_rates_t* r, int64_t max_len) {
  if (r->kp < 0.0 || r->kzi <= 0.0 || r->kzj <= 0.0 || r->ke <= 0.0) return -1;
  double kmin = r->kzi;
  if (r->kzj < kmin) kmin = r->kzj;
  if (r->ke < kmin) kmin = r->ke;
  if (r->kp > 0.0 && r->kp < kmin) kmin = r->kp;
  double d0 = ceil(8.0 / kmin);
  if (d0 >= (double)max_len) return -1;
  float c[GSBN_COEF_STRIDE];
  /* bracket by doubling, then bisect on the monotone predicate */
  int64_t lo = (int64_t)d0, hi = lo;
  gsbn_coeff_entry(r, hi, c);
  if (at_limit(r, c)) return lo;
  while (true) {
    lo = hi;
    hi = hi * 2;
    if (hi >= max_len) return -1;
    gsbn_coeff_entry(r, hi, c);
    if (at_limit(r, c)) break;
  }
  while (hi - lo > 1) {
    int64_t mid = lo + (hi - lo) / 2;
    gsbn_coeff_entry(r, mid, c);
    if (at_limit(r, c))
      hi = mid;
    else
      lo = mid;
  }
  return hi;
}
