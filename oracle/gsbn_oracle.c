/*
 * gsbn_oracle.c -- CPU restatement of ProcUpdLazy (see gsbn_oracle.h).
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, the reference's -O2 of
 * CMakeLists.txt:128).
 *
 * Parity status of each part (what pins it, see tests/):
 *   Philox draw ............ Random123 known-answer vectors (test_oracle_rng.py)
 *   coefficient table ...... 50-digit decimal divided differences, scipy expm of
 *                            the trace cascade, RK4 (test_oracle_coeff.py)
 *   FP64 closed form ....... scipy expm and RK4 (test_oracle_coeff.py)
 *   REF32 expressions ...... FP64 closed form on well-conditioned rates
 *   Pop/Proj step order .... brute-force hand cases, invariants, and the
 *                            REF32/TABLE/F64 variants against each other
 *                            (test_oracle_step.py)
 */
#define _DEFAULT_SOURCE /* posix_memalign, madvise */
#include "gsbn_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>

#include "detmath.h"

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11), independent of the product's copy */

void gor_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  int r;
  for (r = 0; r < 10; r++) {
    uint64_t a = (uint64_t)0xD2511F53u * c0;
    uint64_t b = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)b;
    uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)a;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* uniform from word 0, Box-Muller normal from words 1, 2 */
void gor_draw(uint64_t seed, uint32_t pop, uint32_t step, uint32_t unit, float* u, float* z) {
  uint32_t ctr[4] = {unit, step, pop, 0u};
  uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
  uint32_t w[4];
  gor_philox4x32_10(ctr, key, w);
  *u = (float)(w[0] >> 8) * 5.9604644775390625e-08f;
  double u1 = (double)((w[1] >> 8) + 1u) * 5.9604644775390625e-08;
  double u2 = (double)(w[2] >> 8) * 5.9604644775390625e-08;
  double rad = gsbn_sqrt(-2.0 * gsbn_log(u1));
  *z = (float)(rad * gsbn_cos2pi(u2));
}

/* the same mapping for any counter (structural plasticity's draws) */
void gor_draw_ctr(const uint32_t ctr[4], const uint32_t key[2], float* u, float* z) {
  uint32_t w[4];
  gor_philox4x32_10(ctr, key, w);
  *u = (float)(w[0] >> 8) * 5.9604644775390625e-08f;
  double u1 = (double)((w[1] >> 8) + 1u) * 5.9604644775390625e-08;
  double u2 = (double)(w[2] >> 8) * 5.9604644775390625e-08;
  double rad = gsbn_sqrt(-2.0 * gsbn_log(u1));
  *z = (float)(rad * gsbn_cos2pi(u2));
}

/* ------------------------------------------------------------------ */
/* coefficient table entry: divided differences of x -> exp(-D x) from the
 * 3x3 bidiagonal matrix exponential, scaling and squaring (McCurdy, Ng,
 * Parlett 1984).  Layout: Ezi Ezj Ee Ep Cez_ij Cpe Cpz_ij Cez_i Cpz_i 0 0 0 */

typedef struct {
  double a, b, c, x, y, z; /* diagonal a b c, superdiagonal x (01) y (12), corner z (02) */
} gor_ut3;

static gor_ut3 gor_ut3_mul(gor_ut3 p, gor_ut3 q) {
  gor_ut3 r;
  r.a = p.a * q.a;
  r.b = p.b * q.b;
  r.c = p.c * q.c;
  r.x = p.a * q.x + p.x * q.b;
  r.y = p.b * q.y + p.y * q.c;
  r.z = (p.a * q.z + p.x * q.y) + p.z * q.c;
  return r;
}

static void gor_divdiff(double x0, double x1, double x2, int64_t gap, double* d01, double* d012) {
  double D = (double)gap;
  double lo = x0;
  if (x1 < lo) lo = x1;
  if (x2 < lo) lo = x2;
  double y0 = x0 - lo, y1 = x1 - lo, y2 = x2 - lo;
  double top = y0;
  if (y1 > top) top = y1;
  if (y2 > top) top = y2;
  double nrm = D * top;
  if (nrm < 1.0) nrm = 1.0;
  int halvings = 0;
  double f = 1.0;
  while (nrm > 0.5) {
    nrm *= 0.5;
    f *= 0.5;
    halvings++;
  }
  gor_ut3 m = {-D * y0 * f, -D * y1 * f, -D * y2 * f, -f, -f, 0.0};
  gor_ut3 acc = {1.0, 1.0, 1.0, 0.0, 0.0, 0.0};
  gor_ut3 term = acc;
  int k;
  for (k = 1; k <= 18; k++) {
    term = gor_ut3_mul(term, m);
    double dk = (double)k;
    term.a /= dk;
    term.b /= dk;
    term.c /= dk;
    term.x /= dk;
    term.y /= dk;
    term.z /= dk;
    acc.a += term.a;
    acc.b += term.b;
    acc.c += term.c;
    acc.x += term.x;
    acc.y += term.y;
    acc.z += term.z;
  }
  for (k = 0; k < halvings; k++) acc = gor_ut3_mul(acc, acc);
  double shift = gsbn_exp(-D * lo);
  *d01 = (acc.x * D) * shift;
  *d012 = (acc.z * (D * D)) * shift;
}

void gor_coeff_entry(double kzi, double kzj, double ke, double kp, int64_t gap, float* out) {
  int i;
  for (i = 0; i < 12; i++) out[i] = 0.0f;
  if (gap <= 0) {
    out[0] = out[1] = out[2] = out[3] = 1.0f;
    return;
  }
  double D = (double)gap;
  double s01, s012, t01, t012, u01, u012;
  gor_divdiff(kzi + kzj, ke, kp, gap, &s01, &s012);
  gor_divdiff(ke, kp, kp, gap, &t01, &t012);
  gor_divdiff(kzi, ke, kp, gap, &u01, &u012);
  out[0] = (float)gsbn_exp(-kzi * D);
  out[1] = (float)gsbn_exp(-kzj * D);
  out[2] = (float)gsbn_exp(-ke * D);
  out[3] = (float)gsbn_exp(-kp * D);
  out[4] = (float)(ke * -s01);
  out[5] = (float)(kp * -t01);
  out[6] = (float)((kp * ke) * s012);
  out[7] = (float)(ke * -u01);
  out[8] = (float)((kp * ke) * u012);
}

/* ------------------------------------------------------------------ */
/* exact closed form in long double (independent method: explicit divided
 * differences with expm1 and a confluent expansion) */

typedef long double ld;

/* phi(a,b) = (e^{-aD} - e^{-bD})/(b-a), via e^{-aD} D expm1(-(b-a)D)/(-(b-a)D) */
static ld gor_phi(ld a, ld b, ld D) {
  ld h = (b - a) * D;
  ld ea = expl(-a * D);
  if (h == 0.0L) return D * ea;
  return D * ea * (expm1l(-h) / (-h));
}

/* psi(a,b,c) = second divided difference of x -> exp(-xD), positive */
static ld gor_psi(ld a, ld b, ld c, ld D) {
  ld t;
  if (a > b) { t = a; a = b; b = t; }
  if (b > c) { t = b; b = c; c = t; }
  if (a > b) { t = a; a = b; b = t; }
  ld spread = (c - a) * D;
  if (spread < 1e-7L) {
    ld mid = (a + b + c) / 3.0L;
    return D * D * expl(-mid * D) / 2.0L;
  }
  /* F[a,b,c] = (F[b,c] - F[a,b])/(c - a) and F[x,y] = -phi(x,y) */
  return (gor_phi(a, b, D) - gor_phi(b, c, D)) / (c - a);
}

void gor_closed_form_f64(double kz, double ke, double kp, int64_t gap, const double* in,
                         double* out) {
  ld D = (ld)gap;
  ld p0 = in[0], e0 = in[1], z0 = in[2];
  ld Ep = expl(-(ld)kp * D), Ee = expl(-(ld)ke * D), Ez = expl(-(ld)kz * D);
  ld cez = (ld)ke * gor_phi(kz, ke, D);
  ld cpe = (ld)kp * gor_phi(ke, kp, D);
  ld cpz = (ld)kp * (ld)ke * gor_psi(kz, ke, kp, D);
  out[0] = (double)(Ep * p0 + cpe * e0 + cpz * z0);
  out[1] = (double)(Ee * e0 + cez * z0);
  out[2] = (double)(Ez * z0);
}

/* ------------------------------------------------------------------ */
/* reference literal float expressions, Proj.cpp:241-247 (row) and
 * Proj.cpp:269-277 (synapse), with libm expf as std::exp(float) resolves to */

void gor_ref32_row(float kp, float ke, float kzi, int pdt, float* ppi, float* pei, float* pzi) {
  float pi = *ppi, ei = *pei, zi = *pzi;
  pi = (pi - ((ei * kp * kzi - ei * ke * kp + ke * kp * zi) / (ke - kp) +
              (ke * kp * zi) / (kp - kzi)) / (ke - kzi)) / expf(kp * pdt) +
       ((expf(kp * pdt - ke * pdt) * (ei * kp * kzi - ei * ke * kp + ke * kp * zi)) / (ke - kp) +
        (ke * kp * zi * expf(kp * pdt - kzi * pdt)) / (kp - kzi)) /
           (expf(kp * pdt) * (ke - kzi));
  ei = (ei - (ke * zi) / (ke - kzi)) / expf(ke * pdt) +
       (ke * zi * expf(ke * pdt - kzi * pdt)) / (expf(ke * pdt) * (ke - kzi));
  zi = zi * expf(-kzi * pdt);
  *ppi = pi;
  *pei = ei;
  *pzi = zi;
}

void gor_ref32_syn(float kp, float ke, float kzi, float kzj, int pdt, float* ppij, float* peij,
                   float* pzi2, float* pzj2) {
  float pij = *ppij, eij = *peij, zi2 = *pzi2, zj2 = *pzj2;
  pij = (pij + ((eij * kp * kzi - eij * ke * kp + eij * kp * kzj + ke * kp * zi2 * zj2) / (ke - kp) -
                (ke * kp * zi2 * zj2) / (kzi - kp + kzj)) / (kzi - ke + kzj)) / expf(kp * pdt) -
        ((expf(kp * pdt - ke * pdt) * (eij * kp * kzi - eij * ke * kp + eij * kp * kzj + ke * kp * zi2 * zj2)) /
             (ke - kp) -
         (ke * kp * zi2 * zj2 * expf(kp * pdt - kzi * pdt - kzj * pdt)) / (kzi - kp + kzj)) /
            (expf(kp * pdt) * (kzi - ke + kzj));
  eij = (eij + (ke * zi2 * zj2) / (kzi - ke + kzj)) / expf(ke * pdt) -
        (ke * zi2 * zj2) / (expf(kzi * pdt) * expf(kzj * pdt) * (kzi - ke + kzj));
  zi2 = zi2 * expf(-kzi * pdt);
  zj2 = zj2 * expf(-kzj * pdt);
  *ppij = pij;
  *peij = eij;
  *pzi2 = zi2;
  *pzj2 = zj2;
}

/* ------------------------------------------------------------------ */
/* network state */

typedef struct {
  int id, H, M, hcu_start, mcu_start, dim_proj;
  float taumdt, wtagain, maxfqdt, igain, wgain, lgbias, snoise;
  float *dsup, *act, *rnd_u, *rnd_n;
  void *epsc, *bj; /* [dim_proj][N], float or double */
  int8_t* spike;
  int32_t* counter;
  int ext_nsteps;
  int32_t *ext_off, *ext_units;
} gor_pop_t;

typedef struct {
  float kp;
  int64_t len;
  float* tab; /* len * 12 */
  double dkp;
  double* dtab; /* len * 9: Ezi Ezj Ee Ep Cez_ij Cpe Cpz_ij Cez_i Cpz_i */
} gor_coefcache_t;

#define GOR_MAXCACHE 8

typedef struct {
  int id, src, dest, H, M, C, proj_in_pop, Nsrc;
  float taupdt, tauedt, tauzidt, tauzjdt, tauepscdt, eps, eps2, kfti, kftj, wgain, bgain;
  int32_t *ii, *qi, *di, *ti, *tij;
  int8_t *siq, *si;
  void *pi, *ei, *zi, *pj, *ej, *zj, *pij, *eij, *zi2, *zj2, *wij;
  int32_t *ssi, *ssj;
  int nssi, nssj;
  int store_prec;          /* GOR_STORE_*: rounding of stored Pij / Eij */
  int p_shift, e_shift;    /* binary16 form: x is stored as binary16(x * 2^shift) */
  int ncache;
  gor_coefcache_t cache[GOR_MAXCACHE];
  int mutation;
} gor_proj_t;

struct gor_net {
  int npop, nproj, arith, wta_pop_max;
  gor_pop_t* pops;
  gor_proj_t* projs;
  float dt;
  uint64_t seed;
  int simstep, cycle_flag, lginp_idx, wmask_idx;
  float prn, old_prn;
  const float* lginp;
  int lginp_rows;
  const float* wmask;
  int wmask_rows;
  int total_units, total_hcus;
  int64_t work[4];
  int mutation; /* gor_set_mutation: a planted mistake, 0 = none */
};

static size_t tsize(const gor_net_t* n) { return n->arith == GOR_ARITH_F64 ? sizeof(double) : sizeof(float); }

/* Large arrays are 2 MiB aligned and marked for transparent huge pages:
 * the column updates walk five per-synapse arrays at a stride of M floats
 * (the reference's layout), which with 4 KiB pages misses the TLB on almost
 * every access. */
static void* zalloc(size_t count, size_t sz) {
  const size_t bytes = (count ? count : 1) * sz;
  void* p = NULL;
  if (bytes >= ((size_t)4 << 20) && posix_memalign(&p, (size_t)2 << 20, bytes) == 0) {
    (void)madvise(p, bytes, MADV_HUGEPAGE);
    memset(p, 0, bytes);
    return p;
  }
  p = calloc(count ? count : 1, sz);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

static void fill_t(const gor_net_t* n, void* p, int64_t count, double v) {
  int64_t i;
  if (n->arith == GOR_ARITH_F64)
    for (i = 0; i < count; i++) ((double*)p)[i] = v;
  else
    for (i = 0; i < count; i++) ((float*)p)[i] = (float)v;
}

/* Pop::init_new (Pop.cpp:8-111) and Proj::init_new (Proj.cpp:6-107) */
gor_net_t* gor_net_new(float dt, int npop, const gor_popcfg_t* pc, int nproj,
                       const gor_projcfg_t* jc, uint64_t seed, int wta_pop_max, int arith) {
  gor_net_t* n = (gor_net_t*)zalloc(1, sizeof(gor_net_t));
  int i;
  n->npop = npop;
  n->nproj = nproj;
  n->arith = arith;
  n->dt = dt;
  n->seed = seed;
  n->wta_pop_max = wta_pop_max;
  n->pops = (gor_pop_t*)zalloc(npop, sizeof(gor_pop_t));
  n->projs = (gor_proj_t*)zalloc(nproj, sizeof(gor_proj_t));
  int hcnt = 0, mcnt = 0;
  for (i = 0; i < npop; i++) {
    gor_pop_t* p = &n->pops[i];
    p->id = i;
    p->H = pc[i].hcu;
    p->M = pc[i].mcu;
    p->hcu_start = hcnt;
    p->mcu_start = mcnt;
    hcnt += p->H;
    mcnt += p->H * p->M;
    p->taumdt = dt / pc[i].taum;
    p->wtagain = pc[i].wtagain;
    p->maxfqdt = pc[i].maxfq * dt;
    p->igain = pc[i].igain;
    p->wgain = pc[i].wgain;
    p->lgbias = pc[i].lgbias;
    p->snoise = pc[i].snoise;
    int N = p->H * p->M;
    p->dsup = (float*)zalloc(N, sizeof(float));
    p->act = (float*)zalloc(N, sizeof(float));
    p->rnd_u = (float*)zalloc(N, sizeof(float));
    p->rnd_n = (float*)zalloc(N, sizeof(float));
    p->spike = (int8_t*)zalloc(N, 1);
    p->counter = (int32_t*)zalloc(N, sizeof(int32_t));
    int j;
    for (j = 0; j < N; j++) {
      p->dsup[j] = (float)log(1.0 / p->M);
      p->act[j] = (float)(1.0 / p->M);
    }
  }
  n->total_units = mcnt;
  n->total_hcus = hcnt;
  for (i = 0; i < nproj; i++) {
    gor_proj_t* q = &n->projs[i];
    q->id = i;
    q->src = jc[i].src;
    q->dest = jc[i].dest;
    gor_pop_t* sp = &n->pops[q->src];
    gor_pop_t* dp = &n->pops[q->dest];
    q->H = dp->H;
    q->M = dp->M;
    q->Nsrc = sp->H * sp->M;
    q->C = q->Nsrc;
    if (q->C > jc[i].slot_num) q->C = (int)jc[i].slot_num;
    q->proj_in_pop = dp->dim_proj;
    dp->dim_proj++;
    q->tauzidt = dt / jc[i].tauzi;
    q->tauzjdt = dt / jc[i].tauzj;
    q->tauedt = dt / jc[i].taue;
    q->taupdt = dt / jc[i].taup;
    q->eps = dt / jc[i].taup;
    q->eps2 = (dt / jc[i].taup) * (dt / jc[i].taup);
    q->kfti = 1 / (jc[i].maxfq * jc[i].tauzi);
    q->kftj = 1 / (jc[i].maxfq * jc[i].tauzj);
    q->bgain = jc[i].bgain;
    q->wgain = jc[i].wgain;
    q->tauepscdt = jc[i].has_tauepsc ? dt / jc[i].tauepsc : q->tauzidt;
    int64_t R = (int64_t)q->H * q->C;
    int64_t S = R * q->M;
    int N = q->H * q->M;
    q->ii = (int32_t*)zalloc(R, 4);
    q->qi = (int32_t*)zalloc(R, 4);
    q->di = (int32_t*)zalloc(R, 4);
    q->ti = (int32_t*)zalloc(R, 4);
    q->siq = (int8_t*)zalloc(R, 1);
    q->ssi = (int32_t*)zalloc(R, 4);
    q->ssj = (int32_t*)zalloc(N, 4);
    q->si = (int8_t*)zalloc(q->Nsrc, 1);
    q->tij = (int32_t*)zalloc(S, 4);
    int64_t r;
    for (r = 0; r < R; r++) q->ii[r] = -1;
    size_t ts = tsize(n);
    q->pi = zalloc(R, ts);
    q->ei = zalloc(R, ts);
    q->zi = zalloc(R, ts);
    q->pj = zalloc(N, ts);
    q->ej = zalloc(N, ts);
    q->zj = zalloc(N, ts);
    q->pij = zalloc(S, ts);
    q->eij = zalloc(S, ts);
    q->zi2 = zalloc(S, ts);
    q->zj2 = zalloc(S, ts);
    q->wij = zalloc(S, ts);
    fill_t(n, q->pi, R, 1.0 / q->C);
    fill_t(n, q->pj, N, 1.0 / q->M);
    fill_t(n, q->pij, S, 1.0 / q->C / q->M);
  }
  /* epsc/bj slices per destination population (Proj.cpp:58-59) */
  for (i = 0; i < npop; i++) {
    gor_pop_t* p = &n->pops[i];
    int64_t cnt = (int64_t)(p->dim_proj ? p->dim_proj : 1) * p->H * p->M;
    p->epsc = zalloc(cnt, tsize(n));
    p->bj = zalloc(cnt, tsize(n));
  }
  return n;
}

void gor_net_free(gor_net_t* n) {
  int i, k;
  if (!n) return;
  for (i = 0; i < n->npop; i++) {
    gor_pop_t* p = &n->pops[i];
    free(p->dsup);
    free(p->act);
    free(p->rnd_u);
    free(p->rnd_n);
    free(p->spike);
    free(p->counter);
    free(p->epsc);
    free(p->bj);
    free(p->ext_off);
    free(p->ext_units);
  }
  for (i = 0; i < n->nproj; i++) {
    gor_proj_t* q = &n->projs[i];
    void* arrs[] = {q->ii, q->qi, q->di, q->ti, q->siq, q->ssi, q->ssj, q->si, q->tij, q->pi,
                    q->ei, q->zi, q->pj, q->ej, q->zj, q->pij, q->eij, q->zi2, q->zj2, q->wij};
    for (k = 0; k < (int)(sizeof(arrs) / sizeof(arrs[0])); k++) free(arrs[k]);
    for (k = 0; k < q->ncache; k++) {
      free(q->cache[k].tab);
      free(q->cache[k].dtab);
    }
  }
  free(n->pops);
  free(n->projs);
  free(n);
}

void gor_set_globals(gor_net_t* n, int simstep, float prn, float old_prn, int cycle_flag,
                     int lginp_idx, int wmask_idx) {
  n->simstep = simstep;
  n->prn = prn;
  n->old_prn = old_prn;
  n->cycle_flag = cycle_flag;
  n->lginp_idx = lginp_idx;
  n->wmask_idx = wmask_idx;
}

void gor_set_inputs(gor_net_t* n, const float* lginp, int lginp_rows, const float* wmask,
                    int wmask_rows) {
  n->lginp = lginp;
  n->lginp_rows = lginp_rows;
  n->wmask = wmask;
  n->wmask_rows = wmask_rows;
}

void gor_set_ext_spikes(gor_net_t* n, int pop, int nsteps, const int32_t* off,
                        const int32_t* units) {
  gor_pop_t* p = &n->pops[pop];
  free(p->ext_off);
  free(p->ext_units);
  p->ext_nsteps = nsteps;
  p->ext_off = (int32_t*)zalloc(nsteps + 1, 4);
  memcpy(p->ext_off, off, sizeof(int32_t) * (nsteps + 1));
  p->ext_units = (int32_t*)zalloc(off[nsteps], 4);
  memcpy(p->ext_units, units, sizeof(int32_t) * off[nsteps]);
}

/* ------------------------------------------------------------------ */
/* coefficient caches, filled on first use of each gap */

static gor_coefcache_t* coef_cache(gor_proj_t* q, float kp) {
  int i;
  for (i = 0; i < q->ncache; i++)
    if (q->cache[i].kp == kp) return &q->cache[i];
  if (q->ncache == GOR_MAXCACHE) {
    fprintf(stderr, "oracle: too many distinct learning rates\n");
    abort();
  }
  gor_coefcache_t* c = &q->cache[q->ncache++];
  memset(c, 0, sizeof(*c));
  c->kp = kp;
  c->dkp = (double)kp;
  return c;
}

static void coef_grow(gor_coefcache_t* c, int64_t need, int dbl) {
  if (need <= c->len) return;
  int64_t nl = c->len ? c->len : 64;
  while (nl < need) nl *= 2;
  if (dbl) {
    double* t = (double*)zalloc(nl * 9, sizeof(double));
    if (c->dtab) memcpy(t, c->dtab, sizeof(double) * 9 * c->len);
    int64_t i;
    for (i = c->len; i < nl; i++) t[i * 9] = -1.0; /* not computed */
    free(c->dtab);
    c->dtab = t;
  } else {
    float* t = (float*)zalloc(nl * 12, sizeof(float));
    if (c->tab) memcpy(t, c->tab, sizeof(float) * 12 * c->len);
    int64_t i;
    for (i = c->len; i < nl; i++) t[i * 12 + 11] = -1.0f; /* not computed */
    free(c->tab);
    c->tab = t;
  }
  c->len = nl;
}

static const float* coef_f32(gor_proj_t* q, float kp, int pdt) {
  gor_coefcache_t* c = coef_cache(q, kp);
  coef_grow(c, (int64_t)pdt + 1, 0);
  float* e = c->tab + (int64_t)pdt * 12;
  if (e[11] != 0.0f) {
    gor_coeff_entry((double)q->tauzidt, (double)q->tauzjdt, (double)q->tauedt, (double)kp, pdt, e);
  }
  return e;
}

static const double* coef_f64(gor_proj_t* q, double kp, int pdt) {
  gor_coefcache_t* c = coef_cache(q, (float)kp);
  coef_grow(c, (int64_t)pdt + 1, 1);
  double* e = c->dtab + (int64_t)pdt * 9;
  if (e[0] < 0.0) {
    ld D = (ld)pdt;
    ld kzi = q->tauzidt, kzj = q->tauzjdt, ke = q->tauedt, k = kp;
    e[0] = (double)expl(-kzi * D);
    e[1] = (double)expl(-kzj * D);
    e[2] = (double)expl(-ke * D);
    e[3] = (double)expl(-k * D);
    e[4] = (double)(ke * gor_phi(kzi + kzj, ke, D));
    e[5] = (double)(k * gor_phi(ke, k, D));
    e[6] = (double)(k * ke * gor_psi(kzi + kzj, ke, k, D));
    e[7] = (double)(ke * gor_phi(kzi, ke, D));
    e[8] = (double)(k * ke * gor_psi(kzi, ke, k, D));
  }
  return e;
}

/* table arithmetic (product semantics) */
static void tab_row(gor_proj_t* q, float kp, int pdt, float* pi, float* ei, float* zi) {
  const float* c = coef_f32(q, kp, pdt);
  float p = (c[3] * (*pi) + c[5] * (*ei)) + c[8] * (*zi);
  float e = c[2] * (*ei) + c[7] * (*zi);
  *pi = p;
  *ei = e;
  *zi = c[0] * (*zi);
}

static void tab_syn(gor_proj_t* q, float kp, int pdt, float* pij, float* eij, float* zi2,
                    float* zj2) {
  const float* c = coef_f32(q, kp, pdt);
  float z0 = (*zi2) * (*zj2);
  float p = (c[3] * (*pij) + c[5] * (*eij)) + c[6] * z0;
  float e = c[2] * (*eij) + c[4] * z0;
  *pij = p;
  *eij = e;
  *zi2 = c[0] * (*zi2);
  *zj2 = c[1] * (*zj2);
}

/* exact closed form in double */
static void f64_row(gor_proj_t* q, double kp, int pdt, double* pi, double* ei, double* zi) {
  const double* c = coef_f64(q, kp, pdt);
  double p = c[3] * (*pi) + c[5] * (*ei) + c[8] * (*zi);
  double e = c[2] * (*ei) + c[7] * (*zi);
  *pi = p;
  *ei = e;
  *zi = c[0] * (*zi);
}

static void f64_syn(gor_proj_t* q, double kp, int pdt, double* pij, double* eij, double* zi2,
                    double* zj2) {
  const double* c = coef_f64(q, kp, pdt);
  double z0 = (*zi2) * (*zj2);
  double p = c[3] * (*pij) + c[5] * (*eij) + c[6] * z0;
  double e = c[2] * (*eij) + c[4] * z0;
  *pij = p;
  *eij = e;
  *zi2 = c[0] * (*zi2);
  *zj2 = c[1] * (*zj2);
}

static void r32_row(gor_proj_t* q, float kp, int pdt, float* pi, float* ei, float* zi) {
  gor_ref32_row(kp, q->tauedt, q->tauzidt, pdt, pi, ei, zi);
}

static void r32_syn(gor_proj_t* q, float kp, int pdt, float* pij, float* eij, float* zi2,
                    float* zj2) {
  gor_ref32_syn(kp, q->tauedt, q->tauzidt, q->tauzjdt, pdt, pij, eij, zi2, zj2);
}

/* ------------------------------------------------------------------ */
/* type-independent Proj parts */

/* Proj::update_ssi_cpu, Proj.cpp:621-630 */
static void proj_update_ssi(gor_proj_t* q) {
  int i;
  q->nssi = 0;
  for (i = 0; i < q->C * q->H; i++)
    if (q->siq[i] > 0) q->ssi[q->nssi++] = i;
}

/* Proj::update_ssj_cpu, Proj.cpp:632-643 */
static void proj_update_ssj(gor_net_t* n, gor_proj_t* q) {
  const int8_t* sj = n->pops[q->dest].spike;
  int i;
  q->nssj = 0;
  for (i = 0; i < q->H * q->M; i++)
    if (sj[i] > 0) q->ssj[q->nssj++] = i;
}

/* Proj::update_que_cpu, Proj.cpp:645-664 */
static void proj_update_que(gor_proj_t* q) {
  int i;
  for (i = 0; i < q->C * q->H; i++) {
    if (q->ii[i] < 0) continue;
    q->siq[i] = (int8_t)(q->qi[i] & 0x01);
    q->qi[i] >>= 1;
    int8_t spk = q->si[q->ii[i]];
    if (spk > 0) q->qi[i] |= (0x01 << (q->di[i] - 1));
  }
}

/* ------------------------------------------------------------------ */
/* 16-bit storage of Pij / Eij.  Independent of the product's integer
 * bit manipulation: the value is rounded in double arithmetic to the
 * format's quantum at its binade with nearbyint (round half to even in the
 * default rounding mode).
 *   bfloat16: 8 significant bits, exponent range of float (smallest
 *             quantum 2^-133, overflow above the largest finite value);
 *   binary16: 11 significant bits of x * 2^shift, smallest quantum 2^-24,
 *             results above 65504 overflow to infinity. */
float gor_round_trace(int prec, float x, int shift) {
  if (prec == GOR_STORE_F32 || isnan(x) || isinf(x) || x == 0.0f) return x;
  const int bits = prec == GOR_STORE_BF16 ? 8 : 11;
  const int qmin = prec == GOR_STORE_BF16 ? -133 : -24;
  const double y = prec == GOR_STORE_BF16 ? (double)x : ldexp((double)x, shift);
  int e;
  (void)frexp(y, &e); /* |y| in [2^(e-1), 2^e) */
  int qe = e - bits;
  if (qe < qmin) qe = qmin;
  double r = ldexp(nearbyint(ldexp(y, -qe)), qe);
  if (prec == GOR_STORE_F16) {
    if (fabs(r) > 65504.0) r = copysign(INFINITY, y);
    return (float)ldexp(r, -shift);
  }
  return (float)r; /* above the largest bfloat16 value this is 2^128 -> inf */
}

void gor_round_trace_n(int prec, int shift, int64_t count, const float* in, float* out) {
  for (int64_t i = 0; i < count; i++) out[i] = gor_round_trace(prec, in[i], shift);
}

int gor_set_trace_form(gor_net_t* n, int proj, int prec, int p_shift, int e_shift) {
  if (proj < 0 || proj >= n->nproj || prec < GOR_STORE_F32 || prec > GOR_STORE_F16) return -1;
  if (n->arith == GOR_ARITH_F64) return prec == GOR_STORE_F32 ? 0 : -1;
  gor_proj_t* q = &n->projs[proj];
  q->store_prec = prec;
  q->p_shift = p_shift;
  q->e_shift = e_shift;
  const int64_t S = (int64_t)q->H * q->C * q->M;
  float* pij = (float*)q->pij;
  float* eij = (float*)q->eij;
  for (int64_t i = 0; i < S; i++) {
    pij[i] = gor_round_trace(prec, pij[i], p_shift);
    eij[i] = gor_round_trace(prec, eij[i], e_shift);
  }
  return 0;
}

/* three instantiations of the Proj step */
#define T float
#define SFX(x) tab_##x
#define ROW_CATCHUP(p, kp, pdt, a, b, c) tab_row(p, kp, pdt, a, b, c)
#define SYN_CATCHUP(p, kp, pdt, a, b, c, d) tab_syn(p, kp, pdt, a, b, c, d)
#define LOGT(x) gsbn_logf(x)
#define STORE_P(p, x) gor_round_trace((p)->store_prec, (x), (p)->p_shift)
#define STORE_E(p, x) gor_round_trace((p)->store_prec, (x), (p)->e_shift)
#include "proj_step.inc"
#undef T
#undef SFX
#undef ROW_CATCHUP
#undef SYN_CATCHUP
#undef LOGT
#undef STORE_P
#undef STORE_E

#define T float
#define SFX(x) r32_##x
#define ROW_CATCHUP(p, kp, pdt, a, b, c) r32_row(p, kp, pdt, a, b, c)
#define SYN_CATCHUP(p, kp, pdt, a, b, c, d) r32_syn(p, kp, pdt, a, b, c, d)
#define LOGT(x) logf(x)
#define STORE_P(p, x) gor_round_trace((p)->store_prec, (x), (p)->p_shift)
#define STORE_E(p, x) gor_round_trace((p)->store_prec, (x), (p)->e_shift)
#include "proj_step.inc"
#undef T
#undef SFX
#undef ROW_CATCHUP
#undef SYN_CATCHUP
#undef LOGT
#undef STORE_P
#undef STORE_E

#define T double
#define SFX(x) f64_##x
#define ROW_CATCHUP(p, kp, pdt, a, b, c) f64_row(p, kp, pdt, a, b, c)
#define SYN_CATCHUP(p, kp, pdt, a, b, c, d) f64_syn(p, kp, pdt, a, b, c, d)
#define LOGT(x) log(x)
#define STORE_P(p, x) (x)
#define STORE_E(p, x) (x)
#include "proj_step.inc"
#undef T
#undef SFX
#undef ROW_CATCHUP
#undef SYN_CATCHUP
#undef LOGT
#undef STORE_P
#undef STORE_E

static void proj_step(gor_net_t* n, gor_proj_t* q) {
  if (n->arith == GOR_ARITH_TABLE)
    tab_proj_step(n, q);
  else if (n->arith == GOR_ARITH_REF32)
    r32_proj_step(n, q);
  else
    f64_proj_step(n, q);
}

/* ------------------------------------------------------------------ */
/* Pop (Pop.cpp:220-409); population arithmetic is float in every variant */

/* update_sup_kernel_1_cpu, Pop.cpp:229-261 */
static void pop_kernel_1(gor_net_t* n, gor_pop_t* p, int i, int j, const float* lginp,
                         const float* wmask, const float* epsc, const float* bj) {
  int idx = i * p->M + j;
  float wsup = 0;
  int offset = 0;
  int mcu_num_in_pop = p->H * p->M;
  int m;
  for (m = 0; m < p->dim_proj; m++) {
    wsup += (n->mutation == GOR_MUT_SUP_NO_BJ ? 0.0f : bj[offset + idx]) + epsc[offset + idx];
    offset += mcu_num_in_pop;
  }
  float sup = p->lgbias + p->igain * lginp[idx] + p->rnd_n[idx];
  sup += (p->wgain * wmask[i]) * wsup;
  float dsup = p->dsup[idx];
  p->dsup[idx] += (sup - dsup) * p->taumdt;
}

/* update_sup_kernel_2_cpu, Pop.cpp:263-303; per-HCU maximum unless
 * wta_pop_max selects the CPU seed ptr_dsup[0] (Pop.cpp:270) */
static void pop_kernel_2(gor_net_t* n, gor_pop_t* p, int i) {
  int M = p->M, m;
  float maxdsup = n->wta_pop_max ? p->dsup[0] : p->dsup[i * M];
  for (m = 0; m < M; m++) {
    float dsup = p->dsup[i * M + m];
    if (dsup > maxdsup) maxdsup = dsup;
  }
  float maxact = gsbn_expf(p->wtagain * maxdsup);
  float vsum = 0;
  for (m = 0; m < M; m++) {
    int idx = i * M + m;
    float dsup = p->dsup[idx];
    float act = gsbn_expf(p->wtagain * (dsup - maxdsup));
    if (maxact < 1) act *= maxact;
    vsum += act;
    p->act[idx] = act;
  }
  if (vsum > 1)
    for (m = 0; m < M; m++) p->act[i * M + m] /= vsum;
}

/* update_sup_kernel_3_cpu, Pop.cpp:305-323 */
static void pop_kernel_3(gor_net_t* n, gor_pop_t* p, int i, int j) {
  int idx = i * p->M + j;
  int8_t spk = n->mutation == GOR_MUT_SPIKE_NO_MAXFQ ? (int8_t)(p->rnd_u[idx] < p->act[idx])
                                                     : (int8_t)(p->rnd_u[idx] < p->act[idx] * p->maxfqdt);
  p->spike[idx] = spk;
  p->counter[idx] += spk;
}

static void pop_step(gor_net_t* n, gor_pop_t* p) {
  int N = p->H * p->M, i, j;
  /* update_rnd_cpu (Pop.cpp:220-227): uniform and snoise * normal per unit */
  for (i = 0; i < N; i++) {
    float u, z;
    gor_draw(n->seed, (uint32_t)p->id, (uint32_t)n->simstep, (uint32_t)i, &u, &z);
    p->rnd_u[i] = u;
    p->rnd_n[i] = p->snoise * z;
  }
  /* update_sup_cpu (Pop.cpp:325-387) */
  const float* wmask = n->wmask + (int64_t)n->wmask_idx * n->total_hcus + p->hcu_start;
  const float* lginp = n->lginp + (int64_t)n->lginp_idx * n->total_units + p->mcu_start;
  for (i = 0; i < p->H; i++) {
    for (j = 0; j < p->M; j++)
      pop_kernel_1(n, p, i, j, lginp, wmask, (const float*)p->epsc, (const float*)p->bj);
    pop_kernel_2(n, p, i);
    for (j = 0; j < p->M; j++) pop_kernel_3(n, p, i, j);
  }
  /* fill_spike (Pop.cpp:389-409) */
  if (p->ext_off) {
    memset(p->spike, 0, N);
    int s = n->simstep;
    if (s >= 0 && s < p->ext_nsteps) {
      int k;
      for (k = p->ext_off[s]; k < p->ext_off[s + 1]; k++) p->spike[p->ext_units[k]] = 1;
    }
  }
}

/* ProcUpdLazy::update_cpu (ProcUpdLazy.cpp:88-124) */
void gor_updlazy_step(gor_net_t* n) {
  int i;
  if (n->cycle_flag != 1) return;
  if (n->arith == GOR_ARITH_F64) {
    fprintf(stderr, "oracle: FP64 networks run in replay mode only\n");
    abort();
  }
  for (i = 0; i < n->npop; i++) pop_step(n, &n->pops[i]);
  for (i = 0; i < n->nproj; i++) proj_step(n, &n->projs[i]);
  n->work[3] += 1;
}

void gor_proj_step_replay(gor_net_t* n, int proj, const int8_t* sj) {
  gor_proj_t* q = &n->projs[proj];
  memcpy(n->pops[q->dest].spike, sj, (size_t)q->H * q->M);
  proj_step(n, q);
}

/* ProcExchangeSpike::update_cpu for one process (ProcExchangeSpike.cpp:94-135) */
void gor_exchange_spikes(gor_net_t* n) {
  int i;
  if (n->cycle_flag != 1) return;
  for (i = 0; i < n->nproj; i++) {
    gor_proj_t* q = &n->projs[i];
    memcpy(q->si, n->pops[q->src].spike, (size_t)q->Nsrc);
  }
}

float gor_det_logf(float x) { return gsbn_logf(x); }
float gor_det_expf(float x) { return gsbn_expf(x); }

void gor_work_counters(gor_net_t* n, int64_t* out4) { memcpy(out4, n->work, sizeof(n->work)); }

int gor_set_pop_id(gor_net_t* n, int pop, int id) {
  if (pop < 0 || pop >= n->npop || id < 0) return -1;
  n->pops[pop].id = id;
  return 0;
}

void gor_set_mutation(gor_net_t* n, int code) {
  int i;
  n->mutation = code;
  for (i = 0; i < n->nproj; i++) n->projs[i].mutation = code;
}

/* ------------------------------------------------------------------ */
/* vector access by Database name */

typedef struct {
  void* ptr;
  int64_t count;
  int dtype; /* 0 i8, 2 i32, 5 f32, 6 f64 */
} gor_vec;

static int gor_lookup(gor_net_t* n, const char* name, gor_vec* v) {
  char base[64];
  int id = -1;
  const char* us = strrchr(name, '_');
  if (!us || us == name) return -1;
  size_t bl = (size_t)(us - name);
  if (bl >= sizeof(base)) return -1;
  memcpy(base, name, bl);
  base[bl] = 0;
  id = atoi(us + 1);
  int ft = n->arith == GOR_ARITH_F64 ? 6 : 5;
  if (id >= 0 && id < n->npop) {
    gor_pop_t* p = &n->pops[id];
    int64_t N = (int64_t)p->H * p->M;
    int64_t PN = (int64_t)(p->dim_proj ? p->dim_proj : 1) * N;
    if (!strcmp(base, "dsup")) { v->ptr = p->dsup; v->count = N; v->dtype = 5; return 0; }
    if (!strcmp(base, "act")) { v->ptr = p->act; v->count = N; v->dtype = 5; return 0; }
    if (!strcmp(base, "epsc")) { v->ptr = p->epsc; v->count = PN; v->dtype = ft; return 0; }
    if (!strcmp(base, "bj")) { v->ptr = p->bj; v->count = PN; v->dtype = ft; return 0; }
    if (!strcmp(base, "spike")) { v->ptr = p->spike; v->count = N; v->dtype = 0; return 0; }
    if (!strcmp(base, ".rnd_uniform01")) { v->ptr = p->rnd_u; v->count = N; v->dtype = 5; return 0; }
    if (!strcmp(base, ".rnd_normal")) { v->ptr = p->rnd_n; v->count = N; v->dtype = 5; return 0; }
    if (!strcmp(base, ".counter")) { v->ptr = p->counter; v->count = N; v->dtype = 2; return 0; }
  }
  if (id >= 0 && id < n->nproj) {
    gor_proj_t* q = &n->projs[id];
    int64_t R = (int64_t)q->H * q->C, S = R * q->M, N = (int64_t)q->H * q->M;
    struct { const char* nm; void* p; int64_t c; int t; } tab[] = {
        {"ii", q->ii, R, 2},     {"qi", q->qi, R, 2},     {"di", q->di, R, 2},
        {".siq", q->siq, R, 0},  {"ti", q->ti, R, 2},     {"tij", q->tij, S, 2},
        {"si", q->si, q->Nsrc, 0}, {".ssi", q->ssi, q->nssi, 2}, {".ssj", q->ssj, q->nssj, 2},
        {"pi", q->pi, R, ft},    {"ei", q->ei, R, ft},    {"zi", q->zi, R, ft},
        {"pj", q->pj, N, ft},    {"ej", q->ej, N, ft},    {"zj", q->zj, N, ft},
        {"pij", q->pij, S, ft},  {"eij", q->eij, S, ft},  {"zi2", q->zi2, S, ft},
        {"zj2", q->zj2, S, ft},  {"wij", q->wij, S, ft},
    };
    size_t k;
    for (k = 0; k < sizeof(tab) / sizeof(tab[0]); k++) {
      if (!strcmp(base, tab[k].nm)) {
        v->ptr = tab[k].p;
        v->count = tab[k].c;
        v->dtype = tab[k].t;
        return 0;
      }
    }
  }
  return -1;
}

static size_t dsz(int t) { return t == 0 ? 1 : t == 2 ? 4 : t == 5 ? 4 : 8; }

int64_t gor_vector(gor_net_t* n, const char* name, int* dtype, void* out, int64_t cap) {
  gor_vec v;
  if (gor_lookup(n, name, &v)) return -1;
  if (dtype) *dtype = v.dtype;
  if (out) {
    int64_t c = v.count < cap ? v.count : cap;
    memcpy(out, v.ptr, (size_t)c * dsz(v.dtype));
  }
  return v.count;
}

int gor_vector_write(gor_net_t* n, const char* name, const void* src, int64_t count) {
  gor_vec v;
  if (gor_lookup(n, name, &v)) return -1;
  if (count != v.count) return -2;
  memcpy(v.ptr, src, (size_t)count * dsz(v.dtype));
  return 0;
}
