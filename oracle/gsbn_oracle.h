/*
 * gsbn_oracle.h -- CPU restatement of the reference's ProcUpdLazy procedure.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the product (gsbn_amd)
 * never links or calls it.
 *
 * It restates, in plain C and in the reference's loop order:
 *   Pop  (src/gsbn/procedures/ProcUpdLazy/Pop.cpp:220-409)
 *   Proj (src/gsbn/procedures/ProcUpdLazy/Proj.cpp:6-107, 209-760)
 *   the step order of ProcUpdLazy::update_cpu (ProcUpdLazy.cpp:88-124)
 *   ProcExchangeSpike (ProcExchangeSpike.cpp:94-135, single process)
 *   ProcCheck counting and scoring (ProcCheck.cpp:149-237, ProcCheck/Pop.cpp:35-60)
 * with three documented substitutions, each pinned by its own tests:
 *   1. random numbers: Philox4x32-10 keyed by (seed, pop, step, unit) instead of
 *      std::random_device / cuRAND (the reference is not reproducible);
 *   2. closed-form catch-up: coefficients from a double-precision table per
 *      elapsed gap instead of the float expression of Proj.cpp:241-247,269-277
 *      (GOR_ARITH_TABLE); the literal float expression is kept as
 *      GOR_ARITH_REF32 and an exact double-precision closed form as
 *      GOR_ARITH_F64 for error measurement;
 *   3. exp/log in the float paths use the deterministic routines of
 *      gsbn_amd/csrc/detmath.h (pinned against glibc to <= 1 ulp in
 *      tests/test_detmath.py) so the GPU can reproduce them bit for bit.
 * The WTA soft-max supports both maxima the reference uses: per-HCU
 * (Pop.cu:73-79, default) and the CPU quirk seeded with the population's first
 * unit (Pop.cpp:270).
 */
#ifndef GSBN_ORACLE_H
#define GSBN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GOR_ARITH_TABLE = 0, GOR_ARITH_REF32 = 1, GOR_ARITH_F64 = 2 };

typedef struct {
  int hcu, mcu;
  float taum, wtagain, maxfq, igain, wgain, lgbias, snoise;
} gor_popcfg_t;

typedef struct {
  int src, dest;
  float tauzi, tauzj, taue, taup, maxfq, bgain, wgain, tauepsc;
  int has_tauepsc;
  float slot_num;
} gor_projcfg_t;

typedef struct gor_net gor_net_t;

/* network construction (Pop::init_new, Proj::init_new) */
gor_net_t* gor_net_new(float dt, int npop, const gor_popcfg_t* pops, int nproj,
                       const gor_projcfg_t* projs, uint64_t seed, int wta_pop_max,
                       int arith);
void gor_net_free(gor_net_t* n);

/* global variables read by the procedure (GlobalVar keys simstep, prn,
 * old-prn, cycle-flag, lginp-idx, wmask-idx) */
void gor_set_globals(gor_net_t* n, int simstep, float prn, float old_prn, int cycle_flag,
                     int lginp_idx, int wmask_idx);
/* .lginp and .wmask tables (row-major; ld = total units / total HCUs) */
void gor_set_inputs(gor_net_t* n, const float* lginp, int lginp_rows, const float* wmask,
                    int wmask_rows);
/* external spikes for a population: CSR by simulation step (Pop.cpp:67-106) */
void gor_set_ext_spikes(gor_net_t* n, int pop, int nsteps, const int32_t* off,
                        const int32_t* units);

/* ProcUpdLazy::update_cpu (ProcUpdLazy.cpp:88-124) */
void gor_updlazy_step(gor_net_t* n);
/* ProcExchangeSpike::update_cpu, one process: si_<proj> <- spike_<src> */
void gor_exchange_spikes(gor_net_t* n);

/* Proj-only stepping for replay runs: the destination population's spikes
 * are taken from sj (H*M int8) instead of a Pop update. */
void gor_proj_step_replay(gor_net_t* n, int proj, const int8_t* sj);

/* Storage form of Pij / Eij in the float variants: GOR_STORE_F32 (plain),
 * GOR_STORE_BF16, GOR_STORE_F16 (binary16 of x * 2^shift); every stored
 * value is rounded to nearest, ties to even, Wij uses the unrounded value.
 * Rounds the current state; returns 0, or -1 for bad arguments or a
 * 16-bit form on the double variant. */
enum { GOR_STORE_F32 = 0, GOR_STORE_BF16 = 1, GOR_STORE_F16 = 2 };
int gor_set_trace_form(gor_net_t* n, int proj, int prec, int p_shift, int e_shift);
float gor_round_trace(int prec, float x, int shift);
void gor_round_trace_n(int prec, int shift, int64_t count, const float* in, float* out);

/* vector access by the reference's Database names, e.g. "pij_0", "spike_0".
 * Returns the element count and the element type code (0 i8, 2 i32, 5 f32,
 * 6 f64) or -1 if unknown; copies min(count, cap) elements when out != 0. */
int64_t gor_vector(gor_net_t* n, const char* name, int* dtype, void* out, int64_t cap);
int gor_vector_write(gor_net_t* n, const char* name, const void* src, int64_t count);

/* coefficient entry of the table path (same layout as the product's table) */
void gor_coeff_entry(double kzi, double kzj, double ke, double kp, int64_t gap, float* out);
/* exact closed form in long double for one synapse/row, for pinning tests:
 * state = {p, e, z} advanced by gap with rates (kz, ke, kp) */
void gor_closed_form_f64(double kz, double ke, double kp, int64_t gap, const double* in,
                         double* out);
/* reference literal float expressions (Proj.cpp:241-247 / 269-277) */
void gor_ref32_row(float kp, float ke, float kzi, int pdt, float* pi, float* ei, float* zi);
void gor_ref32_syn(float kp, float ke, float kzi, float kzj, int pdt, float* pij, float* eij,
                   float* zi2, float* zj2);

/* independent Philox4x32-10 and the per-unit draw */
void gor_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void gor_draw(uint64_t seed, uint32_t pop, uint32_t step, uint32_t unit, float* u, float* z);
/* the same uniform / normal mapping for an arbitrary counter */
void gor_draw_ctr(const uint32_t ctr[4], const uint32_t key[2], float* u, float* z);

/* the deterministic exp/log of detmath.h, for host-side constants */
float gor_det_logf(float x);
float gor_det_expf(float x);

/* counters of work done since creation: row synapse updates, column
 * synapse updates, full-sweep synapse updates, steps */
void gor_work_counters(gor_net_t* n, int64_t* out4);

/* the population id that keys the population's random draws (default: its
 * index); a network that holds a few populations of a larger batch uses
 * their ids in the batch so that their draws are the batch's */
int gor_set_pop_id(gor_net_t* n, int pop, int id);

/* Planted mistakes for testing the tests (default 0: none).  Each makes
 * the oracle deviate from the reference in one plausible way, so that a pin
 * that is meant to fix that part of the algorithm can be shown to fail:
 *   SWEEP_NEW_PRN   full sweep with kp = taupdt * prn instead of old-prn
 *   SWEEP_TARGET_T  full sweep to simstep instead of simstep - 1
 *   SWEEP_W_KP0     full sweep recomputes Wij also when kp == 0
 *   SWEEP_OLD_PI    full sweep's Wij of slot j == 0 uses pi before its catch-up
 *   SUP_NO_BJ       support sum leaves out bj
 *   SPIKE_NO_MAXFQ  spike when u < act instead of u < act * maxfq * dt */
enum {
  GOR_MUT_NONE = 0,
  GOR_MUT_SWEEP_NEW_PRN = 1,
  GOR_MUT_SWEEP_TARGET_T = 2,
  GOR_MUT_SWEEP_W_KP0 = 3,
  GOR_MUT_SWEEP_OLD_PI = 4,
  GOR_MUT_SUP_NO_BJ = 5,
  GOR_MUT_SPIKE_NO_MAXFQ = 6
};
void gor_set_mutation(gor_net_t* n, int code);

#ifdef __cplusplus
}
#endif

#endif
