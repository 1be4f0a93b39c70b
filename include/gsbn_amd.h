/*
 * gsbn_amd.h -- C ABI of the MI355X lazy-update BCPNN core.
 *
 * Two levels:
 *
 * 1. Kernel operations.  Each replaces one device-side piece of the
 *    reference's ProcUpdLazy (src/gsbn/procedures/ProcUpdLazy/{Pop,Proj}.cu);
 *    the arguments are plain structs of device pointers, sizes and float
 *    coefficients; `stream` is a hipStream_t (NULL = default stream).  Every
 *    function returns 0 on success or a negative gsbn_status / positive
 *    hipError_t code and never aborts.  Kernels allocate nothing.
 *
 * 2. Solver.  A complete configuration-driven simulator (the reference's
 *    Solver + Upd + procedure plug-ins ProcExtGen, ProcUpdLazy, ProcInitConn,
 *    ProcExchangeSpike, ProcCheck) behind an opaque handle, reading the
 *    reference's prototxt / StimRawData / conn-map formats and exposing every
 *    state vector under the reference's Database names.
 *
 * State layout: per-synapse state is one gsbn_syn_t record per synapse, in
 * the reference's (h*C + k)*M + j order; per-row state one gsbn_rowst_t per
 * row (h*C + k).  The solver presents both as the reference's separate
 * vectors (pij_<id>, eij_<id>, ..., pi_<id>, ...).
 */
#ifndef GSBN_AMD_H
#define GSBN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gsbn_status {
  GSBN_OK = 0,
  GSBN_ERR_ARG = -1,        /* invalid argument or size */
  GSBN_ERR_CONFIG = -2,     /* configuration / file error */
  GSBN_ERR_NAME = -3,       /* unknown vector or global name */
  GSBN_ERR_STATE = -4,      /* operation not valid in the current state */
  GSBN_ERR_UNSUPPORTED = -5 /* outside this build's scope */
};

/* ---- state records (must match gsbn_amd/csrc/gsbn_core.h) ---- */
typedef struct {
  float pij, eij, zi2, zj2, wij;
  int32_t tij;
  int32_t pad0, pad1;
} gsbn_syn_rec;

typedef struct {
  float pi, ei, zi;
  int32_t ti;
} gsbn_row_rec;

/* Reduced-precision trace forms: Pij and Eij stored in 16 bits, all
 * arithmetic in float.  GSBN_PREC_BF16 stores bfloat16(x); GSBN_PREC_F16
 * stores binary16(x * scale) with a power-of-two scale per projection (one
 * for Pij, one for Eij).  Conversions round to nearest, ties to even.  The
 * weights live in a separate float array in these forms. */
#define GSBN_PREC_F32 0
#define GSBN_PREC_BF16 1
#define GSBN_PREC_F16 2
typedef struct {
  uint16_t pij, eij;
  float zi2, zj2;
  int32_t tij;
} gsbn_syn16_rec;

#define GSBN_FUT 64 /* slots of the pending-activation ring (delays <= 31) */
#define GSBN_MAX_PROJ_IN 8
#define GSBN_KP_SLOTS 4 /* distinct learning rates a batched projection can switch between */

/* per-step fields shared by all populations / projections of a batch that
 * run in lock-step (one schedule) */
typedef struct {
  int32_t simstep, exec, do_jxx, kp_slot;
  float prn;
  int32_t prec; /* trace form (GSBN_PREC_*) of every projection of a batched row/column launch */
  int64_t lginp_offset; /* added to every population's lginp pointer */
  int64_t wmask_offset; /* added to every population's wmask pointer */
  /* Row overflow queue (ovf_keep > 0; gsbn_pop_update_batch_wave with
   * gsbn_proj_rowcol_batch_wave): a hypercolumn whose active rows this step
   * number more than ovf_keep keeps the first ovf_keep for its own row/column
   * wave and queues the rest, one task per row, {proj0.bidx, hcu, list
   * position}, in one of 64 sub-queues chosen by hashing (bidx, hcu): task i
   * of sub-queue q at ovf_task[3 * (q * (ovf_cap / 64) + i)], i counted by
   * ovf_cnt[((exec % 2) * 64 + q) * 16] (2 * 64 * 16 words; the population
   * launch of step x clears the counters of parity x + 1; all start at 0).
   * The row/column launch's helper waves run the tasks; when a sub-queue
   * holds more than ovf_cap / 64 rows, its hypercolumns run all of their
   * rows themselves. */
  int32_t ovf_keep, ovf_cap;
  int32_t* ovf_cnt;
  int32_t* ovf_task;
} gsbn_step_args;

/* one projection entering a population, as seen by gsbn_pop_update */
typedef struct {
  /* pending support contributions of the previous step's row update:
   * contrib[(h*C + a)*M + j] for a < pend_cnt[h] (ascending row order) */
  const float* contrib;
  int32_t* pend_cnt;
  int32_t C;
  /* postsynaptic traces (Proj.cpp:310-363) */
  float* pj;
  float* ej;
  float* zj;
  float taupdt, ke, kzj, kepsc, kftj, bgain, eps;
  /* pending activations of this projection's rows (see gsbn_fanout):
   * the population kernel reads and clears slot exec % GSBN_FUT */
  uint32_t* fut;
  int32_t W; /* words per HCU and slot, (C + 31) / 32 */
  /* outputs: this step's active rows per HCU, ascending (the reference's
   * .ssi_<id>, Proj.cpp:621-630), and the exec stamp of each active row */
  int32_t* rows_list;
  int32_t* rows_cnt;
  int32_t* act_stamp;
  /* Deferred column updates (dK > 0, a power of two): the population
   * kernel records each spike of postsynaptic unit u at step s in u's ring,
   * dhist[u*dK + n % dK] = s for its n-th spike (dcnt[u] spikes so far), and
   * the column update of that spike waits: rows replay it when they are
   * next updated (see gsbn_rowcol_args).  When a unit's pending spikes since
   * its last column walk (dfl[u] = dcnt[u] then) reach dK - 1, the unit is
   * listed for a walk this step: walk_list/walk_cnt per HCU (ascending), with
   * dcnt[u] at listing in walk_T. */
  int32_t* dhist;
  int32_t* dcnt;
  int32_t* dfl;
  int32_t* walk_list;
  int32_t* walk_cnt;
  int32_t* walk_T;
  int32_t dK;
  int32_t bidx; /* this projection's index in a batch's row/column argument array
                   (the row overflow queue's tasks name it) */
} gsbn_pop_proj;

/* Delay queue in push form (Proj.cpp:645-664 with ProcExchangeSpike's
 * one-step lag): a spike of source unit n at exec step s makes every row
 * r = h*C + k with ii[r] == n active at exec step s + di[r] + 2.  Pending
 * activations are a ring of bitmaps, one bit per row:
 *   fut[((x % GSBN_FUT) * H + h) * W + (k >> 5)], bit k & 31,
 * set for exec step x by the spike's push and cleared when step x reads it.
 * fwd_off/fwd_row/fwd_del list the rows of each source unit (CSR over the
 * source units, rows ascending) with their delays. */
typedef struct {
  const int32_t* fwd_off; /* [Nsrc + 1] */
  const int32_t* fwd_row; /* r = h*C + k */
  const uint8_t* fwd_del; /* di[r] */
  uint32_t* fut;          /* ring of the destination projection */
  int32_t H, C, W, pad;   /* destination HCUs, slots, words per HCU and slot */
} gsbn_fanout;

/* Fused population step: epsc finalisation, random draws, support, dsup,
 * WTA soft-max, spike sampling and counters (Pop.cpp:220-387, Pop.cu:17-108),
 * external-spike override (Pop.cpp:389-409), and for every incoming
 * projection its postsynaptic update jxx (Proj.cpp:310-363; skipped when
 * do_jxx == 0), its active-column list (.ssj, Proj.cpp:632-643) and its
 * active-row list (.ssi via the delay queue, Proj.cpp:621-630, 645-664).
 * Grid: one workgroup per HCU. */
typedef struct {
  int32_t pop_id, H, M, simstep, exec;
  uint64_t seed;
  float taumdt, wtagain, maxfqdt, igain, wgain, lgbias, snoise, prn;
  int32_t wta_pop_max; /* 1: maximum seeded with the population's first unit (Pop.cpp:270) */
  const float* lginp;  /* N, this population's slice of the current .lginp row */
  const float* wmask;  /* H, this population's slice of the current .wmask row */
  float* dsup;
  float* act;
  int8_t* spike;
  int32_t* counter;
  float* rnd_u;
  float* rnd_n;
  float* epsc; /* [nproj_in][N] */
  float* bj;   /* [nproj_in][N] */
  float* scratch_dsup0; /* device float, used when wta_pop_max (filled by this call) */
  const int32_t* ext_off; /* external spikes CSR by simstep, or NULL */
  const int32_t* ext_units;
  int32_t ext_nsteps;
  int32_t* spk_list; /* [H][M] spiking units of each HCU, ascending */
  int32_t* spk_cnt;  /* [H] */
  /* outgoing projections whose queue push this kernel does itself: when
   * ProcExchangeSpike delivers spike_<pop> to si_<proj> once per step, the
   * push of step x+1 is this step's spikes, so they go straight into the
   * rings (0 outgoing: the push is left to gsbn_proj_que) */
  int32_t nproj_out, pad_out;
  const gsbn_fanout* outs; /* device array */
  int32_t nproj_in, do_jxx;
  int32_t max_slots; /* largest C among the incoming projections (sizes the row bitmap) */
  const gsbn_pop_proj* projs; /* device array, index = position of the projection in the population */
  int64_t* work; /* optional device counters [2*nproj_in]: row and column synapse updates */
  /* copies of projs[0] (when nproj_in > 0) and outs[0] (when nproj_out > 0):
   * the kernel takes the first projection's pointers from its arguments, so
   * its loads do not wait for a read of the device arrays */
  gsbn_pop_proj proj0;
  gsbn_fanout out0;
} gsbn_pop_args;
int gsbn_pop_update(const gsbn_pop_args* a, void* stream);

/* gsbn_pop_update without the workspace argument: scratch_dsup0 is ignored
 * and the workspace (one float, the population's first unit's dsup when
 * wta_pop_max) is taken from a fixed device array of 64 floats that the
 * library holds per GPU until the process ends.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It allocates nothing; every call on one GPU shares that
 * GPU's array, so calls that overlap on one GPU (on different streams) race
 * and must be serialised by the caller. */
int gsbn_pop_update_noalloc(const gsbn_pop_args* a, void* stream);

/* Apply pending support contributions to epsc without a population update
 * (used before host reads of epsc_<pop>). Grid: one workgroup per HCU. */
int gsbn_pop_flush(const gsbn_pop_args* a, void* stream);

/* Full catch-up on a learning-rate change (Proj.cpp:209-308, 538-588):
 * every row and synapse to `target` (= simstep-1) with the table of the old
 * learning rate; Wij recomputed when kp != 0. */
typedef struct {
  int32_t H, M, C, target;
  float kp, wgain, eps, eps2;
  const float* table;
  int32_t table_len;
  gsbn_syn_rec* syn;
  gsbn_row_rec* rowst;
  const float* pj;
  /* reduced-precision form (prec != GSBN_PREC_F32): syn16 + wij replace syn */
  int32_t prec;
  gsbn_syn16_rec* syn16;
  float* wij;
  float p_scale, e_scale; /* GSBN_PREC_F16 only; powers of two */
  /* deferred column updates (dK > 0): the pending column kicks of steps up
   * to `target` are applied first, as the column updates of those steps
   * (slots with ii >= 0) */
  const int32_t* ii;
  const int32_t* dhist;
  const int32_t* dcnt;
  int32_t dK;
  float kftj;
} gsbn_all_args;
int gsbn_proj_all(const gsbn_all_args* a, void* stream);

/* Deferred column updates: apply every pending column kick (steps up to
 * `target`) to every synapse of a slot with ii >= 0, leaving the state the
 * reference's eager column updates would have left (before host reads of
 * the synapse state, snapshots and host writes).  Same arguments as
 * gsbn_proj_all; table = the current learning rate's; kp, wgain, eps, eps2,
 * pj and rowst are not used. */
int gsbn_proj_materialize(const gsbn_all_args* a, void* stream);

/* Postsynaptic Euler update and bias (Proj.cpp:310-363), standalone form. */
typedef struct {
  int32_t N;
  const int8_t* sj;
  float *pj, *ej, *zj, *bj, *epsc;
  float kp, ke, kzj, kepsc, kftj, bgain, eps;
} gsbn_jxx_args;
int gsbn_proj_jxx(const gsbn_jxx_args* a, void* stream);

/* Row and column lazy updates of one step (Proj.cpp:365-536, 667-760).
 * Grid: `parts` workgroups per destination HCU; part q of HCU h takes the
 * active rows q, q+parts, ... of h and the slice q of the C slots of every
 * spiking unit of h.
 * active rows catch up with kfti on zi/zi2, Wij when kp != 0, and write their
 * support contribution; columns of spiking postsynaptic units catch up with
 * kftj on zj2 (slots with ii < 0 skipped).  A synapse in both an active row
 * and an active column gets the row update followed by the column increment,
 * as in the reference's row-then-column order. */
typedef struct {
  int32_t H, M, C, simstep, exec;
  float kp, kfti, kftj, wgain, eps, eps2;
  const float* table;
  int32_t table_len;
  gsbn_syn_rec* syn;
  gsbn_row_rec* rowst;
  const float* pj;
  const int8_t* sj;
  const int32_t* ii;
  const int32_t* rows_list;
  const int32_t* rows_cnt;
  const int32_t* spk_list;
  const int32_t* spk_cnt;
  const int32_t* act_stamp;
  float* contrib;
  int32_t parts; /* workgroups per HCU (0: chosen from H) */
  /* batched form: kp = taupdt * step.prn and the table of step.kp_slot */
  float taupdt;
  const float* tables[GSBN_KP_SLOTS];
  int32_t table_lens[GSBN_KP_SLOTS];
  /* reduced-precision form (prec != GSBN_PREC_F32): syn16 + wij replace syn;
   * Wij is computed from the float value of Pij before it is stored.  All
   * projections of one batched launch must share prec. */
  int32_t prec;
  gsbn_syn16_rec* syn16;
  float* wij;
  float p_scale, e_scale; /* GSBN_PREC_F16 only; powers of two */
  /* deferred column updates (dK > 0, see gsbn_pop_proj): a row update of
   * synapse (r, j) first replays the pending column kicks of unit (h, j)
   * (ring steps s with tij < s < simstep, oldest first, each the column
   * update of step s); the column pass walks the units of spk_list /
   * spk_cnt (then the walk lists), replaying their kicks up to simstep from
   * the ring state walk_T[entry] */
  const int32_t* dhist;
  const int32_t* dcnt;
  const int32_t* walk_T;
  int32_t dK, dpad;
} gsbn_rowcol_args;
int gsbn_proj_rowcol(const gsbn_rowcol_args* a, void* stream);

/* Row/column workgroups of trace form `prec` that the current device holds
 * at once (resident workgroups per CU times CUs): a split whose grid
 * (HCUs x parts) fits starts in one dispatch round.  parts == 0 in
 * gsbn_proj_rowcol picks the largest such split (at most 32). */
int gsbn_rowcol_resident(int32_t prec, int32_t* out);

/* Batched forms: many populations / projections (e.g. independent networks)
 * in one launch.  The argument structs live in device memory (their per-step
 * fields are taken from `step`); wg_pop/wg_hcu (wg_proj/wg_hcu) map each
 * workgroup (HCU) to its population (projection) and local HCU. */
int gsbn_pop_update_batch(const gsbn_pop_args* pops_dev, int32_t npop, const int32_t* wg_pop,
                          const int32_t* wg_hcu, int32_t nwg, int32_t threads, int32_t max_slots,
                          int32_t any_wta_pop_max, const gsbn_step_args* step, void* stream);
int gsbn_proj_rowcol_batch(const gsbn_rowcol_args* projs_dev, const int32_t* wg_proj,
                           const int32_t* wg_hcu, int32_t nhcu, int32_t parts,
                           const gsbn_step_args* step, void* stream);

/* gsbn_pop_update_batch (Pop.cu:8-108 and the queue push, jxx and active-row
 * steps of gsbn_pop_update) with one wave per hypercolumn, units_per_lane
 * (1 or 2) units per lane: same results, for batches whose every population
 * has M <= 64 * units_per_lane, at most one incoming projection with C <=
 * 8192 slots, at most one outgoing projection, no external spikes and no
 * wta_pop_max (the caller checks; gsbn_pop_update_batch covers the rest). */
int gsbn_pop_update_batch_wave(const gsbn_pop_args* pops_dev, const int32_t* wg_pop,
                               const int32_t* wg_hcu, int32_t nwg, int32_t units_per_lane,
                               const gsbn_step_args* step, void* stream);

/* gsbn_proj_rowcol_batch with one 64-thread workgroup (one wave) per
 * hypercolumn of the map: same results, more hypercolumns in flight when
 * each has little work.  With step->ovf_keep > 0 (after a
 * gsbn_pop_update_batch_wave of the same step with the same queue), 4096
 * helper waves ahead of them run the rows of the row overflow queue and each
 * hypercolumn keeps ovf_keep rows (see gsbn_step_args). */
int gsbn_proj_rowcol_batch_wave(const gsbn_rowcol_args* projs_dev, const int32_t* wg_proj,
                                const int32_t* wg_hcu, int32_t nhcu, const gsbn_step_args* step,
                                void* stream);

/* Batched row/column update with a work split per hypercolumn: workgroup w
 * runs part (wg_part[w] & 0xffff) of (wg_part[w] >> 16) parts of hypercolumn
 * wg_hcu[w] of projection wg_proj[w]; every part of a hypercolumn must be in
 * the map once.  Used to split large hypercolumns of a batch more finely
 * than small ones (ProcUpdLazy's split per size bucket).  Replaces, like
 * gsbn_proj_rowcol_batch, the row and column updates of the reference
 * (Proj.cu:488-583) for every projection of a batch. */
int gsbn_proj_rowcol_graded(const gsbn_rowcol_args* projs_dev, const int32_t* wg_proj,
                            const int32_t* wg_hcu, const int32_t* wg_part, int32_t nwg,
                            const gsbn_step_args* step, void* stream);

/* gsbn_proj_rowcol_graded with one 64-thread workgroup (one wave) per map
 * entry (Proj.cu:488-583 for every projection of a batch): the one-wave
 * form with more waves for the hypercolumns with the most slots, whose many
 * active rows in a step would otherwise end the launch alone. */
int gsbn_proj_rowcol_graded_wave(const gsbn_rowcol_args* projs_dev, const int32_t* wg_proj,
                                 const int32_t* wg_hcu, const int32_t* wg_part, int32_t nwg,
                                 const gsbn_step_args* step, void* stream);

/* One launch per executed step without a sweep (ProcUpdLazy.cpp:112-123
 * and 153-165 in one kernel), for populations with one incoming projection
 * and M <= 256: workgroup h runs the population update of hypercolumn h and
 * then the whole row/column update of h (gsbn_proj_rowcol with one part),
 * taking the hypercolumn's spikes, pj and active rows from its own LDS.
 * Results equal the two separate launches.  (Several workgroups per
 * hypercolumn would each need the population's inputs while one of them
 * overwrites them in place, so the fused form has one.)  Single form: one
 * population and its projection.  Batched form: population wg_pop[w],
 * hypercolumn wg_hcu[w] for w < nwg, with the rowcol arguments
 * projs_dev[in_proj[p]] of the projection into population p; max_m is the
 * largest M, max_slots the largest slot count. */
int gsbn_step_fused(const gsbn_pop_args* pop, const gsbn_rowcol_args* proj, void* stream);

/* Chained step (one launch per executed step; one population with its one
 * recurrent projection, M <= 256, no wta_pop_max): workgroup h < H first
 * runs the row updates of the previous executed step for hypercolumn h
 * (`prev`, when has_prev: that step's simstep, exec, kp, table and lists),
 * adding each updated weight to the support sum of its unit in ascending
 * row order, then this step's population update (`pop`); H * prev->parts
 * further workgroups run the previous step's column updates.  This step's
 * row and column updates are left to the next chained launch or to
 * gsbn_proj_rowcol with this step's arguments.  `prev`'s spk_list, spk_cnt
 * and act_stamp must be other arrays than `pop`'s spk_list, spk_cnt and
 * proj0.act_stamp (two sets, alternated by step), since the column updates
 * read them while the population update writes its own.  No counterpart in
 * the reference: its ProcUpdLazy (ProcUpdLazy.cpp:88-160) runs these
 * updates as separate passes within one step. */
int gsbn_step_chained(const gsbn_pop_args* pop, const gsbn_rowcol_args* prev, int32_t has_prev,
                      void* stream);
/* chained-step workgroups the current device holds at once */
int gsbn_chained_resident(int32_t prec, int32_t* out);
int gsbn_step_fused_batch(const gsbn_pop_args* pops_dev, int32_t npop,
                          const gsbn_rowcol_args* projs_dev, const int32_t* wg_pop,
                          const int32_t* wg_hcu, int32_t nwg, const int32_t* in_proj,
                          int32_t max_m, int32_t max_slots, int32_t any_wta_pop_max,
                          const gsbn_step_args* step, void* stream);

/* Delay-queue push (Proj.cpp:645-664) from si: si holds the spikes of exec
 * step exec - 1 (ProcExchangeSpike's copy), each source unit with si > 0
 * pushes its rows.  Grid: one workgroup per source HCU. */
typedef struct {
  int32_t Hsrc, Msrc, exec, pad;
  const int8_t* si;
  gsbn_fanout out;
} gsbn_que_args;
int gsbn_proj_que(const gsbn_que_args* a, void* stream);

/* ProcCheck counting (ProcCheck/Pop.cpp:20-28) and per-HCU first strict
 * argmax with counter reset (ProcCheck/Pop.cpp:35-50). */
int gsbn_check_count(int32_t n, const int8_t* spike, int32_t* count, void* stream);
int gsbn_check_argmax(int32_t H, int32_t M, int32_t* count, int32_t* out_idx, int32_t* out_cnt,
                      void* stream);
/* the same over several populations in one launch; item i writes its H
 * results at out_idx/out_cnt + out_offset */
typedef struct {
  const int8_t* spike;
  int32_t* count;
  int32_t H, M, out_offset, pad;
} gsbn_check_item;
int gsbn_check_count_multi(const gsbn_check_item* items_dev, int32_t nitems, void* stream);
int gsbn_check_argmax_multi(const gsbn_check_item* items_dev, int32_t nitems, int32_t* out_idx,
                            int32_t* out_cnt, void* stream);

/* Structural plasticity, device half (ProcStructuralPlasitcity.cpp:280-300,
 * ProcStructuralPlasticity/Prj.cpp:74-138).
 *
 * gsbn_proj_prune_flags: Prj::prune's test for every row r < R:
 * flags[r] = 1 when ii[r] >= 0 and either ti[r] < t_min or no weight of the
 * row is >= wp or <= wn; else 0.  wij from the FP32 records (prec
 * GSBN_PREC_F32) or from the float weight array of the 16-bit forms. */
typedef struct {
  int64_t R;
  int32_t M, prec;
  const int32_t* ii;
  const gsbn_row_rec* rowst;
  const gsbn_syn_rec* syn;
  const float* wij;
  int32_t t_min;
  float wp, wn;
  int32_t pad;
  uint8_t* flags;
} gsbn_prune_args;
int gsbn_proj_prune_flags(const gsbn_prune_args* a, void* stream);

/* gsbn_proj_clear_rows: Prj::remove_conn's state reset for rows[0..n)
 * (device pointer; entries outside [0, H*C) are skipped): pi, ei, ti = 0 and
 * pij, eij, zj2 of the row's M synapses = 0 (zi, zi2, wij, tij kept), and
 * the row's bit cleared in every slot of the pending-activation ring.  ii/di
 * are the caller's (host vectors). */
typedef struct {
  const int32_t* rows;
  int64_t n;
  int32_t M, prec, H, C, W, pad;
  gsbn_row_rec* rowst;
  gsbn_syn_rec* syn;
  gsbn_syn16_rec* syn16;
  uint32_t* fut;
} gsbn_clear_rows_args;
int gsbn_proj_clear_rows(const gsbn_clear_rows_args* a, void* stream);

/* field gather/scatter between the record arrays and plain vectors
 * field: 0 pij 1 eij 2 zi2 3 zj2 4 wij 5 tij (syn); 0 pi 1 ei 2 zi 3 ti (row) */
int gsbn_syn_field(gsbn_syn_rec* syn, int64_t n, int field, void* buf, int to_records,
                   void* stream);
int gsbn_row_field(gsbn_row_rec* row, int64_t n, int field, void* buf, int to_records,
                   void* stream);

/* the same for the reduced-precision records (field 4 = the wij array);
 * values are converted to / from float with the form's rounding and scale */
int gsbn_syn16_field(gsbn_syn16_rec* syn, float* wij, int64_t n, int field, void* buf,
                     int to_records, int prec, float p_scale, float e_scale, void* stream);

/* fill record / int arrays with one value (state initialisation) */
int gsbn_syn_fill(gsbn_syn_rec* syn, int64_t n, gsbn_syn_rec v, void* stream);
int gsbn_row_fill(gsbn_row_rec* row, int64_t n, gsbn_row_rec v, void* stream);
int gsbn_i32_fill(int32_t* p, int64_t n, int32_t v, void* stream);
int gsbn_syn16_fill(gsbn_syn16_rec* syn, int64_t n, gsbn_syn16_rec v, void* stream);

/* device-to-device copy of `bytes` (a multiple of 16, 16-byte aligned
 * pointers) by a plain streaming kernel; no counterpart in the reference:
 * the calibration point of the memory-traffic counters in measurements */
int gsbn_device_copy(void* dst, const void* src, int64_t bytes, void* stream);

/* host copies of the 16-bit trace conversions (same rounding as the kernels) */
uint16_t gsbn_host_pack_trace(int prec, float x, float scale);
float gsbn_host_unpack_trace(int prec, uint16_t h, float inv_scale);

/* Coefficient table entry (host function): GSBN_COEF_STRIDE = 12 floats
 * Ezi Ezj Ee Ep Cez_ij Cpe Cpz_ij Cez_i Cpz_i 0 0 0 for an elapsed gap. */
void gsbn_coeff_host_entry(double kzi, double kzj, double ke, double kp, int64_t gap,
                           float* out12);
int64_t gsbn_coeff_host_saturation(double kzi, double kzj, double ke, double kp,
                                   int64_t max_len);

/* deterministic host copies of the device math, for tests */
float gsbn_host_expf(float x);
float gsbn_host_logf(float x);
void gsbn_host_draw(uint64_t seed, uint32_t pop, uint32_t step, uint32_t unit, float* u,
                    float* z);
/* evaluate the device math on the GPU for n inputs (tests): op 0 expf,
 * 1 logf, 2 Philox draw, 3 float division, 4 Wij, 5 float -> bfloat16 code,
 * 6 float -> binary16 code, 7 binary16 code -> float */
int gsbn_device_math(int op, int64_t n, const void* in_host, void* out_host);

/* ---- frozen-weight batched recall ----
 * No single counterpart in the reference: B cues of one trained population
 * run the reference's recall-mode dynamics (prn = 0: bj and Wij frozen,
 * Proj.cpp:310-363 and 365-479; Pop.cpp:229-323) in lock-step, each from the
 * population's initial state (Pop.cpp:59-60) with an empty delay queue and
 * epsc = 0, cue b drawing with Philox key seed + b, population 0.
 * GSBN_RECALL_EXACT adds float Wij of the active rows in the reference's
 * ascending row order (a single-cue solver run with the same weights gives
 * the same spikes); GSBN_RECALL_MFMA computes the support of all cues as
 * the product of the 0/1 spike matrix and W in bfloat16 on the matrix cores
 * with float accumulation (one delay for all rows required). */
#define GSBN_RECALL_EXACT 0
#define GSBN_RECALL_MFMA 1
typedef struct {
  int32_t H, M, C;     /* population of H HCUs x M MCUs; C slots per HCU of its recurrent projection */
  const int32_t* ii;   /* [H*C] source unit of each slot, -1 when empty (host memory) */
  const int32_t* di;   /* [H*C] delays (host memory) */
  const float* wij;    /* [H*C*M] frozen weights (host memory) */
  const float* bj;     /* [H*M] frozen bias of the projection (host memory) */
  float taumdt, wtagain, maxfqdt, igain, wgain, lgbias, snoise; /* population parameters */
  float kepsc;         /* support decay per step, dt / tauepsc */
  float lg_on, lg_off; /* .lginp of a clamped MCU and of the others (log(1+eps), log(eps)) */
  uint64_t seed;
  int32_t batch;       /* cues per run, <= 65535 */
  int32_t mode;        /* GSBN_RECALL_EXACT or GSBN_RECALL_MFMA */
} gsbn_recall_desc;
typedef struct gsbn_recall gsbn_recall;
gsbn_recall* gsbn_recall_new(const gsbn_recall_desc* d);
void gsbn_recall_free(gsbn_recall* r);
/* Run `steps` steps for the batch: clamp[b*H + h] = clamped MCU of HCU h in
 * cue b or -1, wmask[b*H + h] the support mask (host memory).  Spikes of steps
 * t > count_from are counted; out_idx/out_cnt[b*H + h] = first strict maximum
 * of the counts and its count (-1 / 0 without spikes, ProcCheck/Pop.cpp:35-50).
 * raster (optional, host, steps*batch*H*M int8) receives every step's spikes;
 * support_ms (optional) the summed time of the support products (mfma). */
int gsbn_recall_run(gsbn_recall* r, const int16_t* clamp, const float* wmask, int32_t steps,
                    int32_t count_from, int32_t* out_idx, int32_t* out_cnt, int8_t* raster,
                    double* support_ms);
/* E[b][n] = beta E[b][n] + sum_k S[b][k] WT[n][k], S and WT bfloat16 with
 * Kp columns (Kp % 64 == 0, zero padded), S with ceil(B/256)*256 rows and WT
 * with ceil(N/256)*256 rows (padding rows zero), E float [B][N]; device
 * pointers. */
int gsbn_spike_gemm(int32_t B, int32_t N, int32_t Kp, const uint16_t* S, const uint16_t* WT,
                    float* E, float beta, void* stream);

/* ---- solver level ---- */
typedef struct gsbn_solver gsbn_solver;

/* Solver(n_path, s_path) (Solver.cpp:10-65): a new network when s_path is
 * NULL or empty, else resumed from the SolverState file s_path
 * (Solver.cpp:44-61, Database.cpp:30-115).  NULL on error, message in
 * gsbn_last_error(). */
gsbn_solver* gsbn_solver_new(const char* net_path, const char* snapshot_path);
void gsbn_solver_free(gsbn_solver* s);
/* run Upd::update up to nsteps times, stopping when cycle-flag < 0;
 * returns the number of updates performed or an error code */
int64_t gsbn_solver_step(gsbn_solver* s, int64_t nsteps);
/* Solver::run: until cycle-flag < 0 */
int64_t gsbn_solver_run(gsbn_solver* s);
int gsbn_solver_sync(gsbn_solver* s);
/* vector access by Database name; dtype codes 0 i8 1 i16 2 i32 3 i64 4 f16 5 f32 6 f64 */
int64_t gsbn_solver_vector_info(gsbn_solver* s, const char* name, int* dtype);
int64_t gsbn_solver_vector_read(gsbn_solver* s, const char* name, void* out, int64_t cap);
int gsbn_solver_vector_write(gsbn_solver* s, const char* name, const void* in, int64_t count);
/* GlobalVar access: type 'i', 'f', 'b' */
int gsbn_solver_global(gsbn_solver* s, const char* key, char type, void* out);
/* ProcCheck results: correct / total windows, per-window flags */
int gsbn_solver_check(gsbn_solver* s, int32_t* correct, int32_t* total, int32_t* flags,
                      int32_t cap);
/* counters: [0] row synapse updates [1] column synapse updates [2] sweep
 * synapse updates [3] executed steps [4] sweeps */
int gsbn_solver_work(gsbn_solver* s, int64_t* out5);
/* per-kernel timing with HIP events on the solver stream.  enable > 0 is a
 * bit mask of the kinds to time (0 pop, 1 rowcol, 2 all, 3 jxx, 4 que, 5
 * check, 6 fused step, 7 unused), 0 stops timing, < 0 also clears the
 * totals; when launches8/ms8 are given the totals {launches, total_ms} per
 * kind are written first. */
int gsbn_solver_timing(gsbn_solver* s, int enable, int64_t* launches8, double* ms8);

/* write the solver's state now, as ProcSnapshot does at its steps
 * (ProcSnapshot.cpp:69-78): timestamp simstep*dt, the current prn */
int gsbn_solver_snapshot(gsbn_solver* s, const char* path);

/* ---- SolverState files (gsbn.proto:100-152), no device needed ----
 * One saved vector: `count` elements of type `dtype` (codes as above; 4 f16
 * is held as float) at `data`, with the leading dimension `ld`. */
typedef struct gsbn_vec_state {
  const char* name;
  int32_t dtype;
  uint32_t ld;
  int64_t count;
  const void* data;
} gsbn_vec_state;
/* SolverState::SerializeToOstream of {timestamp, prn, vectors}
 * (ProcSnapshot.cpp:69-78, Database.cpp:257-295): vectors of one type in
 * name order, types in field order, whatever the order of `vecs` */
int gsbn_state_write(const char* path, float timestamp, float prn, int32_t nvec,
                     const gsbn_vec_state* vecs);
/* SolverState::ParseFromIstream (Solver.cpp:45-51); NULL on error */
typedef struct gsbn_state gsbn_state;
gsbn_state* gsbn_state_read(const char* path);
int32_t gsbn_state_nvec(const gsbn_state* st);
int gsbn_state_header(const gsbn_state* st, float* timestamp, float* prn);
/* vector i in file order; name and data point into `st` */
int gsbn_state_vec(const gsbn_state* st, int32_t i, gsbn_vec_state* out);
void gsbn_state_free(gsbn_state* st);

/* parse a SolverParam prototxt (and its stimulus file) without a device:
 * out4 = {populations, projections, mode_params, proc_params} */
int gsbn_config_check(const char* net_path, int32_t* out4);
/* decode a StimRawData file: dims4 = {data_rows, data_cols, mask_rows, mask_cols} */
int gsbn_stim_read(const char* path, int32_t* dims4, uint32_t* data, int64_t data_cap, float* mask,
                   int64_t mask_cap);

const char* gsbn_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GSBN_AMD_H */
