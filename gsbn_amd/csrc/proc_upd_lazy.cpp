/*
 * proc_upd_lazy.cpp -- the ProcUpdLazy procedure (ProcUpdLazy.cpp:8-180) and
 * its Pop (Pop.cpp:8-409) and Proj (Proj.cpp:6-760) state owners, driving
 * the kernels of kernels.hip through the kernel-level C ABI.
 *
 * Database contract: the same vector names, element types and row-major
 * [hcu][slot][mcu] semantics as the reference (Pop.cpp:47-57,
 * Proj.cpp:64-86).  The per-synapse and per-row traces live in record arrays
 * on the device; their named vectors (pij_<id>, ..., pi_<id>, ...) are views
 * that gather a field on host reads and scatter it back before the next step
 * after host writes.  The delay queue is a ring of per-step bitmaps of
 * pending row activations (a spike pushes its rows into the slot of the step
 * they become active); qi_<id>, .siq_<id> and .ssi_<id> are materialised from
 * it on host reads.
 *
 * Differences from the reference, by design: update_cpu() raises (this
 * build's step runs on the GPU only; there is no silent CPU fallback),
 * random draws are Philox streams keyed by (seed, population, step, unit),
 * and the closed-form catch-up uses the FP64-built coefficient table.
 *
 * Resume (init_copy, Proj.cpp:109-207, Pop.cpp:113-218): the saved vectors
 * are taken over instead of re-created (the reference re-creates act_/ada_,
 * which a snapshot holds, and its CHECK fails), _eps2 is set (the reference
 * leaves it at zero on resume), and the delay queue is rebuilt from qi_ and
 * siq_ (the next step's active rows, saved under that name because the
 * reference's .siq_ is not saved and its resume drops one step of row
 * activations).  A resumed run continues bit for bit as the uninterrupted
 * one: the draws depend only on the step.
 */
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "../../include/gsbn_amd.h"
#include "coeff.h"
#include "detmath.h"
#include "framework.h"
#include "gsbn_core.h"

namespace gsbn {
namespace proc_upd_lazy {

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T* h, size_t count) {
    if (count > n) alloc(count);
    if (count)
      hip_check(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, active_stream()),
                "H2D");
    hip_check(hipStreamSynchronize(active_stream()), "sync");
  }
  void download(T* h, size_t count) const {
    hip_check(hipStreamSynchronize(active_stream()), "sync");
    if (count) hip_check(hipMemcpy(h, p, count * sizeof(T), hipMemcpyDeviceToHost), "D2H");
  }
};

class Proj;

class Pop {
 public:
  void init_new(ProcParam proc_param, PopParam pop_param, Database& db, vector<Pop*>* list_pop,
                int* hcu_cnt, int* mcu_cnt, bool copy = false);
  void flush();
  gsbn_pop_args args(int simstep, int exec, float prn, int lginp_idx, int wmask_idx, bool do_jxx,
                     uint64_t seed, int wta_pop_max);

  int _id = 0, _rank = 0, _dim_proj = 0, _dim_hcu = 0, _dim_mcu = 0, _hcu_start = 0, _mcu_start = 0;
  float _taumdt = 0, _wtagain = 0, _maxfqdt = 0, _igain = 0, _wgain = 0, _lgbias = 0, _snoise = 0;
  SyncVector<float>*_epsc = nullptr, *_bj = nullptr, *_dsup = nullptr, *_act = nullptr, *_ada = nullptr;
  SyncVector<float>*_rnd_uniform01 = nullptr, *_rnd_normal = nullptr, *_wmask = nullptr, *_lginp = nullptr;
  SyncVector<int8_t>* _spike = nullptr;
  SyncVector<int32_t>* _counter = nullptr;
  vector<Proj*> _in; /* incoming projections in proj_in_pop order */
  DevBuf<int32_t> spk_list, spk_cnt, ext_off, ext_units;
  /* second spike-list set of the chained step (steps alternate between the
   * two), and the set the last executed step wrote */
  DevBuf<int32_t> spk_list1, spk_cnt1;
  int spk_side = 0;
  int32_t* spk_list_at(int side) { return side ? spk_list1.p : spk_list.p; }
  int32_t* spk_cnt_at(int side) { return side ? spk_cnt1.p : spk_cnt.p; }
  vector<Proj*> _out;       /* projections with this population as source */
  bool push_direct = false; /* this population's kernel does their queue push */
  DevBuf<gsbn_fanout> outs_dev;
  vector<gsbn_fanout> outs_host;
  DevBuf<float> scratch;
  DevBuf<gsbn_pop_proj> projs_dev;
  DevBuf<int64_t> work;
  vector<gsbn_pop_proj> projs_host;
  int ext_nsteps = 0;
  bool has_ext = false;
  bool flush_pending = false;
};

struct CoefTable {
  DevBuf<float> dev;
  int64_t len = 0;
  bool saturated = false;
};

/* Coefficient tables shared by all projections with the same rates.  A
 * table covers every gap up to the schedule's horizon or up to its
 * saturation gap; `version` changes whenever a table is (re)built so cached
 * launch arguments holding table pointers can be refreshed. */
struct TableCache {
  std::map<std::array<uint32_t, 4>, std::unique_ptr<CoefTable>> m;
  uint64_t version = 0;
  int sim_end = 0;
  const CoefTable& get(float kzi, float kzj, float ke, float kp, int64_t need) {
    std::array<uint32_t, 4> key;
    memcpy(&key[0], &kzi, 4);
    memcpy(&key[1], &kzj, 4);
    memcpy(&key[2], &ke, 4);
    memcpy(&key[3], &kp, 4);
    auto& slot = m[key];
    if (!slot) slot.reset(new CoefTable());
    CoefTable& t = *slot;
    if (t.saturated || t.len >= need) return t;
    gsbn_rates_t r = {(double)kzi, (double)kzj, (double)ke, (double)kp};
    const int64_t cap = (int64_t)1 << 26;
    int64_t want = std::max<int64_t>(need, std::max<int64_t>(2 * t.len, (int64_t)sim_end + 2));
    const int64_t sat = gsbn_coeff_saturation(&r, cap);
    bool saturated = false;
    if (sat >= 0 && sat + 1 <= want) {
      want = sat + 1;
      saturated = true;
    }
    if (want > cap) throw Error(-5, "coefficient table beyond 2^26 gaps");
    vector<float> host((size_t)want * GSBN_COEF_STRIDE);
    for (int64_t d = 0; d < want; ++d) gsbn_coeff_entry(&r, d, host.data() + d * GSBN_COEF_STRIDE);
    t.dev.upload(host.data(), host.size());
    t.len = want;
    t.saturated = saturated;
    ++version;
    return t;
  }
};

class Proj : public ProjRowOps {
 public:
  ~Proj() override;
  void prune_flags(int t_min, float wp, float wn, vector<uint8_t>* flags) override;
  void clear_rows(const vector<int32_t>& rows) override;
  void rows_added() override;
  void init_new(ProcParam proc_param, ProjParam proj_param, Database& db, vector<Proj*>* list_proj,
                vector<Pop*>* list_pop, bool copy = false);
  void prepare(int simstep);
  void restore_queue(int exec);
  void update_all(int simstep, float old_prn);
  void update_jxx(float prn);
  void update_rowcol(int simstep, int exec, float prn);
  void update_que(int exec);
  void mark_views_device();
  gsbn_pop_proj pop_desc();
  const CoefTable& table(float kp, int64_t need) {
    return tcache->get(_tauzidt, _tauzjdt, _tauedt, kp, need);
  }
  gsbn_rowcol_args static_args(const float* slot_prn, int nslots, int64_t need);
  void rebuild_fanout();
  TableCache* tcache = nullptr;

  int _id = 0, _rank = 0, _proj_in_pop = 0, _dim_conn = 0, _dim_hcu = 0, _dim_mcu = 0, _nsrc = 0;
  float _taupdt = 0, _tauedt = 0, _tauzidt = 0, _tauzjdt = 0, _tauepscdt = 0, _eps = 0, _eps2 = 0;
  float _kfti = 0, _kftj = 0, _wgain = 0, _bgain = 0;
  Pop* _ptr_src_pop = nullptr;
  Pop* _ptr_dest_pop = nullptr;
  SyncVector<int32_t>*_ii = nullptr, *_qi = nullptr, *_di = nullptr, *_ssi = nullptr, *_ssj = nullptr,
                     *_ti = nullptr, *_tij = nullptr;
  SyncVector<int8_t>*_siq = nullptr, *_si = nullptr, *_siq_saved = nullptr;
  SyncVector<float>*_pi = nullptr, *_ei = nullptr, *_zi = nullptr, *_pj = nullptr, *_ej = nullptr,
                   *_zj = nullptr, *_pij = nullptr, *_eij = nullptr, *_zi2 = nullptr, *_zj2 = nullptr,
                   *_wij = nullptr;
  vector<VecBase*> views;
  /* device state */
  int prec = GSBN_PREC_F32;     /* trace form of pij/eij */
  int p_shift = 0, e_shift = 0; /* FP16 form: stored value is binary16(x * 2^shift) */
  int64_t nsyn = 0;
  DevBuf<gsbn_syn_rec> syn;     /* FP32 form */
  DevBuf<gsbn_syn16_rec> syn16; /* 16-bit forms */
  DevBuf<float> wij16;          /* 16-bit forms: the weights */
  float p_scale() const { return std::ldexp(1.0f, p_shift); }
  float e_scale() const { return std::ldexp(1.0f, e_shift); }
  void syn_field(int field, void* buf, int to_records) {
    if (prec == GSBN_PREC_F32)
      op_check(gsbn_syn_field(syn.p, nsyn, field, buf, to_records, active_stream()), "gsbn_syn_field");
    else
      op_check(gsbn_syn16_field(syn16.p, wij16.p, nsyn, field, buf, to_records, prec, p_scale(),
                                e_scale(), active_stream()),
               "gsbn_syn16_field");
  }
  DevBuf<gsbn_row_rec> rowst;
  DevBuf<float> contrib;
  DevBuf<int32_t> rows_list, rows_cnt, pend_cnt, act_stamp;
  DevBuf<int32_t> act_stamp1; /* second stamp set of the chained step */
  int32_t* act_stamp_at(int side) { return side ? act_stamp1.p : act_stamp.p; }
  /* completes row/column work left pending by a chained step; run before
   * any read or change of the synapse and row state */
  std::function<void()> flush_hook;
  void flush_deferred() {
    if (flush_hook) flush_hook();
  }
  /* Deferred column updates (dK > 0, see gsbn_pop_proj): each
   * postsynaptic unit's recent spike steps in a ring, the column update of
   * a spike replayed by the row update that next touches a synapse, by a
   * column walk once dK - 1 spikes of a unit are pending, by a sweep, or by
   * materialize_kicks() before the synapse state is read or changed from the
   * host.  Two walk-list sets alternate by step in the chained plan. */
  int dK = 0;
  DevBuf<int32_t> dhist, dcnt, dfl, walk_list[2], walk_cnt[2], walk_T[2];
  /* the device copy of ii_, as the kernels saw it (not synchronised from a
   * newer host copy: pending kicks belong to the connectivity they arose in) */
  const int32_t* ii_dev = nullptr;
  int kick_t = 0;        /* the last executed step (kicks up to it may be pending) */
  float kick_kp = 0.0f;  /* its learning rate */
  bool kicks_pending = false;
  void enable_deferred(int K) {
    const size_t N = (size_t)_dim_hcu * _dim_mcu;
    dK = K;
    ii_dev = _ii->gpu_data();
    dhist.alloc(N * (size_t)K);
    dcnt.alloc(N);
    dfl.alloc(N);
    op_check(gsbn_i32_fill(dcnt.p, (int64_t)N, 0, active_stream()), "fill");
    op_check(gsbn_i32_fill(dfl.p, (int64_t)N, 0, active_stream()), "fill");
    op_check(gsbn_i32_fill(dhist.p, (int64_t)(N * (size_t)K), 0, active_stream()), "fill");
    for (int i = 0; i < 2; ++i) {
      walk_list[i].alloc(N);
      walk_T[i].alloc(N);
      walk_cnt[i].alloc((size_t)_dim_hcu);
      op_check(gsbn_i32_fill(walk_cnt[i].p, _dim_hcu, 0, active_stream()), "fill");
    }
  }
  void materialize_kicks();
  /* deferral ends (the kicks pending are applied first): structural changes
   * of the rows are not replayed; set by ProcUpdLazy */
  std::function<void()> eager_hook;
  DevBuf<uint8_t> sp_flags; /* structural plasticity scratch */
  DevBuf<int32_t> sp_rows;
  /* delay queue: rows of each source unit (CSR) with their delays, and the
   * ring of pending activations, GSBN_FUT slots x H x W words */
  DevBuf<int32_t> fwd_off, fwd_row;
  DevBuf<uint8_t> fwd_del;
  DevBuf<uint32_t> fut;
  int W = 0;
  gsbn_fanout fanout() const {
    gsbn_fanout f;
    memset(&f, 0, sizeof(f));
    f.fwd_off = fwd_off.p;
    f.fwd_row = fwd_row.p;
    f.fwd_del = fwd_del.p;
    f.fut = fut.p;
    f.H = _dim_hcu;
    f.C = _dim_conn;
    f.W = W;
    return f;
  }
  uint64_t ii_version = ~0ull, di_version = ~0ull;
  int last_exec = -1;
  bool direct_que = false; /* the source population's kernel pushes this step's spikes */
  /* resumed from a snapshot: qi_ and siq_ as saved, until the first step
   * rebuilds the delay queue from them */
  bool resume_pending = false;
  vector<int32_t> resume_qi;
  vector<int8_t> resume_siq;
  int sim_end = 0;
  std::vector<std::unique_ptr<VecView>> view_objs;
};

/* ---- views ---- */

/* the host buffer of a vector whose to_host() is in progress */
static void* raw_data(VecBase* v) {
  switch (v->type) {
    case DT_F32: return static_cast<SyncVector<float>*>(v)->raw_host().data();
    case DT_I32: return static_cast<SyncVector<int32_t>*>(v)->raw_host().data();
    case DT_I8: return static_cast<SyncVector<int8_t>*>(v)->raw_host().data();
    default: throw Error(-4, "unexpected view type");
  }
}

struct SynField : VecView {
  Proj* p;
  int field;
  SynField(Proj* q, int f) : p(q), field(f) {}
  void pull(VecBase* v) override {
    p->flush_deferred();
    p->materialize_kicks();
    const size_t n = (size_t)p->nsyn;
    DevBuf<uint32_t> tmp;
    tmp.alloc(n);
    p->syn_field(field, tmp.p, 0);
    tmp.download((uint32_t*)raw_data(v), n);
  }
  void push(VecBase* v) override {
    const size_t n = (size_t)p->nsyn;
    DevBuf<uint32_t> tmp;
    tmp.upload((const uint32_t*)raw_data(v), n);
    p->syn_field(field, tmp.p, 1);
    hip_check(hipStreamSynchronize(active_stream()), "sync");
  }
};

struct RowField : VecView {
  Proj* p;
  int field;
  RowField(Proj* q, int f) : p(q), field(f) {}
  void pull(VecBase* v) override {
    p->flush_deferred();
    const size_t n = p->rowst.n;
    DevBuf<uint32_t> tmp;
    tmp.alloc(n);
    op_check(gsbn_row_field(p->rowst.p, (int64_t)n, field, tmp.p, 0, active_stream()), "gsbn_row_field");
    tmp.download((uint32_t*)raw_data(v), n);
  }
  void push(VecBase* v) override {
    const size_t n = p->rowst.n;
    DevBuf<uint32_t> tmp;
    tmp.upload((const uint32_t*)raw_data(v), n);
    op_check(gsbn_row_field(p->rowst.p, (int64_t)n, field, tmp.p, 1, active_stream()), "gsbn_row_field");
    hip_check(hipStreamSynchronize(active_stream()), "sync");
  }
};

/* views materialised from the queue ring / step lists; read-only */
struct QueueView : VecView {
  Proj* p;
  int what; /* 0 qi, 1 siq, 2 ssi, 3 ssj */
  QueueView(Proj* q, int w) : p(q), what(w) {}
  void push(VecBase* v) override {
    throw Error(-5, "vector " + v->name + " is derived from the device spike queue and is read-only");
  }
  void pull(VecBase* v) override;
};

/* bring the host copy of a view up to date without re-entering to_host */
template <typename T>
static vector<T>& raw(VecBase* v) {
  return static_cast<SyncVector<T>*>(v)->raw_host();
}

void QueueView::pull(VecBase* v) {
  const int H = p->_dim_hcu, C = p->_dim_conn;
  const int64_t R = (int64_t)H * C;
  if (what == 2) {
    vector<int32_t> cnt(H), lst;
    p->rows_cnt.download(cnt.data(), H);
    lst.resize((size_t)R);
    p->rows_list.download(lst.data(), (size_t)R);
    auto& out = raw<int32_t>(v);
    out.clear();
    for (int h = 0; h < H; ++h)
      for (int a = 0; a < cnt[h]; ++a) out.push_back(h * C + lst[(size_t)h * C + a]);
    static_cast<SyncVector<int32_t>*>(v)->set_logical_size(out.size());
    return;
  }
  if (what == 3) {
    Pop* d = p->_ptr_dest_pop;
    const int M = d->_dim_mcu;
    vector<int32_t> cnt(d->_dim_hcu), lst((size_t)d->_dim_hcu * M);
    DevBuf<int32_t>& dc = d->spk_side ? d->spk_cnt1 : d->spk_cnt;
    DevBuf<int32_t>& dl = d->spk_side ? d->spk_list1 : d->spk_list;
    dc.download(cnt.data(), cnt.size());
    dl.download(lst.data(), lst.size());
    auto& out = raw<int32_t>(v);
    out.clear();
    for (int h = 0; h < d->_dim_hcu; ++h)
      for (int a = 0; a < cnt[h]; ++a) out.push_back(h * M + lst[(size_t)h * M + a]);
    static_cast<SyncVector<int32_t>*>(v)->set_logical_size(out.size());
    return;
  }
  /* qi / siq from the ring of pending activations.  After the queue update
   * of exec step x, bit b of the reference's qi[r] means that row r becomes
   * active at exec x + 2 + b (b < di[r]; a bit at b = di[r] would be a spike
   * of step x itself, which the reference pushes only at step x + 1), and
   * siq[r] that it is active at x + 1. */
  const int x = p->last_exec;
  const int W = p->W;
  vector<uint32_t> ring((size_t)GSBN_FUT * H * W);
  hip_check(hipStreamSynchronize(active_stream()), "sync");
  if (!ring.empty())
    hip_check(hipMemcpy(ring.data(), p->fut.p, ring.size() * 4, hipMemcpyDeviceToHost), "D2H");
  auto pending = [&](int step, int64_t r) -> bool {
    const int h = (int)(r / C), k = (int)(r - (int64_t)h * C);
    const uint32_t w = ring[((size_t)(step % GSBN_FUT) * H + h) * W + (k >> 5)];
    return ((w >> (k & 31)) & 1u) != 0;
  };
  const int* ii = p->_ii->cpu_data();
  const int* di = p->_di->cpu_data();
  if (what == 0) {
    auto& out = raw<int32_t>(v);
    out.assign((size_t)R, 0);
    if (x >= 0)
      for (int64_t r = 0; r < R; ++r) {
        if (ii[r] < 0) continue;
        int q = 0;
        for (int b = 0; b < di[r]; ++b)
          if (pending(x + 2 + b, r)) q |= 1 << b;
        out[(size_t)r] = q;
      }
  } else {
    auto& out = raw<int8_t>(v);
    out.assign((size_t)R, 0);
    if (x >= 0)
      for (int64_t r = 0; r < R; ++r)
        if (ii[r] >= 0 && pending(x + 1, r)) out[(size_t)r] = 1;
  }
}

/* ---- Pop ---- */

/* the vector `name`: on resume the one restored from the snapshot when the
 * snapshot holds it, else a new one (NULL if the name is taken) */
template <typename T>
static SyncVector<T>* attach(Database& db, const string& name, bool copy) {
  if (copy && db.restored(name)) return db.get<T>(name);
  return db.create<T>(name);
}

/* on resume a restored vector must have the size this network gives it */
static bool restored_as(Database& db, VecBase* v, size_t n, bool copy) {
  if (!copy || !db.restored(v->name)) return false;
  if (v->size() != n)
    throw Error(-2, "snapshot vector " + v->name + " has " + std::to_string(v->size()) +
                        " elements, this network needs " + std::to_string(n));
  return true;
}

void Pop::init_new(ProcParam proc_param, PopParam pop_param, Database& db, vector<Pop*>* list_pop,
                   int* hcu_cnt, int* mcu_cnt, bool copy) {
  GlobalVar glv;
  _id = (int)list_pop->size();
  list_pop->push_back(this);
  _dim_proj = 0;
  _dim_hcu = pop_param.hcu_num;
  _dim_mcu = pop_param.mcu_num;
  if (_dim_mcu <= 0 || _dim_hcu <= 0) throw Error(-2, "pop_param: hcu_num and mcu_num must be > 0");
  if (_dim_mcu > 1024) throw Error(-5, "mcu_num above 1024 is not supported");
  _hcu_start = *hcu_cnt;
  _mcu_start = *mcu_cnt;
  *hcu_cnt += _dim_hcu;
  *mcu_cnt += _dim_hcu * _dim_mcu;
  float dt = 0;
  int rank = 0;
  GSBN_CHECK(glv.getf("dt", dt), "dt (ProcExtGen must come before ProcUpdLazy)");
  GSBN_CHECK(glv.geti("rank", rank), "rank");
  _rank = pop_param.rank;
  _taumdt = dt / pop_param.taum;
  _wtagain = pop_param.wtagain;
  _maxfqdt = pop_param.maxfq * dt;
  _igain = pop_param.igain;
  _wgain = pop_param.wgain;
  _lgbias = pop_param.lgbias;
  _snoise = pop_param.snoise;
  if (_rank != rank) throw Error(-5, "populations on other ranks are not supported (one process)");
  const string sid = std::to_string(_id);
  _dsup = attach<float>(db, "dsup_" + sid, copy);
  _act = attach<float>(db, "act_" + sid, copy);
  _ada = attach<float>(db, "ada_" + sid, copy);
  _epsc = attach<float>(db, "epsc_" + sid, copy);
  _bj = attach<float>(db, "bj_" + sid, copy);
  _spike = attach<int8_t>(db, "spike_" + sid, copy);
  _rnd_uniform01 = db.create_sync_vector_f32(".rnd_uniform01_" + sid);
  _rnd_normal = db.create_sync_vector_f32(".rnd_normal_" + sid);
  _wmask = db.sync_vector_f32(".wmask");
  _lginp = db.sync_vector_f32(".lginp");
  _counter = db.create_sync_vector_i32(".counter_" + sid);
  GSBN_CHECK(_dsup && _act && _ada && _epsc && _bj && _spike && _rnd_uniform01 && _rnd_normal &&
                 _wmask && _lginp && _counter,
             "Pop vectors");
  const size_t N = (size_t)_dim_hcu * _dim_mcu;
  if (!restored_as(db, _dsup, N, copy)) _dsup->resize(N, (float)log(1.0 / _dim_mcu));
  if (!restored_as(db, _act, N, copy)) _act->resize(N, (float)(1.0 / _dim_mcu));
  if (!restored_as(db, _ada, N, copy)) _ada->resize(N);
  if (!restored_as(db, _spike, N, copy)) _spike->resize(N);
  _rnd_uniform01->resize(N);
  _rnd_normal->resize(N);
  _counter->resize(N);
  spk_list.alloc(N);
  spk_cnt.alloc(_dim_hcu);
  op_check(gsbn_i32_fill(spk_cnt.p, _dim_hcu, 0, active_stream()), "fill");
  scratch.alloc(1);
  /* external spikes for debugging (Pop.cpp:67-110): lines step,pop,idx,... */
  Parser par(proc_param);
  string filename;
  if (par.args("external spike", filename)) {
    std::ifstream f(filename);
    if (!f) throw Error(-2, "Unable to open file to load external spikes");
    std::map<int, vector<int>> ext;
    string line;
    while (std::getline(f, line)) {
      std::stringstream ss(line);
      vector<string> v;
      while (ss.good()) {
        string sub;
        std::getline(ss, sub, ',');
        v.push_back(sub);
      }
      if (v.size() > 2 && std::stoi(v[1]) == _id && std::stoi(v[0]) >= 0) {
        vector<int> spk;
        for (size_t i = 2; i < v.size(); i++) {
          int u = std::stoi(v[i]);
          if (u < 0 || (size_t)u >= N) throw Error(-2, "external spike unit out of range");
          spk.push_back(u);
        }
        ext[std::stoi(v[0])] = spk;
      }
    }
    has_ext = true;
    ext_nsteps = ext.empty() ? 0 : ext.rbegin()->first + 1;
    vector<int32_t> off(ext_nsteps + 1, 0), units;
    for (int s = 0; s < ext_nsteps; ++s) {
      auto it = ext.find(s);
      if (it != ext.end()) units.insert(units.end(), it->second.begin(), it->second.end());
      off[s + 1] = (int32_t)units.size();
    }
    ext_off.upload(off.data(), off.size());
    if (units.empty()) units.push_back(0);
    ext_units.upload(units.data(), units.size());
  }
}

gsbn_pop_args Pop::args(int simstep, int exec, float prn, int lginp_idx, int wmask_idx, bool do_jxx,
                        uint64_t seed, int wta_pop_max) {
  gsbn_pop_args a;
  memset(&a, 0, sizeof(a));
  a.pop_id = _id;
  a.H = _dim_hcu;
  a.M = _dim_mcu;
  a.simstep = simstep;
  a.exec = exec;
  a.seed = seed;
  a.taumdt = _taumdt;
  a.wtagain = _wtagain;
  a.maxfqdt = _maxfqdt;
  a.igain = _igain;
  a.wgain = _wgain;
  a.lgbias = _lgbias;
  a.snoise = _snoise;
  a.prn = prn;
  a.wta_pop_max = wta_pop_max;
  /* .lginp row lginp_idx and .wmask row wmask_idx (Pop.cpp:331-332) */
  const size_t lg_off = _lginp->ld() <= 0 ? (size_t)lginp_idx : (size_t)lginp_idx * _lginp->ld();
  const size_t wm_off = _wmask->ld() <= 0 ? (size_t)wmask_idx : (size_t)wmask_idx * _wmask->ld();
  if (lginp_idx < 0 || lg_off + _mcu_start + (size_t)_dim_hcu * _dim_mcu > _lginp->size())
    throw Error(-2, "lginp-idx out of range");
  if (wmask_idx < 0 || wm_off + _hcu_start + (size_t)_dim_hcu > _wmask->size())
    throw Error(-2, "wmask-idx " + std::to_string(wmask_idx) + " out of range");
  a.lginp = _lginp->gpu_data(lginp_idx) + _mcu_start;
  a.wmask = _wmask->gpu_data(wmask_idx) + _hcu_start;
  a.dsup = _dsup->mutable_gpu_data();
  a.act = _act->mutable_gpu_data();
  a.spike = _spike->mutable_gpu_data();
  a.counter = _counter->mutable_gpu_data();
  a.rnd_u = _rnd_uniform01->mutable_gpu_data();
  a.rnd_n = _rnd_normal->mutable_gpu_data();
  a.epsc = _epsc->mutable_gpu_data();
  a.bj = _bj->mutable_gpu_data();
  a.scratch_dsup0 = scratch.p;
  if (has_ext) {
    a.ext_off = ext_off.p;
    a.ext_units = ext_units.p;
    a.ext_nsteps = ext_nsteps;
  }
  a.spk_list = spk_list.p;
  a.spk_cnt = spk_cnt.p;
  if (push_direct && !_out.empty()) {
    vector<gsbn_fanout> outs;
    for (Proj* q : _out) outs.push_back(q->fanout());
    if (outs.size() != outs_host.size() ||
        memcmp(outs.data(), outs_host.data(), outs.size() * sizeof(gsbn_fanout))) {
      outs_host = outs;
      outs_dev.upload(outs.data(), outs.size());
    }
    a.nproj_out = (int32_t)outs.size();
    a.outs = outs_dev.p;
    a.out0 = outs[0];
  }
  a.nproj_in = (int)_in.size();
  a.do_jxx = do_jxx ? 1 : 0;
  int max_slots = 0;
  vector<gsbn_pop_proj> descs;
  for (Proj* q : _in) {
    descs.push_back(q->pop_desc());
    max_slots = std::max(max_slots, q->_dim_conn);
  }
  a.max_slots = max_slots;
  if (descs.size() != projs_host.size() ||
      (descs.size() && memcmp(descs.data(), projs_host.data(), descs.size() * sizeof(gsbn_pop_proj)))) {
    projs_host = descs;
    projs_dev.upload(descs.data(), descs.size());
  }
  a.projs = projs_dev.p;
  if (!descs.empty()) a.proj0 = descs[0];
  if (!work.p) {
    work.alloc(2 * GSBN_MAX_PROJ_IN);
    vector<int64_t> z(2 * GSBN_MAX_PROJ_IN, 0);
    work.upload(z.data(), z.size());
  }
  a.work = work.p;
  return a;
}

/* apply pending support contributions before a host read of epsc */
void Pop::flush() {
  /* row updates a chained step left pending write the contributions */
  for (Proj* q : _in) q->flush_deferred();
  if (!flush_pending) return;
  flush_pending = false;
  gsbn_pop_args a;
  memset(&a, 0, sizeof(a));
  a.H = _dim_hcu;
  a.M = _dim_mcu;
  a.epsc = _epsc->mutable_gpu_data();
  a.nproj_in = (int)_in.size();
  a.projs = projs_dev.p;
  if (a.nproj_in && a.projs)
    op_check(gsbn_pop_flush(&a, active_stream()), "gsbn_pop_flush");
}

/* ---- Proj ---- */

/* Power-of-two scale of the FP16 form.  A presynaptic trace gains kfti per
 * spike (at most one per step) and decays by exp(-kzi) per step, so
 * zi2 <= kfti / (1 - exp(-kzi)), likewise zj2; Eij and Pij are averages of
 * zi2*zj2 (and of the initial Pij <= 1), so they never exceed the product
 * of the two limits.  The largest 2^k that keeps this bound below the
 * binary16 maximum 65504 maps every reachable value to a finite number and
 * leaves as much range as possible to small values. */
static int fp16_shift(float kzi, float kzj, float kfti, float kftj) {
  const double zi = (double)kfti / -std::expm1(-(double)kzi);
  const double zj = (double)kftj / -std::expm1(-(double)kzj);
  const double bound = std::max(1.0, zi * zj);
  if (!std::isfinite(bound) || kzi <= 0 || kzj <= 0)
    throw Error(-2, "ProcUpdLazy: fp16 trace form needs finite positive z decay rates");
  int k = (int)std::floor(std::log2(65504.0 / bound));
  while (k > -126 && bound * std::ldexp(1.0, k) > 65504.0) --k; /* guard log2 rounding */
  if (k < -100) throw Error(-2, "ProcUpdLazy: fp16 trace form: trace bound out of range");
  return std::min(k, 100);
}

void Proj::init_new(ProcParam proc_param, ProjParam proj_param, Database& db,
                    vector<Proj*>* list_proj, vector<Pop*>* list_pop, bool copy) {
  GlobalVar glv;
  _id = (int)list_proj->size();
  list_proj->push_back(this);
  _ptr_src_pop = (*list_pop)[proj_param.src_pop];
  _ptr_dest_pop = (*list_pop)[proj_param.dest_pop];
  Pop* p = _ptr_dest_pop;
  _dim_hcu = p->_dim_hcu;
  _dim_mcu = p->_dim_mcu;
  _nsrc = _ptr_src_pop->_dim_hcu * _ptr_src_pop->_dim_mcu;
  _dim_conn = _nsrc;
  if (_dim_conn > proj_param.slot_num) _dim_conn = (int)proj_param.slot_num;
  if (_dim_conn <= 0) throw Error(-2, "proj_param: slot_num must be > 0");
  _proj_in_pop = p->_dim_proj;
  p->_dim_proj++;
  if (p->_dim_proj > GSBN_MAX_PROJ_IN)
    throw Error(-5, "more than 8 projections into one population are not supported");
  p->_in.push_back(this);
  _rank = p->_rank;
  float dt = 0;
  GSBN_CHECK(glv.getf("dt", dt), "dt");
  if (!glv.geti("sim-end-step", sim_end)) sim_end = 0;
  _tauzidt = dt / proj_param.tauzi;
  _tauzjdt = dt / proj_param.tauzj;
  _tauedt = dt / proj_param.taue;
  _taupdt = dt / proj_param.taup;
  _eps = dt / proj_param.taup;
  _eps2 = (dt / proj_param.taup) * (dt / proj_param.taup);
  _kfti = 1 / (proj_param.maxfq * proj_param.tauzi);
  _kftj = 1 / (proj_param.maxfq * proj_param.tauzj);
  _bgain = proj_param.bgain;
  _wgain = proj_param.wgain;
  _tauepscdt = proj_param.has_tauepsc() ? dt / proj_param.tauepsc : _tauzidt;

  const size_t N = (size_t)_dim_hcu * _dim_mcu;
  /* the population's epsc_/bj_ hold one slice per incoming projection */
  if (copy && db.restored(p->_epsc->name)) {
    if (p->_epsc->size() < (size_t)p->_dim_proj * N || p->_bj->size() < (size_t)p->_dim_proj * N)
      throw Error(-2, "snapshot vectors epsc_/bj_ are smaller than this network's projections need");
  } else {
    p->_epsc->resize((size_t)p->_dim_proj * N);
    p->_bj->resize((size_t)p->_dim_proj * N);
  }

  const string sid = std::to_string(_id);
  _ii = attach<int32_t>(db, "ii_" + sid, copy);
  _qi = attach<int32_t>(db, "qi_" + sid, copy);
  _di = attach<int32_t>(db, "di_" + sid, copy);
  _ssi = db.create_sync_vector_i32(".ssi_" + sid);
  _siq = db.create_sync_vector_i8(".siq_" + sid);
  _siq_saved = attach<int8_t>(db, "siq_" + sid, copy);
  _pi = attach<float>(db, "pi_" + sid, copy);
  _ei = attach<float>(db, "ei_" + sid, copy);
  _zi = attach<float>(db, "zi_" + sid, copy);
  _ti = attach<int32_t>(db, "ti_" + sid, copy);
  _ssj = db.create_sync_vector_i32(".ssj_" + sid);
  _pj = attach<float>(db, "pj_" + sid, copy);
  _ej = attach<float>(db, "ej_" + sid, copy);
  _zj = attach<float>(db, "zj_" + sid, copy);
  _pij = attach<float>(db, "pij_" + sid, copy);
  _eij = attach<float>(db, "eij_" + sid, copy);
  _zi2 = attach<float>(db, "zi2_" + sid, copy);
  _zj2 = attach<float>(db, "zj2_" + sid, copy);
  _tij = attach<int32_t>(db, "tij_" + sid, copy);
  _wij = attach<float>(db, "wij_" + sid, copy);
  _si = attach<int8_t>(db, "si_" + sid, copy);
  GSBN_CHECK(_ii && _qi && _di && _ssi && _siq && _siq_saved && _pi && _ei && _zi && _ti && _ssj &&
                 _pj && _ej && _zj && _pij && _eij && _zi2 && _zj2 && _tij && _wij && _si,
             "Proj vectors");

  const int64_t R = (int64_t)_dim_hcu * _dim_conn;
  const int64_t S = R * _dim_mcu;
  if (!restored_as(db, _ii, (size_t)R, copy)) _ii->resize(R, -1);
  if (!restored_as(db, _di, (size_t)R, copy)) _di->resize(R);
  if (!restored_as(db, _pj, N, copy)) _pj->resize(N, (float)(1.0 / _dim_mcu));
  if (!restored_as(db, _ej, N, copy)) _ej->resize(N);
  if (!restored_as(db, _zj, N, copy)) _zj->resize(N);
  if (!restored_as(db, _si, (size_t)_nsrc, copy)) _si->resize(_nsrc);

  /* device state, initial values of Proj.cpp:88-105 */
  Parser par(proc_param);
  string form;
  if (par.args("trace-precision", form)) {
    if (form == "fp32")
      prec = GSBN_PREC_F32;
    else if (form == "bf16")
      prec = GSBN_PREC_BF16;
    else if (form == "fp16")
      prec = GSBN_PREC_F16;
    else
      throw Error(-2, "ProcUpdLazy: trace-precision must be fp32, bf16 or fp16, got '" + form + "'");
  }
  nsyn = S;
  rowst.alloc((size_t)R);
  const float pij0 = (float)(1.0 / _dim_conn / _dim_mcu);
  if (prec == GSBN_PREC_F32) {
    syn.alloc((size_t)S);
    gsbn_syn_rec s0;
    memset(&s0, 0, sizeof(s0));
    s0.pij = pij0;
    op_check(gsbn_syn_fill(syn.p, S, s0, active_stream()), "gsbn_syn_fill");
  } else {
    if (prec == GSBN_PREC_F16) {
      p_shift = fp16_shift(_tauzidt, _tauzjdt, _kfti, _kftj);
      e_shift = p_shift;
    }
    syn16.alloc((size_t)S);
    wij16.alloc((size_t)S);
    gsbn_syn16_rec s0;
    memset(&s0, 0, sizeof(s0));
    s0.pij = gsbn_pack_trace(prec, pij0, p_scale());
    op_check(gsbn_syn16_fill(syn16.p, S, s0, active_stream()), "gsbn_syn16_fill");
    op_check(gsbn_i32_fill(reinterpret_cast<int32_t*>(wij16.p), S, 0, active_stream()), "fill");
  }
  {
    GlobalVar g;
    g.puti("trace-precision-" + sid, prec);
    g.puti("trace-p-shift-" + sid, p_shift);
    g.puti("trace-e-shift-" + sid, e_shift);
  }
  gsbn_row_rec r0;
  memset(&r0, 0, sizeof(r0));
  r0.pi = (float)(1.0 / _dim_conn);
  op_check(gsbn_row_fill(rowst.p, R, r0, active_stream()), "gsbn_row_fill");
  contrib.alloc((size_t)S);
  rows_list.alloc((size_t)R);
  rows_cnt.alloc(_dim_hcu);
  pend_cnt.alloc(_dim_hcu);
  act_stamp.alloc((size_t)R);
  W = (_dim_conn + 31) / 32;
  const int64_t nfut = (int64_t)GSBN_FUT * _dim_hcu * W;
  fut.alloc((size_t)nfut);
  op_check(gsbn_i32_fill(rows_cnt.p, _dim_hcu, 0, active_stream()), "fill");
  op_check(gsbn_i32_fill(pend_cnt.p, _dim_hcu, 0, active_stream()), "fill");
  op_check(gsbn_i32_fill(act_stamp.p, R, -1, active_stream()), "fill");
  op_check(gsbn_i32_fill(reinterpret_cast<int32_t*>(fut.p), nfut, 0, active_stream()), "fill");
  _ptr_src_pop->_out.push_back(this);
  proj_row_ops()[_id] = this;

  /* views over the device state */
  struct VB {
    VecBase* v;
    VecView* view;
    size_t n;
  };
  const VB vb[] = {
      {_pij, new SynField(this, 0), (size_t)S}, {_eij, new SynField(this, 1), (size_t)S},
      {_zi2, new SynField(this, 2), (size_t)S}, {_zj2, new SynField(this, 3), (size_t)S},
      {_wij, new SynField(this, 4), (size_t)S}, {_tij, new SynField(this, 5), (size_t)S},
      {_pi, new RowField(this, 0), (size_t)R},  {_ei, new RowField(this, 1), (size_t)R},
      {_zi, new RowField(this, 2), (size_t)R},  {_ti, new RowField(this, 3), (size_t)R},
      {_qi, new QueueView(this, 0), (size_t)R}, {_siq, new QueueView(this, 1), (size_t)R},
      {_siq_saved, new QueueView(this, 1), (size_t)R},
      {_ssi, new QueueView(this, 2), 0},        {_ssj, new QueueView(this, 3), 0},
  };
  for (auto& e : vb) {
    /* on resume the saved host copy is pushed into the device state before
     * the first step (the queue: rebuilt by restore_queue) */
    const bool keep = restored_as(db, e.v, e.n, copy);
    if (keep && e.v == _qi) resume_qi = raw<int32_t>(e.v);
    if (keep && e.v == _siq_saved) resume_siq = raw<int8_t>(e.v);
    view_objs.emplace_back(e.view);
    e.v->view = e.view;
    if (e.v->type == DT_F32) static_cast<SyncVector<float>*>(e.v)->set_logical_size(e.n);
    if (e.v->type == DT_I32) static_cast<SyncVector<int32_t>*>(e.v)->set_logical_size(e.n);
    if (e.v->type == DT_I8) static_cast<SyncVector<int8_t>*>(e.v)->set_logical_size(e.n);
    if (keep)
      e.v->status = (e.v == _qi || e.v == _siq_saved) ? VecBase::SYNCED : VecBase::HOST_NEWER;
    else
      e.v->status = VecBase::DEVICE_NEWER;
    views.push_back(e.v);
  }
  if (copy) {
    /* a snapshot without siq_ (written by the reference) resumes with no
     * row active at the first step, as the reference's resume does */
    resume_qi.resize((size_t)R, 0);
    resume_siq.resize((size_t)R, 0);
    resume_pending = true;
  }
  Pop* dp = _ptr_dest_pop;
  _ptr_dest_pop->_epsc->before_host_read = [dp]() { dp->flush(); };
}

gsbn_pop_proj Proj::pop_desc() {
  gsbn_pop_proj d;
  memset(&d, 0, sizeof(d));
  d.contrib = contrib.p;
  d.pend_cnt = pend_cnt.p;
  d.C = _dim_conn;
  d.pj = _pj->mutable_gpu_data();
  d.ej = _ej->mutable_gpu_data();
  d.zj = _zj->mutable_gpu_data();
  d.taupdt = _taupdt;
  d.ke = _tauedt;
  d.kzj = _tauzjdt;
  d.kepsc = _tauepscdt;
  d.kftj = _kftj;
  d.bgain = _bgain;
  d.eps = _eps;
  d.fut = fut.p;
  d.W = W;
  d.rows_list = rows_list.p;
  d.rows_cnt = rows_cnt.p;
  d.act_stamp = act_stamp.p;
  if (dK) {
    d.dK = dK;
    d.dhist = dhist.p;
    d.dcnt = dcnt.p;
    d.dfl = dfl.p;
    d.walk_list = walk_list[0].p;
    d.walk_cnt = walk_cnt[0].p;
    d.walk_T = walk_T[0].p;
  }
  return d;
}

/* rows of each source unit (ii == n), ascending, with their delays */
void Proj::rebuild_fanout() {
  const int64_t R = (int64_t)_dim_hcu * _dim_conn;
  const int* ii = _ii->cpu_data();
  const int* di = _di->cpu_data();
  vector<int32_t> off((size_t)_nsrc + 1, 0);
  for (int64_t r = 0; r < R; ++r) {
    const int n = ii[r];
    if (n < 0) continue;
    if (n >= _nsrc) throw Error(-2, "ii_" + std::to_string(_id) + ": source unit out of range");
    const int d = di[r];
    if (d < 1 || d > 31) throw Error(-2, "di_" + std::to_string(_id) + ": delay must be in [1, 31]");
    off[(size_t)n + 1]++;
  }
  for (int n = 0; n < _nsrc; ++n) off[(size_t)n + 1] += off[(size_t)n];
  const size_t nnz = (size_t)off[(size_t)_nsrc];
  vector<int32_t> rows(std::max<size_t>(nnz, 1), 0);
  vector<uint8_t> del(std::max<size_t>(nnz, 1), 0);
  vector<int32_t> fill(off.begin(), off.end() - 1);
  for (int64_t r = 0; r < R; ++r) {
    const int n = ii[r];
    if (n < 0) continue;
    const int32_t q = fill[(size_t)n]++;
    rows[(size_t)q] = (int32_t)r;
    del[(size_t)q] = (uint8_t)di[r];
  }
  fwd_off.upload(off.data(), off.size());
  fwd_row.upload(rows.data(), rows.size());
  fwd_del.upload(del.data(), del.size());
}

void Proj::prepare(int simstep) {
  for (VecBase* v : views) {
    if (v->status == VecBase::HOST_NEWER && v->view) {
      v->view->push(v);
      v->status = VecBase::SYNCED;
    }
  }
  if (_ii->host_version != ii_version || _di->host_version != di_version) {
    rebuild_fanout();
    ii_version = _ii->host_version;
    di_version = _di->host_version;
  }
  (void)_ii->gpu_data();
  (void)_di->gpu_data();
  (void)simstep;
}

void Proj::update_all(int simstep, float old_prn) {
  const float kp = _taupdt * old_prn;
  const int target = simstep - 1;
  const CoefTable& t = table(kp, (int64_t)std::max(target, 0) + 2);
  gsbn_all_args a;
  memset(&a, 0, sizeof(a));
  a.H = _dim_hcu;
  a.M = _dim_mcu;
  a.C = _dim_conn;
  a.target = target;
  a.kp = kp;
  a.wgain = _wgain;
  a.eps = _eps;
  a.eps2 = _eps2;
  a.table = t.dev.p;
  a.table_len = (int32_t)t.len;
  a.syn = syn.p;
  a.rowst = rowst.p;
  a.pj = _pj->gpu_data();
  a.prec = prec;
  a.syn16 = syn16.p;
  a.wij = wij16.p;
  a.p_scale = p_scale();
  a.e_scale = e_scale();
  if (dK) {
    a.dK = dK;
    a.ii = _ii->gpu_data();
    a.dhist = dhist.p;
    a.dcnt = dcnt.p;
    a.kftj = _kftj;
  }
  Timing::get().begin(2, active_stream());
  op_check(gsbn_proj_all(&a, active_stream()), "gsbn_proj_all");
  Timing::get().end(2, active_stream());
}

/* every pending column kick, applied (the state the reference's eager
 * column updates leave); the learning rate has not changed since the kicks
 * (a change sweeps them up) */
void Proj::materialize_kicks() {
  if (!dK || !kicks_pending) return;
  kicks_pending = false;
  const CoefTable& t = table(kick_kp, (int64_t)std::max(kick_t, 0) + 2);
  gsbn_all_args a;
  memset(&a, 0, sizeof(a));
  a.H = _dim_hcu;
  a.M = _dim_mcu;
  a.C = _dim_conn;
  a.target = kick_t;
  a.table = t.dev.p;
  a.table_len = (int32_t)t.len;
  a.syn = syn.p;
  a.prec = prec;
  a.syn16 = syn16.p;
  a.wij = wij16.p;
  a.p_scale = p_scale();
  a.e_scale = e_scale();
  a.dK = dK;
  a.ii = ii_dev;
  a.dhist = dhist.p;
  a.dcnt = dcnt.p;
  a.kftj = _kftj;
  op_check(gsbn_proj_materialize(&a, active_stream()), "gsbn_proj_materialize");
}

void Proj::update_jxx(float prn) {
  Pop* dp = _ptr_dest_pop;
  gsbn_jxx_args a;
  memset(&a, 0, sizeof(a));
  const int64_t N = (int64_t)_dim_hcu * _dim_mcu;
  a.N = (int32_t)N;
  a.sj = dp->_spike->gpu_data();
  a.pj = _pj->mutable_gpu_data();
  a.ej = _ej->mutable_gpu_data();
  a.zj = _zj->mutable_gpu_data();
  a.bj = dp->_bj->mutable_gpu_data() + _proj_in_pop * N;
  a.epsc = dp->_epsc->mutable_gpu_data() + _proj_in_pop * N;
  a.kp = _taupdt * prn;
  a.ke = _tauedt;
  a.kzj = _tauzjdt;
  a.kepsc = _tauepscdt;
  a.kftj = _kftj;
  a.bgain = _bgain;
  a.eps = _eps;
  Timing::get().begin(3, active_stream());
  op_check(gsbn_proj_jxx(&a, active_stream()), "gsbn_proj_jxx");
  Timing::get().end(3, active_stream());
}

/* launch arguments that stay fixed between refreshes; the batched and the
 * single launch fill simstep/exec/kp/table per step */
gsbn_rowcol_args Proj::static_args(const float* slot_prn, int nslots, int64_t need) {
  Pop* dp = _ptr_dest_pop;
  gsbn_rowcol_args a;
  memset(&a, 0, sizeof(a));
  a.H = _dim_hcu;
  a.M = _dim_mcu;
  a.C = _dim_conn;
  a.kfti = _kfti;
  a.kftj = _kftj;
  a.wgain = _wgain;
  a.eps = _eps;
  a.eps2 = _eps2;
  a.syn = syn.p;
  a.rowst = rowst.p;
  a.pj = _pj->gpu_data();
  a.sj = dp->_spike->gpu_data();
  a.ii = _ii->gpu_data();
  a.rows_list = rows_list.p;
  a.rows_cnt = rows_cnt.p;
  a.spk_list = dp->spk_list.p;
  a.spk_cnt = dp->spk_cnt.p;
  a.act_stamp = act_stamp.p;
  a.contrib = contrib.p;
  a.taupdt = _taupdt;
  a.prec = prec;
  a.syn16 = syn16.p;
  a.wij = wij16.p;
  a.p_scale = p_scale();
  a.e_scale = e_scale();
  if (dK) {
    /* the column pass walks the listed units instead of every spiking one */
    a.dK = dK;
    a.dhist = dhist.p;
    a.dcnt = dcnt.p;
    a.spk_list = walk_list[0].p;
    a.spk_cnt = walk_cnt[0].p;
    a.walk_T = walk_T[0].p;
  }
  for (int i = 0; i < nslots; ++i) {
    const CoefTable& t = table(_taupdt * slot_prn[i], need);
    a.tables[i] = t.dev.p;
    a.table_lens[i] = (int32_t)t.len;
  }
  return a;
}

void Proj::update_que(int exec) {
  last_exec = exec;
  if (direct_que) return; /* the population kernel already stored si of the next step */
  gsbn_que_args a;
  memset(&a, 0, sizeof(a));
  a.Hsrc = _ptr_src_pop->_dim_hcu;
  a.Msrc = _ptr_src_pop->_dim_mcu;
  a.exec = exec;
  a.si = _si->gpu_data();
  a.out = fanout();
  Timing::get().begin(4, active_stream());
  op_check(gsbn_proj_que(&a, active_stream()), "gsbn_proj_que");
  Timing::get().end(4, active_stream());
}

/* Resume: the ring as it stood after the queue update of exec step
 * x = exec - 1 of the uninterrupted run (the inverse of QueueView::pull):
 * rows of siq_ active at x + 1, bit b of qi_[r] active at x + 2 + b.  When
 * the source population pushes its own spikes, those of step x (spike_ of
 * the snapshot) are still to be pushed, as exec x would have. */
void Proj::restore_queue(int exec) {
  const int H = _dim_hcu, C = _dim_conn;
  const int64_t R = (int64_t)H * C;
  vector<uint32_t> ring((size_t)GSBN_FUT * H * W, 0u);
  auto set = [&](int step, int64_t r) {
    const int h = (int)(r / C), k = (int)(r - (int64_t)h * C);
    ring[((size_t)(step % GSBN_FUT) * H + h) * W + (k >> 5)] |= 1u << (k & 31);
  };
  const int* ii = _ii->cpu_data();
  const int* di = _di->cpu_data();
  for (int64_t r = 0; r < R; ++r) {
    if (ii[r] < 0) continue;
    if (resume_siq[(size_t)r]) set(exec, r);
    for (int b = 0; b < di[r] && b < 31; ++b)
      if ((resume_qi[(size_t)r] >> b) & 1) set(exec + 1 + b, r);
  }
  hip_check(hipStreamSynchronize(active_stream()), "sync");
  if (!ring.empty())
    hip_check(hipMemcpy(fut.p, ring.data(), ring.size() * 4, hipMemcpyHostToDevice), "H2D");
  if (direct_que) {
    gsbn_que_args a;
    memset(&a, 0, sizeof(a));
    a.Hsrc = _ptr_src_pop->_dim_hcu;
    a.Msrc = _ptr_src_pop->_dim_mcu;
    a.exec = exec; /* pushes as exec - 1 */
    a.si = _ptr_src_pop->_spike->gpu_data();
    a.out = fanout();
    op_check(gsbn_proj_que(&a, active_stream()), "gsbn_proj_que");
  }
  resume_pending = false;
  vector<int32_t>().swap(resume_qi);
  vector<int8_t>().swap(resume_siq);
}

Proj::~Proj() {
  auto& ops = proj_row_ops();
  auto it = ops.find(_id);
  if (it != ops.end() && it->second == this) ops.erase(it);
}

void Proj::prune_flags(int t_min, float wp, float wn, vector<uint8_t>* flags) {
  flush_deferred();
  if (eager_hook) eager_hook();
  const int64_t R = (int64_t)_dim_hcu * _dim_conn;
  prepare(0); /* host writes of ii and the traces reach the device first */
  gsbn_prune_args a;
  memset(&a, 0, sizeof(a));
  a.R = R;
  a.M = _dim_mcu;
  a.prec = prec;
  a.ii = _ii->gpu_data();
  a.rowst = rowst.p;
  a.syn = syn.p;
  a.wij = wij16.p;
  a.t_min = t_min;
  a.wp = wp;
  a.wn = wn;
  if ((int64_t)sp_flags.n < R) sp_flags.alloc((size_t)R);
  a.flags = sp_flags.p;
  op_check(gsbn_proj_prune_flags(&a, active_stream()), "gsbn_proj_prune_flags");
  flags->resize((size_t)R);
  sp_flags.download(flags->data(), (size_t)R);
}

void Proj::clear_rows(const vector<int32_t>& rows) {
  flush_deferred();
  if (eager_hook) eager_hook();
  if (rows.empty()) return;
  const int64_t R = (int64_t)_dim_hcu * _dim_conn;
  for (int32_t r : rows)
    if (r < 0 || r >= R) throw Error(-4, "clear_rows: row out of range");
  prepare(0);
  sp_rows.upload(rows.data(), rows.size());
  gsbn_clear_rows_args a;
  memset(&a, 0, sizeof(a));
  a.rows = sp_rows.p;
  a.n = (int64_t)rows.size();
  a.M = _dim_mcu;
  a.prec = prec;
  a.H = _dim_hcu;
  a.C = _dim_conn;
  a.W = W;
  a.rowst = rowst.p;
  a.syn = syn.p;
  a.syn16 = syn16.p;
  a.fut = fut.p;
  op_check(gsbn_proj_clear_rows(&a, active_stream()), "gsbn_proj_clear_rows");
  /* host copies of the traces and the queue views are stale now */
  mark_views_device();
}

void Proj::rows_added() {
  flush_deferred();
  if (eager_hook) eager_hook();
  /* the fan-out lists follow ii/di now */
  prepare(0);
  /* with the population kernel doing the queue push, the spikes of the last
   * step went out over the old fan-out lists; pushing them again over the
   * new ones adds the new rows' bits (setting a bit twice changes nothing) */
  if (!direct_que || last_exec < 0) return;
  gsbn_que_args a;
  memset(&a, 0, sizeof(a));
  a.Hsrc = _ptr_src_pop->_dim_hcu;
  a.Msrc = _ptr_src_pop->_dim_mcu;
  a.exec = last_exec + 1; /* pushes as exec - 1 */
  a.si = _ptr_src_pop->_spike->gpu_data();
  a.out = fanout();
  op_check(gsbn_proj_que(&a, active_stream()), "gsbn_proj_que");
  mark_views_device();
}

void Proj::mark_views_device() {
  for (VecBase* v : views) v->status = VecBase::DEVICE_NEWER;
}

/* ---- the procedure ---- */

class ProcUpdLazy : public ProcedureBase {
  REGISTER(ProcUpdLazy)
 public:
  ~ProcUpdLazy() override;
  void init_new(SolverParam solver_param, Database& db) override { init(solver_param, db, false); }
  void init_copy(SolverParam solver_param, Database& db) override { init(solver_param, db, true); }
  void update_cpu() override;
  void update_gpu() override;
  int64_t work(int i);
  int64_t sweep_synapses = 0, sweeps = 0, execs = 0;

 private:
  void init(SolverParam solver_param, Database& db, bool copy);
  int kp_slot(float prn);
  void refresh_args(int simstep);
  void build_batch_maps();
  vector<Proj*> _list_proj;
  vector<Pop*> _list_pop;
  GlobalVar _glv;
  uint64_t _seed = 0;
  int _wta_pop_max = 0;
  int _exec = 0;
  bool _configured = false;
  TableCache _tables;
  /* cached launch arguments */
  bool _batch = false;
  bool _args_dirty = true;
  uint64_t _epoch_seen = ~0ull, _tables_seen = ~0ull, _write_epoch_seen = ~0ull;
  vector<gsbn_pop_args> _pop_args;
  vector<gsbn_rowcol_args> _proj_args;
  float _slot_prn[GSBN_KP_SLOTS] = {0, 0, 0, 0};
  int _nslots = 0, _next_evict = 0;
  int _total_units = 0, _total_hcus = 0;
  /* batched launch state */
  DevBuf<gsbn_pop_args> _pop_dev;
  DevBuf<gsbn_rowcol_args> _proj_dev;
  vector<uint8_t> _pop_sent, _proj_sent; /* what the two arrays hold on the device */
  DevBuf<int32_t> _wg_pop, _wg_pop_hcu, _wg_proj, _wg_proj_hcu, _in_proj;
  int _pop_threads = 64, _max_slots = 0, _any_quirk = 0, _nwg_pop = 0, _nhcu_proj = 0, _parts = 1;
  /* pop-wave 1 (default): a batch's population update with one wave per
   * hypercolumn (gsbn_pop_update_batch_wave, _pop_upl units per lane) when
   * every population qualifies; 0: the block kernel */
  int _pop_wave = 1, _pop_upl = 0;
  /* row overflow queue of the one-wave row/column launch (gsbn_step_args
   * ovf_*): ovf-keep rows per hypercolumn in its own wave (0: no queue),
   * ovf-cap queued rows at most per step (a busier step runs every row in
   * the hypercolumns' own waves); _ovf_exec: the last step whose population
   * launch cleared the next step's counter */
  int _ovf_keep = 2, _ovf_cap = 1 << 18;
  static constexpr int kOvfCounters = 2 * 64 * 16; /* 2 parities x 64 sub-queues, 64-byte stride */
  int64_t _ovf_exec = -2;
  DevBuf<int32_t> _ovf_cnt, _ovf_task;
  vector<int32_t> _in_proj_host;

  int _max_m = 0, _parts0 = 1;
  /* Chained step (plan -2): one launch per step runs the previous step's
   * row and column updates with this step's population update
   * (gsbn_step_chained); this step's are pending until the next launch or a
   * flush.  Possible for one population with its one recurrent projection
   * (direct queue push, M <= 256, no wta_pop_max). */
  /* Graded splits of a batch (plans -11, -12, -13): the row/column
   * hypercolumns fall into four size buckets by the quartiles of their work
   * (slots x MCUs); plan -10-g splits bucket b (0 smallest) into
   * 2^max(0, b - 3 + g) parts, so g = 3 gives 1, 2, 4, 8 parts.  One map per
   * plan, built on first use. */
  int _graded_fixed = 0;
  vector<int64_t> _hcu_work; /* per row/column HCU, in launch order */
  vector<int32_t> _proj_map_host, _hcu_map_host;
  int64_t _bucket_hi[3] = {0, 0, 0};
  DevBuf<int32_t> _gmap_proj[4], _gmap_hcu[4], _gmap_part[4];
  int _gmap_n[4] = {0, 0, 0, 0};
  void build_graded_map(int g);
  bool _chainable = false, _pending = false;
  int _chain_fixed = 0, _chain_parts = 1, _chain_warm = 0;
  gsbn_rowcol_args _pend_ra;
  void flush_deferred();
  /* every population has M <= 256 and one incoming projection, and the
   * fused step's LDS (active-row bitmap and list) fits */
  bool _fusable = false;
  /* Step plan: the row/column work split (workgroups per destination HCU),
   * or -1 for the fused launch (population and row/column update of an HCU
   * in one workgroup, one launch per step).  Fixed by argi rowcol-parts, or
   * chosen online: from step kTuneStart on, the candidates take turns step
   * by step (so that each sees the same mix of busy and quiet steps; results
   * do not depend on the plan), a step without a sweep timed on the device
   * from its first launch to its last; after kTuneSamples steps each, the
   * plan with the smallest median step time is kept.  (Runs of several
   * consecutive steps per sample were tried: they met different spiking
   * activity per candidate and chose worse plans.) */
  static constexpr int kTuneStart = 20, kTuneRun = 1, kTuneSamples = 10;
  int _parts_fixed = 0, _fused_fixed = 0;
  /* rowcol-wave 1: a batch's row/column update with one wave per
   * hypercolumn (plan -20); 2, 3, 4: one wave per part of the graded split
   * g = 1, 2, 3 (plans -21, -22, -23) */
  int _wave_fixed = 0;
  /* deferred-columns K: column updates deferred through rings of K spike
   * steps per unit (0: eager); see Proj::dK.  Same results for every K, a
   * fraction of the column traffic.  Default where it applies: 8 for one
   * network (its row update replays a ring of 8 from registers), 16 for a
   * batch, whose one-wave-per-hypercolumn row/column launch is dominated by
   * column walks at 8 (measured on 512 networks: 50.8 us at 8, 38.9 us at
   * 16, 39.1 us at 32 and 64). */
  static constexpr int kDeferDefault = 8, kDeferBatch = 16;
  int _defer_arg = -1; /* the argument, -1 when not given */
  int _defer_k = kDeferDefault;
  void end_deferral();
  bool _tuned = false;
  int _tune_pos = 0, _run_len = 0;
  vector<int> _cand;
  vector<hipEvent_t> _tev;
  int rowcol_parts(hipStream_t s, bool* timed);
  void finish_tuning();
};
REGISTERIMPL(ProcUpdLazy);

ProcUpdLazy::~ProcUpdLazy() {
  for (hipEvent_t e : _tev) (void)hipEventDestroy(e);
  for (auto* p : _list_proj) delete p;
  for (auto* p : _list_pop) delete p;
}

/* the row and column updates a chained step left pending, as a row/column
 * launch of that step */
void ProcUpdLazy::flush_deferred() {
  if (!_pending) return;
  _pending = false;
  Proj* q = _list_proj[0];
  gsbn_rowcol_args r = _pend_ra;
  const CoefTable& t = q->table(r.kp, (int64_t)r.simstep + 2);
  r.table = t.dev.p;
  r.table_len = (int32_t)t.len;
  r.parts = _parts0;
  hipStream_t s = active_stream();
  Timing::get().begin(1, s);
  op_check(gsbn_proj_rowcol(&r, s), "gsbn_proj_rowcol");
  Timing::get().end(1, s);
}

/* column updates become eager again (a structural change of the rows is
 * coming, which the replay of pending kicks cannot follow): the pending
 * row/column work and every pending kick are applied first */
void ProcUpdLazy::end_deferral() {
  if (!_defer_k) return;
  flush_deferred();
  for (auto* q : _list_proj) {
    q->materialize_kicks();
    q->dK = 0;
  }
  _defer_k = 0;
  _args_dirty = true;
  _glv.puti("deferred-columns", 0);
}

/* the plan of this step; *timed while tuning (the step belongs to run
 * _tune_pos, timed by events _tev[2 _tune_pos] and _tev[2 _tune_pos + 1]) */
int ProcUpdLazy::rowcol_parts(hipStream_t s, bool* timed) {
  *timed = false;
  if (_fused_fixed) return -1;
  if (_chain_fixed > 0) return -2;
  if (_graded_fixed > 0) return -10 - _graded_fixed;
  if (_wave_fixed > 0) return -19 - _wave_fixed;
  if (_parts_fixed > 0) return _parts_fixed;
  if (_tuned || execs < kTuneStart) return _parts;
  if (_cand.empty()) {
    for (int c : {1, 2, 4, 8, 16, 32})
      if ((int64_t)c * std::max(_nhcu_proj, 1) <= 262144) _cand.push_back(c);
    /* and the largest split whose grid the device holds at once (a grid
     * beyond it waits for a second dispatch round) */
    int32_t res = 0;
    const int prec = _list_proj.empty() ? GSBN_PREC_F32 : _list_proj[0]->prec;
    if (gsbn_rowcol_resident(prec, &res) == 0) {
      const int fit = res / std::max(_nhcu_proj, 1);
      if (fit > 1 && fit < 32 && std::find(_cand.begin(), _cand.end(), fit) == _cand.end())
        _cand.push_back(fit);
    }
    if (_cand.empty()) _cand.push_back(1);
    if (_fusable) _cand.push_back(-1);
    if (_chainable) _cand.push_back(-2);
    if (_batch && !_hcu_work.empty()) _cand.push_back(-20);
    /* graded splits; their maps are built here, before any sample is timed
     * (their one-wave forms, plans -21 to -23, are no candidates: on 512
     * networks the best of them, -21, ran within 2 % of plan -20, and every
     * candidate more spreads the tuner's samples thinner) */
    if (_batch && !_hcu_work.empty())
      for (int g = 1; g <= 3; ++g) {
        _cand.push_back(-10 - g);
        if (!_gmap_n[g]) build_graded_map(g);
      }
    _tev.resize(2 * _cand.size() * kTuneSamples);
    for (auto& e : _tev) hip_check(hipEventCreate(&e), "hipEventCreate");
  }
  (void)s;
  const int c = _cand[_tune_pos % _cand.size()];
  /* a chained sample is the second of two chained steps (the first one
   * leaves the work that a chained launch takes over) */
  if (c == -2 && !_chain_warm) {
    _chain_warm = 1;
    return -2;
  }
  _chain_warm = 0;
  *timed = true;
  return c;
}

void ProcUpdLazy::finish_tuning() {
  const size_t nc = _cand.size();
  hip_check(hipEventSynchronize(_tev.back()), "hipEventSynchronize");
  float best = 0.0f;
  for (size_t c = 0; c < nc; ++c) {
    vector<float> t;
    for (int k = 0; k < kTuneSamples; ++k) {
      const size_t i = (size_t)k * nc + c;
      float ms = 0.0f;
      hip_check(hipEventElapsedTime(&ms, _tev[2 * i], _tev[2 * i + 1]), "hipEventElapsedTime");
      t.push_back(ms / kTuneRun);
    }
    std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
    const float med = t[t.size() / 2];
    /* the measured median step time of each plan, for inspection */
    _glv.putf("step-plan-ms" + std::to_string(_cand[c]), med);
    if (c == 0 || med < best) {
      best = med;
      _parts = _cand[c];
    }
  }
  _tuned = true;
  _glv.puti("rowcol-parts", _parts);
  /* the plan is final from the next step on */
  _glv.puti("plan-tuned", 1);
}

/* init_new (ProcUpdLazy.cpp:8-46) and init_copy (ProcUpdLazy.cpp:48-86) */
void ProcUpdLazy::init(SolverParam solver_param, Database& db, bool copy) {
  NetParam net_param = solver_param.net_param;
  ProcParam proc_param = get_proc_param(solver_param);
  Parser par(proc_param);
  int lo = 0, hi = 0;
  par.argi("seed-lo", lo);
  par.argi("seed-hi", hi);
  _seed = (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 31);
  par.argi("wta-pop-max", _wta_pop_max);
  par.argi("rowcol-parts", _parts_fixed);
  par.argi("fused-step", _fused_fixed);
  /* chained-step: 1 always, -1 never, 0 (default) one of the tuned plans */
  par.argi("chained-step", _chain_fixed);
  /* rowcol-graded g (1-3): always the graded split g of a batch */
  par.argi("rowcol-graded", _graded_fixed);
  par.argi("rowcol-wave", _wave_fixed);
  par.argi("pop-wave", _pop_wave);
  par.argi("ovf-keep", _ovf_keep);
  par.argi("ovf-cap", _ovf_cap);
  if (_ovf_keep < 0 || _ovf_cap < 0)
    throw Error(-2, "ProcUpdLazy: ovf-keep and ovf-cap must be >= 0");
  if (_wave_fixed < 0 || _wave_fixed > 4)
    throw Error(-2, "ProcUpdLazy: rowcol-wave must be in [0, 4]");
  par.argi("deferred-columns", _defer_arg);
  if (_defer_arg != -1 &&
      (_defer_arg < 0 || _defer_arg == 1 || _defer_arg > 64 || (_defer_arg & (_defer_arg - 1))))
    throw Error(-2, "ProcUpdLazy: deferred-columns must be 0 or a power of two in [2, 64]");
  if (_graded_fixed < 0 || _graded_fixed > 3)
    throw Error(-2, "ProcUpdLazy: rowcol-graded must be in [0, 3]");
  if (_chain_fixed && _fused_fixed > 0 && _chain_fixed > 0)
    throw Error(-2, "ProcUpdLazy: fused-step and chained-step exclude each other");
  if (_parts_fixed < 0 || _parts_fixed > 1024)
    throw Error(-2, "ProcUpdLazy: rowcol-parts must be in [0, 1024]");
  int hcu_cnt = 0, mcu_cnt = 0;
  for (auto& pp : net_param.pop_param) {
    for (uint32_t j = 0; j < pp.pop_num; j++) {
      Pop* p = new Pop();
      try {
        p->init_new(proc_param, pp, db, &_list_pop, &hcu_cnt, &mcu_cnt, copy);
      } catch (...) {
        if (_list_pop.empty() || _list_pop.back() != p) delete p;
        throw;
      }
    }
  }
  const int total_pop_num = (int)_list_pop.size();
  _total_units = mcu_cnt;
  _total_hcus = hcu_cnt;
  if (!_glv.geti("sim-end-step", _tables.sim_end)) _tables.sim_end = 0;
  for (auto& jp : net_param.proj_param) {
    if ((int)jp.src_pop < total_pop_num && (int)jp.dest_pop < total_pop_num) {
      Proj* proj = new Proj();
      proj->tcache = &_tables;
      try {
        proj->init_new(proc_param, jp, db, &_list_proj, &_list_pop, copy);
      } catch (...) {
        if (_list_proj.empty() || _list_proj.back() != proj) delete proj;
        throw;
      }
    }
  }
  hip_check(hipStreamSynchronize(active_stream()), "sync");
  _glv.puti("updlazy-ready", 1);
}

void ProcUpdLazy::update_cpu() {
  throw Error(-5, "ProcUpdLazy::update_cpu: this build runs the lazy update on the GPU only");
}

int ProcUpdLazy::kp_slot(float prn) {
  for (int i = 0; i < _nslots; ++i)
    if (!memcmp(&_slot_prn[i], &prn, 4)) return i;
  _args_dirty = true;
  if (_nslots < GSBN_KP_SLOTS) {
    _slot_prn[_nslots] = prn;
    return _nslots++;
  }
  const int i = _next_evict;
  _next_evict = (i + 1) % GSBN_KP_SLOTS;
  _slot_prn[i] = prn;
  return i;
}

void ProcUpdLazy::build_batch_maps() {
  vector<int32_t> wp, wh, qp, qh;
  _pop_threads = 64;
  _any_quirk = _wta_pop_max;
  /* one launch for all populations, blocks sized for the largest M (one
   * launch per block size was measured slower on the 512 networks) */
  for (size_t i = 0; i < _list_pop.size(); ++i) {
    Pop* p = _list_pop[i];
    for (int h = 0; h < p->_dim_hcu; ++h) {
      wp.push_back((int32_t)i);
      wh.push_back(h);
    }
    _pop_threads = std::max(_pop_threads, ((p->_dim_mcu + 63) / 64) * 64);
  }
  _pop_upl = 0;
  if (_pop_wave && !_wta_pop_max) {
    bool ok = true;
    int maxm = 0;
    for (Pop* p : _list_pop) {
      maxm = std::max(maxm, p->_dim_mcu);
      if (p->_dim_mcu > 128 || p->_in.size() > 1 || p->_out.size() > 1 || p->ext_off.p) ok = false;
      if (!p->_in.empty() && p->_in[0]->_dim_conn > 8192) ok = false;
    }
    if (ok) _pop_upl = maxm <= 64 ? 1 : 2;
  }
  /* row/column workgroups of the largest hypercolumns (slots x MCUs, the
   * size of a column pass and of a row) first: the long ones start early
   * instead of ending the launch alone */
  _max_slots = 0;
  vector<int> order(_list_proj.size());
  for (size_t i = 0; i < _list_proj.size(); ++i) {
    order[i] = (int)i;
    _max_slots = std::max(_max_slots, _list_proj[i]->_dim_conn);
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    const Proj* a = _list_proj[x];
    const Proj* b = _list_proj[y];
    return (int64_t)a->_dim_conn * a->_dim_mcu > (int64_t)b->_dim_conn * b->_dim_mcu;
  });
  _hcu_work.clear();
  for (int i : order)
    for (int h = 0; h < _list_proj[i]->_dim_hcu; ++h) {
      qp.push_back((int32_t)i);
      qh.push_back(h);
      _hcu_work.push_back((int64_t)_list_proj[i]->_dim_conn * _list_proj[i]->_dim_mcu);
    }
  if (!_hcu_work.empty()) {
    vector<int64_t> w = _hcu_work;
    std::sort(w.begin(), w.end());
    for (int b = 0; b < 3; ++b) _bucket_hi[b] = w[(w.size() * (size_t)(b + 1)) / 4];
  }
  _nwg_pop = (int)wp.size();
  _nhcu_proj = (int)qp.size();
  _proj_map_host = qp;
  _hcu_map_host = qh;
  /* the projection into each population (fused step: exactly one) */
  vector<int32_t> in;
  for (Pop* p : _list_pop)
    in.push_back(p->_in.empty() ? 0
                                : (int32_t)(std::find(_list_proj.begin(), _list_proj.end(),
                                                      p->_in[0]) - _list_proj.begin()));
  _in_proj.upload(in.data(), in.size());
  _in_proj_host = in;
  if (_pop_upl && _ovf_keep > 0) {
    _ovf_cnt.alloc(kOvfCounters);
    _ovf_task.alloc((size_t)3 * std::max(_ovf_cap, 1));
    op_check(gsbn_i32_fill(_ovf_cnt.p, kOvfCounters, 0, active_stream()), "fill");
  }
  _wg_pop.upload(wp.data(), wp.size());
  _wg_pop_hcu.upload(wh.data(), wh.size());
  if (!qp.empty()) {
    _wg_proj.upload(qp.data(), qp.size());
    _wg_proj_hcu.upload(qh.data(), qh.size());
  }
}

void ProcUpdLazy::build_graded_map(int g) {
  vector<int32_t> mp, mh, mq;
  for (size_t i = 0; i < _hcu_work.size(); ++i) {
    int b = 0;
    while (b < 3 && _hcu_work[i] > _bucket_hi[b]) ++b;
    const int parts = 1 << std::max(0, b - 3 + g);
    for (int part = 0; part < parts; ++part) {
      mp.push_back(_proj_map_host[i]);
      mh.push_back(_hcu_map_host[i]);
      mq.push_back(part | (parts << 16));
    }
  }
  _gmap_proj[g].upload(mp.data(), mp.size());
  _gmap_hcu[g].upload(mh.data(), mh.size());
  _gmap_part[g].upload(mq.data(), mq.size());
  _gmap_n[g] = (int)mp.size();
}

void ProcUpdLazy::refresh_args(int simstep) {
  _pop_args.clear();
  for (auto* p : _list_pop) _pop_args.push_back(p->args(0, 0, 0.0f, 0, 0, true, _seed, _wta_pop_max));
  /* the incoming projection's index in the batch's row/column argument
   * array, which the overflow queue's tasks name */
  if (_batch)
    for (size_t i = 0; i < _pop_args.size() && i < _in_proj_host.size(); ++i)
      _pop_args[i].proj0.bidx = _in_proj_host[i];
  _proj_args.clear();
  for (auto* q : _list_proj) _proj_args.push_back(q->static_args(_slot_prn, _nslots, (int64_t)simstep + 2));
  /* a host write (a new stimulus row, say) leaves these unchanged: the
   * upload, which waits for the queued steps, is skipped then */
  if (_batch) {
    const size_t pb = _pop_args.size() * sizeof(gsbn_pop_args);
    if (_pop_sent.size() != pb || memcmp(_pop_sent.data(), _pop_args.data(), pb)) {
      _pop_dev.upload(_pop_args.data(), _pop_args.size());
      _pop_sent.assign(reinterpret_cast<const uint8_t*>(_pop_args.data()),
                       reinterpret_cast<const uint8_t*>(_pop_args.data()) + pb);
    }
    const size_t qb = _proj_args.size() * sizeof(gsbn_rowcol_args);
    if (!_proj_args.empty() &&
        (_proj_sent.size() != qb || memcmp(_proj_sent.data(), _proj_args.data(), qb))) {
      _proj_dev.upload(_proj_args.data(), _proj_args.size());
      _proj_sent.assign(reinterpret_cast<const uint8_t*>(_proj_args.data()),
                        reinterpret_cast<const uint8_t*>(_proj_args.data()) + qb);
    }
  }
  _args_dirty = false;
}

void ProcUpdLazy::update_gpu() {
  int cycle_flag = 0;
  GSBN_CHECK(_glv.geti("cycle-flag", cycle_flag), "cycle-flag");
  if (cycle_flag != 1) return;
  int simstep = 0, lginp_idx = 0, wmask_idx = 0;
  float prn = 0, old_prn = 0;
  GSBN_CHECK(_glv.geti("simstep", simstep), "simstep");
  GSBN_CHECK(_glv.getf("prn", prn), "prn");
  GSBN_CHECK(_glv.getf("old-prn", old_prn), "old-prn");
  GSBN_CHECK(_glv.geti("lginp-idx", lginp_idx), "lginp-idx");
  GSBN_CHECK(_glv.geti("wmask-idx", wmask_idx), "wmask-idx");
  hipStream_t stream = active_stream();
  if (!_configured) {
    /* ProcExchangeSpike initialised after this procedure delivers
     * spike_<src> to si_<proj> once per executed step, so the queue push
     * of step x+1 is the spikes of step x: the population kernel pushes
     * them itself */
    int direct = 0;
    _glv.geti("exchange-spike-direct", direct);
    for (auto* q : _list_proj) {
      q->direct_que = direct != 0;
      if (q->direct_que) q->_ptr_src_pop->push_direct = true;
    }
    _batch = _list_pop.size() > 1 || _list_proj.size() > 1;
    if (_batch) build_batch_maps();
    _nhcu_proj = 0;
    for (auto* q : _list_proj) _nhcu_proj += q->_dim_hcu;
    _max_m = 0;
    _fusable = !_list_proj.empty();
    for (auto* p : _list_pop) {
      _max_m = std::max(_max_m, p->_dim_mcu);
      if (p->_dim_mcu > 256 || p->_in.size() != 1) _fusable = false;
    }
    if (!_batch) {
      _max_slots = 0;
      for (auto* q : _list_proj) _max_slots = std::max(_max_slots, q->_dim_conn);
    }
    if ((3 * 256 + (int64_t)(_max_slots + 31) / 32 + _max_slots) * 4 > 48 * 1024) _fusable = false;
    if (_fused_fixed && !_fusable)
      throw Error(-2, "ProcUpdLazy: fused-step needs populations of at most 256 MCUs with one "
                      "incoming projection each, and the slot bitmap and list in 48 KB of LDS");
    /* starting split until the tuning has run: the largest that the device
     * holds at once */
    {
      int32_t res = 0;
      const int prec = _list_proj.empty() ? GSBN_PREC_F32 : _list_proj[0]->prec;
      if (gsbn_rowcol_resident(prec, &res) != 0 || res <= 0) res = 1536;
      _parts0 = std::max(1, std::min(32, res / std::max(_nhcu_proj, 1)));
    }
    _chainable = !_batch && _list_pop.size() == 1 && _list_proj.size() == 1 &&
                 _list_proj[0]->direct_que && _list_pop[0]->_dim_mcu <= 256 && !_wta_pop_max &&
                 (3 * 256 + (int64_t)(_max_slots + 31) / 32) * 4 <= 48 * 1024;
    if (_chain_fixed > 0 && !_chainable)
      throw Error(-2, "ProcUpdLazy: chained-step needs one population of at most 256 MCUs with "
                      "one recurrent projection, the population kernel's queue push and no "
                      "wta-pop-max");
    if (_chain_fixed < 0) _chainable = false;
    if (_chainable) {
      Pop* dp = _list_pop[0];
      Proj* q = _list_proj[0];
      const int H = dp->_dim_hcu, M = dp->_dim_mcu;
      dp->spk_list1.alloc((size_t)H * M);
      dp->spk_cnt1.alloc((size_t)H);
      op_check(gsbn_i32_fill(dp->spk_cnt1.p, H, 0, stream), "fill");
      const int64_t R = (int64_t)q->_dim_hcu * q->_dim_conn;
      q->act_stamp1.alloc((size_t)R);
      op_check(gsbn_i32_fill(q->act_stamp1.p, R, -1, stream), "fill");
      int32_t res = 0;
      if (gsbn_chained_resident(q->prec, &res) != 0 || res <= 0) res = 768;
      /* column workgroups per HCU: the workgroups of HCUs without a spike
       * return at once, so slices finer than one dispatch round holds
       * shorten the busy steps; about 128 slots per workgroup, at most 64
       * parts and about 16 k column workgroups, and no fewer than one round
       * holds (up to 32) */
      const int fit = std::min(32, (res - H) / std::max(H, 1));
      const int by_slots = (q->_dim_conn + 127) / 128;
      const int budget = std::max(1, 16384 / std::max(H, 1));
      _chain_parts = std::max(1, std::min(64, std::max(fit, std::min(by_slots, budget))));
      q->flush_hook = [this]() { flush_deferred(); };
    }
    /* deferred column updates: every population with exactly one incoming
     * projection (the population kernel keeps the first projection's rings) */
    _defer_k = _defer_arg >= 0 ? _defer_arg : _batch ? kDeferBatch : kDeferDefault;
    if (_defer_k) {
      bool ok = !_list_proj.empty();
      for (auto* p : _list_pop)
        if (p->_in.size() != 1) ok = false;
      if (!ok) _defer_k = 0;
    }
    /* the fused launch does eager column updates: asking for it asks for them */
    if (_fused_fixed) _defer_k = 0;
    if (_defer_k) {
      _fusable = false;
      for (auto* q : _list_proj) {
        q->enable_deferred(_defer_k);
        q->eager_hook = [this]() { end_deferral(); };
      }
    }
    _glv.puti("deferred-columns", _defer_k);
    /* GlobalVar pop-wave: units per lane of the one-wave population launch
     * of a batch (0: the block kernel, or no batch) */
    _glv.puti("pop-wave", _batch ? _pop_upl : 0);
    if (_graded_fixed && !_batch)
      throw Error(-2, "ProcUpdLazy: rowcol-graded needs a batch of populations or projections");
    if (_wave_fixed > 0 && !_batch)
      throw Error(-2, "ProcUpdLazy: rowcol-wave needs a batch of populations or projections");
    _parts = _fused_fixed ? -1 : (_chain_fixed > 0 ? -2 : _parts0);
    _glv.puti("rowcol-parts", _fused_fixed ? -1
                              : _chain_fixed > 0 ? -2
                              : _graded_fixed > 0 ? -10 - _graded_fixed
                              : _wave_fixed > 0 ? -19 - _wave_fixed
                              : _parts_fixed > 0 ? _parts_fixed : _parts);
    /* GlobalVar plan-tuned: 1 once the plan no longer changes */
    _glv.puti("plan-tuned", (_fused_fixed || _chain_fixed > 0 || _graded_fixed > 0 ||
                             _wave_fixed > 0 || _parts_fixed > 0 || _list_proj.empty()) ? 1 : 0);
    _configured = true;
    _args_dirty = true;
  }
  const int slot = kp_slot(prn);
  for (auto* q : _list_proj) (void)q->table(q->_taupdt * prn, (int64_t)simstep + 2);
  const bool touched = VecBase::host_epoch() != _epoch_seen;
  /* host writes reach the device below: pending work of a chained step
   * runs first, on the state it was computed from (reads need no flush:
   * the views of the synapse and row state flush when they are read) */
  if (VecBase::host_write_epoch() != _write_epoch_seen) {
    if (_pending) flush_deferred();
    for (auto* q : _list_proj) q->materialize_kicks();
  }
  _write_epoch_seen = VecBase::host_write_epoch();
  if (touched)
    for (auto* q : _list_proj) q->prepare(simstep);
  for (auto* q : _list_proj)
    if (q->resume_pending) q->restore_queue(_exec);
  if (touched || _args_dirty || _tables.version != _tables_seen) refresh_args(simstep);

  /* current .lginp / .wmask rows (Pop.cpp:331-332) */
  Pop* p0 = _list_pop.front();
  const int64_t lg_off = p0->_lginp->ld() <= 0 ? (int64_t)lginp_idx : (int64_t)lginp_idx * p0->_lginp->ld();
  const int64_t wm_off = p0->_wmask->ld() <= 0 ? (int64_t)wmask_idx : (int64_t)wmask_idx * p0->_wmask->ld();
  if (lginp_idx < 0 || lg_off + _total_units > (int64_t)p0->_lginp->size())
    throw Error(-2, "lginp-idx out of range");
  if (wmask_idx < 0 || wm_off + _total_hcus > (int64_t)p0->_wmask->size())
    throw Error(-2, "wmask-idx " + std::to_string(wmask_idx) + " out of range");

  const bool sweep = old_prn != prn;
  gsbn_step_args st;
  memset(&st, 0, sizeof(st));
  st.simstep = simstep;
  st.exec = _exec;
  st.do_jxx = sweep ? 0 : 1;
  st.kp_slot = slot;
  st.prn = prn;
  st.prec = _list_proj.empty() ? GSBN_PREC_F32 : _list_proj[0]->prec;
  st.lginp_offset = lg_off;
  st.wmask_offset = wm_off;

  /* the step plan (row/column parts, or -1: the fused launch); sweep steps
   * run the sweep between the two launches and are not tuning samples */
  bool timed = false;
  int plan = _list_proj.empty() ? 1 : rowcol_parts(stream, &timed);
  if (sweep) {
    if (timed) _run_len = 0; /* the run restarts after the sweep */
    timed = false;
    _chain_warm = 0;
    if (plan == -1) plan = _parts0;
  }
  if (plan != -2 && _pending) flush_deferred();
  /* the row overflow queue: passed to every one-wave population launch (it
   * clears the next step's counter), used by the one-wave row/column plan */
  const bool ovf = _batch && _pop_upl && _ovf_cnt.p;
  if (ovf) {
    st.ovf_cnt = _ovf_cnt.p;
    st.ovf_task = _ovf_task.p;
    st.ovf_cap = _ovf_cap;
    st.ovf_keep = plan == -20 ? _ovf_keep : 0;
    if (_ovf_exec != (int64_t)_exec - 1)
      op_check(gsbn_i32_fill(_ovf_cnt.p, kOvfCounters, 0, stream), "fill"); /* not cleared by the last step */
  }
  const size_t ev = 2 * (size_t)_tune_pos;
  if (timed && _run_len == 0) hip_check(hipEventRecord(_tev[ev], stream), "hipEventRecord");

  gsbn_pop_args pa;
  if (!_batch) {
    pa = _pop_args[0];
    pa.simstep = simstep;
    pa.exec = _exec;
    pa.prn = prn;
    pa.do_jxx = st.do_jxx;
    pa.lginp += lg_off;
    pa.wmask += wm_off;
  }
  gsbn_rowcol_args ra;
  if (!_batch && !_list_proj.empty()) {
    ra = _proj_args[0];
    ra.simstep = simstep;
    ra.exec = _exec;
    ra.kp = ra.taupdt * prn;
    ra.table = ra.tables[slot];
    ra.table_len = ra.table_lens[slot];
    ra.parts = plan < 0 ? 1 : plan;
  }
  if (plan == -2) {
    /* chained: this step's lists and stamps go to the set of its exec
     * parity; the previous step's (the other set) feed its column updates */
    Pop* dp = _list_pop[0];
    Proj* q = _list_proj[0];
    const int side = _exec & 1;
    pa.spk_list = dp->spk_list_at(side);
    pa.spk_cnt = dp->spk_cnt_at(side);
    pa.proj0.act_stamp = q->act_stamp_at(side);
    if (q->dK) {
      pa.proj0.walk_list = q->walk_list[side].p;
      pa.proj0.walk_cnt = q->walk_cnt[side].p;
      pa.proj0.walk_T = q->walk_T[side].p;
    }
    if (_pending) {
      const CoefTable& t = q->table(_pend_ra.kp, (int64_t)_pend_ra.simstep + 2);
      _pend_ra.table = t.dev.p;
      _pend_ra.table_len = (int32_t)t.len;
    } else {
      _pend_ra = ra; /* not read without a previous step */
    }
    Timing::get().begin(6, stream);
    op_check(gsbn_step_chained(&pa, &_pend_ra, _pending ? 1 : 0, stream), "gsbn_step_chained");
    Timing::get().end(6, stream);
    dp->flush_pending = true;
    dp->spk_side = side;
    ra.spk_list = pa.spk_list;
    ra.spk_cnt = pa.spk_cnt;
    if (q->dK) {
      ra.spk_list = q->walk_list[side].p;
      ra.spk_cnt = q->walk_cnt[side].p;
      ra.walk_T = q->walk_T[side].p;
    }
    ra.act_stamp = pa.proj0.act_stamp;
    ra.parts = _chain_parts;
    _pend_ra = ra;
    _pending = true;
    if (sweep) {
      q->update_all(simstep, old_prn);
      sweep_synapses += (int64_t)q->_dim_hcu * q->_dim_conn * q->_dim_mcu;
      sweeps++;
      q->update_jxx(prn);
      flush_deferred();
    }
  } else if (plan == -1) {
    for (auto* p : _list_pop) p->spk_side = 0;
    Timing::get().begin(6, stream);
    if (_batch)
      op_check(gsbn_step_fused_batch(_pop_dev.p, (int32_t)_pop_args.size(), _proj_dev.p, _wg_pop.p,
                                     _wg_pop_hcu.p, _nwg_pop, _in_proj.p, _max_m, _max_slots,
                                     _any_quirk, &st, stream),
               "gsbn_step_fused_batch");
    else
      op_check(gsbn_step_fused(&pa, &ra, stream), "gsbn_step_fused");
    Timing::get().end(6, stream);
    for (auto* p : _list_pop) p->flush_pending = true;
  } else {
    for (auto* p : _list_pop) p->spk_side = 0;
    Timing::get().begin(0, stream);
    if (_batch && _pop_upl) {
      op_check(gsbn_pop_update_batch_wave(_pop_dev.p, _wg_pop.p, _wg_pop_hcu.p, _nwg_pop, _pop_upl,
                                          &st, stream),
               "gsbn_pop_update_batch_wave");
      if (ovf) _ovf_exec = _exec;
    }
    else if (_batch)
      op_check(gsbn_pop_update_batch(_pop_dev.p, (int32_t)_pop_args.size(), _wg_pop.p,
                                     _wg_pop_hcu.p, _nwg_pop, _pop_threads, _max_slots, _any_quirk,
                                     &st, stream),
               "gsbn_pop_update_batch");
    else
      op_check(gsbn_pop_update(&pa, stream), "gsbn_pop_update");
    Timing::get().end(0, stream);
    for (auto* p : _list_pop) p->flush_pending = true;
    if (sweep) {
      for (auto* q : _list_proj) {
        q->update_all(simstep, old_prn);
        sweep_synapses += (int64_t)q->_dim_hcu * q->_dim_conn * q->_dim_mcu;
      }
      sweeps++;
      for (auto* q : _list_proj) q->update_jxx(prn);
    }
    if (!_list_proj.empty()) {
      Timing::get().begin(1, stream);
      if (_batch && plan <= -21 && plan >= -23) {
        const int g = -20 - plan;
        if (!_gmap_n[g]) build_graded_map(g);
        op_check(gsbn_proj_rowcol_graded_wave(_proj_dev.p, _gmap_proj[g].p, _gmap_hcu[g].p,
                                              _gmap_part[g].p, _gmap_n[g], &st, stream),
                 "gsbn_proj_rowcol_graded_wave");
      } else if (_batch && plan == -20) {
        op_check(gsbn_proj_rowcol_batch_wave(_proj_dev.p, _wg_proj.p, _wg_proj_hcu.p, _nhcu_proj,
                                             &st, stream),
                 "gsbn_proj_rowcol_batch_wave");
      } else if (_batch && plan <= -10) {
        const int g = -10 - plan;
        if (!_gmap_n[g]) build_graded_map(g);
        op_check(gsbn_proj_rowcol_graded(_proj_dev.p, _gmap_proj[g].p, _gmap_hcu[g].p,
                                         _gmap_part[g].p, _gmap_n[g], &st, stream),
                 "gsbn_proj_rowcol_graded");
      } else if (_batch)
        op_check(gsbn_proj_rowcol_batch(_proj_dev.p, _wg_proj.p, _wg_proj_hcu.p, _nhcu_proj, plan,
                                        &st, stream),
                 "gsbn_proj_rowcol_batch");
      else
        op_check(gsbn_proj_rowcol(&ra, stream), "gsbn_proj_rowcol");
      Timing::get().end(1, stream);
    }
  }
  for (auto* q : _list_proj) q->update_que(_exec);
  for (auto* q : _list_proj)
    if (q->dK) {
      q->kick_t = simstep;
      q->kick_kp = q->_taupdt * prn;
      q->kicks_pending = true;
    }
  if (timed && ++_run_len == kTuneRun) {
    hip_check(hipEventRecord(_tev[ev + 1], stream), "hipEventRecord");
    _run_len = 0;
    if (++_tune_pos == (int)(_cand.size() * kTuneSamples)) finish_tuning();
  }
  if (touched)
    for (auto* q : _list_proj) q->mark_views_device();
  _epoch_seen = VecBase::host_epoch();
  _tables_seen = _tables.version;
  _exec++;
  execs++;
}

/* counters: 0 row synapse updates, 1 column synapse updates */
int64_t ProcUpdLazy::work(int i) {
  int64_t total = 0;
  for (auto* p : _list_pop) {
    if (!p->work.p) continue;
    vector<int64_t> w(2 * GSBN_MAX_PROJ_IN);
    p->work.download(w.data(), w.size());
    for (size_t m = 0; m < p->_in.size(); ++m) total += w[2 * m + i];
  }
  return total;
}

}  // namespace proc_upd_lazy

/* accessors used by the C ABI */
bool updlazy_work(Upd& upd, int64_t* out5) {
  for (auto* p : upd.procedures()) {
    auto* u = dynamic_cast<proc_upd_lazy::ProcUpdLazy*>(p);
    if (u) {
      out5[0] = u->work(0);
      out5[1] = u->work(1);
      out5[2] = u->sweep_synapses;
      out5[3] = u->execs;
      out5[4] = u->sweeps;
      return true;
    }
  }
  return false;
}

}  // namespace gsbn
