/*
 * structural_plasticity.cpp -- ProcStructuralPlasticity
 * (ProcStructuralPlasitcity.cpp:8-576, ProcStructuralPlasticity/Pop.cpp:5-95,
 * ProcStructuralPlasticity/Prj.cpp:5-138): pruning of unused connections and
 * growth of new ones between active source units and destination
 * hypercolumns with free slots.
 *
 * Steps executed with plasticity on:
 *   every `pruning-period` steps   rows with ti < simstep - t-th, or with no
 *                                  weight >= wp or <= wn, are removed (ii = -1,
 *                                  di = 0; pi, ei, ti, pij, eij, zj2 = 0) and
 *                                  their source units regain one fan-out each
 *   every `period` steps           each source population hands its active
 *                                  units (spike count over the period above
 *                                  int(0.2 * period * 100 * dt), counters then
 *                                  reset) to destination hypercolumns whose
 *                                  e-trace sum exceeds 0.5, in random order,
 *                                  filling free slots, one fan-out per new row
 *   at init with `enable-init-conn` the same growth with every unit and every
 *                                  hypercolumn, no thresholds
 *
 * The row test and the state reset run on the device (ProjRowOps of the
 * projection's owner); the list work is host code, as in the reference.
 *
 * Differences from the reference, by design:
 *   - draws: Philox4x32-10 keyed by argi seed-lo/seed-hi, counter (block,
 *     simstep, lane, 1), instead of random_device-seeded std::mt19937 and
 *     the Random class (ProcStructuralPlasitcity.cpp:186-218,461,578-608);
 *     shuffles are Fisher-Yates from the last element down with
 *     j = (word * (i + 1)) >> 32, a projection is picked as
 *     int(u * count) with u = (word >> 8) * 2^-24;
 *   - a destination hypercolumn counts as having a free slot when one of its
 *     own slots is free (Prj.cpp:100 tests the slots of hypercolumn 0 for
 *     every hypercolumn);
 *   - a removed row leaves the delay queue (in the reference its qi/siq stay
 *     frozen, Proj.cpp:651-653, so a row removed while active keeps being
 *     updated every step);
 *   - delays from the geometry are at least 1 and at most 31 steps;
 *   - the fan-out counters are saved in snapshots (as fanout_<pop>) and
 *     resume does not grow the initial connections a second time.
 */
#include <algorithm>
#include <cmath>

#include "../../include/gsbn_amd.h"
#include "detmath.h"
#include "framework.h"
#include "philox.h"

namespace gsbn {
namespace proc_structural_plasticity {

/* the draws of one (step, lane): Philox blocks consumed word by word */
class Draws {
 public:
  Draws(uint64_t seed, uint32_t step, uint32_t lane) : _seed(seed), _step(step), _lane(lane) {}
  uint32_t word() {
    if (_pos == 4) refill();
    return _w.v[_pos++];
  }
  float uniform() { return (float)(word() >> 8) * 5.9604644775390625e-08f; }
  /* integer in [0, n) */
  uint32_t below(uint32_t n) { return (uint32_t)(((uint64_t)word() * (uint64_t)n) >> 32); }
  template <typename T>
  void shuffle(vector<T>& v) {
    for (size_t i = v.size(); i > 1; --i) std::swap(v[i - 1], v[below((uint32_t)i)]);
  }
  /* a normal of mean m and deviation s from one fresh block: Box-Muller of
   * words 1 and 2 in double, rounded to float (as the population draws) */
  float normal(float m, float s) {
    refill();
    _pos = 4;
    const double inv24 = 5.9604644775390625e-08;
    const double u1 = (double)((_w.v[1] >> 8) + 1u) * inv24;
    const double u2 = (double)(_w.v[2] >> 8) * inv24;
    const float z = (float)(gsbn_sqrt(-2.0 * gsbn_log(u1)) * gsbn_cos2pi(u2));
    return m + s * z;
  }

 private:
  void refill() {
    gsbn_u32x4 c;
    c.v[0] = _block++;
    c.v[1] = _step;
    c.v[2] = _lane;
    c.v[3] = 1u; /* population draws use 0 here */
    _w = gsbn_philox4x32_10(c, (uint32_t)_seed, (uint32_t)(_seed >> 32));
    _pos = 0;
  }
  uint64_t _seed;
  uint32_t _step, _lane, _block = 0;
  gsbn_u32x4 _w{};
  int _pos = 4;
};

enum { LANE_GROW = 0, LANE_DELAY = 1, LANE_INIT = 2, LANE_INIT_DELAY = 3 };

/* a saved alias of a host-only vector (the snapshot sees it under a name
 * without the leading '.') */
struct AliasView : VecView {
  SyncVector<int32_t>* target;
  explicit AliasView(SyncVector<int32_t>* t) : target(t) {}
  void pull(VecBase* v) override {
    const vector<int32_t>* src = target->cpu_vector();
    auto& dst = static_cast<SyncVector<int32_t>*>(v)->raw_host();
    dst = *src;
  }
  void push(VecBase* v) override {
    auto& src = static_cast<SyncVector<int32_t>*>(v)->raw_host();
    vector<int32_t>* dst = target->mutable_cpu_vector();
    if (src.size() != dst->size()) throw Error(-2, v->name + ": size mismatch");
    *dst = src;
  }
};

struct Pop {
  int id, rank, H, M, hcu_start, mcu_start;
  vector<uint32_t> shape;
  vector<int32_t> position;
  SyncVector<int32_t>* counter = nullptr;
  SyncVector<int32_t>* fanout = nullptr;
  SyncVector<int32_t>* saved = nullptr; /* fanout_<id>, the snapshot's alias */
  vector<int> avail_prj; /* projections with this population as source */

  /* Pop::get_avail_active_mcu_list (Pop.cpp:52-63): units with a count above
   * the threshold, ascending; every count is reset */
  vector<int> active_units(int threshold) {
    vector<int> v;
    int32_t* c = counter->mutable_cpu_data();
    for (int i = 0; i < H * M; i++) {
      if (c[i] > threshold) v.push_back(i);
      c[i] = 0;
    }
    return v;
  }
  /* Pop::hcu_coor (Pop.cpp:65-93) */
  vector<int> hcu_coor(int hcu) const {
    vector<int> coor(shape.size());
    int index = hcu;
    for (size_t i = 0; i < shape.size(); i++) {
      const int d = (int)shape[i];
      GSBN_CHECK(d > 0, "pop_param shape > 0");
      coor[i] = index % d;
      index /= d;
    }
    for (size_t i = 0; i < position.size(); i++) {
      if (i < coor.size())
        coor[i] += position[i];
      else
        coor.push_back(position[i]);
    }
    return coor;
  }
};

struct Prj {
  int id, src, dest, H, M, C;
  SyncVector<int32_t>* ii = nullptr;
  SyncVector<int32_t>* di = nullptr;
  SyncVector<float>* ej = nullptr;
  ProjRowOps* ops = nullptr;
  vector<int32_t> local;    /* proposed ii: current rows plus new ones */
  vector<int> avail_hcu;    /* destination hypercolumns open to growth */

  /* Prj::get_avail_active_hcu_list (Prj.cpp:84-110): e-trace sum above the
   * threshold and a free slot of the hypercolumn's own */
  void open_hcus(float threshold) {
    avail_hcu.clear();
    const float* e = ej->cpu_data();
    const int32_t* p = ii->cpu_data();
    for (int h = 0; h < H; h++) {
      float acc = 0;
      for (int j = 0; j < M; j++) acc += e[h * M + j];
      bool free_slot = false;
      for (int k = 0; k < C; k++)
        if (p[(int64_t)h * C + k] < 0) {
          free_slot = true;
          break;
        }
      if (acc > threshold && free_slot) avail_hcu.push_back(h);
    }
  }
};

class ProcStructuralPlasticity : public ProcedureBase {
  REGISTER(ProcStructuralPlasticity)
 public:
  void init_new(SolverParam solver_param, Database& db) override { init(solver_param, db, false); }
  void init_copy(SolverParam solver_param, Database& db) override { init(solver_param, db, true); }
  void update_cpu() override { update_gpu(); }
  void update_gpu() override;

 private:
  void init(SolverParam solver_param, Database& db, bool copy);
  void prune(int simstep);
  void grow(uint32_t step, uint32_t lane, uint32_t delay_lane, int mcu_threshold,
            float hcu_threshold);
  int delay_cycle(Draws& d, const Prj& prj, int src_unit, int dest_hcu);
  GlobalVar _glv;
  float _dt = 0, _d_norm = 0.75f, _v_cond = 0.2f, _wp = 0.5f, _wn = -0.5f;
  int _t_th = 100, _period = 100, _pruning_period = 1000, _enable_geometry = 0,
      _enable_init_conn = 0;
  uint64_t _seed = 0;
  vector<Pop> _pops;
  vector<Prj> _prjs;
  vector<std::unique_ptr<VecView>> _views;
};
REGISTERIMPL(ProcStructuralPlasticity);

void ProcStructuralPlasticity::init(SolverParam solver_param, Database& db, bool copy) {
  NetParam net_param = solver_param.net_param;
  GSBN_CHECK(_glv.getf("dt", _dt), "dt");
  int rank = 0;
  GSBN_CHECK(_glv.geti("rank", rank), "rank");
  Parser par(get_proc_param(solver_param));
  /* ProcStructuralPlasitcity.cpp:17-66: defaults and corrections */
  if (par.argf("d-norm", _d_norm) && _d_norm < 0) _d_norm = 0;
  if (par.argf("v-cond", _v_cond) && _v_cond <= 0) _v_cond = 1;
  if (par.argi("t-th", _t_th) && _t_th <= 0) _t_th = 100;
  if (par.argi("period", _period) && _period <= 0) _period = 100;
  if (par.argi("pruning-period", _pruning_period) && _pruning_period <= 0) _pruning_period = 1000;
  if (par.argi("enable-geometry", _enable_geometry) && _enable_geometry != 0) _enable_geometry = 1;
  if (par.argi("enable-init-conn", _enable_init_conn) && _enable_init_conn != 0)
    _enable_init_conn = 1;
  par.argf("wp", _wp);
  par.argf("wn", _wn);
  int lo = 0, hi = 0;
  par.argi("seed-lo", lo);
  par.argi("seed-hi", hi);
  _seed = (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 31);

  int hcu_cnt = 0, mcu_cnt = 0;
  for (auto& pp : net_param.pop_param) {
    for (uint32_t k = 0; k < pp.pop_num; ++k) {
      Pop p;
      p.id = (int)_pops.size();
      p.rank = (int)pp.rank;
      p.H = (int)pp.hcu_num;
      p.M = (int)pp.mcu_num;
      p.hcu_start = hcu_cnt;
      p.mcu_start = mcu_cnt;
      p.shape = pp.shape;
      p.position = pp.position;
      hcu_cnt += p.H;
      mcu_cnt += p.H * p.M;
      const string sid = std::to_string(p.id);
      p.counter = db.sync_vector_i32(".counter_" + sid);
      GSBN_CHECK(p.counter, ".counter_" + sid + " (ProcUpdLazy must come first)");
      /* Pop.cpp:34-40 */
      p.fanout = db.create_sync_vector_i32(".fanout_" + sid);
      GSBN_CHECK(p.fanout, ".fanout_" + sid);
      p.fanout->resize((size_t)p.H * p.M, (int32_t)pp.fanout_num);
      /* saved alias: restored counters on resume */
      SyncVector<int32_t>* saved =
          copy && db.restored("fanout_" + sid) ? db.get<int32_t>("fanout_" + sid) : nullptr;
      if (saved) {
        if (saved->size() != p.fanout->size())
          throw Error(-2, "snapshot vector fanout_" + sid + " has the wrong size");
        *p.fanout->mutable_cpu_vector() = *saved->cpu_vector();
      } else {
        saved = db.create_sync_vector_i32("fanout_" + sid);
        GSBN_CHECK(saved, "fanout_" + sid);
      }
      _views.emplace_back(new AliasView(p.fanout));
      p.saved = saved;
      saved->view = _views.back().get();
      saved->set_logical_size(p.fanout->size());
      saved->status = VecBase::DEVICE_NEWER;
      _pops.push_back(p);
    }
  }
  for (auto& jp : net_param.proj_param) {
    if ((int)jp.src_pop >= (int)_pops.size() || (int)jp.dest_pop >= (int)_pops.size()) continue;
    Prj q;
    q.id = (int)_prjs.size();
    q.src = (int)jp.src_pop;
    q.dest = (int)jp.dest_pop;
    q.H = _pops[q.dest].H;
    q.M = _pops[q.dest].M;
    q.C = _pops[q.src].H * _pops[q.src].M;
    if (q.C > jp.slot_num) q.C = (int)jp.slot_num;
    const string sid = std::to_string(q.id);
    q.ii = db.sync_vector_i32("ii_" + sid);
    q.di = db.sync_vector_i32("di_" + sid);
    q.ej = db.sync_vector_f32("ej_" + sid);
    GSBN_CHECK(q.ii && q.di && q.ej, "ii_/di_/ej_" + sid + " (ProcUpdLazy must come first)");
    auto it = proj_row_ops().find(q.id);
    if (it == proj_row_ops().end())
      throw Error(-5, "ProcStructuralPlasticity: projection " + sid +
                          " has no device row operations (ProcUpdLazy must come first)");
    q.ops = it->second;
    _pops[q.src].avail_prj.push_back(q.id);
    _prjs.push_back(q);
  }
  /* ProcStructuralPlasitcity.cpp:101-103 (init_conn); not again on resume */
  bool restored_conn = false;
  for (auto& q : _prjs) restored_conn = restored_conn || db.restored("ii_" + std::to_string(q.id));
  if (_enable_init_conn && !(copy && restored_conn)) grow(0, LANE_INIT, LANE_INIT_DELAY, -1, -1.0f);
}

void ProcStructuralPlasticity::update_gpu() {
  int simstep = 0;
  GSBN_CHECK(_glv.geti("simstep", simstep), "simstep");
  if (simstep != 0) {
    int cycle_flag = 0;
    GSBN_CHECK(_glv.geti("cycle-flag", cycle_flag), "cycle-flag");
    if (cycle_flag != 1) return;
  }
  bool plasticity = false;
  GSBN_CHECK(_glv.getb("plasticity", plasticity), "plasticity");
  if (!plasticity && simstep != 0) return;
  const bool do_prune = simstep % _pruning_period == 0, do_grow = simstep % _period == 0;
  if (!do_prune && !do_grow) return;
  /* host writes of fanout_<id> reach the counters */
  for (auto& p : _pops)
    if (p.saved->status == VecBase::HOST_NEWER) {
      p.saved->view->push(p.saved);
      p.saved->status = VecBase::SYNCED;
    }
  if (do_prune) prune(simstep);
  if (do_grow) {
    const int maxfq = 100;
    grow((uint32_t)simstep, LANE_GROW, LANE_DELAY, (int)(0.2 * _period * maxfq * _dt), 0.5f);
  }
  for (auto& p : _pops) p.saved->status = VecBase::DEVICE_NEWER;
}

/* STAGE 1-2 (ProcStructuralPlasitcity.cpp:280-352) */
void ProcStructuralPlasticity::prune(int simstep) {
  GSBN_CHECK(_wp > _wn, "w+ should be larger than w- !!");
  for (auto& q : _prjs) {
    vector<uint8_t> flags;
    q.ops->prune_flags(simstep - _t_th, _wp, _wn, &flags);
    vector<int32_t> rows;
    for (size_t r = 0; r < flags.size(); ++r)
      if (flags[r]) rows.push_back((int32_t)r);
    if (rows.empty()) continue;
    q.ops->clear_rows(rows);
    int32_t* ii = q.ii->mutable_cpu_data();
    int32_t* di = q.di->mutable_cpu_data();
    int32_t* fan = _pops[q.src].fanout->mutable_cpu_data();
    for (int32_t r : rows) {
      fan[ii[r]] += 1;
      ii[r] = -1;
      di[r] = 0;
    }
  }
}

/* Pop hcu distance -> delay (ProcStructuralPlasitcity.cpp:578-608) */
int ProcStructuralPlasticity::delay_cycle(Draws& d, const Prj& prj, int src_unit, int dest_hcu) {
  const Pop& p0 = _pops[prj.src];
  const Pop& p1 = _pops[prj.dest];
  vector<int> a = p0.hcu_coor(src_unit / p0.M);
  vector<int> b = p1.hcu_coor(dest_hcu);
  while (a.size() < b.size()) a.push_back(0);
  while (b.size() < a.size()) b.push_back(0);
  float d2 = 0;
  for (size_t i = 0; i < a.size(); i++) d2 += (float)((a[i] - b[i]) * (a[i] - b[i]));
  const float distance = std::sqrt(d2);
  const float mean = _d_norm * distance / _v_cond + 1;
  const float tij = d.normal(mean, 0.1f * mean);
  const int delay = (int)std::ceil((double)tij * 0.001 / (double)_dt);
  return std::min(31, std::max(1, delay));
}

/* STAGE 3-5 (ProcStructuralPlasitcity.cpp:354-576) and init_conn (:109-258) */
void ProcStructuralPlasticity::grow(uint32_t step, uint32_t lane, uint32_t delay_lane,
                                    int mcu_threshold, float hcu_threshold) {
  for (auto& q : _prjs) {
    q.local = *q.ii->cpu_vector();
    q.open_hcus(hcu_threshold);
  }
  Draws draws(_seed, step, lane);
  for (auto& pop : _pops) {
    int32_t* fan = pop.fanout->mutable_cpu_data();
    vector<int> prjs = pop.avail_prj;
    vector<int> active = pop.active_units(mcu_threshold);
    vector<uint8_t> taken((size_t)pop.H * pop.M);
    while (!prjs.empty() && !active.empty()) {
      size_t idx = (size_t)(int)(draws.uniform() * (float)prjs.size());
      if (idx >= prjs.size()) idx = prjs.size() - 1;
      Prj& q = _prjs[prjs[idx]];
      vector<int> hcus = q.avail_hcu;
      if (hcus.empty()) break;
      draws.shuffle(hcus);
      for (int h : hcus) {
        int32_t* row = q.local.data() + (int64_t)h * q.C;
        for (int k = 0; k < q.C; k++)
          if (row[k] >= 0) taken[row[k]] = 1;
        vector<int> cand;
        for (int u : active)
          if (!taken[u]) cand.push_back(u);
        for (int k = 0; k < q.C; k++)
          if (row[k] >= 0) taken[row[k]] = 0;
        if (cand.empty()) continue;
        draws.shuffle(cand);
        size_t l = 0;
        for (int k = 0; k < q.C && l < cand.size(); k++)
          if (row[k] < 0) row[k] = cand[l++];
        for (size_t k = 0; k < l; k++) {
          const int u = cand[k];
          if (--fan[u] <= 0) active.erase(std::remove(active.begin(), active.end(), u), active.end());
        }
      }
      prjs.erase(prjs.begin() + idx);
    }
  }
  /* STAGE 5: the new rows.  Growth fills a hypercolumn's free slots from the
   * lowest one up, so add_row's first free slot (ProcStructuralPlasitcity.cpp:
   * 566-576) is the slot the new row was proposed for. */
  Draws delays(_seed, step, delay_lane);
  for (auto& q : _prjs) {
    const int32_t* cur = q.ii->cpu_data();
    bool any = false;
    for (size_t i = 0; i < q.local.size(); i++)
      if (cur[i] < 0 && q.local[i] >= 0) any = true;
    if (!any) continue;
    int32_t* ii = q.ii->mutable_cpu_data();
    int32_t* di = q.di->mutable_cpu_data();
    for (int h = 0; h < q.H; h++)
      for (int k = 0; k < q.C; k++) {
        const int64_t index = (int64_t)h * q.C + k;
        if (ii[index] >= 0 || q.local[index] < 0) continue;
        ii[index] = q.local[index];
        di[index] = _enable_geometry ? delay_cycle(delays, q, q.local[index], h) : 1;
      }
    q.ops->rows_added();
  }
}

}  // namespace proc_structural_plasticity
}  // namespace gsbn
