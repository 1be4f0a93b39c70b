/*
 * procedures.cpp -- the procedures around ProcUpdLazy on its data path, with
 * the reference's names, Database/GlobalVar contracts and error cases:
 *
 *   ProcExtGen         simulation clock, mode schedule, .lginp/.wmask/.lgidx
 *                      (ProcExtGen.cpp:8-153)
 *   ProcInitConn       ii/di from a connection map (ProcInitConn.cpp:8-103)
 *   ProcExchangeSpike  si_<proj> <- spike_<src> once per executed step; one
 *                      process, so the exchange is a device copy
 *                      (ProcExchangeSpike.cpp:94-135)
 *   ProcCheck          recall scoring (ProcCheck.cpp:8-237,
 *                      ProcCheck/Pop.cpp:35-60); counting and the per-HCU
 *                      argmax run on the device
 *   ProcSnapshot       SolverState file of every saved vector each `period`
 *                      steps from `offset` (ProcSnapshot.cpp:8-80)
 *   ProcSpkRec         spike raster and active-row/unit CSV files
 *                      (ProcSpkRec.cpp:10-210, ProcSpkRec/Pop.cpp:30-52,
 *                      ProcSpkRec/Prj.cpp:95-137)
 */
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "../../include/gsbn_amd.h"
#include "detmath.h"
#include "framework.h"

namespace gsbn {

/* ------------------------------------------------------------------ */
class ProcExtGen : public ProcedureBase {
  REGISTER(ProcExtGen)
 public:
  struct mode_t {
    int begin_step, end_step, lgidx_id, lgexp_id, wmask_id;
    float prn;
    int plasticity;
  };
  void init_new(SolverParam solver_param, Database& db) override;
  void init_copy(SolverParam solver_param, Database& db) override { init_new(solver_param, db); }
  void update_cpu() override;
  void update_gpu() override { update_cpu(); }

 private:
  GlobalVar _glv;
  vector<mode_t> _list_mode;
  SyncVector<int32_t>* _lgidx = nullptr;
  SyncVector<float>* _wmask = nullptr;
  SyncVector<float>* _lginp = nullptr;
  vector<int> _mcu_in_hcu;
  float _eps = 0;
  int _cursor = 0;
  /* every row's input resident in .lginp (selected by lginp-idx), or one
   * row rebuilt per mode change when that would exceed kResidentMax floats */
  static constexpr int64_t kResidentMax = (int64_t)1 << 28; /* 1 GiB */
  bool _resident = true;
  int _built_row = -1;
  void build_row(const uint32_t* lg, float* out) const;
};
REGISTERIMPL(ProcExtGen);

void ProcExtGen::init_new(SolverParam solver_param, Database& db) {
  GenParam gen_param = solver_param.gen_param;
  _eps = gen_param.eps;
  float dt = gen_param.dt;
  _glv.putf("dt", dt);
  int max_step = -1;
  int last_end = 0;
  for (auto& mp : gen_param.mode_param) {
    int begin_step = int(mp.begin_time / dt);
    int end_step = int(mp.end_time / dt);
    if (begin_step < max_step)
      throw Error(-2, "Order of modes is wrong or there is overlapping time range, abort!");
    if (end_step < begin_step) throw Error(-2, "Time range is wrong, abort!");
    int lg = mp.begin_lgidx_id, lx = mp.begin_lgexp_id, wm = mp.begin_wmask_id;
    while (begin_step < end_step) {
      mode_t m;
      m.begin_step = begin_step;
      m.end_step = begin_step + (int)mp.time_step;
      if (m.end_step >= end_step) m.end_step = end_step;
      m.lgidx_id = lg;
      m.lgexp_id = lx;
      m.wmask_id = wm;
      lg += mp.lgidx_step;
      lx += mp.lgexp_step;
      wm += mp.wmask_step;
      m.prn = mp.prn;
      m.plasticity = mp.plasticity;
      _list_mode.push_back(m);
      begin_step = m.end_step;
    }
    if (end_step > last_end) last_end = end_step;
  }
  /* horizon of the schedule, used to size the coefficient tables */
  _glv.puti("sim-end-step", last_end);

  _lgidx = db.create_sync_vector_i32(".lgidx");
  _wmask = db.create_sync_vector_f32(".wmask");
  _lginp = db.create_sync_vector_f32(".lginp");
  GSBN_CHECK(_lgidx && _wmask && _lginp, "ProcExtGen vectors already exist");

  StimRawData stim = read_stim_raw_data(gen_param.stim_file);
  if (stim.data.size() != (size_t)stim.data_rows * stim.data_cols)
    throw Error(-2, "Bad stimuli!!! (data size)");
  if (stim.mask.size() != (size_t)stim.mask_rows * stim.mask_cols)
    throw Error(-2, "Bad stimuli!!! (mask size)");
  _lgidx->set_ld(stim.data_cols);
  _wmask->set_ld(stim.mask_cols);
  auto* vd = _lgidx->mutable_cpu_vector();
  vd->clear();
  for (uint32_t v : stim.data) vd->push_back((int32_t)v);
  _lgidx->resize(vd->size());
  auto* vm = _wmask->mutable_cpu_vector();
  vm->clear();
  for (float v : stim.mask) vm->push_back(v);
  _wmask->resize(vm->size());

  int mcu_num = 0;
  for (auto& pp : solver_param.net_param.pop_param) {
    mcu_num += pp.hcu_num * pp.mcu_num * pp.pop_num;
    for (uint32_t k = 0; k < pp.pop_num; ++k)
      for (uint32_t j = 0; j < pp.hcu_num; ++j) _mcu_in_hcu.push_back(pp.mcu_num);
  }
  /* every stimulus row's input (ProcExtGen.cpp:141-150: log(1 + eps) for
   * the MCU the row names in each HCU, log(eps) elsewhere), built once: a
   * mode selects its row with lginp-idx, as .wmask's rows are selected by
   * wmask-idx, so a mode change writes nothing to the device */
  const int nrows = stim.data_rows;
  if (nrows > 0 && stim.data_cols < (int)_mcu_in_hcu.size())
    throw Error(-2, "Bad stimuli!!! (fewer columns than hypercolumns)");
  /* beyond kResidentMax floats the rows are built one at a time instead, at
   * each mode change, into a one-row .lginp (the reference's way) */
  _resident = (int64_t)std::max(nrows, 1) * mcu_num <= kResidentMax;
  /* argi resident-stimulus 0 asks for the one-row form at any size */
  for (auto& pp : solver_param.proc_param)
    if (pp.name == proc_name()) {
      int32_t v = 1;
      if (Parser(pp).argi("resident-stimulus", v) && v == 0) _resident = false;
    }
  auto* vl = _lginp->mutable_cpu_vector();
  vl->assign((size_t)(_resident ? std::max(nrows, 1) : 1) * mcu_num, 0.0f);
  if (_resident)
    for (int r = 0; r < nrows; ++r) build_row(stim.data.data() + (size_t)r * stim.data_cols,
                                             vl->data() + (size_t)r * mcu_num);
  _lginp->resize(vl->size());
  _lginp->set_ld(mcu_num);
  _cursor = 0;
  _built_row = -1;
}

void ProcExtGen::build_row(const uint32_t* lg, float* out) const {
  const float on = gsbn_logf(1 + _eps), off = gsbn_logf(0 + _eps);
  for (size_t i = 0; i < _mcu_in_hcu.size(); ++i)
    for (int j = 0; j < _mcu_in_hcu[i]; ++j) *out++ = ((int32_t)lg[i] == j) ? on : off;
}

void ProcExtGen::update_cpu() {
  int current_step = 0;
  GSBN_CHECK(_glv.geti("simstep", current_step), "simstep");
  current_step++;
  _glv.puti("simstep", current_step);
  int r = (int)_list_mode.size();
  for (; _cursor < r; _cursor++) {
    mode_t m = _list_mode[_cursor];
    if (current_step > m.begin_step && current_step <= m.end_step) {
      float old_prn = 0;
      GSBN_CHECK(_glv.getf("prn", old_prn), "prn");
      _glv.putf("old-prn", old_prn);
      _glv.putf("prn", m.prn);
      _glv.puti("wmask-idx", m.wmask_id);
      _glv.putb("plasticity", (bool)m.plasticity);
      if (m.lgidx_id < 0 || (size_t)(m.lgidx_id + 1) * _lgidx->ld() > _lgidx->size())
        throw Error(-2, "stimulus row " + std::to_string(m.lgidx_id) + " out of range");
      if (_resident) {
        _glv.puti("lginp-idx", m.lgidx_id);
      } else {
        if (_built_row != m.lgidx_id) {
          _built_row = m.lgidx_id;
          build_row(reinterpret_cast<const uint32_t*>(_lgidx->cpu_data(m.lgidx_id)),
                    _lginp->mutable_cpu_data());
        }
        _glv.puti("lginp-idx", 0);
      }
      _glv.puti("cycle-flag", 1);
      return;
    } else if (current_step <= m.begin_step) {
      _glv.puti("cycle-flag", 0);
      return;
    }
  }
  _glv.puti("cycle-flag", -1);
}

/* ------------------------------------------------------------------ */
class ProcInitConn : public ProcedureBase {
  REGISTER(ProcInitConn)
 public:
  void init_new(SolverParam solver_param, Database& db) override;
  /* ProcInitConn.cpp:79-81 re-runs init_new, which adds the map's rows again
   * to the free slots of the restored ii_ (duplicating every row of a
   * sparse map); a snapshot's ii_/di_ are kept as saved instead */
  void init_copy(SolverParam solver_param, Database& db) override {
    for (size_t i = 0; i < solver_param.net_param.proj_param.size(); ++i)
      if (db.restored("ii_" + std::to_string(i))) return;
    init_new(solver_param, db);
  }
  void update_cpu() override {}
  void update_gpu() override {}
};
REGISTERIMPL(ProcInitConn);

void ProcInitConn::init_new(SolverParam solver_param, Database& db) {
  struct Prj {
    int dim_conn = 0, dim_hcu = 0, nsrc = 0;
    SyncVector<int32_t>* ii = nullptr;
    SyncVector<int32_t>* di = nullptr;
    vector<int> cursor; /* per destination HCU: no free slot below this index */
  };
  vector<std::pair<int, int>> pops;  // hcu, mcu
  for (auto& pp : solver_param.net_param.pop_param)
    for (uint32_t k = 0; k < pp.pop_num; ++k) pops.push_back({(int)pp.hcu_num, (int)pp.mcu_num});
  vector<Prj> prjs;
  for (auto& jp : solver_param.net_param.proj_param) {
    if (jp.src_pop >= pops.size() || jp.dest_pop >= pops.size()) continue;
    Prj p;
    p.dim_hcu = pops[jp.dest_pop].first;
    p.nsrc = pops[jp.src_pop].first * pops[jp.src_pop].second;
    p.dim_conn = p.nsrc;
    if (p.dim_conn > jp.slot_num) p.dim_conn = (int)jp.slot_num;
    int id = (int)prjs.size();
    p.ii = db.sync_vector_i32("ii_" + std::to_string(id));
    p.di = db.sync_vector_i32("di_" + std::to_string(id));
    prjs.push_back(p);
  }
  Parser par(get_proc_param(solver_param));
  string conn_map_file;
  if (!par.args("conn-map-file", conn_map_file))
    throw Error(-2, "Connection map is not specified! Abort");
  std::ifstream f(conn_map_file);
  if (!f) throw Error(-2, "cannot open connection map " + conn_map_file);
  string line;
  while (std::getline(f, line)) {
    std::stringstream ss(line);
    vector<string> v;
    while (ss.good()) {
      string sub;
      std::getline(ss, sub, ',');
      v.push_back(sub);
    }
    if (v.size() <= 2) continue;
    int prj_id = std::stoi(v[0]);
    if (prj_id < 0 || prj_id >= (int)prjs.size()) throw Error(-2, "conn-map: bad projection id");
    Prj& p = prjs[prj_id];
    if (!p.ii || !p.di) throw Error(-4, "conn-map: ii/di vectors missing (ProcUpdLazy must come first)");
    int hcu_id = std::stoi(v[1]);
    if (hcu_id < 0 || hcu_id >= p.dim_hcu) throw Error(-2, "conn-map: bad destination hcu");
    if ((v.size() - 2) % 2) throw Error(-2, "conn-map: unpaired source/delay");
    int* pii = p.ii->mutable_cpu_data() + (int64_t)hcu_id * p.dim_conn;
    int* pdi = p.di->mutable_cpu_data() + (int64_t)hcu_id * p.dim_conn;
    /* ProcInitConn::add_row: first free slot (ii < 0) of the hypercolumn,
     * or nothing when all are taken.  Slots only go from free to taken
     * here, so the first free slot never moves backwards: a per-HCU cursor
     * finds it without rescanning the taken prefix. */
    if (p.cursor.empty()) p.cursor.assign(p.dim_hcu, 0);
    int& k = p.cursor[hcu_id];
    for (size_t i = 2; i + 1 < v.size(); i += 2) {
      int ii = std::stoi(v[i]);
      int di = std::stoi(v[i + 1]);
      if (ii < 0 || ii >= p.nsrc) throw Error(-2, "conn-map: source unit out of range");
      if (di < 1 || di > 31) throw Error(-2, "conn-map: delay must be in [1, 31]");
      while (k < p.dim_conn && pii[k] >= 0) ++k;
      if (k < p.dim_conn) {
        pii[k] = ii;
        pdi[k] = di;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
class ProcExchangeSpike : public ProcedureBase {
  REGISTER(ProcExchangeSpike)
 public:
  void init_new(SolverParam solver_param, Database& db) override;
  void init_copy(SolverParam solver_param, Database& db) override { init_new(solver_param, db); }
  void update_cpu() override { update_gpu(); }
  void update_gpu() override;

 private:
  GlobalVar _glv;
  struct Link {
    SyncVector<int8_t>* sj;
    SyncVector<int8_t>* si;
  };
  vector<Link> _links;
  bool _direct = false;
};
REGISTERIMPL(ProcExchangeSpike);

void ProcExchangeSpike::init_new(SolverParam solver_param, Database& db) {
  int rank = 0, num_rank = 1;
  GSBN_CHECK(_glv.geti("rank", rank), "rank");
  GSBN_CHECK(_glv.geti("num-rank", num_rank), "num-rank");
  if (num_rank != 1) throw Error(-5, "ProcExchangeSpike: one process only in this build");
  int npop = 0;
  for (auto& pp : solver_param.net_param.pop_param) npop += pp.pop_num;
  int id = 0;
  for (auto& jp : solver_param.net_param.proj_param) {
    if ((int)jp.src_pop >= npop || (int)jp.dest_pop >= npop) continue;
    Link l;
    l.sj = db.sync_vector_i8("spike_" + std::to_string(jp.src_pop));
    l.si = db.sync_vector_i8("si_" + std::to_string(id));
    GSBN_CHECK(l.sj && l.si, "ProcExchangeSpike: spike/si vectors missing");
    GSBN_CHECK(l.si->size() == l.sj->size(), "ProcExchangeSpike: si size");
    _links.push_back(l);
    ++id;
  }
  /* initialised after ProcUpdLazy: every executed step delivers the spikes
   * of that step to si after the lazy update, so the next step's queue push
   * sees them; ProcUpdLazy then takes the push from its own spike history */
  int ready = 0, allow = 1;
  Parser par(get_proc_param(solver_param));
  par.argi("direct-history", allow); /* 0: keep the copy and the separate queue push */
  if (allow && _glv.geti("updlazy-ready", ready) && ready) {
    _glv.puti("exchange-spike-direct", 1);
    _direct = true;
    /* si_<proj> is then only materialised when the host reads it: it equals
     * the source population's spike vector of the last executed step */
    for (auto& l : _links) {
      (void)l.si->mutable_gpu_data();
      SyncVector<int8_t>* si = l.si;
      SyncVector<int8_t>* sj = l.sj;
      si->before_host_read = [si, sj]() {
        if (si->size())
          hip_check(hipMemcpyAsync(si->mutable_gpu_data(), sj->gpu_data(), si->size(),
                                   hipMemcpyDeviceToDevice, active_stream()),
                    "exchange copy");
      };
    }
  }
}

void ProcExchangeSpike::update_gpu() {
  int cycle_flag = 0;
  GSBN_CHECK(_glv.geti("cycle-flag", cycle_flag), "cycle-flag");
  if (cycle_flag != 1) return;
  if (_direct) {
    for (auto& l : _links) l.si->mark_device();
    return;
  }
  for (auto& l : _links) {
    const int8_t* src = l.sj->gpu_data();
    int8_t* dst = l.si->mutable_gpu_data();
    if (l.sj->size())
      hip_check(hipMemcpyAsync(dst, src, l.sj->size(), hipMemcpyDeviceToDevice, active_stream()),
                "exchange copy");
  }
}

/* ------------------------------------------------------------------ */
/* the directory <log-dir><procedure name>, created when missing */
static string record_dir(GlobalVar& glv, const string& proc) {
  string log_dir;
  GSBN_CHECK(glv.gets("log-dir", log_dir) && !log_dir.empty(), "log-dir");
  const string dir = log_dir + proc;
  struct stat info;
  if (!(stat(dir.c_str(), &info) == 0 && (info.st_mode & S_IFDIR))) {
    const string cmd = "mkdir -p '" + dir + "'";
    if (system(cmd.c_str()) != 0) throw Error(-2, "Cannot create directory for state records!");
  }
  return dir;
}

/* offset (>= 0, default 0) and period (> 0, default 1) of a recorder */
static void record_schedule(Parser& par, int* offset, int* period) {
  if (!par.argi("offset", *offset)) *offset = 0;
  GSBN_CHECK(*offset >= 0, "offset >= 0");
  if (!par.argi("period", *period)) *period = 1;
  GSBN_CHECK(*period > 0, "period > 0");
}

class ProcSnapshot : public ProcedureBase {
  REGISTER(ProcSnapshot)
 public:
  void init_new(SolverParam solver_param, Database& db) override;
  void init_copy(SolverParam solver_param, Database& db) override { init_new(solver_param, db); }
  void update_cpu() override { update_gpu(); }
  void update_gpu() override;

 private:
  GlobalVar _glv;
  Database* _db = nullptr;
  string _directory;
  int _offset = 0, _period = 1;
};
REGISTERIMPL(ProcSnapshot);

void ProcSnapshot::init_new(SolverParam solver_param, Database& db) {
  _db = &db;
  _directory = record_dir(_glv, proc_name());
  Parser par(get_proc_param(solver_param));
  record_schedule(par, &_offset, &_period);
}

void ProcSnapshot::update_gpu() {
  int cycle_flag = 0, simstep = 0;
  float dt = 0, prn = 0;
  GSBN_CHECK(_glv.geti("cycle-flag", cycle_flag), "cycle-flag");
  if (cycle_flag != 1) return;
  GSBN_CHECK(_glv.geti("simstep", simstep), "simstep");
  GSBN_CHECK(_glv.getf("dt", dt), "dt");
  GSBN_CHECK(_glv.getf("prn", prn), "prn");
  GSBN_CHECK(simstep >= 0 && dt >= 0, "simstep, dt");
  if (simstep < _offset || simstep % _period != 0) return;
  const float t = simstep * dt;
  _db->save_state(_directory + "/Snapshot_" + std::to_string(t) + ".bin", t, prn);
}

/* ------------------------------------------------------------------ */
class ProcSpkRec : public ProcedureBase {
  REGISTER(ProcSpkRec)
 public:
  void init_new(SolverParam solver_param, Database& db) override;
  void init_copy(SolverParam solver_param, Database& db) override { init_new(solver_param, db); }
  void update_cpu() override { update_gpu(); }
  void update_gpu() override;

 private:
  struct PopRec {
    int id, H, M;
    SyncVector<int8_t>* spike;
    int buffer = 1; /* spike_ rows kept (ld > 0), the row of step t is t % buffer */
  };
  struct PrjRec {
    int id;
    SyncVector<int32_t>* ssi;
    SyncVector<int32_t>* ssj;
  };
  GlobalVar _glv;
  string _directory;
  int _offset = 0, _period = 1;
  vector<PopRec> _pops;
  vector<PrjRec> _prjs;
};
REGISTERIMPL(ProcSpkRec);

void ProcSpkRec::init_new(SolverParam solver_param, Database& db) {
  _directory = record_dir(_glv, proc_name());
  Parser par(get_proc_param(solver_param));
  record_schedule(par, &_offset, &_period);
  float dt = 0;
  GSBN_CHECK(_glv.getf("dt", dt), "dt");
  int id = 0;
  for (auto& pp : solver_param.net_param.pop_param) {
    for (uint32_t k = 0; k < pp.pop_num; ++k, ++id) {
      PopRec r;
      r.id = id;
      r.H = (int)pp.hcu_num;
      r.M = (int)pp.mcu_num;
      r.spike = db.sync_vector_i8("spike_" + std::to_string(id));
      GSBN_CHECK(r.spike, "spike_" + std::to_string(id) + " (ProcUpdLazy must come first)");
      if (r.spike->ld() > 0) r.buffer = (int)(r.spike->size() / r.spike->ld());
      _pops.push_back(r);
      std::fstream out(_directory + "/spk_pop_" + std::to_string(id) + ".csv",
                       std::ios::out | std::ofstream::trunc);
      out << r.H << "," << r.M << "," << dt << "," << pp.maxfq << std::endl;
    }
  }
  for (size_t i = 0; i < solver_param.net_param.proj_param.size(); ++i) {
    PrjRec r;
    r.id = (int)i;
    r.ssi = db.sync_vector_i32(".ssi_" + std::to_string(i));
    r.ssj = db.sync_vector_i32(".ssj_" + std::to_string(i));
    GSBN_CHECK(r.ssi && r.ssj, ".ssi_/.ssj_ of projection " + std::to_string(i));
    _prjs.push_back(r);
    for (const char* f : {"/ssi_pop_", "/ssj_pop_"})
      std::fstream(_directory + f + std::to_string(i) + ".csv", std::ios::out | std::ofstream::trunc);
  }
}

void ProcSpkRec::update_gpu() {
  int cycle_flag = 0, simstep = 0;
  GSBN_CHECK(_glv.geti("cycle-flag", cycle_flag), "cycle-flag");
  if (cycle_flag != 1) return;
  GSBN_CHECK(_glv.geti("simstep", simstep), "simstep");
  GSBN_CHECK(simstep >= 0, "simstep >= 0");
  if (simstep < _offset || simstep % _period != 0) return;
  for (auto& p : _pops) {
    /* Pop::record (ProcSpkRec/Pop.cpp:30-52): "step,unit,unit,..." when
     * the step has spikes */
    const int row = p.buffer > 1 ? simstep % p.buffer : 0;
    const int8_t* spk = p.spike->cpu_data() + (size_t)row * p.H * p.M;
    std::ostringstream line;
    bool any = false;
    for (int i = 0; i < p.H * p.M; i++)
      if (spk[i] > 0) {
        if (!any) line << simstep;
        any = true;
        line << "," << i;
      }
    if (any) {
      std::fstream out(_directory + "/spk_pop_" + std::to_string(p.id) + ".csv",
                       std::ios::out | std::ios::app);
      out << line.str() << std::endl;
    }
  }
  /* Prj::record_ssj / record_ssi (ProcSpkRec/Prj.cpp:95-137): "step,1"
   * when unit / row 9 is in the step's active list, else "step,0" */
  for (auto& r : _prjs) {
    const std::pair<const char*, SyncVector<int32_t>*> files[2] = {{"/ssj_pop_", r.ssj},
                                                                   {"/ssi_pop_", r.ssi}};
    for (auto& f : files) {
      const int32_t* d = f.second->cpu_data();
      bool active = false;
      for (size_t i = 0; i < f.second->size(); i++)
        if (d[i] == 9) {
          active = true;
          break;
        }
      std::fstream out(_directory + f.first + std::to_string(r.id) + ".csv",
                       std::ios::out | std::ios::app);
      out << simstep << "," << (active ? 1 : 0) << std::endl;
    }
  }
}

/* ------------------------------------------------------------------ */
class ProcCheck : public ProcedureBase {
  REGISTER(ProcCheck)
 public:
  ~ProcCheck() override;
  void init_new(SolverParam solver_param, Database& db) override;
  void init_copy(SolverParam solver_param, Database& db) override { init_new(solver_param, db); }
  void update_cpu() override { update_gpu(); }
  void update_gpu() override;
  void finish() override { report(false); }
  void results(int* correct, int* total, vector<int>* flags);

 private:
  struct mode_t {
    int begin_step, end_step, lgexp_id;
  };
  struct PopInfo {
    int H, M;
    SyncVector<int8_t>* spike;
    int32_t* count;
  };
  void report(bool final_line);
  GlobalVar _glv;
  vector<mode_t> _list_mode;
  SyncVector<int32_t>* _lgidx = nullptr;
  vector<PopInfo> _pops;
  int _total_hcu = 0;
  int _threshold = 0;
  int _cursor = 0;
  bool _updated_flag = false;
  string _logfile;
  int32_t* _res_idx = nullptr; /* [windows][total_hcu] */
  int32_t* _res_cnt = nullptr;
  vector<int> _window_mode;    /* mode index of each scored window */
  int _reported = 0;
  int _correct = 0, _total = 0;
  vector<int> _flags;
  vector<gsbn_check_item> _items;
  gsbn_check_item* _items_dev = nullptr;
  bool _final_done = false;
};
REGISTERIMPL(ProcCheck);

ProcCheck::~ProcCheck() {
  for (auto& p : _pops)
    if (p.count) (void)hipFree(p.count);
  if (_res_idx) (void)hipFree(_res_idx);
  if (_items_dev) (void)hipFree(_items_dev);
  if (_res_cnt) (void)hipFree(_res_cnt);
}

void ProcCheck::init_new(SolverParam solver_param, Database& db) {
  string log_dir;
  GSBN_CHECK(_glv.gets("log-dir", log_dir) && !log_dir.empty(), "log-dir");
  string dir = log_dir + proc_name();
  string cmd = "mkdir -p '" + dir + "'";
  if (system(cmd.c_str()) != 0) throw Error(-2, "Cannot create directory for state records!");
  _logfile = dir + "/check_result.txt";
  GenParam gen_param = solver_param.gen_param;
  float dt = gen_param.dt;
  int max_step = -1;
  for (auto& mp : gen_param.mode_param) {
    int begin_step = int(mp.begin_time / dt);
    int end_step = int(mp.end_time / dt);
    if (begin_step < max_step) throw Error(-2, "Order of modes is wrong");
    if (end_step < begin_step) throw Error(-2, "Time range is wrong, abort!");
    if (mp.prn != 0.0f) continue;
    int lx = mp.begin_lgexp_id;
    while (begin_step < end_step) {
      mode_t m;
      m.begin_step = begin_step;
      m.end_step = begin_step + (int)mp.time_step;
      m.lgexp_id = lx;
      lx += mp.lgexp_step;
      _list_mode.push_back(m);
      begin_step += mp.time_step;
    }
  }
  _lgidx = db.sync_vector_i32(".lgidx");
  GSBN_CHECK(_lgidx, ".lgidx missing (ProcExtGen must come first)");
  int pid = 0;
  for (auto& pp : solver_param.net_param.pop_param) {
    for (uint32_t k = 0; k < pp.pop_num; ++k) {
      PopInfo pi;
      pi.H = pp.hcu_num;
      pi.M = pp.mcu_num;
      pi.spike = db.sync_vector_i8("spike_" + std::to_string(pid));
      GSBN_CHECK(pi.spike, "spike vector missing");
      hip_check(hipMalloc(&pi.count, sizeof(int32_t) * pi.H * pi.M), "hipMalloc");
      op_check(gsbn_i32_fill(pi.count, (int64_t)pi.H * pi.M, 0, active_stream()), "gsbn_i32_fill");
      _total_hcu += pi.H;
      _pops.push_back(pi);
      ++pid;
    }
  }
  Parser par(get_proc_param(solver_param));
  if (!par.argi("threashold", _threshold)) _threshold = 0;
  size_t nwin = _list_mode.size() ? _list_mode.size() : 1;
  hip_check(hipMalloc(&_res_idx, sizeof(int32_t) * nwin * _total_hcu), "hipMalloc");
  hip_check(hipMalloc(&_res_cnt, sizeof(int32_t) * nwin * _total_hcu), "hipMalloc");
  std::fstream out(_logfile, std::ios::out | std::ofstream::trunc);
  out << "result | reference pattern | recorded pattern | recorded pattern spike count" << std::endl;
  _cursor = 0;
  _updated_flag = false;
}

void ProcCheck::update_gpu() {
  int cycle_flag = 0;
  GSBN_CHECK(_glv.geti("cycle-flag", cycle_flag), "cycle-flag");
  if (cycle_flag == 0) return;
  if (cycle_flag < 0) {
    report(true);
    return;
  }
  if (_cursor >= (int)_list_mode.size()) return;
  int timestep = 0;
  GSBN_CHECK(_glv.geti("simstep", timestep), "simstep");
  const mode_t& m = _list_mode[_cursor];
  if (timestep <= m.begin_step) return;
  if (timestep < m.end_step) {
    bool multi = _pops.size() > 1;
    for (auto& p : _pops)
      if (p.H > 1024) multi = false; /* gsbn_check_argmax_multi covers 1024 HCUs */
    if (!multi) {
      for (auto& p : _pops)
        op_check(gsbn_check_count(p.H * p.M, p.spike->gpu_data(), p.count, active_stream()),
                 "gsbn_check_count");
    } else {
      if (_items.empty()) {
        vector<gsbn_check_item> it;
        int hoff = 0;
        for (auto& p : _pops) {
          gsbn_check_item c;
          c.spike = p.spike->gpu_data();
          c.count = p.count;
          c.H = p.H;
          c.M = p.M;
          c.out_offset = hoff;
          hoff += p.H;
          it.push_back(c);
        }
        _items = it;
        hip_check(hipMalloc(&_items_dev, sizeof(gsbn_check_item) * it.size()), "hipMalloc");
        hip_check(hipMemcpy(_items_dev, it.data(), sizeof(gsbn_check_item) * it.size(),
                            hipMemcpyHostToDevice), "H2D");
      }
      op_check(gsbn_check_count_multi(_items_dev, (int32_t)_items.size(), active_stream()),
               "gsbn_check_count_multi");
    }
    _updated_flag = true;
    return;
  }
  if (!_updated_flag) {
    while (_cursor < (int)_list_mode.size() && timestep >= _list_mode[_cursor].end_step) _cursor++;
    return;
  }
  const int w = (int)_window_mode.size();
  if (_items_dev) {
    op_check(gsbn_check_argmax_multi(_items_dev, (int32_t)_items.size(),
                                     _res_idx + (int64_t)w * _total_hcu,
                                     _res_cnt + (int64_t)w * _total_hcu, active_stream()),
             "gsbn_check_argmax_multi");
  } else {
    int hoff = 0;
    for (auto& p : _pops) {
      op_check(gsbn_check_argmax(p.H, p.M, p.count, _res_idx + (int64_t)w * _total_hcu + hoff,
                                 _res_cnt + (int64_t)w * _total_hcu + hoff, active_stream()),
               "gsbn_check_argmax");
      hoff += p.H;
    }
  }
  _window_mode.push_back(_cursor);
  _cursor++;
}

/* scores windows not yet written (ProcCheck.cpp:193-223) and appends them to
 * check_result.txt; the summary line (ProcCheck.cpp:227-231) on the final call */
void ProcCheck::report(bool final_line) {
  const int nw = (int)_window_mode.size();
  if (nw > _reported) {
    hip_check(hipStreamSynchronize(active_stream()), "sync");
    vector<int32_t> idx((size_t)nw * _total_hcu), cnt((size_t)nw * _total_hcu);
    hip_check(hipMemcpy(idx.data(), _res_idx, idx.size() * 4, hipMemcpyDeviceToHost), "D2H");
    hip_check(hipMemcpy(cnt.data(), _res_cnt, cnt.size() * 4, hipMemcpyDeviceToHost), "D2H");
    std::fstream out(_logfile, std::ios::out | std::ios::app);
    for (int w = _reported; w < nw; ++w) {
      const mode_t& m = _list_mode[_window_mode[w]];
      if ((size_t)(m.lgexp_id + 1) * _lgidx->ld() > _lgidx->size())
        throw Error(-2, "expected pattern row out of range");
      const int* exp = _lgidx->cpu_data(m.lgexp_id);
      const int32_t* wi = idx.data() + (size_t)w * _total_hcu;
      const int32_t* wc = cnt.data() + (size_t)w * _total_hcu;
      bool flag = true;
      for (int i = 0; i < _total_hcu; i++) {
        if (((wc[i] < _threshold || wi[i] != exp[i]) && exp[i] >= 0) ||
            (wc[i] > _threshold && exp[i] < 0)) {
          flag = false;
          break;
        }
      }
      out << flag << "|[";
      for (int i = 0; i < _total_hcu; i++) out << exp[i] << ",";
      out << "]|[";
      for (int i = 0; i < _total_hcu; i++) out << wi[i] << ",";
      out << "]|[";
      for (int i = 0; i < _total_hcu; i++) out << wc[i] << ",";
      out << "]" << std::endl;
      _total++;
      if (flag) _correct++;
      _flags.push_back(flag ? 1 : 0);
    }
    _reported = nw;
  }
  if (final_line && !_final_done) {
    _final_done = true;
    const double pct = _total ? _correct * 100.0 / (float)_total : 0.0;
    std::fstream out(_logfile, std::ios::out | std::ios::app);
    out << "correct pattern: " << _correct << "/" << _total << "(" << pct << "%)" << std::endl;
  }
}

void ProcCheck::results(int* correct, int* total, vector<int>* flags) {
  report(false);
  *correct = _correct;
  *total = _total;
  if (flags) *flags = _flags;
}

/* lookup used by the C ABI */
bool check_results(Upd& upd, int* correct, int* total, vector<int>* flags) {
  for (auto* p : upd.procedures()) {
    auto* c = dynamic_cast<ProcCheck*>(p);
    if (c) {
      c->results(correct, total, flags);
      return true;
    }
  }
  return false;
}

}  // namespace gsbn
