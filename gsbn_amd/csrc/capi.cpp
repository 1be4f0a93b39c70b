/*
 * capi.cpp -- solver-level C ABI (include/gsbn_amd.h) and host helpers.
 * Exceptions never cross the ABI: each entry point converts gsbn::Error into
 * its status code and keeps the message for gsbn_last_error().
 */
#include <cstring>
#include <string>

#include "../../include/gsbn_amd.h"
#include "coeff.h"
#include "detmath.h"
#include "framework.h"
#include "gsbn_core.h"
#include "philox.h"

namespace gsbn {
bool check_results(Upd& upd, int* correct, int* total, vector<int>* flags);
bool updlazy_work(Upd& upd, int64_t* out5);
}  // namespace gsbn

using gsbn::Error;

struct gsbn_solver {
  gsbn::Solver* s = nullptr;
};

struct gsbn_state {
  gsbn::SolverStateData d;
};

namespace {
thread_local std::string g_err;

int fail(const std::exception& e) {
  g_err = e.what();
  const Error* ge = dynamic_cast<const Error*>(&e);
  return ge ? (ge->code ? ge->code : GSBN_ERR_STATE) : GSBN_ERR_STATE;
}

template <typename F>
int guard(F f) {
  try {
    return f();
  } catch (const std::exception& e) {
    return fail(e);
  }
}

}  // namespace

namespace gsbn {
/* message for gsbn_last_error() from the other C ABI sources */
void set_last_error(const std::string& m) { g_err = m; }
}  // namespace gsbn

namespace {
bool dynamic_size(const std::string& n) {
  return n.rfind(".ssi_", 0) == 0 || n.rfind(".ssj_", 0) == 0;
}
}  // namespace

extern "C" {

const char* gsbn_last_error(void) { return g_err.c_str(); }

int gsbn_solver_snapshot(gsbn_solver* s, const char* path) {
  return guard([&] {
    if (!s || !s->s || !path) throw Error(GSBN_ERR_ARG, "null argument");
    gsbn::GlobalVar glv;
    int simstep = 0;
    float dt = 0, prn = 0;
    glv.geti("simstep", simstep);
    glv.getf("dt", dt);
    glv.getf("prn", prn);
    gsbn::active_stream() = s->s->stream();
    s->s->db().save_state(path, simstep * dt, prn);
    return 0;
  });
}

int gsbn_state_write(const char* path, float timestamp, float prn, int32_t nvec,
                     const gsbn_vec_state* vecs) {
  return guard([&] {
    if (!path || nvec < 0 || (nvec && !vecs)) throw Error(GSBN_ERR_ARG, "null argument");
    std::vector<gsbn::VecSpan> spans;
    for (int32_t i = 0; i < nvec; ++i) {
      const gsbn_vec_state& v = vecs[i];
      if (!v.name || v.dtype < 0 || v.dtype > 6 || v.count < 0 || (v.count && !v.data))
        throw Error(GSBN_ERR_ARG, "bad vector state " + std::to_string(i));
      gsbn::VecSpan sp;
      sp.name = v.name;
      sp.type = (gsbn::DType)v.dtype;
      sp.ld = (int)v.ld;
      sp.count = (size_t)v.count;
      sp.data = v.data;
      spans.push_back(sp);
    }
    gsbn::write_solver_state(path, timestamp, prn, std::move(spans));
    return 0;
  });
}

gsbn_state* gsbn_state_read(const char* path) {
  try {
    if (!path) throw Error(GSBN_ERR_ARG, "null path");
    auto* st = new gsbn_state();
    try {
      st->d = gsbn::read_solver_state(path);
    } catch (...) {
      delete st;
      throw;
    }
    return st;
  } catch (const std::exception& e) {
    fail(e);
    return nullptr;
  }
}

int32_t gsbn_state_nvec(const gsbn_state* st) { return st ? (int32_t)st->d.vecs.size() : 0; }

int gsbn_state_header(const gsbn_state* st, float* timestamp, float* prn) {
  return guard([&] {
    if (!st) throw Error(GSBN_ERR_ARG, "null state");
    if (timestamp) *timestamp = st->d.timestamp;
    if (prn) *prn = st->d.prn;
    return 0;
  });
}

int gsbn_state_vec(const gsbn_state* st, int32_t i, gsbn_vec_state* out) {
  return guard([&] {
    if (!st || !out) throw Error(GSBN_ERR_ARG, "null argument");
    if (i < 0 || i >= (int32_t)st->d.vecs.size()) throw Error(GSBN_ERR_ARG, "vector index out of range");
    const gsbn::VecState& v = st->d.vecs[(size_t)i];
    out->name = v.name.c_str();
    out->dtype = (int32_t)v.type;
    out->ld = v.ld;
    out->count = (int64_t)v.count();
    out->data = v.bytes.empty() ? nullptr : v.bytes.data();
    return 0;
  });
}

void gsbn_state_free(gsbn_state* st) { delete st; }

gsbn_solver* gsbn_solver_new(const char* net_path, const char* snapshot_path) {
  try {
    if (!net_path) throw Error(GSBN_ERR_ARG, "net_path is NULL");
    std::string sp = snapshot_path ? snapshot_path : "";
    auto* h = new gsbn_solver();
    try {
      h->s = new gsbn::Solver(sp.empty() ? gsbn::Solver::NEW_SOLVER : gsbn::Solver::COPY_SOLVER,
                              net_path, sp);
      gsbn::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    } catch (...) {
      delete h;
      throw;
    }
    return h;
  } catch (const std::exception& e) {
    fail(e);
    return nullptr;
  }
}

void gsbn_solver_free(gsbn_solver* s) {
  if (!s) return;
  try {
    delete s->s;
  } catch (...) {
  }
  delete s;
}

int64_t gsbn_solver_step(gsbn_solver* s, int64_t nsteps) {
  try {
    if (!s || !s->s) throw Error(GSBN_ERR_ARG, "null solver");
    int64_t n = 0;
    gsbn::GlobalVar glv;
    int f = 0;
    glv.geti("cycle-flag", f);
    while (n < nsteps && f >= 0) {
      f = s->s->step();
      ++n;
    }
    return n;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int64_t gsbn_solver_run(gsbn_solver* s) { return gsbn_solver_step(s, INT64_MAX); }

int gsbn_solver_sync(gsbn_solver* s) {
  return guard([&] {
    if (!s || !s->s) throw Error(GSBN_ERR_ARG, "null solver");
    gsbn::hip_check(hipStreamSynchronize(s->s->stream()), "hipStreamSynchronize");
    return 0;
  });
}

int64_t gsbn_solver_vector_info(gsbn_solver* s, const char* name, int* dtype) {
  try {
    if (!s || !s->s || !name) throw Error(GSBN_ERR_ARG, "null argument");
    gsbn::active_stream() = s->s->stream(); /* host copies sync on this solver's stream */
    gsbn::VecBase* v = s->s->db().find(name);
    if (!v) throw Error(GSBN_ERR_NAME, std::string("no vector named ") + name);
    if (dynamic_size(name)) (void)v->host_ptr_sync();
    if (dtype) *dtype = (int)v->type;
    return (int64_t)v->size();
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int64_t gsbn_solver_vector_read(gsbn_solver* s, const char* name, void* out, int64_t cap) {
  try {
    if (!s || !s->s || !name) throw Error(GSBN_ERR_ARG, "null argument");
    gsbn::active_stream() = s->s->stream(); /* host copies sync on this solver's stream */
    gsbn::VecBase* v = s->s->db().find(name);
    if (!v) throw Error(GSBN_ERR_NAME, std::string("no vector named ") + name);
    void* p = v->host_ptr_sync();
    int64_t n = (int64_t)v->size();
    if (out && cap > 0 && n > 0) memcpy(out, p, (size_t)std::min(n, cap) * v->elem_size());
    return n;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int gsbn_solver_vector_write(gsbn_solver* s, const char* name, const void* in, int64_t count) {
  return guard([&] {
    if (!s || !s->s || !name || (!in && count)) throw Error(GSBN_ERR_ARG, "null argument");
    gsbn::active_stream() = s->s->stream(); /* host copies sync on this solver's stream */
    gsbn::VecBase* v = s->s->db().find(name);
    if (!v) throw Error(GSBN_ERR_NAME, std::string("no vector named ") + name);
    if (count != (int64_t)v->size())
      throw Error(GSBN_ERR_ARG, std::string("size mismatch writing ") + name);
    void* p = v->host_ptr_mutable();
    if (count) memcpy(p, in, (size_t)count * v->elem_size());
    return 0;
  });
}

int gsbn_solver_global(gsbn_solver* s, const char* key, char type, void* out) {
  return guard([&] {
    if (!s || !key || !out) throw Error(GSBN_ERR_ARG, "null argument");
    gsbn::GlobalVar g;
    bool ok = false;
    if (type == 'i') ok = g.geti(key, *(int*)out);
    else if (type == 'f') ok = g.getf(key, *(float*)out);
    else if (type == 'b') ok = g.getb(key, *(bool*)out);
    else throw Error(GSBN_ERR_ARG, "type must be i, f or b");
    if (!ok) throw Error(GSBN_ERR_NAME, std::string("no global ") + key);
    return 0;
  });
}

int gsbn_solver_check(gsbn_solver* s, int32_t* correct, int32_t* total, int32_t* flags,
                      int32_t cap) {
  return guard([&] {
    if (!s || !s->s || !correct || !total) throw Error(GSBN_ERR_ARG, "null argument");
    std::vector<int> f;
    int c = 0, t = 0;
    if (!gsbn::check_results(s->s->upd(), &c, &t, &f)) throw Error(GSBN_ERR_STATE, "no ProcCheck");
    *correct = c;
    *total = t;
    for (int i = 0; flags && i < cap && i < (int)f.size(); ++i) flags[i] = f[i];
    return 0;
  });
}

int gsbn_solver_work(gsbn_solver* s, int64_t* out5) {
  return guard([&] {
    if (!s || !s->s || !out5) throw Error(GSBN_ERR_ARG, "null argument");
    if (!gsbn::updlazy_work(s->s->upd(), out5)) throw Error(GSBN_ERR_STATE, "no ProcUpdLazy");
    return 0;
  });
}

int gsbn_solver_timing(gsbn_solver* s, int enable, int64_t* launches8, double* ms8) {
  return guard([&] {
    if (!s || !s->s) throw Error(GSBN_ERR_ARG, "null solver");
    auto& t = gsbn::Timing::get();
    if (launches8 && ms8) t.collect(launches8, ms8, 8);
    if (enable < 0) t.reset();
    t.enable(enable > 0 ? enable : 0);
    return 0;
  });
}

int gsbn_config_check(const char* net_path, int32_t* out4) {
  return guard([&] {
    if (!net_path || !out4) throw Error(GSBN_ERR_ARG, "null argument");
    gsbn::SolverParam sp = gsbn::parse_solver_param_file(net_path);
    int npop = 0;
    for (auto& p : sp.net_param.pop_param) npop += p.pop_num;
    out4[0] = npop;
    out4[1] = (int32_t)sp.net_param.proj_param.size();
    out4[2] = (int32_t)sp.gen_param.mode_param.size();
    out4[3] = (int32_t)sp.proc_param.size();
    if (!sp.gen_param.stim_file.empty()) {
      gsbn::StimRawData d = gsbn::read_stim_raw_data(sp.gen_param.stim_file);
      (void)d;
    }
    return 0;
  });
}

int gsbn_stim_read(const char* path, int32_t* dims4, uint32_t* data, int64_t data_cap, float* mask,
                   int64_t mask_cap) {
  return guard([&] {
    if (!path || !dims4) throw Error(GSBN_ERR_ARG, "null argument");
    gsbn::StimRawData d = gsbn::read_stim_raw_data(path);
    dims4[0] = (int32_t)d.data_rows;
    dims4[1] = (int32_t)d.data_cols;
    dims4[2] = (int32_t)d.mask_rows;
    dims4[3] = (int32_t)d.mask_cols;
    for (int64_t i = 0; data && i < data_cap && i < (int64_t)d.data.size(); ++i) data[i] = d.data[i];
    for (int64_t i = 0; mask && i < mask_cap && i < (int64_t)d.mask.size(); ++i) mask[i] = d.mask[i];
    return 0;
  });
}

void gsbn_coeff_host_entry(double kzi, double kzj, double ke, double kp, int64_t gap, float* out12) {
  gsbn_rates_t r = {kzi, kzj, ke, kp};
  gsbn_coeff_entry(&r, gap, out12);
}

int64_t gsbn_coeff_host_saturation(double kzi, double kzj, double ke, double kp, int64_t max_len) {
  gsbn_rates_t r = {kzi, kzj, ke, kp};
  return gsbn_coeff_saturation(&r, max_len);
}

float gsbn_host_expf(float x) { return gsbn_expf(x); }
float gsbn_host_logf(float x) { return gsbn_logf(x); }
void gsbn_host_draw(uint64_t seed, uint32_t pop, uint32_t step, uint32_t unit, float* u, float* z) {
  gsbn_draw_unit(seed, pop, step, unit, u, z);
}
uint16_t gsbn_host_pack_trace(int prec, float x, float scale) {
  return gsbn_pack_trace(prec, x, scale);
}
float gsbn_host_unpack_trace(int prec, uint16_t h, float inv_scale) {
  return gsbn_unpack_trace(prec, h, inv_scale);
}

}  // extern "C"
