/*
 * framework.cpp -- GlobalVar, Parser, procedure registry, Upd, Solver, the
 * prototxt reader and the StimRawData decoder.
 */
#include "framework.h"

#include <sys/stat.h>

#include <cctype>
#include <cmath>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace gsbn {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error((int)e, string(what) + ": " + hipGetErrorString(e));
}

void op_check(int rc, const char* what) {
  if (rc != 0) {
    string msg = string(what) + " failed with code " + std::to_string(rc);
    if (rc > 0) msg += string(" (") + hipGetErrorString((hipError_t)rc) + ")";
    throw Error(rc, msg);
  }
}

hipStream_t& active_stream() {
  static hipStream_t s = nullptr;
  return s;
}

/* ------------------------------------------------------------------ */
/* GlobalVar (GlobalVar.cpp): one process-wide pool */

map<string, GlobalVar::glvar_t>& GlobalVar::pool() {
  static map<string, glvar_t> p;
  return p;
}
void GlobalVar::clear() { pool().clear(); }

bool GlobalVar::get(const string key, glvar_t& val) {
  auto it = pool().find(key);
  if (it == pool().end()) return false;
  val = it->second;
  return true;
}
bool GlobalVar::getb(const string key, bool& val) {
  glvar_t v;
  if (!get(key, v)) return false;
  val = v.b;
  return true;
}
bool GlobalVar::geti(const string key, int& val) {
  glvar_t v;
  if (!get(key, v)) return false;
  val = v.i;
  return true;
}
bool GlobalVar::getf(const string key, float& val) {
  glvar_t v;
  if (!get(key, v)) return false;
  val = v.f;
  return true;
}
bool GlobalVar::gets(const string key, string& val) {
  glvar_t v;
  if (!get(key, v)) return false;
  val = string(v.s);
  return true;
}
bool GlobalVar::put(const string key, const glvar_t val) {
  bool existed = pool().count(key) > 0;
  pool()[key] = val;
  return existed;
}
bool GlobalVar::putb(const string key, const bool val) {
  glvar_t v;
  memset(&v, 0, sizeof(v));
  v.b = val;
  return put(key, v);
}
bool GlobalVar::puti(const string key, const int val) {
  glvar_t v;
  memset(&v, 0, sizeof(v));
  v.i = val;
  return put(key, v);
}
bool GlobalVar::putf(const string key, const float val) {
  glvar_t v;
  memset(&v, 0, sizeof(v));
  v.f = val;
  return put(key, v);
}
bool GlobalVar::puts(const string key, const string val) {
  glvar_t v;
  memset(&v, 0, sizeof(v));
  strncpy(v.s, val.c_str(), sizeof(v.s) - 1);
  return put(key, v);
}

Database::~Database() {}

map<int, ProjRowOps*>& proj_row_ops() {
  static map<int, ProjRowOps*> m;
  return m;
}

/* ------------------------------------------------------------------ */
/* Parser (Parser.cpp:12-43) */

bool Parser::argi(const string key, int32_t& val) {
  for (auto& kv : _proc_param.argi)
    if (kv.first == key) {
      val = kv.second;
      return true;
    }
  return false;
}
bool Parser::argf(const string key, float& val) {
  for (auto& kv : _proc_param.argf)
    if (kv.first == key) {
      val = kv.second;
      return true;
    }
  return false;
}
bool Parser::args(const string key, string& val) {
  for (auto& kv : _proc_param.args)
    if (kv.first == key) {
      val = kv.second;
      return true;
    }
  return false;
}

/* ------------------------------------------------------------------ */
/* ProcedureFactory (ProcedureFactory.cpp:5-33) */

ProcedureCreator::ProcedureCreator(const string& classname) {
  ProcedureFactory::registerit(classname, this);
}
void ProcedureFactory::registerit(const std::string& classname, ProcedureCreator* creator) {
  get_table()[classname] = creator;
}
ProcedureBase* ProcedureFactory::create(const std::string& classname) {
  auto i = get_table().find(classname);
  if (i != get_table().end()) return i->second->create();
  return nullptr;
}
std::map<std::string, ProcedureCreator*>& ProcedureFactory::get_table() {
  static std::map<std::string, ProcedureCreator*> table;
  return table;
}

/* ------------------------------------------------------------------ */
/* Upd (Upd.cpp:12-40); this build has only the GPU mode */

Upd::~Upd() {
  for (auto* p : _list_proc) delete p;
}

void Upd::init(SolverParam solver_param, Database& db, bool initialized_db) {
  for (size_t i = 0; i < solver_param.proc_param.size(); i++) {
    const string& proc_name = solver_param.proc_param[i].name;
    ProcedureBase* proc = ProcedureFactory::create(proc_name);
    if (!proc) throw Error(-2, "unknown procedure " + proc_name);
    _list_proc.push_back(proc);
    if (!initialized_db)
      proc->init_new(solver_param, db);
    else
      proc->init_copy(solver_param, db);
  }
}

void Upd::update() {
  for (auto* p : _list_proc) p->update_gpu();
}

void Upd::finish() {
  for (auto* p : _list_proc) p->finish();
}

/* ------------------------------------------------------------------ */
/* Solver (Solver.cpp:10-92) */

Solver::Solver(type_t type, string n_path, string s_path) {
  GlobalVar::clear();
  SolverParam solver_param = parse_solver_param_file(n_path);
  hip_check(hipStreamCreateWithFlags(&_stream, hipStreamNonBlocking), "hipStreamCreate");
  active_stream() = _stream;
  string log_dir = solver_param.rec_param.directory;
  if (log_dir.empty()) log_dir = "./";
  if (log_dir.back() != '/') log_dir += "/";
  struct stat info;
  if (!(stat(log_dir.c_str(), &info) == 0 && (info.st_mode & S_IFDIR))) {
    string cmd = "mkdir -p '" + log_dir + "'";
    if (system(cmd.c_str()) != 0) throw Error(-2, "cannot create directory " + log_dir);
  }
  _glv.puts("log-dir", log_dir);
  /* single process, one device (Common.cpp:141-216) */
  _glv.puti("rank", 0);
  _glv.puti("rank-local", 0);
  _glv.puti("num-rank", 1);
  _glv.puti("num-gpu", 1);
  if (type == NEW_SOLVER) {
    _upd.init(solver_param, _database, false);
    _glv.puti("simstep", 0);
    _glv.puti("cycle-flag", 0);
    _glv.putf("prn", 0);
    _glv.putf("old-prn", 0);
  } else {
    /* Solver.cpp:44-61: restore the vectors, then init_copy every procedure */
    SolverStateData st = read_solver_state(s_path);
    _database.init_copy(st);
    st.vecs.clear();
    _upd.init(solver_param, _database, true);
    float dt = 0;
    GSBN_CHECK(_glv.getf("dt", dt) && dt > 0, "dt");
    /* the step the snapshot was taken at.  The reference truncates
     * timestamp/dt computed in float (Solver.cpp:57), which gives one step
     * less for about 1 in 130 steps (t = 1 s at dt = 1 ms resumes at 999);
     * rounding to nearest is exact while the step count is below 2^23 */
    _glv.puti("simstep", (int)std::lround((double)st.timestamp / (double)dt));
    _glv.puti("cycle-flag", 0);
    _glv.putf("prn", st.prn);
    _glv.putf("old-prn", st.prn);
  }
}

Solver::~Solver() {
  if (_stream) {
    (void)hipStreamSynchronize(_stream);
    (void)hipStreamDestroy(_stream);
  }
  if (active_stream() == _stream) active_stream() = nullptr;
}

int Solver::step() {
  active_stream() = _stream;
  _upd.update();
  int f = 0;
  _glv.geti("cycle-flag", f);
  return f;
}

void Solver::run() {
  int f = 0;
  _glv.geti("cycle-flag", f);
  while (f >= 0) f = step();
}

/* ------------------------------------------------------------------ */
/* Timing */

Timing& Timing::get() {
  static Timing t;
  return t;
}

hipEvent_t Timing::take() {
  if (!_pool.empty()) {
    hipEvent_t e = _pool.back();
    _pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hip_check(hipEventCreate(&e), "hipEventCreate");
  return e;
}

/* begin/end bracket one kernel launch: the launch records the events as
 * part of its dispatch; if nothing was launched they are recorded here */
void Timing::begin(int kind, hipStream_t s) {
  (void)s;
  if (!_on || !((_mask >> kind) & 1)) return;
  Rec r;
  r.kind = kind;
  r.a = take();
  r.b = take();
  set_launch_events(r.a, r.b);
  _recs.push_back(r);
}

void Timing::end(int kind, hipStream_t s) {
  if (!_on || !((_mask >> kind) & 1) || _recs.empty()) return;
  Rec& r = _recs.back();
  if (r.kind != kind) return;
  if (launch_events_pending()) {
    set_launch_events(nullptr, nullptr);
    hip_check(hipEventRecord(r.a, s), "hipEventRecord");
    hip_check(hipEventRecord(r.b, s), "hipEventRecord");
  }
}

void Timing::collect(int64_t* launches, double* ms, int kinds) {
  for (auto& r : _recs) {
    hip_check(hipEventSynchronize(r.b), "hipEventSynchronize");
    float t = 0.0f;
    hip_check(hipEventElapsedTime(&t, r.a, r.b), "hipEventElapsedTime");
    if (r.kind >= 0 && r.kind < 8) {
      _launch[r.kind] += 1;
      _ms[r.kind] += t;
    }
    _pool.push_back(r.a);
    _pool.push_back(r.b);
  }
  _recs.clear();
  for (int k = 0; k < kinds && k < 8; ++k) {
    launches[k] = _launch[k];
    ms[k] = _ms[k];
  }
}

void Timing::reset() {
  set_launch_events(nullptr, nullptr);
  for (auto& r : _recs) {
    (void)hipEventSynchronize(r.b);
    _pool.push_back(r.a);
    _pool.push_back(r.b);
  }
  _recs.clear();
  for (int k = 0; k < 8; ++k) {
    _launch[k] = 0;
    _ms[k] = 0;
  }
}

/* ------------------------------------------------------------------ */
/* protobuf text format, the subset SolverParam needs */

namespace {

struct TNode;
struct TValue {
  bool is_msg = false;
  string scalar;
  bool quoted = false;
  std::shared_ptr<TNode> msg;
};
struct TNode {
  vector<std::pair<string, TValue>> fields;
};

class TextReader {
 public:
  explicit TextReader(const string& t) : s(t) {}
  std::shared_ptr<TNode> parse_message(char close) {
    auto node = std::make_shared<TNode>();
    while (true) {
      skip();
      if (pos >= s.size()) {
        if (close) fail("unexpected end of input");
        return node;
      }
      if (close && s[pos] == close) {
        ++pos;
        return node;
      }
      string name = ident();
      skip();
      bool colon = false;
      if (pos < s.size() && s[pos] == ':') {
        colon = true;
        ++pos;
        skip();
      }
      if (pos < s.size() && (s[pos] == '{' || s[pos] == '<')) {
        char c = s[pos] == '{' ? '}' : '>';
        ++pos;
        TValue v;
        v.is_msg = true;
        v.msg = parse_message(c);
        node->fields.push_back({name, v});
      } else {
        if (!colon) fail("expected ':' after " + name);
        if (pos < s.size() && s[pos] == '[') {
          ++pos;
          while (true) {
            skip();
            if (pos < s.size() && s[pos] == ']') {
              ++pos;
              break;
            }
            node->fields.push_back({name, scalar()});
            skip();
            if (pos < s.size() && s[pos] == ',') ++pos;
          }
        } else {
          node->fields.push_back({name, scalar()});
        }
      }
      skip();
      if (pos < s.size() && (s[pos] == ',' || s[pos] == ';')) ++pos;
    }
  }

 private:
  const string& s;
  size_t pos = 0;
  int line = 1;
  [[noreturn]] void fail(const string& m) {
    throw Error(-2, "prototxt line " + std::to_string(line) + ": " + m);
  }
  void skip() {
    while (pos < s.size()) {
      char c = s[pos];
      if (c == '\n') {
        ++line;
        ++pos;
      } else if (isspace((unsigned char)c)) {
        ++pos;
      } else if (c == '#') {
        while (pos < s.size() && s[pos] != '\n') ++pos;
      } else {
        break;
      }
    }
  }
  string ident() {
    size_t b = pos;
    while (pos < s.size() && (isalnum((unsigned char)s[pos]) || s[pos] == '_')) ++pos;
    if (b == pos) fail("expected a field name");
    return s.substr(b, pos - b);
  }
  TValue scalar() {
    TValue v;
    if (pos < s.size() && (s[pos] == '"' || s[pos] == '\'')) {
      char q = s[pos++];
      string out;
      while (pos < s.size() && s[pos] != q) {
        char c = s[pos++];
        if (c == '\\' && pos < s.size()) {
          char e = s[pos++];
          switch (e) {
            case 'n': out += '\n'; break;
            case 't': out += '\t'; break;
            default: out += e;
          }
        } else {
          out += c;
        }
      }
      if (pos >= s.size()) fail("unterminated string");
      ++pos;
      v.scalar = out;
      v.quoted = true;
      /* adjacent string literals concatenate */
      skip();
      if (pos < s.size() && (s[pos] == '"' || s[pos] == '\'')) {
        TValue w = scalar();
        v.scalar += w.scalar;
      }
      return v;
    }
    size_t b = pos;
    while (pos < s.size() && (isalnum((unsigned char)s[pos]) || s[pos] == '.' || s[pos] == '-' ||
                              s[pos] == '+' || s[pos] == '_'))
      ++pos;
    if (b == pos) fail("expected a value");
    v.scalar = s.substr(b, pos - b);
    return v;
  }
};

float to_f(const TValue& v, const string& name) {
  if (v.is_msg) throw Error(-2, "field " + name + " expects a number");
  string t = v.scalar;
  if (!t.empty() && (t.back() == 'f' || t.back() == 'F') && t.find_first_of("xX") == string::npos &&
      t != "inf" && t != "-inf")
    t.pop_back();
  char* end = nullptr;
  errno = 0;
  float f = strtof(t.c_str(), &end);
  if (end == t.c_str() || *end) throw Error(-2, "field " + name + ": bad number '" + v.scalar + "'");
  return f;
}
int64_t to_i(const TValue& v, const string& name) {
  if (v.is_msg) throw Error(-2, "field " + name + " expects an integer");
  char* end = nullptr;
  long long x = strtoll(v.scalar.c_str(), &end, 0);
  if (end == v.scalar.c_str() || *end) throw Error(-2, "field " + name + ": bad integer '" + v.scalar + "'");
  return x;
}
uint32_t to_u(const TValue& v, const string& name) {
  int64_t x = to_i(v, name);
  if (x < 0 || x > 0xffffffffLL) throw Error(-2, "field " + name + ": out of range");
  return (uint32_t)x;
}
string to_s(const TValue& v, const string& name) {
  if (v.is_msg) throw Error(-2, "field " + name + " expects a string");
  return v.scalar;
}
const TNode& to_m(const TValue& v, const string& name) {
  if (!v.is_msg) throw Error(-2, "field " + name + " expects a message");
  return *v.msg;
}
[[noreturn]] void unknown(const string& msg, const string& f) {
  throw Error(-2, "unknown field " + f + " in " + msg);
}

ModeParam read_mode(const TNode& n) {
  ModeParam m;
  bool bt = false, et = false;
  for (auto& kv : n.fields) {
    const string& f = kv.first;
    const TValue& v = kv.second;
    if (f == "begin_time") { m.begin_time = to_f(v, f); bt = true; }
    else if (f == "end_time") { m.end_time = to_f(v, f); et = true; }
    else if (f == "begin_lgidx_id") m.begin_lgidx_id = to_u(v, f);
    else if (f == "begin_lgexp_id") m.begin_lgexp_id = to_u(v, f);
    else if (f == "begin_wmask_id") m.begin_wmask_id = to_u(v, f);
    else if (f == "time_step") m.time_step = to_u(v, f);
    else if (f == "lgidx_step") m.lgidx_step = to_u(v, f);
    else if (f == "lgexp_step") m.lgexp_step = to_u(v, f);
    else if (f == "wmask_step") m.wmask_step = to_u(v, f);
    else if (f == "prn") m.prn = to_f(v, f);
    else if (f == "plasticity") m.plasticity = to_u(v, f);
    else unknown("mode_param", f);
  }
  if (!bt || !et) throw Error(-2, "mode_param requires begin_time and end_time");
  return m;
}

PopParam read_pop(const TNode& n) {
  PopParam p;
  for (auto& kv : n.fields) {
    const string& f = kv.first;
    const TValue& v = kv.second;
    if (f == "pop_num") p.pop_num = to_u(v, f);
    else if (f == "hcu_num") p.hcu_num = to_u(v, f);
    else if (f == "mcu_num") p.mcu_num = to_u(v, f);
    else if (f == "fanout_num") p.fanout_num = to_u(v, f);
    else if (f == "taum") p.taum = to_f(v, f);
    else if (f == "wtagain") p.wtagain = to_f(v, f);
    else if (f == "maxfq") p.maxfq = to_f(v, f);
    else if (f == "igain") p.igain = to_f(v, f);
    else if (f == "wgain") p.wgain = to_f(v, f);
    else if (f == "lgbias") p.lgbias = to_f(v, f);
    else if (f == "snoise") p.snoise = to_f(v, f);
    else if (f == "adgain") p.adgain = to_f(v, f);
    else if (f == "taua") p.taua = to_f(v, f);
    else if (f == "position") p.position.push_back((int32_t)to_i(v, f));
    else if (f == "shape") p.shape.push_back(to_u(v, f));
    else if (f == "type") p.type = to_u(v, f);
    else if (f == "rank") p.rank = to_u(v, f);
    else unknown("pop_param", f);
  }
  return p;
}

ProjParam read_proj(const TNode& n) {
  ProjParam p;
  for (auto& kv : n.fields) {
    const string& f = kv.first;
    const TValue& v = kv.second;
    if (f == "src_pop") p.src_pop = to_u(v, f);
    else if (f == "dest_pop") p.dest_pop = to_u(v, f);
    else if (f == "tauzi") p.tauzi = to_f(v, f);
    else if (f == "tauzj") p.tauzj = to_f(v, f);
    else if (f == "taue") p.taue = to_f(v, f);
    else if (f == "taup") p.taup = to_f(v, f);
    else if (f == "maxfq") p.maxfq = to_f(v, f);
    else if (f == "bgain") p.bgain = to_f(v, f);
    else if (f == "wgain") p.wgain = to_f(v, f);
    else if (f == "tauepsc") { p.tauepsc = to_f(v, f); p.has_tauepsc_ = true; }
    else if (f == "slot_num") p.slot_num = to_f(v, f);
    else if (f == "distance") p.distance = to_f(v, f);
    else unknown("proj_param", f);
  }
  return p;
}

ProcParam read_proc(const TNode& n) {
  ProcParam p;
  for (auto& kv : n.fields) {
    const string& f = kv.first;
    const TValue& v = kv.second;
    if (f == "name") {
      p.name = to_s(v, f);
    } else if (f == "argi" || f == "argf" || f == "args") {
      const TNode& a = to_m(v, f);
      string key;
      const TValue* val = nullptr;
      for (auto& kv2 : a.fields) {
        if (kv2.first == "key") key = to_s(kv2.second, "key");
        else if (kv2.first == "val") val = &kv2.second;
        else unknown(f, kv2.first);
      }
      if (!val) throw Error(-2, f + " needs key and val");
      if (f == "argi") p.argi.push_back({key, (int32_t)to_i(*val, "val")});
      else if (f == "argf") p.argf.push_back({key, to_f(*val, "val")});
      else p.args.push_back({key, to_s(*val, "val")});
    } else {
      unknown("proc_param", f);
    }
  }
  return p;
}

}  // namespace

SolverParam parse_solver_param_text(const string& text) {
  TextReader r(text);
  auto root = r.parse_message(0);
  SolverParam sp;
  bool has_gen = false, has_net = false;
  for (auto& kv : root->fields) {
    const string& f = kv.first;
    if (f == "gen_param") {
      has_gen = true;
      const TNode& g = to_m(kv.second, f);
      bool has_dt = false;
      for (auto& gv : g.fields) {
        if (gv.first == "stim_file") sp.gen_param.stim_file = to_s(gv.second, gv.first);
        else if (gv.first == "dt") { sp.gen_param.dt = to_f(gv.second, gv.first); has_dt = true; }
        else if (gv.first == "eps") sp.gen_param.eps = to_f(gv.second, gv.first);
        else if (gv.first == "mode_param") sp.gen_param.mode_param.push_back(read_mode(to_m(gv.second, gv.first)));
        else if (gv.first == "procedure") (void)to_s(gv.second, gv.first);
        else unknown("gen_param", gv.first);
      }
      if (!has_dt) throw Error(-2, "gen_param requires dt");
    } else if (f == "net_param") {
      has_net = true;
      const TNode& n = to_m(kv.second, f);
      for (auto& nv : n.fields) {
        if (nv.first == "pop_param") sp.net_param.pop_param.push_back(read_pop(to_m(nv.second, nv.first)));
        else if (nv.first == "proj_param") sp.net_param.proj_param.push_back(read_proj(to_m(nv.second, nv.first)));
        else unknown("net_param", nv.first);
      }
    } else if (f == "rec_param") {
      const TNode& n = to_m(kv.second, f);
      for (auto& nv : n.fields) {
        if (nv.first == "directory") sp.rec_param.directory = to_s(nv.second, nv.first);
        else unknown("rec_param", nv.first);
      }
    } else if (f == "proc_param") {
      sp.proc_param.push_back(read_proc(to_m(kv.second, f)));
    } else {
      unknown("SolverParam", f);
    }
  }
  if (!has_gen || !has_net) throw Error(-2, "SolverParam requires gen_param and net_param");
  return sp;
}

SolverParam parse_solver_param_file(const string& path) {
  std::ifstream f(path);
  if (!f) throw Error(-2, "cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_solver_param_text(ss.str());
}

/* ------------------------------------------------------------------ */
/* StimRawData wire format */

StimRawData read_stim_raw_data(const string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Error(-2, "File not found! " + path);
  std::string buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  const unsigned char* p = (const unsigned char*)buf.data();
  size_t n = buf.size(), pos = 0;
  auto varint = [&](size_t& q) -> uint64_t {
    uint64_t v = 0;
    int sh = 0;
    while (true) {
      if (q >= n || sh > 63) throw Error(-2, "Parse file error! truncated varint in " + path);
      uint8_t b = p[q++];
      v |= (uint64_t)(b & 0x7f) << sh;
      sh += 7;
      if (!(b & 0x80)) return v;
    }
  };
  StimRawData d;
  bool r1 = false, r2 = false, r4 = false, r5 = false;
  while (pos < n) {
    uint64_t tag = varint(pos);
    int fno = (int)(tag >> 3), wt = (int)(tag & 7);
    if (wt == 0) {
      uint64_t v = varint(pos);
      if (fno == 1) { d.data_rows = (uint32_t)v; r1 = true; }
      else if (fno == 2) { d.data_cols = (uint32_t)v; r2 = true; }
      else if (fno == 3) d.data.push_back((uint32_t)v);
      else if (fno == 4) { d.mask_rows = (uint32_t)v; r4 = true; }
      else if (fno == 5) { d.mask_cols = (uint32_t)v; r5 = true; }
    } else if (wt == 2) {
      uint64_t len = varint(pos);
      if (pos + len > n) throw Error(-2, "Parse file error! truncated field in " + path);
      size_t end = pos + len;
      if (fno == 3) {
        while (pos < end) d.data.push_back((uint32_t)varint(pos));
      } else if (fno == 6) {
        if (len % 4) throw Error(-2, "Parse file error! bad packed float in " + path);
        for (size_t q = pos; q < end; q += 4) {
          float fv;
          memcpy(&fv, p + q, 4);
          d.mask.push_back(fv);
        }
      }
      pos = end;
    } else if (wt == 5) {
      if (pos + 4 > n) throw Error(-2, "Parse file error! truncated fixed32 in " + path);
      if (fno == 6) {
        float fv;
        memcpy(&fv, p + pos, 4);
        d.mask.push_back(fv);
      }
      pos += 4;
    } else if (wt == 1) {
      pos += 8;
    } else {
      throw Error(-2, "Parse file error! unsupported wire type in " + path);
    }
  }
  if (!(r1 && r2 && r4 && r5)) throw Error(-2, "Parse file error! missing required field in " + path);
  return d;
}

}  // namespace gsbn
