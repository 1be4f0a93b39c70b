/*
 * framework.h -- the procedure framework the lazy-update path plugs into,
 * with the reference's names and contracts:
 *
 *   GlobalVar           process-wide key/value pool (GlobalVar.hpp:23-143)
 *   SyncVector<T>       named vector with a host and a device copy and lazy
 *                       synchronisation (SyncVector.hpp:181-188)
 *   Database            typed named vectors, create_* returns NULL when the
 *                       name exists, sync_vector_* NULL when missing
 *                       (Database.hpp:53-158)
 *   SolverParam & co    the subset of src/gsbn/proto/gsbn.proto this path reads
 *   Parser              argi/argf/args lookup in a ProcParam (Parser.cpp:12-43)
 *   ProcedureBase, ProcedureFactory, REGISTER, REGISTERIMPL
 *                       plug-in registry (ProcedureFactory.hpp:17-85)
 *   Upd, Solver         procedure list and main loop (Upd.cpp:12-40,
 *                       Solver.cpp:10-92)
 *
 * Errors: where the reference aborts through CHECK/LOG(FATAL), this code
 * throws gsbn::Error; the C ABI converts it into a status code.
 */
#ifndef GSBN_FRAMEWORK_H
#define GSBN_FRAMEWORK_H

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace gsbn {

using std::map;
using std::string;
using std::vector;

struct Error : std::runtime_error {
  int code;
  Error(int c, const string& m) : std::runtime_error(m), code(c) {}
};

#define GSBN_CHECK(cond, msg)                                                          \
  do {                                                                                 \
    if (!(cond)) throw ::gsbn::Error(-4, std::string("check failed: ") + (msg));       \
  } while (0)

void hip_check(hipError_t e, const char* what);
void op_check(int rc, const char* what);
/* the stream all framework copies and kernels use (the current solver's) */
hipStream_t& active_stream();

/* ------------------------------------------------------------------ */
class GlobalVar {
 public:
  typedef union {
    bool b;
    int i;
    float f;
    char s[256];
  } glvar_t;
  bool get(const string key, glvar_t& val);
  bool getb(const string key, bool& val);
  bool geti(const string key, int& val);
  bool getf(const string key, float& val);
  bool gets(const string key, string& val);
  bool put(const string key, const glvar_t val);
  bool putb(const string key, const bool val);
  bool puti(const string key, const int val);
  bool putf(const string key, const float val);
  bool puts(const string key, const string val);
  static void clear();

 private:
  static map<string, glvar_t>& pool();
};

/* ------------------------------------------------------------------ */
enum DType { DT_I8 = 0, DT_I16 = 1, DT_I32 = 2, DT_I64 = 3, DT_F16 = 4, DT_F32 = 5, DT_F64 = 6 };

template <typename T>
struct dtype_of;
template <> struct dtype_of<int8_t> { static const DType v = DT_I8; };
template <> struct dtype_of<int16_t> { static const DType v = DT_I16; };
template <> struct dtype_of<int32_t> { static const DType v = DT_I32; };
template <> struct dtype_of<int64_t> { static const DType v = DT_I64; };
template <> struct dtype_of<float> { static const DType v = DT_F32; };
template <> struct dtype_of<double> { static const DType v = DT_F64; };

class VecBase;

/* A vector whose authoritative storage is device-side state owned by a
 * procedure (record arrays, queues): the view copies it in and out. */
struct VecView {
  virtual ~VecView() {}
  virtual void pull(VecBase* v) = 0; /* device state -> host copy */
  virtual void push(VecBase* v) = 0; /* host copy -> device state */
};

class VecBase {
 public:
  enum Status { SYNCED, HOST_NEWER, DEVICE_NEWER };
  explicit VecBase(const string& n, DType t) : name(n), type(t) {}
  virtual ~VecBase() {}
  virtual size_t size() const = 0;
  virtual size_t elem_size() const = 0;
  virtual void* host_ptr_sync() = 0;     /* host copy, brought up to date */
  virtual void* host_ptr_mutable() = 0;  /* same, then host becomes authoritative */
  virtual void resize_raw(size_t n) = 0;
  string name;
  DType type;
  int ld_ = 0;
  VecView* view = nullptr;
  std::function<void()> before_host_read; /* e.g. flush pending device work */
  uint64_t host_version = 0;              /* bumped on every mutable host access */
  Status status = SYNCED;
  /* process-wide count of host-side accesses that moved or invalidated data;
   * procedures compare it between steps to skip re-validating their vectors */
  static uint64_t& host_epoch() {
    static uint64_t e = 0;
    return e;
  }
  /* the same for host writes only */
  static uint64_t& host_write_epoch() {
    static uint64_t e = 0;
    return e;
  }
};

template <typename T>
class SyncVector : public VecBase {
 public:
  explicit SyncVector(const string& n) : VecBase(n, dtype_of<T>::v) {}
  ~SyncVector() override {
    if (stage_ev_) {
      (void)hipEventSynchronize(stage_ev_);
      (void)hipEventDestroy(stage_ev_);
    }
    if (stage_) (void)hipHostFree(stage_);
    if (d_) (void)hipFree(d_);
  }
  size_t size() const override { return n_; }
  size_t elem_size() const override { return sizeof(T); }
  int ld() const { return ld_; }
  void set_ld(int l) { ld_ = l; }

  void resize(size_t n, T v = T()) {
    to_host();
    h_.resize(n, v);
    n_ = n;
    mark_host();
  }
  /* logical size for view-backed vectors (no host storage until read) */
  void set_logical_size(size_t n) { n_ = n; }

  const vector<T>* cpu_vector() {
    to_host();
    return &h_;
  }
  vector<T>* mutable_cpu_vector() {
    to_host();
    mark_host();
    return &h_;
  }
  const T* cpu_data(int i = 0) {
    to_host();
    return h_.empty() ? nullptr : h_.data() + offset(i);
  }
  T* mutable_cpu_data(int i = 0) {
    to_host();
    mark_host();
    return h_.empty() ? nullptr : h_.data() + offset(i);
  }
  const T* gpu_data(int i = 0) {
    to_device();
    return d_ ? d_ + offset(i) : nullptr;
  }
  T* mutable_gpu_data(int i = 0) {
    to_device();
    if (d_) status = DEVICE_NEWER;
    return d_ ? d_ + offset(i) : nullptr;
  }
  /* the device copy was modified by a kernel */
  void mark_device() { status = DEVICE_NEWER; }
  void mark_host() {
    status = HOST_NEWER;
    ++host_version;
    ++host_epoch();
    ++host_write_epoch();
  }

  void* host_ptr_sync() override {
    to_host();
    return h_.data();
  }
  void* host_ptr_mutable() override {
    to_host();
    mark_host();
    return h_.data();
  }
  void resize_raw(size_t n) override { resize(n); }

  static hipStream_t stream() { return active_stream(); }

  void to_host() {
    if (status != DEVICE_NEWER) return;
    ++host_epoch();
    if (before_host_read) before_host_read();
    if (view) {
      h_.resize(n_);
      view->pull(this);
    } else {
      h_.resize(n_);
      if (n_) {
        hip_check(hipStreamSynchronize(stream()), "sync");
        hip_check(hipMemcpy(h_.data(), d_, n_ * sizeof(T), hipMemcpyDeviceToHost), "D2H");
      }
    }
    status = SYNCED;
  }
  void to_device() {
    if (view) {
      if (status == HOST_NEWER) {
        view->push(this);
        status = SYNCED;
      }
      return;
    }
    if (n_ != d_n_) {
      /* (re)allocate; keep the host copy authoritative */
      if (status == DEVICE_NEWER) to_host();
      if (d_) hip_check(hipFree(d_), "free");
      d_ = nullptr;
      if (n_) hip_check(hipMalloc(&d_, n_ * sizeof(T)), "hipMalloc");
      d_n_ = n_;
      if (status == SYNCED) status = HOST_NEWER;
    }
    if (status == HOST_NEWER) {
      if (n_) {
        if (h_.size() < n_) h_.resize(n_);
        upload();
      }
      status = SYNCED;
    }
  }
  vector<T>& raw_host() { return h_; }

 private:
  /* host -> device copy in stream order.  Vectors of up to kStageMax bytes
   * (stimulus rows of every network of a batch, masks, counters) go through
   * a pinned staging buffer of their own, so the call returns without
   * waiting for the device work queued before it; the buffer is refilled
   * only after its previous copy has run (an event).  Larger vectors wait
   * for their copy. */
  static constexpr size_t kStageMax = (size_t)64 << 20;
  void upload() {
    const size_t bytes = n_ * sizeof(T);
    if (bytes <= kStageMax) {
      if (stage_ev_)
        hip_check(hipEventSynchronize(stage_ev_), "staging event");
      else
        hip_check(hipEventCreateWithFlags(&stage_ev_, hipEventDisableTiming), "hipEventCreate");
      if (stage_n_ < bytes) {
        if (stage_) hip_check(hipHostFree(stage_), "hipHostFree");
        stage_ = nullptr;
        stage_n_ = 0;
        hip_check(hipHostMalloc(&stage_, bytes, hipHostMallocDefault), "hipHostMalloc");
        stage_n_ = bytes;
      }
      memcpy(stage_, h_.data(), bytes);
      hip_check(hipMemcpyAsync(d_, stage_, bytes, hipMemcpyHostToDevice, stream()), "H2D");
      hip_check(hipEventRecord(stage_ev_, stream()), "hipEventRecord");
      return;
    }
    hip_check(hipMemcpyAsync(d_, h_.data(), bytes, hipMemcpyHostToDevice, stream()), "H2D");
    hip_check(hipStreamSynchronize(stream()), "sync");
  }
  void* stage_ = nullptr;
  size_t stage_n_ = 0;
  hipEvent_t stage_ev_ = nullptr;

  size_t offset(int i) const {
    if (i <= 0) return 0;
    return ld_ <= 0 ? (size_t)i : (size_t)i * (size_t)ld_;
  }
  vector<T> h_;
  T* d_ = nullptr;
  size_t n_ = 0, d_n_ = 0;
};

/* SolverState (gsbn.proto:100-152): one saved vector.  `bytes` holds the
 * elements in the Database type (float for DT_F16, as VectorStateF16). */
struct VecState {
  string name;
  DType type = DT_F32;
  uint32_t ld = 1;
  vector<uint8_t> bytes;
  size_t count() const;
};
struct SolverStateData {
  float timestamp = 0, prn = 0;
  vector<VecState> vecs;
};
/* a vector to save: `count` elements of `type` at `data` (host memory) */
struct VecSpan {
  string name;
  DType type = DT_F32;
  int ld = 0;
  size_t count = 0;
  const void* data = nullptr;
};
/* snapshot.cpp: the SolverState wire encoding */
void write_solver_state(const string& path, float timestamp, float prn, vector<VecSpan> vecs);
SolverStateData read_solver_state(const string& path);

class Database {
 public:
  ~Database();
  /* Database::init_copy (Database.cpp:30-115): one vector per saved state */
  void init_copy(const SolverStateData& st);
  /* state_to_proto (Database.cpp:257-295) serialised to `path`: every
   * vector whose name does not start with '.' */
  void save_state(const string& path, float timestamp, float prn);
  /* the vector was created by init_copy from a snapshot */
  bool restored(const string& name) const { return _restored.count(name) > 0; }
  SyncVector<int8_t>* create_sync_vector_i8(const string name) { return create<int8_t>(name); }
  SyncVector<int16_t>* create_sync_vector_i16(const string name) { return create<int16_t>(name); }
  SyncVector<int32_t>* create_sync_vector_i32(const string name) { return create<int32_t>(name); }
  SyncVector<int64_t>* create_sync_vector_i64(const string name) { return create<int64_t>(name); }
  SyncVector<float>* create_sync_vector_f32(const string name) { return create<float>(name); }
  SyncVector<double>* create_sync_vector_f64(const string name) { return create<double>(name); }
  SyncVector<int8_t>* sync_vector_i8(const string name) { return get<int8_t>(name); }
  SyncVector<int16_t>* sync_vector_i16(const string name) { return get<int16_t>(name); }
  SyncVector<int32_t>* sync_vector_i32(const string name) { return get<int32_t>(name); }
  SyncVector<int64_t>* sync_vector_i64(const string name) { return get<int64_t>(name); }
  SyncVector<float>* sync_vector_f32(const string name) { return get<float>(name); }
  SyncVector<double>* sync_vector_f64(const string name) { return get<double>(name); }
  VecBase* find(const string& name) {
    auto it = _vecs.find(name);
    return it == _vecs.end() ? nullptr : it->second.get();
  }
  const map<string, std::unique_ptr<VecBase>>& all() const { return _vecs; }

  /* typed create / lookup: NULL when the name exists / is missing or has
   * another element type */
  template <typename T>
  SyncVector<T>* create(const string& name) {
    if (_vecs.count(name)) return nullptr;
    auto* v = new SyncVector<T>(name);
    _vecs[name].reset(v);
    return v;
  }
  template <typename T>
  SyncVector<T>* get(const string& name) {
    auto it = _vecs.find(name);
    if (it == _vecs.end() || it->second->type != dtype_of<T>::v) return nullptr;
    return static_cast<SyncVector<T>*>(it->second.get());
  }

 private:
  map<string, std::unique_ptr<VecBase>> _vecs;
  std::set<string> _restored;
};

/* ------------------------------------------------------------------ */
/* SolverParam subset (gsbn.proto:4-98) with the schema's defaults */
struct ModeParam {
  float begin_time = 0, end_time = 0;
  uint32_t begin_lgidx_id = 0, begin_lgexp_id = 0, begin_wmask_id = 0, time_step = 1;
  uint32_t lgidx_step = 1, lgexp_step = 1, wmask_step = 1;
  float prn = 1.0f;
  uint32_t plasticity = 1;
};
struct GenParam {
  string stim_file;
  float dt = 0;
  float eps = 0.001f;
  vector<ModeParam> mode_param;
};
struct PopParam {
  uint32_t pop_num = 0, hcu_num = 0, mcu_num = 0, fanout_num = 0;
  float taum = 0.01f, wtagain = 4, maxfq = 100, igain = 1, wgain = 1, lgbias = 0, snoise = 0;
  float adgain = 1, taua = 0.05f;
  vector<int32_t> position;
  vector<uint32_t> shape;
  uint32_t type = 0, rank = 0;
};
struct ProjParam {
  uint32_t src_pop = 0, dest_pop = 0;
  float tauzi = 0, tauzj = 0, taue = 0, taup = 0, maxfq = 100, bgain = 0, wgain = 0;
  bool has_tauepsc_ = false;
  float tauepsc = 0;
  float slot_num = 0, distance = 0;
  bool has_tauepsc() const { return has_tauepsc_; }
};
struct NetParam {
  vector<PopParam> pop_param;
  vector<ProjParam> proj_param;
};
struct RecParam {
  string directory = ".";
};
struct ProcParam {
  string name;
  vector<std::pair<string, int32_t>> argi;
  vector<std::pair<string, float>> argf;
  vector<std::pair<string, string>> args;
};
struct SolverParam {
  GenParam gen_param;
  NetParam net_param;
  RecParam rec_param;
  vector<ProcParam> proc_param;
};

/* parse the protobuf text format of a SolverParam (the subset above) */
SolverParam parse_solver_param_file(const string& path);
SolverParam parse_solver_param_text(const string& text);

/* StimRawData wire-format decoder (gsbn.proto:153-160) */
struct StimRawData {
  uint32_t data_rows = 0, data_cols = 0, mask_rows = 0, mask_cols = 0;
  vector<uint32_t> data;
  vector<float> mask;
};
StimRawData read_stim_raw_data(const string& path);

class Parser {
 public:
  explicit Parser(ProcParam proc_param) : _proc_param(proc_param) {}
  bool argi(const string key, int32_t& val);
  bool argf(const string key, float& val);
  bool args(const string key, string& val);

 private:
  ProcParam _proc_param;
};

/* ------------------------------------------------------------------ */
/* Row operations on a projection's device-resident synapse state, offered
 * by the procedure that owns the state (ProcUpdLazy) to procedures that
 * edit connectivity (ProcStructuralPlasticity), by projection id. */
struct ProjRowOps {
  virtual ~ProjRowOps() {}
  /* Prj::prune's test per row (ProcStructuralPlasticity/Prj.cpp:112-138) */
  virtual void prune_flags(int t_min, float wp, float wn, vector<uint8_t>* flags) = 0;
  /* Prj::remove_conn's state reset except ii/di (Prj.cpp:78-107); the rows
   * also leave the delay queue */
  virtual void clear_rows(const vector<int32_t>& rows) = 0;
  /* rows were added to ii/di on the host after a step: those whose source
   * unit spiked in that step are queued with that spike, as the next queue
   * update would queue it (Proj.cpp:655-661) */
  virtual void rows_added() = 0;
};
map<int, ProjRowOps*>& proj_row_ops();

/* ------------------------------------------------------------------ */
class ProcedureBase {
 public:
  ProcedureBase() {}
  virtual ~ProcedureBase() {}
  virtual void init_new(SolverParam net_param, Database& db) = 0;
  virtual void init_copy(SolverParam net_param, Database& db) = 0;
  virtual void update_cpu() = 0;
  virtual void update_gpu() = 0;
  /* device work enqueued by the procedure is complete (host reads follow) */
  virtual void finish() {}
};

class ProcedureCreator {
 public:
  explicit ProcedureCreator(const string& classname);
  virtual ~ProcedureCreator() {}
  virtual ProcedureBase* create() = 0;
};

template <class T>
class ProcedureCreatorImpl : public ProcedureCreator {
 public:
  explicit ProcedureCreatorImpl(const std::string& classname) : ProcedureCreator(classname) {}
  ProcedureBase* create() override { return new T; }
};

class ProcedureFactory {
 public:
  static ProcedureBase* create(const std::string& classname);
  static void registerit(const std::string& classname, ProcedureCreator* creator);

 private:
  static std::map<std::string, ProcedureCreator*>& get_table();
};

#define REGISTER(classname)                                   \
 private:                                                     \
  static const ProcedureCreatorImpl<classname> creator;       \
  static const string proc_name();                            \
  static const ProcParam get_proc_param(SolverParam solver_param);

#define REGISTERIMPL(classname)                                                     \
  const ProcedureCreatorImpl<classname> classname::creator(#classname);             \
  const string classname::proc_name() { return #classname; }                        \
  const ProcParam classname::get_proc_param(SolverParam solver_param) {            \
    for (size_t i = 0; i < solver_param.proc_param.size(); i++)                     \
      if (solver_param.proc_param[i].name == proc_name()) return solver_param.proc_param[i]; \
    throw ::gsbn::Error(-2, "No parameters specified for `" + proc_name() + "' !"); \
  }

class Upd {
 public:
  ~Upd();
  void init(SolverParam solver_param, Database& db, bool initialized_db = false);
  void update();
  void finish();
  const vector<ProcedureBase*>& procedures() const { return _list_proc; }

 private:
  vector<ProcedureBase*> _list_proc;
};

class Solver {
 public:
  enum type_t { NEW_SOLVER, COPY_SOLVER };
  Solver(type_t type, string n_path, string s_path);
  ~Solver();
  void run();
  /* one Upd::update; returns the cycle-flag after it */
  int step();
  Database& db() { return _database; }
  Upd& upd() { return _upd; }
  hipStream_t stream() const { return _stream; }

 private:
  Upd _upd;
  Database _database;
  GlobalVar _glv;
  hipStream_t _stream = nullptr;
};

/* kernel timing with HIP events on the solver stream */
/* kernels.hip: the next timed kernel launch records these events itself */
void set_launch_events(hipEvent_t a, hipEvent_t b);
bool launch_events_pending();

class Timing {
 public:
  static Timing& get();
  void enable(int mask) { _mask = mask; _on = mask != 0; }
  bool on() const { return _on; }
  void begin(int kind, hipStream_t s);
  void end(int kind, hipStream_t s);
  void collect(int64_t* launches, double* ms, int kinds);
  void reset();

 private:
  struct Rec {
    int kind;
    hipEvent_t a, b;
  };
  bool _on = false;
  int _mask = 0;
  vector<Rec> _recs;
  vector<hipEvent_t> _pool;
  int64_t _launch[8] = {0};
  double _ms[8] = {0};
  hipEvent_t take();
};

}  // namespace gsbn

#endif
