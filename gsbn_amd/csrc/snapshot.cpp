/*
 * snapshot.cpp -- SolverState files and the Database's half of snapshot and
 * resume.
 *
 * A snapshot file is a SolverState message (gsbn.proto:100-152) in protobuf's
 * binary wire encoding, written and parsed here without a protobuf library:
 *
 *   SolverState  1 timestamp (fixed32 float)  2 prn (fixed32 float)
 *                3 table_state (skipped on read, never written)
 *                4..10 vector_state_{i8,i16,i32,i64,f16,f32,f64} (messages)
 *   VectorState  1 name (bytes)  2 ld (varint, uint32)  3 data (packed:
 *                int32 / int64 varints, float fixed32, double fixed64)
 *
 * Fields are written in field-number order and the vectors of one element
 * type in the order of their Database key (name + type suffix, as
 * Database.cpp:257-295 walks its map), so the bytes equal what the
 * reference's state_to_proto + SerializeToOstream produce for the same
 * vectors.  Reading accepts any field order, packed or unpacked repeated
 * data and unknown fields (skipped), like a protobuf parser.  Unlike the
 * reference, files above protobuf's 2 GiB message limit are written and
 * read (the 100x100 dense network's state is 2.4 GB).
 */
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>

#include "framework.h"

namespace gsbn {

namespace {

const char* const kSuffix[7] = {"_i08", "_i16", "_i32", "_i64", "_f16", "_f32", "_f64"};

size_t elem_bytes(DType t) {
  switch (t) {
    case DT_I8: return 1;
    case DT_I16: return 2;
    case DT_I32: return 4;
    case DT_I64: return 8;
    case DT_F16: return 4; /* held as float, like VectorStateF16 */
    case DT_F32: return 4;
    case DT_F64: return 8;
  }
  return 0;
}

bool is_varint_type(DType t) { return t == DT_I8 || t == DT_I16 || t == DT_I32 || t == DT_I64; }

size_t varint_len(uint64_t v) {
  size_t n = 1;
  while (v >= 0x80) {
    v >>= 7;
    ++n;
  }
  return n;
}

/* element i of an integer vector as the int64 whose varint protobuf writes
 * (int32 fields sign-extend to 64 bits) */
inline uint64_t int_elem(DType t, const void* d, size_t i) {
  switch (t) {
    case DT_I8: return (uint64_t)(int64_t)((const int8_t*)d)[i];
    case DT_I16: return (uint64_t)(int64_t)((const int16_t*)d)[i];
    case DT_I32: return (uint64_t)(int64_t)((const int32_t*)d)[i];
    default: return (uint64_t)((const int64_t*)d)[i];
  }
}

size_t payload_len(const VecSpan& v) {
  if (!is_varint_type(v.type)) return v.count * elem_bytes(v.type);
  size_t n = 0;
  for (size_t i = 0; i < v.count; ++i) n += varint_len(int_elem(v.type, v.data, i));
  return n;
}

class Writer {
 public:
  explicit Writer(const string& path) : _path(path) {
    _f = fopen(path.c_str(), "wb");
    if (!_f) throw Error(-2, "Failed to write states: cannot open " + path);
    _buf.reserve(kCap);
  }
  ~Writer() {
    if (_f) fclose(_f);
  }
  void bytes(const void* p, size_t n) {
    const char* c = (const char*)p;
    if (n >= kCap) {
      flush();
      if (fwrite(c, 1, n, _f) != n) fail();
      return;
    }
    if (_buf.size() + n > kCap) flush();
    _buf.insert(_buf.end(), c, c + n);
  }
  void byte(uint8_t b) {
    if (_buf.size() + 1 > kCap) flush();
    _buf.push_back((char)b);
  }
  void varint(uint64_t v) {
    while (v >= 0x80) {
      byte((uint8_t)(v | 0x80));
      v >>= 7;
    }
    byte((uint8_t)v);
  }
  void close() {
    flush();
    if (fclose(_f) != 0) {
      _f = nullptr;
      fail();
    }
    _f = nullptr;
  }

 private:
  static constexpr size_t kCap = 1 << 22;
  void flush() {
    if (!_buf.empty() && fwrite(_buf.data(), 1, _buf.size(), _f) != _buf.size()) fail();
    _buf.clear();
  }
  [[noreturn]] void fail() { throw Error(-2, "Failed to write states to " + _path); }
  string _path;
  FILE* _f = nullptr;
  vector<char> _buf;
};

/* a read-only mapping of a whole file */
struct Mapped {
  const uint8_t* p = nullptr;
  size_t n = 0;
  explicit Mapped(const string& path) {
    int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) throw Error(-2, "File not found, abort! " + path);
    struct stat st;
    if (fstat(fd, &st) != 0) {
      close(fd);
      throw Error(-2, "cannot stat " + path);
    }
    n = (size_t)st.st_size;
    if (n) {
      void* m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
      if (m == MAP_FAILED) {
        close(fd);
        throw Error(-2, "cannot map " + path);
      }
      p = (const uint8_t*)m;
    }
    close(fd);
  }
  ~Mapped() {
    if (p) munmap((void*)p, n);
  }
};

struct Cursor {
  const uint8_t* p;
  size_t pos, end;
  const string* path;
  [[noreturn]] void bad(const char* what) const {
    throw Error(-2, string("Parse file error, abort! ") + what + " in " + *path);
  }
  uint64_t varint() {
    uint64_t v = 0;
    for (int sh = 0; sh < 64; sh += 7) {
      if (pos >= end) bad("truncated varint");
      const uint8_t b = p[pos++];
      v |= (uint64_t)(b & 0x7f) << sh;
      if (!(b & 0x80)) return v;
    }
    bad("varint longer than 10 bytes");
  }
  void need(size_t n) const {
    if (n > end - pos) bad("truncated field");
  }
  void skip(int wt) {
    switch (wt) {
      case 0: (void)varint(); return;
      case 1: need(8); pos += 8; return;
      case 2: {
        const uint64_t n = varint();
        need(n);
        pos += n;
        return;
      }
      case 5: need(4); pos += 4; return;
      default: bad("unsupported wire type");
    }
  }
};

/* append one element, read as protobuf would for the field's type, to the
 * vector's bytes in the Database type (int32 values truncate to int8/int16
 * as the reference's push_back does, Database.cpp:36-58) */
void append_int(VecState& v, uint64_t raw) {
  switch (v.type) {
    case DT_I8: { int8_t x = (int8_t)(int32_t)raw; v.bytes.insert(v.bytes.end(), (uint8_t*)&x, (uint8_t*)&x + 1); break; }
    case DT_I16: { int16_t x = (int16_t)(int32_t)raw; v.bytes.insert(v.bytes.end(), (uint8_t*)&x, (uint8_t*)&x + 2); break; }
    case DT_I32: { int32_t x = (int32_t)raw; v.bytes.insert(v.bytes.end(), (uint8_t*)&x, (uint8_t*)&x + 4); break; }
    default: { int64_t x = (int64_t)raw; v.bytes.insert(v.bytes.end(), (uint8_t*)&x, (uint8_t*)&x + 8); break; }
  }
}

VecState parse_vector(Cursor c, DType type) {
  VecState v;
  v.type = type;
  v.ld = 1; /* [default=1] */
  bool has_name = false;
  const size_t eb = elem_bytes(type);
  while (c.pos < c.end) {
    const uint64_t tag = c.varint();
    const int fno = (int)(tag >> 3), wt = (int)(tag & 7);
    if (fno == 1 && wt == 2) {
      const uint64_t n = c.varint();
      c.need(n);
      v.name.assign((const char*)c.p + c.pos, (size_t)n);
      c.pos += n;
      has_name = true;
    } else if (fno == 2 && wt == 0) {
      v.ld = (uint32_t)c.varint();
    } else if (fno == 3 && is_varint_type(type) && (wt == 0 || wt == 2)) {
      if (wt == 0) {
        append_int(v, c.varint());
      } else {
        const uint64_t n = c.varint();
        c.need(n);
        Cursor s{c.p, c.pos, c.pos + (size_t)n, c.path};
        if (v.bytes.empty()) v.bytes.reserve((size_t)n * eb);
        while (s.pos < s.end) append_int(v, s.varint());
        c.pos += n;
      }
    } else if (fno == 3 && !is_varint_type(type) && (wt == 2 || wt == (eb == 8 ? 1 : 5))) {
      size_t n = eb;
      if (wt == 2) {
        n = (size_t)c.varint();
        if (n % eb) c.bad("packed data length not a multiple of the element size");
      }
      c.need(n);
      v.bytes.insert(v.bytes.end(), c.p + c.pos, c.p + c.pos + n);
      c.pos += n;
    } else {
      c.skip(wt);
    }
  }
  if (!has_name) c.bad("vector state without a name");
  return v;
}

}  // namespace

size_t VecState::count() const { return bytes.size() / elem_bytes(type); }

void write_solver_state(const string& path, float timestamp, float prn, vector<VecSpan> vecs) {
  for (auto& v : vecs) {
    if ((int)v.type < 0 || (int)v.type > 6) throw Error(-2, "snapshot: bad element type of " + v.name);
    if (v.count && !v.data) throw Error(-2, "snapshot: no data for " + v.name);
  }
  /* field-number order, then the Database's key order within a type */
  std::stable_sort(vecs.begin(), vecs.end(), [](const VecSpan& a, const VecSpan& b) {
    if (a.type != b.type) return a.type < b.type;
    return a.name + kSuffix[a.type] < b.name + kSuffix[b.type];
  });
  Writer w(path);
  w.byte((1 << 3) | 5);
  w.bytes(&timestamp, 4);
  w.byte((2 << 3) | 5);
  w.bytes(&prn, 4);
  for (auto& v : vecs) {
    const uint64_t ld = (uint64_t)(uint32_t)v.ld;
    const size_t pl = v.count ? payload_len(v) : 0;
    size_t sub = 1 + varint_len(v.name.size()) + v.name.size() + 1 + varint_len(ld);
    if (v.count) sub += 1 + varint_len(pl) + pl;
    w.varint(((uint64_t)(4 + (int)v.type) << 3) | 2);
    w.varint(sub);
    w.byte((1 << 3) | 2);
    w.varint(v.name.size());
    w.bytes(v.name.data(), v.name.size());
    w.byte((2 << 3) | 0);
    w.varint(ld);
    if (!v.count) continue;
    w.byte((3 << 3) | 2);
    w.varint(pl);
    if (is_varint_type(v.type)) {
      for (size_t i = 0; i < v.count; ++i) w.varint(int_elem(v.type, v.data, i));
    } else {
      /* little-endian host: the in-memory floats are the wire bytes */
      w.bytes(v.data, pl);
    }
  }
  w.close();
}

SolverStateData read_solver_state(const string& path) {
  Mapped m(path);
  Cursor c{m.p, 0, m.n, &path};
  SolverStateData st;
  bool has_ts = false, has_prn = false;
  while (c.pos < c.end) {
    const uint64_t tag = c.varint();
    const int fno = (int)(tag >> 3), wt = (int)(tag & 7);
    if ((fno == 1 || fno == 2) && wt == 5) {
      c.need(4);
      float f;
      memcpy(&f, c.p + c.pos, 4);
      c.pos += 4;
      if (fno == 1) {
        st.timestamp = f;
        has_ts = true;
      } else {
        st.prn = f;
        has_prn = true;
      }
    } else if (fno >= 4 && fno <= 10 && wt == 2) {
      const uint64_t n = c.varint();
      c.need(n);
      st.vecs.push_back(parse_vector(Cursor{c.p, c.pos, c.pos + (size_t)n, &path}, (DType)(fno - 4)));
      c.pos += n;
    } else {
      c.skip(wt);
    }
  }
  if (!has_ts || !has_prn) c.bad("missing timestamp or prn");
  return st;
}

/* ------------------------------------------------------------------ */
/* Database side */

template <typename T>
static void restore_into(SyncVector<T>* v, const VecState& s) {
  const size_t n = s.count();
  v->resize(n);
  if (n) memcpy(v->mutable_cpu_data(), s.bytes.data(), n * sizeof(T));
  v->set_ld((int)s.ld);
}

void Database::init_copy(const SolverStateData& st) {
  for (const VecState& s : st.vecs) {
    VecBase* made = nullptr;
    switch (s.type) {
      case DT_I8: { auto* v = create<int8_t>(s.name); if (v) restore_into(v, s); made = v; break; }
      case DT_I16: { auto* v = create<int16_t>(s.name); if (v) restore_into(v, s); made = v; break; }
      case DT_I32: { auto* v = create<int32_t>(s.name); if (v) restore_into(v, s); made = v; break; }
      case DT_I64: { auto* v = create<int64_t>(s.name); if (v) restore_into(v, s); made = v; break; }
      case DT_F32: { auto* v = create<float>(s.name); if (v) restore_into(v, s); made = v; break; }
      case DT_F64: { auto* v = create<double>(s.name); if (v) restore_into(v, s); made = v; break; }
      case DT_F16:
        throw Error(-5, "snapshot vector " + s.name + ": half-precision vectors are not supported");
    }
    /* a name saved twice: the reference's create returns NULL and its
     * CHECK aborts (Database.cpp:40) */
    if (!made) throw Error(-4, "check failed: snapshot holds vector " + s.name + " twice");
    _restored.insert(s.name);
  }
}

void Database::save_state(const string& path, float timestamp, float prn) {
  vector<VecSpan> spans;
  for (auto& kv : _vecs) {
    VecBase* v = kv.second.get();
    if (kv.first.empty() || kv.first[0] == '.') continue; /* Database.cpp:261 */
    VecSpan s;
    s.name = kv.first;
    s.type = v->type;
    s.ld = v->ld_;
    s.count = v->size();
    s.data = s.count ? v->host_ptr_sync() : nullptr;
    spans.push_back(s);
  }
  write_solver_state(path, timestamp, prn, std::move(spans));
}

}  // namespace gsbn
