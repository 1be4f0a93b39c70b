"""SolverState files (gsbn.proto:100-152) without a device.

The writer and reader are pinned against an independent protobuf
implementation: the installed protobuf runtime with message classes built
from a descriptor that restates gsbn.proto's SolverState, TableState and
VectorState* messages (field numbers, types, labels, packed options and the
ld default of 1).  Checked:

  * bytes: the library's file equals protobuf's serialisation of the same
    message with the vectors in the reference's order (field number, then the
    Database key name + type suffix, Database.cpp:257-295), whatever order
    the vectors are passed in;
  * values: protobuf parses the library's file back to the same values, and
    the library reads protobuf-written files, also with the repeated data
    unpacked, with unknown fields and with table_state entries present;
  * errors: truncated and malformed files are refused, never half-read.
"""

import os

import numpy as np
import pytest
from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

from gsbn_amd import GsbnError
from gsbn_amd import snapshot as S

F = descriptor_pb2.FieldDescriptorProto
KINDS = [("I8", F.TYPE_INT32), ("I16", F.TYPE_INT32), ("I32", F.TYPE_INT32),
         ("I64", F.TYPE_INT64), ("F16", F.TYPE_FLOAT), ("F32", F.TYPE_FLOAT),
         ("F64", F.TYPE_DOUBLE)]
SUFFIX = ["_i08", "_i16", "_i32", "_i64", "_f16", "_f32", "_f64"]
NP = [np.int8, np.int16, np.int32, np.int64, np.float32, np.float32, np.float64]


def _classes(tag: str, packed: bool = True, extra_field: bool = False):
    """SolverState message classes from a descriptor restating
    gsbn.proto:100-152; `packed` False writes the data unpacked, and
    `extra_field` adds fields the reference's schema does not have."""
    fd = descriptor_pb2.FileDescriptorProto(name="state_%s.proto" % tag, package="t" + tag,
                                            syntax="proto2")

    def field(m, name, num, typ, label=F.LABEL_OPTIONAL, type_name=None, default=None,
              pack=None):
        f = m.field.add(name=name, number=num, type=typ, label=label)
        if type_name:
            f.type_name = type_name
        if default is not None:
            f.default_value = default
        if pack is not None:
            f.options.packed = pack

    ts = fd.message_type.add(name="TableState")
    field(ts, "name", 1, F.TYPE_STRING, F.LABEL_REQUIRED)
    field(ts, "desc", 2, F.TYPE_BYTES, F.LABEL_REQUIRED)
    field(ts, "data", 3, F.TYPE_BYTES, F.LABEL_REQUIRED)
    for kind, typ in KINDS:
        m = fd.message_type.add(name="VectorState" + kind)
        field(m, "name", 1, F.TYPE_STRING, F.LABEL_REQUIRED)
        field(m, "ld", 2, F.TYPE_UINT32, default="1")
        field(m, "data", 3, typ, F.LABEL_REPEATED, pack=packed)
        if extra_field:
            field(m, "note", 15, F.TYPE_STRING)
    st = fd.message_type.add(name="SolverState")
    field(st, "timestamp", 1, F.TYPE_FLOAT, F.LABEL_REQUIRED)
    field(st, "prn", 2, F.TYPE_FLOAT, F.LABEL_REQUIRED)
    field(st, "table_state", 3, F.TYPE_MESSAGE, F.LABEL_REPEATED, ".t%s.TableState" % tag)
    for i, (kind, _) in enumerate(KINDS):
        field(st, "vector_state_" + kind.lower(), 4 + i, F.TYPE_MESSAGE, F.LABEL_REPEATED,
              ".t%s.VectorState%s" % (tag, kind))
    if extra_field:
        field(st, "comment", 40, F.TYPE_STRING)
        field(st, "counter", 41, F.TYPE_FIXED64)
    pool = descriptor_pool.DescriptorPool()
    pool.Add(fd)
    return message_factory.GetMessageClass(pool.FindMessageTypeByName("t%s.SolverState" % tag))


PB = _classes("a")


def _vectors(rng):
    """Vectors of every saved type, with negative integers (ten-byte
    varints), signed zeros, infinities, NaN payloads, an empty vector and
    names whose order differs between name and name + suffix."""
    f32 = rng.normal(0, 1, 300).astype(np.float32)
    f32[:4] = [0.0, -0.0, np.inf, -np.inf]
    f32 = f32.view(np.uint32)
    f32[4] = 0x7FC12345  # a NaN payload survives as bits
    f32 = f32.view(np.float32)
    return {
        "spike_0": rng.integers(-128, 128, 257).astype(np.int8),
        "h16": rng.integers(-32768, 32768, 99).astype(np.int16),
        "ii_0": np.concatenate([np.full(5, -1), rng.integers(-2**31, 2**31, 200),
                                [2**31 - 1, -2**31]]).astype(np.int32),
        "a": rng.integers(0, 1000, 7).astype(np.int32),
        "a_b": rng.integers(0, 1000, 3).astype(np.int32),
        "big": np.array([0, 1, -1, 2**62, -2**63, 2**63 - 1], np.int64),
        "pij_0": f32,
        "empty": np.zeros(0, np.float32),
        "d": rng.normal(0, 1e300, 50).astype(np.float64),
    }


def _canonical(vecs):
    """(type code, name) in the reference's serialisation order."""
    code = {np.dtype(np.int8): 0, np.dtype(np.int16): 1, np.dtype(np.int32): 2,
            np.dtype(np.int64): 3, np.dtype(np.float32): 5, np.dtype(np.float64): 6}
    keys = [(code[a.dtype], n) for n, a in vecs.items()]
    return sorted(keys, key=lambda k: (k[0], k[1] + SUFFIX[k[0]]))


def _message(cls, ts, prn, vecs, lds, order):
    m = cls(timestamp=ts, prn=prn)
    for t, n in order:
        kind = KINDS[t][0].lower()
        v = getattr(m, "vector_state_" + kind).add(name=n, ld=lds.get(n, 0))
        v.data.extend(vecs[n].tolist())
    return m


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))
    return np.array_equal(a, b)


def test_bytes_equal_protobuf_serialisation(tmp_path):
    rng = np.random.default_rng(1)
    vecs = _vectors(rng)
    lds = {"ii_0": 3, "pij_0": 300, "a": 0, "big": 2**32 - 1}
    order = _canonical(vecs)
    want = _message(PB, 1.5, 0.25, vecs, lds, order).SerializeToString()
    names = list(vecs)
    for k in range(3):
        rng.shuffle(names)
        p = str(tmp_path / ("s%d.bin" % k))
        S.write_state(p, 1.5, 0.25, {n: vecs[n] for n in names}, lds)
        assert open(p, "rb").read() == want


def test_protobuf_parses_library_file(tmp_path):
    rng = np.random.default_rng(2)
    vecs = _vectors(rng)
    p = str(tmp_path / "s.bin")
    S.write_state(p, np.float32(0.1), np.float32(1.0), vecs, {"spike_0": 10})
    m = PB()
    m.ParseFromString(open(p, "rb").read())
    assert m.timestamp == np.float32(0.1) and m.prn == 1.0
    got = {}
    for kind, _ in KINDS:
        for v in getattr(m, "vector_state_" + kind.lower()):
            got[v.name] = (np.array(v.data, NP[[k for k, _ in KINDS].index(kind)]), v.ld)
    assert set(got) == set(vecs)
    for n, a in vecs.items():
        assert _same(got[n][0].astype(a.dtype), a), n
        assert got[n][1] == (10 if n == "spike_0" else 0)


@pytest.mark.parametrize("packed,extra", [(True, False), (False, False), (True, True)])
def test_library_reads_protobuf_files(tmp_path, packed, extra):
    rng = np.random.default_rng(3)
    vecs = _vectors(rng)
    cls = PB if (packed and not extra) else _classes("p%d%d" % (packed, extra), packed, extra)
    order = list(reversed(_canonical(vecs)))
    m = _message(cls, 2.0, 0.0, vecs, {"d": 5}, order)
    m.table_state.add(name="tbl", desc=b"\x01\x02", data=b"abc")
    if extra:
        m.comment = "not in the schema"
        m.counter = 7
        m.vector_state_i32[0].note = "x"
    del m.vector_state_f32[:]  # ld absent on a vector: the default 1
    m.vector_state_f32.add(name="pij_0").data.extend(vecs["pij_0"].tolist())
    p = str(tmp_path / "s.bin")
    open(p, "wb").write(m.SerializeToString())
    ts, prn, got = S.read_state(p)
    assert (ts, prn) == (2.0, 0.0)
    assert set(got) == set(vecs) - {"empty"}
    for n in got:
        assert _same(got[n][0], vecs[n]), n
    assert got["d"][1] == 5 and got["pij_0"][1] == 1 and got["a"][1] == 0


def test_half_vectors_read_as_float(tmp_path):
    m = PB(timestamp=0.0, prn=1.0)
    m.vector_state_f16.add(name="h", ld=2, data=[0.5, -2.0, 65504.0])
    p = str(tmp_path / "s.bin")
    open(p, "wb").write(m.SerializeToString())
    _, _, got = S.read_state(p)
    assert _same(got["h"][0], np.array([0.5, -2.0, 65504.0], np.float32)) and got["h"][1] == 2


def test_malformed_files_are_refused(tmp_path):
    rng = np.random.default_rng(4)
    p = str(tmp_path / "s.bin")
    S.write_state(p, 1.0, 1.0, _vectors(rng))
    raw = open(p, "rb").read()
    bad = str(tmp_path / "b.bin")
    for cut in (1, 5, 12, len(raw) // 2, len(raw) - 1):
        open(bad, "wb").write(raw[:cut])
        with pytest.raises(GsbnError, match="Parse file error"):
            S.read_state(bad)
    # no timestamp / prn
    m = PB(timestamp=0.0, prn=0.0)
    m.vector_state_i8.add(name="x", data=[1])
    body = m.SerializeToString()[10:]  # drop fields 1 and 2 (5 bytes each)
    open(bad, "wb").write(body)
    with pytest.raises(GsbnError, match="missing timestamp or prn"):
        S.read_state(bad)
    # a vector without a name
    open(bad, "wb").write(bytes([0x0D, 0, 0, 0, 0, 0x15, 0, 0, 0, 0, 0x22, 2, 0x10, 1]))
    with pytest.raises(GsbnError, match="without a name"):
        S.read_state(bad)
    with pytest.raises(GsbnError, match="not found"):
        S.read_state(str(tmp_path / "missing.bin"))


def test_snapshot_name_follows_to_string_of_float_time():
    # std::to_string(float) prints "%f" of the float product simstep * dt
    assert S.snapshot_name(500, 0.001) == "Snapshot_0.500000.bin"
    assert S.snapshot_name(1000, 0.001) == "Snapshot_1.000000.bin"
    assert S.snapshot_name(3, 0.001) == "Snapshot_0.003000.bin"
