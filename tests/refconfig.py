"""Reader of the reference's SolverParam text format into a NetSpec for the
oracle (test infrastructure).

The product parses the same files in C++ (framework.cpp); this is a second,
independent reading in Python so that a test can run the reference's own
configuration file through both and compare.  It covers the subset of
gsbn.proto:4-98 that the shipped data/10x10_rnd.prototxt uses, with the
schema's defaults (gsbn.proto:19-31 for mode_param, :41-58 pop_param,
:61-74 proj_param), and decodes the StimRawData file with the schema of
gsbn.proto:153-160 (protobuf wire format, read here field by field).
"""

from __future__ import annotations

import re
import struct
from typing import Dict, List

import numpy as np

from gsbn_amd import configs

_TOKEN = re.compile(r'"(?:[^"\\]|\\.)*"|#[^\n]*|[{}:]|[^\s{}:#"]+')


def _tokens(text: str) -> List[str]:
    """Strings, braces, colons and bare words; comments (# to the end of the
    line, outside strings) dropped."""
    return [t for t in _TOKEN.findall(text) if not t.startswith("#")]


def parse_text(text: str) -> Dict[str, list]:
    """Protobuf text format to nested dicts; every field is a list of its
    occurrences (scalars as strings, strings unquoted)."""
    toks = _tokens(text)
    pos = 0

    def message(end_brace: bool):
        nonlocal pos
        msg: Dict[str, list] = {}
        while pos < len(toks):
            t = toks[pos]
            if t == "}":
                if not end_brace:
                    raise ValueError("unbalanced }")
                pos += 1
                return msg
            name = t
            pos += 1
            if toks[pos] == ":":
                pos += 1
            if toks[pos] == "{":
                pos += 1
                val = message(True)
            else:
                val = toks[pos]
                if val.startswith('"'):
                    val = val[1:-1]
                pos += 1
            msg.setdefault(name, []).append(val)
        if end_brace:
            raise ValueError("missing }")
        return msg

    return message(False)


def _get(m, key, default=None, conv=float):
    v = m.get(key)
    return conv(v[-1]) if v else default


def decode_stim(buf: bytes):
    """StimRawData (gsbn.proto:153-160): 1 data_rows, 2 data_cols, 3 packed
    uint32 data, 4 mask_rows, 5 mask_cols, 6 packed float mask."""
    pos = 0
    fields: Dict[int, list] = {}

    def varint():
        nonlocal pos
        v, sh = 0, 0
        while True:
            b = buf[pos]
            pos += 1
            v |= (b & 0x7F) << sh
            sh += 7
            if not b & 0x80:
                return v

    while pos < len(buf):
        key = varint()
        num, wt = key >> 3, key & 7
        if wt == 0:
            fields.setdefault(num, []).append(varint())
        elif wt == 2:
            n = varint()
            chunk = buf[pos:pos + n]
            pos += n
            if num == 3:
                p2, vals = 0, []
                while p2 < len(chunk):
                    v, sh = 0, 0
                    while True:
                        b = chunk[p2]
                        p2 += 1
                        v |= (b & 0x7F) << sh
                        sh += 7
                        if not b & 0x80:
                            break
                    vals.append(v)
                fields.setdefault(3, []).extend(vals)
            elif num == 6:
                fields.setdefault(6, []).extend(struct.unpack("<%df" % (n // 4), chunk))
        elif wt == 5:
            fields.setdefault(num, []).append(struct.unpack("<f", buf[pos:pos + 4])[0])
            pos += 4
        else:
            raise ValueError("unexpected wire type %d" % wt)
    dr, dc = fields[1][0], fields[2][0]
    mr, mc = fields[4][0], fields[5][0]
    data = np.array(fields.get(3, []), np.uint32).reshape(dr, dc)
    mask = np.array(fields.get(6, []), np.float32).reshape(mr, mc)
    return data, mask


def spec_from_prototxt(text: str, stim_bytes: bytes) -> configs.NetSpec:
    """The NetSpec the oracle runs for a reference configuration file: pops,
    projections, modes, stimulus rows and, when listed, the
    ProcStructuralPlasticity arguments; no connection map (the shipped file
    lists no ProcInitConn, so every slot starts empty)."""
    sp = parse_text(text)
    gen = sp["gen_param"][0]
    dt = _get(gen, "dt")
    modes = []
    for m in gen.get("mode_param", []):
        modes.append(configs.ModeSpec(
            begin_time=_get(m, "begin_time"), end_time=_get(m, "end_time"),
            time_step=_get(m, "time_step", 1, int),
            begin_lgidx_id=_get(m, "begin_lgidx_id", 0, int),
            begin_lgexp_id=_get(m, "begin_lgexp_id", 0, int),
            begin_wmask_id=_get(m, "begin_wmask_id", 0, int),
            lgidx_step=_get(m, "lgidx_step", 1, int), lgexp_step=_get(m, "lgexp_step", 1, int),
            wmask_step=_get(m, "wmask_step", 1, int), prn=_get(m, "prn", 1.0),
            plasticity=_get(m, "plasticity", 1, int)))
    net = sp["net_param"][0]
    pops = []
    for p in net.get("pop_param", []):
        for _ in range(_get(p, "pop_num", 1, int)):
            pops.append(configs.PopSpec(
                hcu=_get(p, "hcu_num", conv=int), mcu=_get(p, "mcu_num", conv=int),
                taum=_get(p, "taum", 0.01), wtagain=_get(p, "wtagain", 4.0),
                maxfq=_get(p, "maxfq", 100.0), igain=_get(p, "igain", 1.0),
                wgain=_get(p, "wgain", 1.0), lgbias=_get(p, "lgbias", 0.0),
                snoise=_get(p, "snoise", 0.0), fanout=_get(p, "fanout_num", conv=int),
                shape=tuple(int(v) for v in p.get("shape", [])),
                position=tuple(int(v) for v in p.get("position", []))))
    projs = []
    for q in net.get("proj_param", []):
        projs.append(configs.ProjSpec(
            src=_get(q, "src_pop", conv=int), dest=_get(q, "dest_pop", conv=int),
            tauzi=_get(q, "tauzi"), tauzj=_get(q, "tauzj"), taue=_get(q, "taue"),
            taup=_get(q, "taup"), maxfq=_get(q, "maxfq", 100.0), bgain=_get(q, "bgain", 0.0),
            wgain=_get(q, "wgain", 0.0), tauepsc=_get(q, "tauepsc", None),
            slot_num=_get(q, "slot_num")))
    stim, mask = decode_stim(stim_bytes)
    spec = configs.NetSpec(dt=dt, pops=pops, projs=projs, modes=modes, stim=stim, mask=mask,
                           conn=[[] for _ in projs], eps=_get(gen, "eps", 0.001), name="ref")
    for proc in sp.get("proc_param", []):
        name = proc["name"][0]
        argi = {a["key"][0]: int(a["val"][0]) for a in proc.get("argi", [])}
        argf = {a["key"][0]: float(a["val"][0]) for a in proc.get("argf", [])}
        if name == "ProcStructuralPlasticity":
            spec.structural = configs.StructuralSpec(
                period=argi.get("period", 100), pruning_period=argi.get("pruning-period", 1000),
                t_th=argi.get("t-th", 100), wp=argf.get("wp", 0.5), wn=argf.get("wn", -0.5),
                enable_geometry=argi.get("enable-geometry", 0),
                enable_init_conn=argi.get("enable-init-conn", 0),
                d_norm=argf.get("d-norm", 0.75), v_cond=argf.get("v-cond", 0.2),
                seed=(argi.get("seed-lo", 0) & 0x7FFFFFFF) |
                     ((argi.get("seed-hi", 0) & 0x7FFFFFFF) << 31))
        elif name == "ProcUpdLazy":
            spec.seed = ((argi.get("seed-lo", 0) & 0x7FFFFFFF) |
                         ((argi.get("seed-hi", 0) & 0x7FFFFFFF) << 31))
            spec.wta_pop_max = argi.get("wta-pop-max", 0)
        elif name == "ProcCheck":
            spec.check_threshold = argi.get("threashold", 0)
    return spec
