"""Network descriptions for the lazy-update path: prototxt, stimulus and
connection-map writers, and the benchmark workloads.

The files follow the reference's formats:

* the ``SolverParam`` text format of ``src/gsbn/proto/gsbn.proto:4-98``
  (``gen_param``/``net_param``/``rec_param``/``proc_param``);
* the ``StimRawData`` protobuf wire format (``gsbn.proto:153-160``):
  ``data_rows, data_cols, packed uint32 data, mask_rows, mask_cols, packed
  float mask``; ``0x7fffffff`` marks an unclamped hypercolumn, as written by
  ``tools/stimgen/gen_rnd.py:61-78``;
* the connection map read by ``ProcInitConn`` (``ProcInitConn.cpp:48-75``):
  ``prj_id,dest_hcu,src_mcu,delay[,src_mcu,delay...]``, slots filled in file
  order.

Workloads (times in steps of dt = 1 ms):

``recall_net(hcu, mcu, ...)`` builds one population with a dense recurrent
projection without self-hypercolumn connections (slot k of destination
hypercolumn h holds the k-th source unit outside h, ascending, delay 1),
``n_pat`` random one-MCU-per-HCU patterns trained for ``train`` steps each
followed by ``quiet`` unclamped steps, then ``n_pat`` cues with a fraction of
the hypercolumns clamped, recalled for ``recall`` steps each with the
recurrent support enabled (``wmask = 1``) on the unclamped hypercolumns and
scored by ProcCheck.
"""

from __future__ import annotations

import dataclasses
import os
import struct
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

UNCLAMPED = 0x7FFFFFFF


@dataclasses.dataclass
class PopSpec:
    hcu: int
    mcu: int
    taum: float = 0.01
    wtagain: float = 4.0
    maxfq: float = 100.0
    igain: float = 1.0
    wgain: float = 1.0
    lgbias: float = 0.0
    snoise: float = 0.0
    fanout: Optional[int] = None  # fanout_num (structural plasticity); default mcu
    shape: Tuple[int, ...] = ()     # hypercolumn grid and offset (geometric delays)
    position: Tuple[int, ...] = ()


@dataclasses.dataclass
class ProjSpec:
    src: int
    dest: int
    tauzi: float
    tauzj: float
    taue: float
    taup: float
    maxfq: float = 100.0
    bgain: float = 1.0
    wgain: float = 1.0
    tauepsc: Optional[float] = None
    slot_num: float = 0.0


@dataclasses.dataclass
class ModeSpec:
    begin_time: float
    end_time: float
    time_step: int = 1
    begin_lgidx_id: int = 0
    begin_lgexp_id: int = 0
    begin_wmask_id: int = 0
    lgidx_step: int = 1
    lgexp_step: int = 1
    wmask_step: int = 1
    prn: float = 1.0
    plasticity: int = 0


@dataclasses.dataclass
class StructuralSpec:
    """ProcStructuralPlasticity's arguments (ProcStructuralPlasitcity.cpp:17-66
    defaults); `seed` keys its draws."""
    period: int = 100
    pruning_period: int = 1000
    t_th: int = 100
    wp: float = 0.5
    wn: float = -0.5
    enable_geometry: int = 0
    enable_init_conn: int = 0
    d_norm: float = 0.75
    v_cond: float = 0.2
    seed: int = 0


@dataclasses.dataclass
class NetSpec:
    dt: float
    pops: List[PopSpec]
    projs: List[ProjSpec]
    modes: List[ModeSpec]
    stim: np.ndarray  # uint32 [rows, total_hcu]
    mask: np.ndarray  # float32 [rows, total_hcu]
    conn: List[List[Tuple[int, List[Tuple[int, int]]]]]  # per proj: (dest_hcu, [(src, delay)])
    eps: float = 0.001
    seed: int = 0
    wta_pop_max: int = 0
    check_threshold: int = 0
    ext_spikes: Optional[Dict[int, Dict[int, List[int]]]] = None  # pop -> step -> units
    name: str = "net"
    # False: ProcExchangeSpike copies spikes into si_<proj> every step and the
    # delay-queue push runs as its own kernel (argi direct-history 0)
    exchange_direct: bool = True
    conn_text: Optional[str] = None  # verbatim connection-map text (overrides conn)
    # per projection (ii, di) arrays of shape [H*C], written through the
    # Database API instead of a connection-map file (no ProcInitConn)
    conn_arrays: Optional[List[Tuple[np.ndarray, np.ndarray]]] = None
    trace_precision: str = "fp32"  # storage of Pij/Eij: fp32, bf16 or fp16
    # (offset, period) of ProcSnapshot / ProcSpkRec at the end of the list
    snapshot: Optional[Tuple[int, int]] = None
    spike_record: Optional[Tuple[int, int]] = None
    structural: Optional[StructuralSpec] = None  # ProcStructuralPlasticity after the exchange
    # ProcUpdLazy step plan: fused_step 1 runs the population and row/column
    # updates as one launch per step; rowcol_parts > 0 fixes the split
    # launches' work split; both 0: chosen online
    fused_step: int = 0
    rowcol_parts: int = 0
    # chained_step 1: one launch per step that runs the previous step's
    # row/column updates with this step's population update; -1 never;
    # 0: one of the plans chosen online
    chained_step: int = 0
    # False: ProcExtGen builds one stimulus row per mode change into a
    # one-row .lginp instead of keeping every row resident (argi
    # resident-stimulus 0); the form it falls back to beyond 2^28 floats
    resident_stimulus: bool = True
    # rowcol_graded g (1-3): batches split their row/column work per size
    # bucket of hypercolumns (plan -10-g); 0: chosen online
    rowcol_graded: int = 0
    # rowcol_wave 1: batches run their row/column update with one wave per
    # hypercolumn (plan -20); 2-4: one wave per part of the graded split
    # g = rowcol_wave - 1 (plans -21 to -23); 0: chosen online
    rowcol_wave: int = 0
    # pop_wave 0: batches run their population update with the block kernel;
    # None: ProcUpdLazy's default (one wave per hypercolumn where every
    # population qualifies)
    pop_wave: Optional[int] = None
    # ovf_keep / ovf_cap: the one-wave row/column plan's row overflow queue
    # (rows a hypercolumn keeps for its own wave, queued rows per step at
    # most; ovf_keep 0: no queue); None: ProcUpdLazy's defaults (2, 2^18)
    ovf_keep: Optional[int] = None
    ovf_cap: Optional[int] = None
    # deferred_columns K (a power of two, 2..64): column updates deferred
    # through rings of K spike steps per postsynaptic unit; 0: eager; None:
    # ProcUpdLazy's default (where it applies: 8 for one network, 16 for a batch)
    deferred_columns: Optional[int] = None

    @property
    def total_hcu(self) -> int:
        return sum(p.hcu for p in self.pops)

    @property
    def total_units(self) -> int:
        return sum(p.hcu * p.mcu for p in self.pops)


def step_time(step: int, dt: float) -> float:
    """A time in seconds that ProcExtGen maps to ``step`` (ProcExtGen.cpp:22-23
    truncates ``float(time)/float(dt)``; half a step keeps the quotient away
    from an integer)."""
    return (step + 0.5) * dt


def _varint(v: int) -> bytes:
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def encode_stim(stim: np.ndarray, mask: np.ndarray) -> bytes:
    """StimRawData wire encoding (gsbn.proto:153-160), fields in number order."""
    stim = np.asarray(stim, dtype=np.uint32)
    mask = np.asarray(mask, dtype=np.float32)
    out = bytearray()
    out += b"\x08" + _varint(stim.shape[0])
    out += b"\x10" + _varint(stim.shape[1])
    payload = b"".join(_varint(int(v)) for v in stim.reshape(-1))
    out += b"\x1a" + _varint(len(payload)) + payload
    out += b"\x20" + _varint(mask.shape[0])
    out += b"\x28" + _varint(mask.shape[1])
    fpay = mask.reshape(-1).astype("<f4").tobytes()
    out += b"\x32" + _varint(len(fpay)) + fpay
    return bytes(out)


def decode_stim(buf: bytes):
    """Inverse of encode_stim for the field subset StimRawData uses."""
    pos = 0
    fields: Dict[int, object] = {}
    data: List[int] = []
    mask: List[float] = []

    def rd_varint(p):
        shift = 0
        val = 0
        while True:
            b = buf[p]
            p += 1
            val |= (b & 0x7F) << shift
            shift += 7
            if not b & 0x80:
                return val, p

    while pos < len(buf):
        tag, pos = rd_varint(pos)
        fno, wt = tag >> 3, tag & 7
        if wt == 0:
            v, pos = rd_varint(pos)
            if fno == 3:
                data.append(v)
            else:
                fields[fno] = v
        elif wt == 2:
            ln, pos = rd_varint(pos)
            end = pos + ln
            if fno == 3:
                while pos < end:
                    v, pos = rd_varint(pos)
                    data.append(v)
            elif fno == 6:
                mask.extend(struct.unpack("<%df" % (ln // 4), buf[pos:end]))
                pos = end
            else:
                pos = end
        elif wt == 5:
            if fno == 6:
                mask.append(struct.unpack("<f", buf[pos:pos + 4])[0])
            pos += 4
        else:
            raise ValueError("unsupported wire type %d" % wt)
    drows, dcols = fields.get(1, 0), fields.get(2, 0)
    mrows, mcols = fields.get(4, 0), fields.get(5, 0)
    return (np.array(data, dtype=np.uint32).reshape(drows, dcols),
            np.array(mask, dtype=np.float32).reshape(mrows, mcols))


def _fmt(v: float) -> str:
    return repr(float(v))


def prototxt(spec: NetSpec, stim_path: str, conn_path: str, rec_dir: str,
             ext_spike_path: Optional[str] = None) -> str:
    """SolverParam text (gsbn.proto:4-98) with the procedure list
    ProcExtGen, ProcUpdLazy, ProcInitConn, ProcExchangeSpike,
    [ProcStructuralPlasticity], ProcCheck, [ProcSpkRec], [ProcSnapshot]."""
    L = []
    L.append("gen_param {")
    L.append('  stim_file: "%s"' % stim_path)
    L.append("  dt: %s" % _fmt(spec.dt))
    L.append("  eps: %s" % _fmt(spec.eps))
    for m in spec.modes:
        L.append("  mode_param {")
        for f in dataclasses.fields(m):
            v = getattr(m, f.name)
            L.append("    %s: %s" % (f.name, _fmt(v) if isinstance(v, float) else v))
        L.append("  }")
    L.append("}")
    L.append("net_param {")
    for p in spec.pops:
        L.append("  pop_param {")
        L.append("    pop_num: 1")
        L.append("    hcu_num: %d" % p.hcu)
        L.append("    mcu_num: %d" % p.mcu)
        L.append("    fanout_num: %d" % (p.mcu if p.fanout is None else p.fanout))
        for v in p.position:
            L.append("    position: %d" % v)
        for v in p.shape:
            L.append("    shape: %d" % v)
        for f in ("taum", "wtagain", "maxfq", "igain", "wgain", "lgbias", "snoise"):
            L.append("    %s: %s" % (f, _fmt(getattr(p, f))))
        L.append("    rank: 0")
        L.append("  }")
    for q in spec.projs:
        L.append("  proj_param {")
        L.append("    src_pop: %d" % q.src)
        L.append("    dest_pop: %d" % q.dest)
        for f in ("tauzi", "tauzj", "taue", "taup", "maxfq", "bgain", "wgain"):
            L.append("    %s: %s" % (f, _fmt(getattr(q, f))))
        if q.tauepsc is not None:
            L.append("    tauepsc: %s" % _fmt(q.tauepsc))
        L.append("    slot_num: %s" % _fmt(q.slot_num))
        L.append("  }")
    L.append("}")
    L.append('rec_param { directory: "%s" }' % rec_dir)
    if spec.resident_stimulus:
        L.append('proc_param { name: "ProcExtGen" }')
    else:
        L.append('proc_param { name: "ProcExtGen" argi { key: "resident-stimulus" val: 0 } }')
    L.append("proc_param {")
    L.append('  name: "ProcUpdLazy"')
    L.append('  argi { key: "seed-lo" val: %d }' % (spec.seed & 0x7FFFFFFF))
    L.append('  argi { key: "seed-hi" val: %d }' % ((spec.seed >> 31) & 0x7FFFFFFF))
    L.append('  argi { key: "wta-pop-max" val: %d }' % spec.wta_pop_max)
    if ext_spike_path:
        L.append('  args { key: "external spike" val: "%s" }' % ext_spike_path)
    if spec.trace_precision != "fp32":
        L.append('  args { key: "trace-precision" val: "%s" }' % spec.trace_precision)
    if spec.fused_step:
        L.append('  argi { key: "fused-step" val: 1 }')
    if spec.rowcol_parts:
        L.append('  argi { key: "rowcol-parts" val: %d }' % spec.rowcol_parts)
    if spec.chained_step:
        L.append('  argi { key: "chained-step" val: %d }' % spec.chained_step)
    if spec.rowcol_graded:
        L.append('  argi { key: "rowcol-graded" val: %d }' % spec.rowcol_graded)
    if spec.rowcol_wave:
        L.append('  argi { key: "rowcol-wave" val: %d }' % spec.rowcol_wave)
    if spec.deferred_columns is not None:
        L.append('  argi { key: "deferred-columns" val: %d }' % spec.deferred_columns)
    if spec.pop_wave is not None:
        L.append('  argi { key: "pop-wave" val: %d }' % spec.pop_wave)
    if spec.ovf_keep is not None:
        L.append('  argi { key: "ovf-keep" val: %d }' % spec.ovf_keep)
    if spec.ovf_cap is not None:
        L.append('  argi { key: "ovf-cap" val: %d }' % spec.ovf_cap)
    L.append("}")
    if spec.conn_arrays is None:
        L.append("proc_param {")
        L.append('  name: "ProcInitConn"')
        L.append('  args { key: "conn-map-file" val: "%s" }' % conn_path)
        L.append("}")
    if spec.exchange_direct:
        L.append('proc_param { name: "ProcExchangeSpike" }')
    else:
        L.append('proc_param { name: "ProcExchangeSpike" argi { key: "direct-history" val: 0 } }')
    sp = spec.structural
    if sp is not None:
        L.append("proc_param {")
        L.append('  name: "ProcStructuralPlasticity"')
        for key, val in (("period", sp.period), ("pruning-period", sp.pruning_period),
                         ("t-th", sp.t_th), ("enable-geometry", sp.enable_geometry),
                         ("enable-init-conn", sp.enable_init_conn),
                         ("seed-lo", sp.seed & 0x7FFFFFFF), ("seed-hi", (sp.seed >> 31) & 0x7FFFFFFF)):
            L.append('  argi { key: "%s" val: %d }' % (key, val))
        for key, val in (("wp", sp.wp), ("wn", sp.wn), ("d-norm", sp.d_norm), ("v-cond", sp.v_cond)):
            L.append('  argf { key: "%s" val: %s }' % (key, _fmt(val)))
        L.append("}")
    L.append("proc_param {")
    L.append('  name: "ProcCheck"')
    L.append('  argi { key: "threashold" val: %d }' % spec.check_threshold)
    L.append("}")
    for name, sched in (("ProcSpkRec", spec.spike_record), ("ProcSnapshot", spec.snapshot)):
        if sched is not None:
            L.append('proc_param { name: "%s" argi { key: "offset" val: %d } '
                     'argi { key: "period" val: %d } }' % (name, sched[0], sched[1]))
    return "\n".join(L) + "\n"


def conn_map_text(spec: NetSpec) -> str:
    if spec.conn_text is not None:
        return spec.conn_text
    lines = []
    for pid, rows in enumerate(spec.conn):
        for dest_hcu, pairs in rows:
            if not pairs:
                continue
            parts = [str(pid), str(dest_hcu)]
            for s, d in pairs:
                parts.append(str(s))
                parts.append(str(d))
            lines.append(",".join(parts))
    return "\n".join(lines) + "\n"


def write(spec: NetSpec, directory: str) -> str:
    """Write prototxt + stimulus + connection map (+ external spikes) into
    ``directory``; return the prototxt path."""
    os.makedirs(directory, exist_ok=True)
    stim_path = os.path.join(directory, spec.name + "_stim.bin")
    conn_path = os.path.join(directory, spec.name + "_conn.map")
    rec_dir = os.path.join(directory, spec.name + "_rec")
    with open(stim_path, "wb") as f:
        f.write(encode_stim(spec.stim, spec.mask))
    if spec.conn_arrays is None:
        with open(conn_path, "w") as f:
            f.write(conn_map_text(spec))
    ext_path = None
    if spec.ext_spikes:
        ext_path = os.path.join(directory, spec.name + "_ext.csv")
        with open(ext_path, "w") as f:
            for pop, steps in spec.ext_spikes.items():
                for step in sorted(steps):
                    units = steps[step]
                    f.write(",".join([str(step), str(pop)] + [str(u) for u in units]) + "\n")
    path = os.path.join(directory, spec.name + ".prototxt")
    with open(path, "w") as f:
        f.write(prototxt(spec, stim_path, conn_path, rec_dir, ext_path))
    return path


def parse_conn_map(text: str, nproj: int = 1):
    """Connection-map text -> per projection [(dest_hcu, [(src, delay), ...])]."""
    out: List[List[Tuple[int, List[Tuple[int, int]]]]] = [[] for _ in range(nproj)]
    for line in text.splitlines():
        v = line.strip().split(",")
        if len(v) <= 2:
            continue
        pid, h = int(v[0]), int(v[1])
        pairs = [(int(v[i]), int(v[i + 1])) for i in range(2, len(v) - 1, 2)]
        out[pid].append((h, pairs))
    return out


def dense_conn(hcu: int, mcu: int, delay: int = 1, self_hcu: bool = False):
    """Dense recurrent map: destination HCU h receives every source unit outside
    h (or every unit if self_hcu), ascending, all with the same delay."""
    rows = []
    n = hcu * mcu
    for h in range(hcu):
        pairs = [(s, delay) for s in range(n) if self_hcu or s // mcu != h]
        rows.append((h, pairs))
    return rows


def random_patterns(rng: np.random.Generator, n_pat: int, hcu: int, mcu: int) -> np.ndarray:
    """Distinct one-MCU-per-HCU patterns (tools/stimgen/gen_rnd.py:38-58)."""
    pats: List[List[int]] = []
    seen = set()
    while len(pats) < n_pat:
        p = [int(v) for v in rng.integers(0, mcu, size=hcu)]
        t = tuple(p)
        if t in seen:
            continue
        seen.add(t)
        pats.append(p)
    return np.array(pats, dtype=np.uint32)


def recall_net(hcu: int, mcu: int, n_pat: int, train: int, quiet: int, recall: int,
               clamp_frac: float, *, seed: int = 0, stim_seed: int = 0, snoise: float = 0.1,
               wtagain: float = 4.0, pop_wgain: float = 8.0, tauzi: float = 0.01,
               tauzj: float = 0.01,
               taue: float = 0.02, taup: float = 1.0, tauepsc: Optional[float] = None,
               delay: int = 1, dt: float = 0.001, reset: int = 0, reset_prn: float = 1e-30,
               name: str = "net") -> NetSpec:
    rng = np.random.Generator(np.random.Philox(stim_seed))
    pats = random_patterns(rng, n_pat, hcu, mcu)
    quiet_row = np.full((1, hcu), UNCLAMPED, dtype=np.uint32)
    cues = pats.copy()
    masks = np.zeros((n_pat, hcu), dtype=np.float32)
    n_clamp = int(round(clamp_frac * hcu))
    for i in range(n_pat):
        clamped = set(int(v) for v in rng.permutation(hcu)[:n_clamp])
        for h in range(hcu):
            if h not in clamped:
                cues[i, h] = UNCLAMPED
                masks[i, h] = 1.0
    stim = np.concatenate([pats, quiet_row, cues], axis=0)
    mask = np.concatenate([np.zeros((1, hcu), dtype=np.float32), masks], axis=0)
    modes = recall_schedule(n_pat, train, quiet, recall, reset, reset_prn, dt)
    pop = PopSpec(hcu=hcu, mcu=mcu, wtagain=wtagain, wgain=pop_wgain, snoise=snoise)
    proj = ProjSpec(src=0, dest=0, tauzi=tauzi, tauzj=tauzj, taue=taue, taup=taup,
                    tauepsc=tauepsc, slot_num=float((hcu - 1) * mcu))
    return NetSpec(dt=dt, pops=[pop], projs=[proj], modes=modes, stim=stim, mask=mask,
                   conn=[dense_conn(hcu, mcu, delay)], seed=seed, name=name)


def recall_schedule(n_pat: int, train: int, quiet: int, recall: int, reset: int = 0,
                    reset_prn: float = 1e-30, dt: float = 0.001) -> List[ModeSpec]:
    """Stimulus rows: 0..n_pat-1 patterns, n_pat quiet, n_pat+1.. cues; mask
    rows: 0 training, 1.. one per cue."""
    modes: List[ModeSpec] = []
    t = 0
    for i in range(n_pat):
        modes.append(ModeSpec(step_time(t, dt), step_time(t + train, dt), time_step=train,
                              begin_lgidx_id=i, begin_wmask_id=0, prn=1.0))
        t += train
        if quiet > 0:
            modes.append(ModeSpec(step_time(t, dt), step_time(t + quiet, dt), time_step=quiet,
                                  begin_lgidx_id=n_pat, begin_wmask_id=0, prn=1.0))
            t += quiet
    if reset <= 0:
        modes.append(ModeSpec(step_time(t, dt), step_time(t + n_pat * recall, dt),
                              time_step=recall, begin_lgidx_id=n_pat + 1, lgidx_step=1,
                              begin_wmask_id=1, wmask_step=1, begin_lgexp_id=0, lgexp_step=1,
                              prn=0.0))
    else:
        # each cue is preceded by `reset` unclamped steps without recurrent
        # support and with a negligible, non-zero learning rate, so that the
        # previous attractor decays and ProcCheck (which scores prn == 0
        # modes only) does not score the reset
        for i in range(n_pat):
            modes.append(ModeSpec(step_time(t, dt), step_time(t + reset, dt), time_step=reset,
                                  begin_lgidx_id=n_pat, begin_wmask_id=0, prn=reset_prn))
            t += reset
            modes.append(ModeSpec(step_time(t, dt), step_time(t + recall, dt), time_step=recall,
                                  begin_lgidx_id=n_pat + 1 + i, begin_wmask_id=1 + i,
                                  begin_lgexp_id=i, prn=0.0))
            t += recall
    return modes


def config0(seed: int = 0) -> NetSpec:
    """10 HCU x 10 MCU, 10 patterns x 100 steps + 10 quiet steps, 10 cues with
    50% of the HCUs clamped, 100 recall steps each after a 50-step reset."""
    return recall_net(10, 10, 10, 100, 10, 100, 0.5, seed=seed, reset=50, name="cfg0")


def config1(seed: int = 0) -> NetSpec:
    """100 HCU x 100 MCU (9.9e7 synapses), 100 patterns x 50 steps + 10 quiet
    steps, 100 cues with 30% of the HCUs clamped, 50 recall steps each after a
    50-step reset."""
    return recall_net(100, 100, 100, 50, 10, 50, 0.3, seed=seed, reset=50, name="cfg1")


def apply_conn_arrays(target, spec: NetSpec):
    """Write ii_<id>/di_<id> of every projection (a gsbn_amd.Solver or an
    oracle network: anything with write(name, array))."""
    for pid, (ii, di) in enumerate(spec.conn_arrays or []):
        target.write("ii_%d" % pid, ii)
        target.write("di_%d" % pid, di)


def sparse_conn_arrays(rng: np.random.Generator, hcu: int, mcu: int, p: float, delay: int = 1):
    """Bernoulli(p) recurrent connectivity without self-HCU connections: slot k
    of HCU h holds the k-th selected source unit (ascending), as ProcInitConn
    fills slots in file order.  Returns (ii, di, C)."""
    n = hcu * mcu
    src = np.arange(n)
    masks = rng.random((hcu, n)) < p
    for h in range(hcu):
        masks[h, h * mcu:(h + 1) * mcu] = False
    counts = masks.sum(axis=1)
    C = max(1, int(counts.max()))
    ii = np.full((hcu, C), -1, np.int32)
    di = np.zeros((hcu, C), np.int32)
    for h in range(hcu):
        sel = src[masks[h]]
        ii[h, :len(sel)] = sel
        di[h, :len(sel)] = delay
    return ii.reshape(-1), di.reshape(-1), C


def batched_nets(n_net: int = 512, n_pat: int = 20, train: int = 50, quiet: int = 10,
                 recall: int = 50, reset: int = 50, clamp_frac: float = 0.5, *,
                 hcu_range=(8, 64), mcu_range=(16, 128), probs=(0.1, 0.5, 0.9), seed: int = 0,
                 snoise: float = 0.1, wtagain: float = 4.0, pop_wgain: float = 8.0,
                 dt: float = 0.001, name: str = "batch") -> NetSpec:
    """n_net independent networks, each its own population and sparse
    recurrent projection, stepped in lock-step by one schedule.  Network i has
    H ~ U{hcu_range}, M ~ U{mcu_range}, connection probability from `probs`,
    its own random streams (population id) and its own patterns."""
    rng = np.random.Generator(np.random.Philox(seed))
    pops, projs, conns, stim_cols, mask_cols = [], [], [], [], []
    for i in range(n_net):
        H = int(rng.integers(hcu_range[0], hcu_range[1] + 1))
        M = int(rng.integers(mcu_range[0], mcu_range[1] + 1))
        p = float(probs[int(rng.integers(0, len(probs)))])
        ii, di, C = sparse_conn_arrays(rng, H, M, p)
        pops.append(PopSpec(hcu=H, mcu=M, wtagain=wtagain, wgain=pop_wgain, snoise=snoise))
        projs.append(ProjSpec(src=i, dest=i, tauzi=0.01, tauzj=0.01, taue=0.02, taup=1.0,
                              slot_num=float(C)))
        conns.append((ii, di))
        pats = random_patterns(rng, n_pat, H, M)
        cues = pats.copy()
        masks = np.zeros((n_pat, H), np.float32)
        nc = int(round(clamp_frac * H))
        for k in range(n_pat):
            clamped = set(int(v) for v in rng.permutation(H)[:nc])
            for h in range(H):
                if h not in clamped:
                    cues[k, h] = UNCLAMPED
                    masks[k, h] = 1.0
        stim_cols.append(np.concatenate([pats, np.full((1, H), UNCLAMPED, np.uint32), cues]))
        mask_cols.append(np.concatenate([np.zeros((1, H), np.float32), masks]))
    stim = np.concatenate(stim_cols, axis=1)
    mask = np.concatenate(mask_cols, axis=1)
    modes = recall_schedule(n_pat, train, quiet, recall, reset, dt=dt)
    return NetSpec(dt=dt, pops=pops, projs=projs, modes=modes, stim=stim, mask=mask,
                   conn=[[] for _ in projs], conn_arrays=conns, seed=seed, name=name)
