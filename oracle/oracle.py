"""Python driver of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this module.  It wraps liboracle.so (the C restatement of
ProcUpdLazy, see gsbn_oracle.h) and restates in Python the two small host
procedures around it:

* ``ExtGen``: ProcExtGen's mode schedule (ProcExtGen.cpp:8-107 init,
  :113-153 update): chunking of each mode_param into ``time_step`` pieces,
  ``simstep`` increment, ``prn``/``old-prn``/``wmask-idx``/``cycle-flag``
  updates and the ``.lginp`` rebuild (``log(1+eps)`` for the clamped MCU,
  ``log(eps)`` otherwise).
* ``Check``: ProcCheck (ProcCheck.cpp:8-143 init, :149-237 update;
  ProcCheck/Pop.cpp:35-60): per-window spike counts for prn==0 modes, first
  strict argmax per HCU, the correctness criterion of ProcCheck.cpp:196.
* ``StructuralPlasticity``: ProcStructuralPlasticity
  (ProcStructuralPlasitcity.cpp:8-608, ProcStructuralPlasticity/Pop.cpp,
  Prj.cpp) with this build's seeded draws; see the class for the pins.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ARITH_TABLE, ARITH_REF32, ARITH_F64 = 0, 1, 2
# planted mistakes (gsbn_oracle.h GOR_MUT_*), for showing that a pin fails
MUTATIONS = {"sweep_new_prn": 1, "sweep_target_t": 2, "sweep_w_kp0": 3, "sweep_old_pi": 4,
             "sup_no_bj": 5, "spike_no_maxfq": 6}
_DT = {0: np.int8, 2: np.int32, 5: np.float32, 6: np.float64}


class PopCfg(ctypes.Structure):
    _fields_ = [("hcu", ctypes.c_int), ("mcu", ctypes.c_int), ("taum", ctypes.c_float),
                ("wtagain", ctypes.c_float), ("maxfq", ctypes.c_float),
                ("igain", ctypes.c_float), ("wgain", ctypes.c_float),
                ("lgbias", ctypes.c_float), ("snoise", ctypes.c_float)]


class ProjCfg(ctypes.Structure):
    _fields_ = [("src", ctypes.c_int), ("dest", ctypes.c_int), ("tauzi", ctypes.c_float),
                ("tauzj", ctypes.c_float), ("taue", ctypes.c_float), ("taup", ctypes.c_float),
                ("maxfq", ctypes.c_float), ("bgain", ctypes.c_float),
                ("wgain", ctypes.c_float), ("tauepsc", ctypes.c_float),
                ("has_tauepsc", ctypes.c_int), ("slot_num", ctypes.c_float)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, f32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int64
        L.gor_net_new.restype = vp
        L.gor_net_new.argtypes = [f32, i32, ctypes.POINTER(PopCfg), i32, ctypes.POINTER(ProjCfg),
                                  ctypes.c_uint64, i32, i32]
        L.gor_net_free.argtypes = [vp]
        L.gor_set_globals.argtypes = [vp, i32, f32, f32, i32, i32, i32]
        L.gor_set_inputs.argtypes = [vp, vp, i32, vp, i32]
        L.gor_set_ext_spikes.argtypes = [vp, i32, i32, vp, vp]
        L.gor_updlazy_step.argtypes = [vp]
        L.gor_exchange_spikes.argtypes = [vp]
        L.gor_proj_step_replay.argtypes = [vp, i32, vp]
        L.gor_vector.restype = i64
        L.gor_vector.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), vp, i64]
        L.gor_vector_write.restype = i32
        L.gor_vector_write.argtypes = [vp, ctypes.c_char_p, vp, i64]
        L.gor_coeff_entry.argtypes = [ctypes.c_double] * 4 + [i64, vp]
        L.gor_closed_form_f64.argtypes = [ctypes.c_double] * 3 + [i64, vp, vp]
        L.gor_ref32_row.argtypes = [f32, f32, f32, i32, vp, vp, vp]
        L.gor_ref32_syn.argtypes = [f32, f32, f32, f32, i32, vp, vp, vp, vp]
        L.gor_philox4x32_10.argtypes = [vp, vp, vp]
        L.gor_draw_ctr.argtypes = [vp, vp, vp, vp]
        L.gor_draw.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.c_uint32, vp, vp]
        L.gor_work_counters.argtypes = [vp, vp]
        L.gor_set_mutation.argtypes = [vp, i32]
        L.gor_set_pop_id.restype = i32
        L.gor_set_pop_id.argtypes = [vp, i32, i32]
        L.gor_det_logf.restype = f32
        L.gor_det_logf.argtypes = [f32]
        L.gor_det_expf.restype = f32
        L.gor_det_expf.argtypes = [f32]
        L.gor_set_trace_form.restype = i32
        L.gor_set_trace_form.argtypes = [vp, i32, i32, i32, i32]
        L.gor_round_trace.restype = f32
        L.gor_round_trace.argtypes = [i32, f32, i32]
        L.gor_round_trace_n.argtypes = [i32, i32, i64, vp, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


TRACE_FORMS = {"fp32": 0, "bf16": 1, "fp16": 2}


def round_trace(prec: int, x, shift: int = 0) -> np.ndarray:
    """The oracle's 16-bit storage rounding applied to an array."""
    a = np.ascontiguousarray(x, dtype=np.float32).ravel()
    o = np.empty_like(a)
    lib().gor_round_trace_n(prec, shift, a.size, _p(a), _p(o))
    return o.reshape(np.shape(x))


def fp16_shift(q, dt: float) -> int:
    """Scale exponent of the binary16 trace form for a projection: the
    largest k with zmax_i * zmax_j * 2^k <= 65504, where a z trace kicked by
    kft at most once per step and decaying by exp(-kz) per step never
    exceeds kft / (1 - exp(-kz)).  Rates in float as the procedure computes
    them (Proj.cpp:38-52)."""
    f = np.float32
    kzi = float(f(dt) / f(q.tauzi))
    kzj = float(f(dt) / f(q.tauzj))
    kfti = float(f(1) / (f(q.maxfq) * f(q.tauzi)))
    kftj = float(f(1) / (f(q.maxfq) * f(q.tauzj)))
    bound = max(1.0, kfti / -np.expm1(-kzi) * kftj / -np.expm1(-kzj))
    k = int(np.floor(np.log2(65504.0 / bound)))
    while bound * 2.0 ** k > 65504.0:
        k -= 1
    return k


def philox(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().gor_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def draw(seed: int, pop: int, step: int, unit: int):
    u = np.zeros(1, np.float32)
    z = np.zeros(1, np.float32)
    lib().gor_draw(seed, pop, step, unit, _p(u), _p(z))
    return float(u[0]), float(z[0])


def det_logf(x) -> np.float32:
    return np.float32(lib().gor_det_logf(np.float32(x)))


def det_expf(x) -> np.float32:
    return np.float32(lib().gor_det_expf(np.float32(x)))


def coeff_entry(kzi, kzj, ke, kp, gap) -> np.ndarray:
    o = np.zeros(12, np.float32)
    lib().gor_coeff_entry(float(kzi), float(kzj), float(ke), float(kp), int(gap), _p(o))
    return o


def closed_form_f64(kz, ke, kp, gap, state) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.float64)
    o = np.zeros(3, np.float64)
    lib().gor_closed_form_f64(float(kz), float(ke), float(kp), int(gap), _p(s), _p(o))
    return o


def ref32_row(kp, ke, kzi, pdt, pi, ei, zi):
    a = np.array([pi, ei, zi], dtype=np.float32)
    lib().gor_ref32_row(kp, ke, kzi, int(pdt), _p(a[0:]), _p(a[1:]), _p(a[2:]))
    return a


def ref32_syn(kp, ke, kzi, kzj, pdt, pij, eij, zi2, zj2):
    a = np.array([pij, eij, zi2, zj2], dtype=np.float32)
    lib().gor_ref32_syn(kp, ke, kzi, kzj, int(pdt), _p(a[0:]), _p(a[1:]), _p(a[2:]), _p(a[3:]))
    return a


class ExtGen:
    """ProcExtGen restated (ProcExtGen.cpp:8-153)."""

    def __init__(self, spec):
        dt = np.float32(spec.dt)
        self.eps = np.float32(spec.eps)
        self.modes: List[dict] = []
        for m in spec.modes:
            begin = int(np.float32(m.begin_time) / dt)
            end = int(np.float32(m.end_time) / dt)
            lg, lx, wm = m.begin_lgidx_id, m.begin_lgexp_id, m.begin_wmask_id
            while begin < end:
                e = begin + m.time_step
                if e >= end:
                    e = end
                self.modes.append(dict(begin=begin, end=e, lgidx=lg, lgexp=lx, wmask=wm,
                                       prn=float(np.float32(m.prn)),
                                       plasticity=bool(getattr(m, "plasticity", 1))))
                lg += m.lgidx_step
                lx += m.lgexp_step
                wm += m.wmask_step
                begin = e
        self.stim = np.asarray(spec.stim, dtype=np.uint32).astype(np.int64)
        self.mask = np.ascontiguousarray(spec.mask, dtype=np.float32)
        self.mcu_in_hcu = [p.mcu for p in spec.pops for _ in range(p.hcu)]
        self.lginp = np.zeros(sum(self.mcu_in_hcu), dtype=np.float32)
        self.cursor = 0
        self.old_lgidx = -1
        self.simstep = 0
        self.prn = 0.0
        self.old_prn = 0.0
        self.cycle_flag = 0
        self.wmask_idx = 0
        self.lginp_idx = 0
        self.plasticity = False
        self.l_on = det_logf(np.float32(1) + self.eps)
        self.l_off = det_logf(np.float32(0) + self.eps)

    def update(self):
        self.simstep += 1
        step = self.simstep
        while self.cursor < len(self.modes):
            m = self.modes[self.cursor]
            if m["begin"] < step <= m["end"]:
                self.old_prn = self.prn
                self.prn = m["prn"]
                self.wmask_idx = m["wmask"]
                self.plasticity = m["plasticity"]
                self.lginp_idx = 0
                self.cycle_flag = 1
                if self.old_lgidx != m["lgidx"]:
                    self.old_lgidx = m["lgidx"]
                    row = self.stim[m["lgidx"]]
                    k = 0
                    for i, nm in enumerate(self.mcu_in_hcu):
                        for j in range(nm):
                            self.lginp[k] = self.l_on if row[i] == j else self.l_off
                            k += 1
                return
            elif step <= m["begin"]:
                self.cycle_flag = 0
                return
            self.cursor += 1
        self.cycle_flag = -1


class Check:
    """ProcCheck restated (ProcCheck.cpp:8-237, ProcCheck/Pop.cpp:35-60)."""

    def __init__(self, spec, threshold: int = 0):
        dt = np.float32(spec.dt)
        self.modes = []
        for m in spec.modes:
            if np.float32(m.prn) != 0.0:
                continue
            begin = int(np.float32(m.begin_time) / dt)
            end = int(np.float32(m.end_time) / dt)
            lx = m.begin_lgexp_id
            while begin < end:
                self.modes.append(dict(begin=begin, end=begin + m.time_step, lgexp=lx))
                lx += m.lgexp_step
                begin += m.time_step
        self.stim = np.asarray(spec.stim, dtype=np.uint32).astype(np.int64)
        self.hcu = [(p.hcu, p.mcu) for p in spec.pops]
        self.counters = [np.zeros((h, m), dtype=np.int64) for h, m in self.hcu]
        self.threshold = threshold
        self.cursor = 0
        self.updated = False
        self.results: List[dict] = []
        self.correct = 0
        self.total = 0

    def update(self, cycle_flag: int, simstep: int, spikes: List[np.ndarray]):
        if cycle_flag == 0 or cycle_flag < 0:
            return
        if self.cursor >= len(self.modes):
            return
        m = self.modes[self.cursor]
        if simstep <= m["begin"]:
            return
        if simstep < m["end"]:
            for c, s in zip(self.counters, spikes):
                c += (s.reshape(c.shape) > 0)
            self.updated = True
            return
        if not self.updated:
            while self.cursor < len(self.modes) and simstep >= self.modes[self.cursor]["end"]:
                self.cursor += 1
            return
        idx_all, cnt_all = [], []
        for c in self.counters:
            for row in c:
                mx_i, mx_c = -1, 0
                for j, v in enumerate(row):
                    if v > mx_c:
                        mx_i, mx_c = j, int(v)
                idx_all.append(mx_i)
                cnt_all.append(mx_c)
            c[:] = 0
        exp = self.stim[m["lgexp"]]
        exp = np.where(exp > 0x7FFFFFFF, exp - (1 << 32), exp)
        ok = True
        for i in range(len(idx_all)):
            e = int(exp[i])
            if ((cnt_all[i] < self.threshold or idx_all[i] != e) and e >= 0) or \
                    (cnt_all[i] > self.threshold and e < 0):
                ok = False
                break
        self.results.append(dict(ok=ok, expected=[int(v) for v in exp], idx=idx_all,
                                 cnt=cnt_all))
        self.total += 1
        self.correct += int(ok)
        self.cursor += 1


class SpDraws:
    """The structural-plasticity draws of one (step, lane): Philox4x32-10
    blocks, key = seed, counter = (block, step, lane, 1), consumed word by
    word."""

    def __init__(self, seed: int, step: int, lane: int):
        self.key = np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], np.uint32)
        self.step, self.lane, self.block = step & 0xFFFFFFFF, lane, 0
        self.words: List[int] = []

    def _ctr(self):
        c = np.array([self.block, self.step, self.lane, 1], np.uint32)
        self.block += 1
        return c

    def word(self) -> int:
        if not self.words:
            self.words = [int(w) for w in philox(self._ctr(), self.key)]
        return self.words.pop(0)

    def uniform(self) -> np.float32:
        return np.float32(self.word() >> 8) * np.float32(2.0 ** -24)

    def below(self, n: int) -> int:
        return (self.word() * n) >> 32

    def shuffle(self, v: list):
        for i in range(len(v), 1, -1):
            j = self.below(i)
            v[i - 1], v[j] = v[j], v[i - 1]

    def normal(self, m: np.float32, s: np.float32) -> np.float32:
        self.words = []
        u = np.zeros(1, np.float32)
        z = np.zeros(1, np.float32)
        lib().gor_draw_ctr(_p(self._ctr()), _p(self.key), _p(u), _p(z))
        return np.float32(np.float32(m) + np.float32(np.float32(s) * z[0]))


class StructuralPlasticity:
    """ProcStructuralPlasticity restated for the oracle.

    Follows ProcStructuralPlasitcity.cpp:8-608 and the helper classes of
    ProcStructuralPlasticity/Pop.cpp:5-95 and Prj.cpp:5-138 on one rank (the
    MPI window exchange becomes plain array access), with this build's
    choices where the reference is not reproducible or is wrong:
    the draws of ``SpDraws`` instead of random_device-seeded std::mt19937 and
    the Random class; a hypercolumn's own slots decide whether it has a free
    slot (Prj.cpp:100 looks at hypercolumn 0's); a removed row leaves the
    delay queue (qi = siq = 0; the reference's update_que_cpu skips rows with
    ii < 0, Proj.cpp:651-653, freezing them); geometric delays are clamped to
    [1, 31].

    Pinned by tests/test_oracle_structural.py: brute-force hand cases of the
    prune test and the state reset, the fan-out bookkeeping, the growth
    order, and invariants (no duplicate source in a hypercolumn, fan-out
    never below zero for grown rows, counts conserved).
    """

    LANE_GROW, LANE_DELAY, LANE_INIT, LANE_INIT_DELAY = 0, 1, 2, 3

    def __init__(self, spec, net: "OracleNet"):
        sp = spec.structural
        self.net = net
        self.dt = np.float32(spec.dt)
        self.period = sp.period if sp.period > 0 else 100
        self.pruning_period = sp.pruning_period if sp.pruning_period > 0 else 1000
        self.t_th = sp.t_th if sp.t_th > 0 else 100
        self.wp, self.wn = np.float32(sp.wp), np.float32(sp.wn)
        self.d_norm = np.float32(max(sp.d_norm, 0.0))
        self.v_cond = np.float32(sp.v_cond if sp.v_cond > 0 else 1.0)
        self.geometry = 1 if sp.enable_geometry else 0
        self.seed = sp.seed
        self.pops = []
        for i, p in enumerate(spec.pops):
            fan = p.mcu if p.fanout is None else p.fanout
            self.pops.append(dict(id=i, H=p.hcu, M=p.mcu, shape=list(p.shape),
                                  position=list(p.position),
                                  fanout=np.full(p.hcu * p.mcu, fan, np.int64), prjs=[]))
        self.prjs = []
        for i, q in enumerate(spec.projs):
            sp_, dp = spec.pops[q.src], spec.pops[q.dest]
            C = sp_.hcu * sp_.mcu
            if C > q.slot_num:
                C = int(q.slot_num)
            self.prjs.append(dict(id=i, src=q.src, dest=q.dest, H=dp.hcu, M=dp.mcu, C=C))
            self.pops[q.src]["prjs"].append(i)
        if sp.enable_init_conn:
            self.grow(0, self.LANE_INIT, self.LANE_INIT_DELAY, -1, np.float32(-1.0))

    def update(self, simstep: int, cycle_flag: int, plasticity: bool):
        """ProcStructuralPlasitcity.cpp:260-290."""
        if simstep != 0 and cycle_flag != 1:
            return
        if not plasticity and simstep != 0:
            return
        if simstep % self.pruning_period == 0:
            self.prune(simstep)
        if simstep % self.period == 0:
            thr = int(0.2 * self.period * 100 * float(self.dt))
            self.grow(simstep, self.LANE_GROW, self.LANE_DELAY, thr, np.float32(0.5))

    def prune(self, simstep: int):
        """STAGE 1-2 (ProcStructuralPlasitcity.cpp:292-352); Prj::prune and
        Prj::remove_conn (Prj.cpp:78-138)."""
        assert self.wp > self.wn, "w+ should be larger than w- !!"
        net = self.net
        for q in self.prjs:
            s = "_%d" % q["id"]
            ii, di, ti = net.vector("ii" + s), net.vector("di" + s), net.vector("ti" + s)
            wij = net.vector("wij" + s).reshape(-1, q["M"])
            t_min = simstep - self.t_th
            removed = []
            for r in range(ii.size):
                if ii[r] < 0:
                    continue
                if ti[r] < t_min:
                    removed.append(r)
                    continue
                if not np.any((wij[r] >= self.wp) | (wij[r] <= self.wn)):
                    removed.append(r)
            if not removed:
                continue
            fan = self.pops[q["src"]]["fanout"]
            for r in removed:
                fan[ii[r]] += 1
            rows = np.array(removed, np.int64)
            ii[rows] = -1
            di[rows] = 0
            ti[rows] = 0
            net.write("ii" + s, ii)
            net.write("di" + s, di)
            net.write("ti" + s, ti)
            for name in ("pi", "ei", "qi", ".siq"):
                v = net.vector(name + s)
                v[rows] = 0
                net.write(name + s, v)
            for name in ("pij", "eij", "zj2"):
                v = net.vector(name + s).reshape(-1, q["M"])
                v[rows] = 0
                net.write(name + s, v.ravel())

    def _hcu_coor(self, pop, hcu: int):
        """Pop::hcu_coor (ProcStructuralPlasticity/Pop.cpp:65-93)."""
        coor = []
        index = hcu
        for d in pop["shape"]:
            coor.append(index % d)
            index //= d
        for i, p in enumerate(pop["position"]):
            if i < len(coor):
                coor[i] += p
            else:
                coor.append(p)
        return coor

    def _delay(self, draws: SpDraws, q, src_unit: int, dest_hcu: int) -> int:
        """delay_cycle (ProcStructuralPlasitcity.cpp:578-608), clamped to [1, 31]."""
        p0, p1 = self.pops[q["src"]], self.pops[q["dest"]]
        a = self._hcu_coor(p0, src_unit // p0["M"])
        b = self._hcu_coor(p1, dest_hcu)
        n = max(len(a), len(b))
        a += [0] * (n - len(a))
        b += [0] * (n - len(b))
        d2 = np.float32(0)
        for x, y in zip(a, b):
            d2 = np.float32(d2 + np.float32((x - y) * (x - y)))
        dist = np.float32(np.sqrt(d2))
        mean = np.float32(np.float32(np.float32(self.d_norm * dist) / self.v_cond) + np.float32(1))
        tij = draws.normal(mean, np.float32(np.float32(0.1) * mean))
        delay = int(np.ceil(float(tij) * 0.001 / float(self.dt)))
        return min(31, max(1, delay))

    def grow(self, step: int, lane: int, delay_lane: int, mcu_thr: int, hcu_thr: np.float32):
        """STAGE 3-5 (ProcStructuralPlasitcity.cpp:354-576), init_conn
        (:109-258); Pop::get_avail_active_mcu_list (Pop.cpp:52-63),
        Prj::get_avail_active_hcu_list (Prj.cpp:84-110)."""
        net = self.net
        local, avail = {}, {}
        for q in self.prjs:
            s = "_%d" % q["id"]
            ii = net.vector("ii" + s)
            local[q["id"]] = ii.copy()
            ej = net.vector("ej" + s)
            hs = []
            for h in range(q["H"]):
                acc = np.float32(0)
                for j in range(q["M"]):
                    acc = np.float32(acc + ej[h * q["M"] + j])
                free = bool(np.any(ii[h * q["C"]:(h + 1) * q["C"]] < 0))
                if acc > hcu_thr and free:
                    hs.append(h)
            avail[q["id"]] = hs
        draws = SpDraws(self.seed, step, lane)
        for pop in self.pops:
            cnt = net.vector(".counter_%d" % pop["id"])
            active = [u for u in range(cnt.size) if cnt[u] > mcu_thr]
            net.write(".counter_%d" % pop["id"], np.zeros_like(cnt))
            fan = pop["fanout"]
            prjs = list(pop["prjs"])
            while prjs and active:
                idx = int(draws.uniform() * np.float32(len(prjs)))
                idx = min(idx, len(prjs) - 1)
                q = self.prjs[prjs[idx]]
                hcus = list(avail[q["id"]])
                if not hcus:
                    break
                draws.shuffle(hcus)
                buf = local[q["id"]]
                C = q["C"]
                for h in hcus:
                    row = buf[h * C:(h + 1) * C]
                    conn = set(int(c) for c in row if c >= 0)
                    cand = [u for u in active if u not in conn]
                    if not cand:
                        continue
                    draws.shuffle(cand)
                    l = 0
                    for k in range(C):
                        if l >= len(cand):
                            break
                        if row[k] < 0:
                            row[k] = cand[l]
                            l += 1
                    for u in cand[:l]:
                        fan[u] -= 1
                        if fan[u] <= 0:
                            active = [a for a in active if a != u]
                prjs.pop(idx)
        delays = SpDraws(self.seed, step, delay_lane)
        for q in self.prjs:
            s = "_%d" % q["id"]
            ii, di = net.vector("ii" + s), net.vector("di" + s)
            new = (ii < 0) & (local[q["id"]] >= 0)
            if not new.any():
                continue
            for index in np.nonzero(new)[0]:
                h = int(index) // q["C"]
                src = int(local[q["id"]][index])
                # add_row (ProcStructuralPlasitcity.cpp:566-576): first free slot
                k0 = int(np.nonzero(ii[h * q["C"]:(h + 1) * q["C"]] < 0)[0][0])
                ii[h * q["C"] + k0] = src
                di[h * q["C"] + k0] = self._delay(delays, q, src, h) if self.geometry else 1
            net.write("ii" + s, ii)
            net.write("di" + s, di)


class OracleNet:
    """A network in the oracle, built from a gsbn_amd.configs.NetSpec."""

    def __init__(self, spec, arith: int = ARITH_TABLE):
        L = lib()
        self.spec = spec
        pops = (PopCfg * len(spec.pops))()
        for i, p in enumerate(spec.pops):
            pops[i] = PopCfg(p.hcu, p.mcu, p.taum, p.wtagain, p.maxfq, p.igain, p.wgain,
                             p.lgbias, p.snoise)
        projs = (ProjCfg * max(1, len(spec.projs)))()
        for i, q in enumerate(spec.projs):
            projs[i] = ProjCfg(q.src, q.dest, q.tauzi, q.tauzj, q.taue, q.taup, q.maxfq,
                               q.bgain, q.wgain, q.tauepsc if q.tauepsc is not None else 0.0,
                               1 if q.tauepsc is not None else 0, q.slot_num)
        self.h = L.gor_net_new(spec.dt, len(spec.pops), pops, len(spec.projs), projs,
                               spec.seed, spec.wta_pop_max, arith)
        for i, pid in enumerate(getattr(spec, "pop_ids", None) or []):
            if L.gor_set_pop_id(self.h, i, pid) != 0:
                raise ValueError("oracle: bad population id")
        self.arith = arith
        self.extgen = ExtGen(spec)
        self.check = Check(spec, spec.check_threshold)
        self._mask = np.ascontiguousarray(spec.mask, dtype=np.float32)
        self._keep = []
        self._init_conn()
        self.sp = StructuralPlasticity(spec, self) if getattr(spec, "structural", None) else None
        self.trace_shifts = []
        prec = TRACE_FORMS[getattr(spec, "trace_precision", "fp32")]
        if prec and arith != ARITH_F64:
            for pid, q in enumerate(spec.projs):
                k = fp16_shift(q, spec.dt) if prec == 2 else 0
                self.trace_shifts.append(k)
                if L.gor_set_trace_form(self.h, pid, prec, k, k) != 0:
                    raise ValueError("oracle: bad trace form")
        if spec.ext_spikes:
            for pop, steps in spec.ext_spikes.items():
                nsteps = max(steps) + 1 if steps else 0
                off = np.zeros(nsteps + 1, np.int32)
                units: List[int] = []
                for s in range(nsteps):
                    units.extend(steps.get(s, []))
                    off[s + 1] = len(units)
                u = np.array(units, dtype=np.int32)
                L.gor_set_ext_spikes(self.h, pop, nsteps, _p(off), _p(u))

    def _init_conn(self):
        """ProcInitConn::add_row (ProcInitConn.cpp:92-103): fill the first free
        slot of the destination HCU, in file order."""
        if self.spec.conn_arrays is not None:
            for pid, (ii, di) in enumerate(self.spec.conn_arrays):
                self.write("ii_%d" % pid, ii)
                self.write("di_%d" % pid, di)
            return
        for pid, rows in enumerate(self.spec.conn):
            ii = self.vector("ii_%d" % pid)
            di = self.vector("di_%d" % pid)
            q = self.spec.projs[pid]
            dp = self.spec.pops[q.dest]
            sp = self.spec.pops[q.src]
            C = sp.hcu * sp.mcu
            if C > q.slot_num:  # Proj.cpp:22-26
                C = int(q.slot_num)
            # first free slot of the HCU, or none: slots only go from free to
            # taken, so a per-HCU cursor never has to look behind itself
            cursor = {}
            for dest_hcu, pairs in rows:
                base = dest_hcu * C
                k = cursor.get(dest_hcu, 0)
                for s, d in pairs:
                    while k < C and ii[base + k] >= 0:
                        k += 1
                    if k < C:
                        ii[base + k] = s
                        di[base + k] = d
                cursor[dest_hcu] = k
            self.write("ii_%d" % pid, ii)
            self.write("di_%d" % pid, di)
            del dp

    def vector(self, name: str) -> np.ndarray:
        L = lib()
        t = ctypes.c_int(0)
        n = L.gor_vector(self.h, name.encode(), ctypes.byref(t), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n, dtype=_DT[t.value])
        if n:
            L.gor_vector(self.h, name.encode(), ctypes.byref(t), _p(out), n)
        return out

    def write(self, name: str, arr: np.ndarray):
        L = lib()
        t = ctypes.c_int(0)
        n = L.gor_vector(self.h, name.encode(), ctypes.byref(t), None, 0)
        a = np.ascontiguousarray(arr, dtype=_DT[t.value])
        rc = L.gor_vector_write(self.h, name.encode(), _p(a), a.size)
        if rc != 0:
            raise ValueError("write %s failed (%d, expected %d elements)" % (name, rc, n))

    def step(self) -> int:
        """One Upd::update: ExtGen, ProcUpdLazy, ProcExchangeSpike,
        [ProcStructuralPlasticity], ProcCheck.  Returns cycle-flag."""
        L = lib()
        eg = self.extgen
        eg.update()
        L.gor_set_globals(self.h, eg.simstep, eg.prn, eg.old_prn, eg.cycle_flag, eg.lginp_idx,
                          eg.wmask_idx)
        L.gor_set_inputs(self.h, _p(eg.lginp), 1, _p(self._mask), self._mask.shape[0])
        if eg.cycle_flag == 1:
            L.gor_updlazy_step(self.h)
            L.gor_exchange_spikes(self.h)
            spikes = [self.vector("spike_%d" % i) for i in range(len(self.spec.pops))]
        else:
            spikes = []
        if self.sp is not None:
            self.sp.update(eg.simstep, eg.cycle_flag, eg.plasticity)
        self.check.update(eg.cycle_flag, eg.simstep, spikes)
        return eg.cycle_flag

    def run(self, max_steps: Optional[int] = None, record: bool = False):
        raster: Dict[int, np.ndarray] = {}
        n = 0
        while True:
            f = self.step()
            if f < 0:
                break
            if record and f == 1:
                raster[self.extgen.simstep] = np.nonzero(self.vector("spike_0"))[0]
            n += 1
            if max_steps is not None and n >= max_steps:
                break
        return raster

    def plant(self, mistake: Optional[str]):
        """Make this network deviate from the reference in one planted way
        (MUTATIONS), or restore it (None); only tests of the tests use this."""
        lib().gor_set_mutation(self.h, MUTATIONS[mistake] if mistake else 0)

    def work(self) -> np.ndarray:
        o = np.zeros(4, np.int64)
        lib().gor_work_counters(self.h, _p(o))
        return o

    def close(self):
        if self.h:
            lib().gor_net_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
