"""Frozen-weight batched recall (binding of gsbn_recall_* in include/gsbn_amd.h).

A trained population's recurrent weights Wij and biases bj are frozen and B
partial cues run in lock-step through the reference's recall-mode dynamics
(prn = 0), either with float weights in the reference's summation order
("exact") or with the support as one bfloat16 product of the spike matrix
and W on the matrix cores ("mfma").
"""

from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

from . import GsbnError, _p, configs, host_logf, lib

EXACT, MFMA = 0, 1


class RecallDesc(ctypes.Structure):
    _fields_ = [("H", ctypes.c_int32), ("M", ctypes.c_int32), ("C", ctypes.c_int32),
                ("ii", ctypes.c_void_p), ("di", ctypes.c_void_p), ("wij", ctypes.c_void_p),
                ("bj", ctypes.c_void_p),
                ("taumdt", ctypes.c_float), ("wtagain", ctypes.c_float),
                ("maxfqdt", ctypes.c_float), ("igain", ctypes.c_float), ("wgain", ctypes.c_float),
                ("lgbias", ctypes.c_float), ("snoise", ctypes.c_float), ("kepsc", ctypes.c_float),
                ("lg_on", ctypes.c_float), ("lg_off", ctypes.c_float),
                ("seed", ctypes.c_uint64), ("batch", ctypes.c_int32), ("mode", ctypes.c_int32)]


def _setup(L):
    if getattr(L, "_recall_ready", False):
        return
    L.gsbn_recall_new.restype = ctypes.c_void_p
    L.gsbn_recall_new.argtypes = [ctypes.POINTER(RecallDesc)]
    L.gsbn_recall_free.argtypes = [ctypes.c_void_p]
    L.gsbn_recall_run.restype = ctypes.c_int
    L.gsbn_recall_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                  ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.POINTER(ctypes.c_double)]
    L.gsbn_spike_gemm.restype = ctypes.c_int
    L.gsbn_spike_gemm.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p]
    L._recall_ready = True


def population_params(spec, pop: int = 0, proj: int = 0) -> dict:
    """The population and projection constants as Pop::init_new and
    Proj::init_new compute them in float (Pop.cpp:8-111, Proj.cpp:38-52)."""
    f = np.float32
    p = spec.pops[pop]
    q = spec.projs[proj]
    dt = f(spec.dt)
    tauzidt = dt / f(q.tauzi)
    kepsc = dt / f(q.tauepsc) if q.tauepsc is not None else tauzidt
    eps = f(spec.eps)
    return dict(taumdt=float(dt / f(p.taum)), wtagain=float(f(p.wtagain)),
                maxfqdt=float(f(p.maxfq) * dt), igain=float(f(p.igain)), wgain=float(f(p.wgain)),
                lgbias=float(f(p.lgbias)), snoise=float(f(p.snoise)), kepsc=float(kepsc),
                lg_on=float(host_logf(f(1) + eps)), lg_off=float(host_logf(f(0) + eps)))


class Recall:
    """gsbn_recall: B cues of one frozen population."""

    def __init__(self, H: int, M: int, C: int, ii, di, wij, bj, params: dict, batch: int,
                 mode: int = EXACT, seed: int = 0):
        L = lib()
        _setup(L)
        self._keep = [np.ascontiguousarray(ii, np.int32), np.ascontiguousarray(di, np.int32),
                      np.ascontiguousarray(wij, np.float32), np.ascontiguousarray(bj, np.float32)]
        d = RecallDesc()
        d.H, d.M, d.C = H, M, C
        d.ii, d.di, d.wij, d.bj = (a.ctypes.data for a in self._keep)
        for k, v in params.items():
            setattr(d, k, v)
        d.seed = seed
        d.batch = batch
        d.mode = mode
        self.H, self.M, self.B, self.mode = H, M, batch, mode
        self._h = L.gsbn_recall_new(ctypes.byref(d))
        self._keep = None
        if not self._h:
            raise GsbnError("gsbn_recall_new: " + L.gsbn_last_error().decode())

    @classmethod
    def from_solver(cls, solver, spec, batch: int, mode: int = EXACT, seed: int = 0,
                    pop: int = 0, proj: int = 0) -> "Recall":
        """Freeze the current weights of a solver's population `pop` and its
        recurrent projection `proj`."""
        p = spec.pops[pop]
        ii = solver.vector("ii_%d" % proj)
        C = ii.size // p.hcu
        bj = solver.vector("bj_%d" % pop)[: p.hcu * p.mcu]
        return cls(p.hcu, p.mcu, C, ii, solver.vector("di_%d" % proj),
                   solver.vector("wij_%d" % proj), bj, population_params(spec, pop, proj),
                   batch, mode, seed)

    def close(self):
        if getattr(self, "_h", None):
            lib().gsbn_recall_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, clamp: np.ndarray, wmask: np.ndarray, steps: int, count_from: int = 0,
            raster: bool = False, time_support: bool = False
            ) -> Tuple[np.ndarray, np.ndarray, Optional[np.ndarray], float]:
        """clamp[B, H] (clamped MCU or -1), wmask[B, H].  Returns the per-cue
        per-HCU winner and its count, the optional raster [steps, B, H*M] and,
        with time_support, the summed support-product time in ms (mfma; the
        timing waits for every product, so it slows the run)."""
        B, H, M = self.B, self.H, self.M
        c = np.ascontiguousarray(clamp, np.int16).reshape(B, H)
        w = np.ascontiguousarray(wmask, np.float32).reshape(B, H)
        idx = np.zeros((B, H), np.int32)
        cnt = np.zeros((B, H), np.int32)
        ras = np.zeros((steps, B, H * M), np.int8) if raster else None
        ms = ctypes.c_double(0.0)
        rc = lib().gsbn_recall_run(self._h, _p(c), _p(w), steps, count_from, _p(idx), _p(cnt),
                                   _p(ras) if raster else None,
                                   ctypes.byref(ms) if time_support else None)
        if rc != 0:
            raise GsbnError("gsbn_recall_run: " + lib().gsbn_last_error().decode())
        return idx, cnt, ras, ms.value


def cues(patterns: np.ndarray, batch: int, clamp_frac: float, seed: int = 0):
    """`batch` partial cues over the patterns (cue b shows pattern b mod P):
    a random clamp_frac of the HCUs clamped to the pattern's MCU, the support
    mask 1 on the others (as the recall windows of configs.recall_net)."""
    rng = np.random.Generator(np.random.Philox(seed))
    P, H = patterns.shape
    n_clamp = int(round(clamp_frac * H))
    clamp = np.full((batch, H), -1, np.int16)
    wmask = np.ones((batch, H), np.float32)
    want = np.zeros((batch, H), np.int32)
    for b in range(batch):
        pat = patterns[b % P]
        hs = rng.permutation(H)[:n_clamp]
        clamp[b, hs] = pat[hs]
        wmask[b, hs] = 0.0
        want[b] = pat
    return clamp, wmask, want


def recall_correct(idx: np.ndarray, want: np.ndarray) -> np.ndarray:
    """ProcCheck's criterion with threshold 0 for a full expected pattern
    (ProcCheck.cpp:196): every HCU's winner is the pattern's MCU."""
    return np.all(idx == want, axis=1)


def oracle_recall_spec(spec, clamp_row, wmask_row, steps: int, seed: int, ii, di):
    """A NetSpec that runs one cue in recall mode for `steps` steps from the
    initial state (for the oracle): one prn = 0 mode, the cue as stimulus and
    mask row, connectivity as given."""
    p = spec.pops[0]
    stim = np.where(np.asarray(clamp_row) < 0, configs.UNCLAMPED,
                    np.asarray(clamp_row)).astype(np.uint32)[None, :]
    mask = np.asarray(wmask_row, np.float32)[None, :]
    modes = [configs.ModeSpec(configs.step_time(0, spec.dt), configs.step_time(steps, spec.dt),
                              time_step=steps, begin_lgidx_id=0, begin_wmask_id=0, prn=0.0,
                              lgidx_step=0, wmask_step=0)]
    out = configs.NetSpec(dt=spec.dt, pops=[p], projs=[spec.projs[0]], modes=modes, stim=stim,
                          mask=mask, conn=[[]], eps=spec.eps, seed=seed, name="recall")
    out.conn_arrays = [(np.asarray(ii, np.int32), np.asarray(di, np.int32))]
    return out
