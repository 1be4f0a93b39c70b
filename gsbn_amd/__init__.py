"""gsbn_amd -- MI355X-native lazy-update BCPNN core.

Python binding of libgsbn_amd.so (the C ABI declared in include/gsbn_amd.h).
``Solver`` mirrors the reference's ``gsbn::Solver`` (Solver.cpp:10-92): it is
built from a SolverParam prototxt, ``run()`` steps the procedure list until
``cycle-flag`` < 0, and every state vector is reachable by its Database name
(``pij_0``, ``spike_0``, ...).  The library is required: importing works
without it, but creating a Solver or calling a kernel entry point raises if
``libgsbn_amd.so`` has not been built (``python gsbn_amd/build.py``).
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgsbn_amd.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include", "gsbn_amd.h")

DTYPES = {0: np.int8, 1: np.int16, 2: np.int32, 3: np.int64, 4: np.float16, 5: np.float32,
          6: np.float64}

_lib = None


class GsbnError(RuntimeError):
    pass


def lib():
    """The loaded C library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GsbnError("libgsbn_amd.so not built: run python gsbn_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, f32, f64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_float,
                                  ctypes.c_double)
        L.gsbn_last_error.restype = ctypes.c_char_p
        L.gsbn_solver_new.restype = vp
        L.gsbn_solver_new.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.gsbn_solver_free.argtypes = [vp]
        L.gsbn_solver_step.restype = i64
        L.gsbn_solver_step.argtypes = [vp, i64]
        L.gsbn_solver_run.restype = i64
        L.gsbn_solver_run.argtypes = [vp]
        L.gsbn_solver_sync.argtypes = [vp]
        L.gsbn_solver_vector_info.restype = i64
        L.gsbn_solver_vector_info.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        L.gsbn_solver_vector_read.restype = i64
        L.gsbn_solver_vector_read.argtypes = [vp, ctypes.c_char_p, vp, i64]
        L.gsbn_solver_vector_write.argtypes = [vp, ctypes.c_char_p, vp, i64]
        L.gsbn_solver_global.argtypes = [vp, ctypes.c_char_p, ctypes.c_char, vp]
        L.gsbn_solver_check.argtypes = [vp, vp, vp, vp, i32]
        L.gsbn_solver_work.argtypes = [vp, vp]
        L.gsbn_solver_timing.argtypes = [vp, i32, vp, vp]
        L.gsbn_coeff_host_entry.argtypes = [f64, f64, f64, f64, i64, vp]
        L.gsbn_coeff_host_saturation.restype = i64
        L.gsbn_coeff_host_saturation.argtypes = [f64, f64, f64, f64, i64]
        L.gsbn_host_expf.restype = f32
        L.gsbn_host_expf.argtypes = [f32]
        L.gsbn_host_logf.restype = f32
        L.gsbn_host_logf.argtypes = [f32]
        L.gsbn_host_draw.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_uint32, vp, vp]
        L.gsbn_device_math.argtypes = [i32, i64, vp, vp]
        _lib = L
    return _lib


def _err(prefix: str) -> GsbnError:
    return GsbnError("%s: %s" % (prefix, lib().gsbn_last_error().decode()))


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Solver:
    """The reference's Solver over this library (Solver.cpp:10-92); with a
    snapshot path it resumes from that SolverState file (Solver.cpp:44-61)."""

    def __init__(self, net_path: str, snapshot_path: Optional[str] = None):
        L = lib()
        self._h = L.gsbn_solver_new(net_path.encode(),
                                    snapshot_path.encode() if snapshot_path else None)
        if not self._h:
            raise _err("Solver")

    def close(self):
        if getattr(self, "_h", None):
            lib().gsbn_solver_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, n: int = 1) -> int:
        r = lib().gsbn_solver_step(self._h, n)
        if r < 0:
            raise _err("step")
        return r

    def run(self) -> int:
        r = lib().gsbn_solver_run(self._h)
        if r < 0:
            raise _err("run")
        return r

    def sync(self):
        if lib().gsbn_solver_sync(self._h) != 0:
            raise _err("sync")

    def vector(self, name: str) -> np.ndarray:
        L = lib()
        t = ctypes.c_int(0)
        n = L.gsbn_solver_vector_info(self._h, name.encode(), ctypes.byref(t))
        if n < 0:
            raise _err("vector %s" % name)
        out = np.zeros(n, dtype=DTYPES[t.value])
        r = L.gsbn_solver_vector_read(self._h, name.encode(), _p(out), n)
        if r < 0:
            raise _err("vector %s" % name)
        return out[:r] if r < n else out

    def write(self, name: str, data: np.ndarray):
        L = lib()
        t = ctypes.c_int(0)
        n = L.gsbn_solver_vector_info(self._h, name.encode(), ctypes.byref(t))
        if n < 0:
            raise _err("vector %s" % name)
        a = np.ascontiguousarray(data, dtype=DTYPES[t.value])
        if L.gsbn_solver_vector_write(self._h, name.encode(), _p(a), a.size) != 0:
            raise _err("write %s" % name)

    def glob(self, key: str, typ: str = "i"):
        v = {"i": ctypes.c_int(0), "f": ctypes.c_float(0), "b": ctypes.c_bool(False)}[typ]
        if lib().gsbn_solver_global(self._h, key.encode(), typ.encode(), ctypes.byref(v)) != 0:
            raise _err("global %s" % key)
        return v.value

    def check(self) -> Tuple[int, int, np.ndarray]:
        c, t = ctypes.c_int32(0), ctypes.c_int32(0)
        flags = np.zeros(4096, np.int32)
        if lib().gsbn_solver_check(self._h, ctypes.byref(c), ctypes.byref(t), _p(flags), 4096) != 0:
            raise _err("check")
        return c.value, t.value, flags[:t.value]

    def work(self) -> np.ndarray:
        o = np.zeros(5, np.int64)
        if lib().gsbn_solver_work(self._h, _p(o)) != 0:
            raise _err("work")
        return o

    def snapshot(self, path: str):
        """Write the current state as a SolverState file (what ProcSnapshot
        writes at its steps); ``Solver(net_path, path)`` resumes from it."""
        from . import snapshot as _snap
        L = lib()
        _snap._setup(L)
        if L.gsbn_solver_snapshot(self._h, path.encode()) != 0:
            raise _err("snapshot")

    def timing(self, enable: int, collect: bool = True):
        """Per-kind launch counts and total ms (kinds: 0 pop, 1 rowcol, 2 all,
        3 jxx, 4 que, 5 check, 6 fused step)."""
        n = np.zeros(8, np.int64)
        ms = np.zeros(8, np.float64)
        if lib().gsbn_solver_timing(self._h, enable, _p(n) if collect else None,
                                    _p(ms) if collect else None) != 0:
            raise _err("timing")
        return n, ms


def coeff_entry(kzi, kzj, ke, kp, gap) -> np.ndarray:
    o = np.zeros(12, np.float32)
    lib().gsbn_coeff_host_entry(float(kzi), float(kzj), float(ke), float(kp), int(gap), _p(o))
    return o


def coeff_saturation(kzi, kzj, ke, kp, max_len=1 << 26) -> int:
    return int(lib().gsbn_coeff_host_saturation(float(kzi), float(kzj), float(ke), float(kp),
                                                int(max_len)))


def host_expf(x) -> np.float32:
    return np.float32(lib().gsbn_host_expf(np.float32(x)))


def host_logf(x) -> np.float32:
    return np.float32(lib().gsbn_host_logf(np.float32(x)))


def device_math(op: int, inputs: np.ndarray) -> np.ndarray:
    """Evaluate the device's deterministic math on the GPU (tests):
    op 0 expf, 1 logf (uint32 bit patterns in/out), 2 Philox draw
    (rows of seed_lo, seed_hi, step, unit -> u, z bits), 3 float division,
    4 Wij of (pij, pi, pj), 5 float -> bfloat16 code, 6 float -> binary16 code,
    7 binary16 code -> float."""
    a = np.ascontiguousarray(inputs, dtype=np.uint32)
    in_w = {0: 1, 1: 1, 2: 4, 3: 2, 4: 3, 5: 1, 6: 1, 7: 1}[op]
    out_w = 2 if op == 2 else 1
    n = a.size // in_w
    out = np.zeros(n * out_w, np.uint32)
    rc = lib().gsbn_device_math(op, n, _p(a), _p(out))
    if rc != 0:
        raise GsbnError("device_math failed with code %d" % rc)
    return out
