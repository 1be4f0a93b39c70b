"""SolverState files (binding of gsbn_state_* in include/gsbn_amd.h).

The reference's snapshot format (gsbn.proto:100-152): a timestamp, the
learning rate prn and every saved vector by name, written by ProcSnapshot
(ProcSnapshot.cpp:69-78) and read back by ``Solver(net, snapshot)``
(Solver.cpp:44-61).  These helpers read and write such files without a
device, for inspection and for preparing a resume.
"""

from __future__ import annotations

import ctypes
from typing import Dict, Tuple

import numpy as np

from . import GsbnError, lib

_CODES = {np.dtype(np.int8): 0, np.dtype(np.int16): 1, np.dtype(np.int32): 2,
          np.dtype(np.int64): 3, np.dtype(np.float32): 5, np.dtype(np.float64): 6}
_TYPES = {0: np.int8, 1: np.int16, 2: np.int32, 3: np.int64, 4: np.float32, 5: np.float32,
          6: np.float64}


class VecState(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("dtype", ctypes.c_int32), ("ld", ctypes.c_uint32),
                ("count", ctypes.c_int64), ("data", ctypes.c_void_p)]


def _setup(L):
    if getattr(L, "_state_ready", False):
        return
    L.gsbn_state_write.restype = ctypes.c_int
    L.gsbn_state_write.argtypes = [ctypes.c_char_p, ctypes.c_float, ctypes.c_float, ctypes.c_int32,
                                   ctypes.POINTER(VecState)]
    L.gsbn_state_read.restype = ctypes.c_void_p
    L.gsbn_state_read.argtypes = [ctypes.c_char_p]
    L.gsbn_state_nvec.restype = ctypes.c_int32
    L.gsbn_state_nvec.argtypes = [ctypes.c_void_p]
    L.gsbn_state_header.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                    ctypes.POINTER(ctypes.c_float)]
    L.gsbn_state_vec.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(VecState)]
    L.gsbn_state_free.argtypes = [ctypes.c_void_p]
    L.gsbn_solver_snapshot.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    L._state_ready = True


def write_state(path: str, timestamp: float, prn: float,
                vectors: Dict[str, np.ndarray], lds: Dict[str, int] = None) -> None:
    """Write a SolverState file.  The element type of each vector is its
    numpy dtype (int8/16/32/64, float32, float64); `lds` gives leading
    dimensions (default 0, as a SyncVector without set_ld)."""
    L = lib()
    _setup(L)
    keep = []
    arr = (VecState * max(len(vectors), 1))()
    for i, (name, a) in enumerate(vectors.items()):
        a = np.ascontiguousarray(a).ravel()
        if a.dtype not in _CODES:
            raise GsbnError("write_state: unsupported element type %s of %s" % (a.dtype, name))
        nb = name.encode()
        keep += [a, nb]
        arr[i] = VecState(nb, _CODES[a.dtype], int((lds or {}).get(name, 0)) & 0xFFFFFFFF, a.size,
                          a.ctypes.data if a.size else None)
    rc = L.gsbn_state_write(path.encode(), timestamp, prn, len(vectors), arr)
    if rc != 0:
        raise GsbnError("gsbn_state_write: " + L.gsbn_last_error().decode())


def read_state(path: str) -> Tuple[float, float, Dict[str, Tuple[np.ndarray, int]]]:
    """(timestamp, prn, {name: (array, ld)}) of a SolverState file."""
    L = lib()
    _setup(L)
    h = L.gsbn_state_read(path.encode())
    if not h:
        raise GsbnError("gsbn_state_read: " + L.gsbn_last_error().decode())
    try:
        ts, prn = ctypes.c_float(), ctypes.c_float()
        L.gsbn_state_header(h, ctypes.byref(ts), ctypes.byref(prn))
        out = {}
        for i in range(L.gsbn_state_nvec(h)):
            v = VecState()
            if L.gsbn_state_vec(h, i, ctypes.byref(v)) != 0:
                raise GsbnError("gsbn_state_vec: " + L.gsbn_last_error().decode())
            t = np.dtype(_TYPES[v.dtype])
            a = np.zeros(v.count, t)
            if v.count:
                ctypes.memmove(a.ctypes.data, v.data, v.count * t.itemsize)
            out[v.name.decode()] = (a, int(v.ld))
        return float(ts.value), float(prn.value), out
    finally:
        L.gsbn_state_free(h)


def snapshot_name(simstep: int, dt: float) -> str:
    """ProcSnapshot's file name for a step: Snapshot_<to_string(simstep*dt)>.bin
    (float product, printed with six decimals)."""
    t = np.float32(np.float32(simstep) * np.float32(dt))
    return "Snapshot_%f.bin" % float(t)
