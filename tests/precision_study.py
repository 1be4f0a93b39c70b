"""Error of the trace storage forms against the exact double-precision
oracle (test infrastructure).

A network runs with Pij/Eij stored in fp32, bf16 or fp16: either the product
on the GPU (--engine gpu) or the oracle's table arithmetic with the same
storage rounding (--engine oracle, bit-identical to the product by the GPU
parity tests).  After every executed step its spikes are replayed through the
oracle's double-precision projection (closed form in long double, no
rounding of stored values), so both see the same spike trains and differ only
by arithmetic and storage.  At the end the errors of Pij, Eij, Pi, bj
(relative) and Wij (absolute) are reported with ProcCheck's recall result.

Relative errors are taken over entries whose exact value is at least a floor
(eps^2 = 1e-6 for Pij/Eij, eps = 1e-3 for Pi): below it a value is dominated
by the eps^2 term of Wij = log((Pij + eps^2) / ...) and its relative error says
nothing about the network.

    python tests/precision_study.py --form bf16 --point default --steps 10000 \
        --engine gpu --out result.json
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from gsbn_amd import configs  # noqa: E402


def point_spec(point: str, hcu: int, mcu: int):
    """The 100x100 workload (or a smaller one of the same shape) at one of the
    parameter points: default (taup 1 s, taue 20 ms), taup5 (taup 5 s) and
    close (taup 5 s, taue = taup - 1 ms, the cancellation case)."""
    if (hcu, mcu) == (100, 100):
        spec = configs.config1()
    else:
        spec = configs.recall_net(hcu, mcu, max(4, hcu), 50, 10, 50, 0.3, reset=50,
                                  name="study")
    q = spec.projs[0]
    if point == "taup5":
        q.taup = 5.0
    elif point == "close":
        q.taup = 5.0
        q.taue = 4.999
    elif point != "default":
        raise SystemExit("unknown point " + point)
    return spec


def rel_stats(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = np.abs(b) >= floor
    if not m.any():
        return {"max": None, "mean": None, "n": 0}
    r = np.abs(a[m] - b[m]) / np.abs(b[m])
    return {"max": float(r.max()), "mean": float(r.mean()), "n": int(m.sum()),
            "max_abs_all": float(np.abs(a - b).max())}


def abs_stats(a, b):
    d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    return {"max": float(d.max()), "mean": float(d.mean())}


def run(form: str, point: str, steps: int, engine: str, hcu: int = 100, mcu: int = 100,
        log=print) -> dict:
    spec = point_spec(point, hcu, mcu)
    spec.trace_precision = form
    L = O.lib()
    f64 = O.OracleNet(spec, arith=O.ARITH_F64)
    eg = O.ExtGen(spec)
    tmp = None
    if engine == "gpu":
        from gsbn_amd import Solver
        tmp = tempfile.TemporaryDirectory()
        net = Solver(configs.write(spec, tmp.name))
        configs.apply_conn_arrays(net, spec)
    else:
        net = O.OracleNet(spec, arith=O.ARITH_TABLE)
    t0 = time.time()
    done = 0
    for s in range(steps):
        if engine == "gpu":
            net.step(1)
            eg.update()
            flag = eg.cycle_flag
        else:
            flag = net.step()
            eg.update()
        if flag < 0:
            break
        if engine == "gpu" and net.glob("simstep") != eg.simstep:
            raise RuntimeError("schedule mismatch at step %d" % s)
        L.gor_set_globals(f64.h, eg.simstep, eg.prn, eg.old_prn, eg.cycle_flag, 0, eg.wmask_idx)
        if eg.cycle_flag == 1:
            sj = np.ascontiguousarray(net.vector("spike_0"), np.int8)
            L.gor_proj_step_replay(f64.h, 0, O._p(sj))
            L.gor_exchange_spikes(f64.h)
        done += 1
        if done % 500 == 0:
            log("%s %s step %d  %.0f s" % (form, point, done, time.time() - t0))
    res = {"form": form, "point": point, "engine": engine, "steps": done,
           "hcu": hcu, "mcu": mcu, "seconds": time.time() - t0,
           "taup": spec.projs[0].taup, "taue": spec.projs[0].taue}
    res["pij"] = rel_stats(net.vector("pij_0"), f64.vector("pij_0"), 1e-6)
    res["eij"] = rel_stats(net.vector("eij_0"), f64.vector("eij_0"), 1e-6)
    res["pi"] = rel_stats(net.vector("pi_0"), f64.vector("pi_0"), 1e-3)
    res["bj"] = rel_stats(net.vector("bj_0"), f64.vector("bj_0"), 0.0)
    res["wij"] = abs_stats(net.vector("wij_0"), f64.vector("wij_0"))
    if engine == "gpu":
        c, t, _ = net.check()
        net.close()
        tmp.cleanup()
    else:
        c, t = net.check.correct, net.check.total
    res["recall"] = {"correct": int(c), "total": int(t)}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--form", default="bf16", choices=["fp32", "bf16", "fp16"])
    ap.add_argument("--point", default="default", choices=["default", "taup5", "close"])
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--engine", default="gpu", choices=["gpu", "oracle"])
    ap.add_argument("--hcu", type=int, default=100)
    ap.add_argument("--mcu", type=int, default=100)
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    r = run(a.form, a.point, a.steps, a.engine, a.hcu, a.mcu,
            log=lambda m: print(m, flush=True))
    line = json.dumps(r)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
