#!/usr/bin/env python
"""Benchmark of the lazy-update hot path on one MI355X.

A step is one simulated timestep (1 ms of model time) of the 100 HCU x 100
MCU network with its dense recurrent projection (9.9e7 synapses, FP32 state):
ProcExtGen, ProcUpdLazy (population update, row and column lazy updates, delay
queue), ProcExchangeSpike and ProcCheck, exactly as the solver runs them.
Synthetic input: seeded one-MCU-per-HCU patterns.  The timed steps are taken
from the training phase after the warm-up steps (the first step carries the
full catch-up sweep of the learning-rate switch and is part of the warm-up).

Step plan: the online plan tuner first runs to completion on a solver of its
own (same network and schedule); the measured solvers then run its choice
from their first step, so no tuning sample falls inside a timed window.

Printed value: synapse-trace updates per second, counted as the reference's
loops do (each active row updates M synapses, each active column C synapses,
plus any full-sweep synapses), with simulated steps per second beside it.

roofline: the kernel with the largest share of kernel time in the timed
window (the population kernel at 100 x 100), from a second solver that
replays the same steps with HIP events recorded by every launch (results are
deterministic, so the replay runs the same work; the timed solver itself
runs uninstrumented).  Algorithmic bytes: 44 per synapse update for the
row/column kernel (read pij, eij, zi2, zj2, tij; write them and wij; 36 in
the 16-bit forms), 73 per unit for the population kernel (which fuses the
postsynaptic trace update) plus 4 per support contribution it reads.
`traffic`: HBM bytes per launch from rocprofv3 request counters of the same
kernel and plan (profiles/traffic_<workload>.json, see gsbn_amd.traffic),
scaled to this run's launches.  `kernels` lists every timed kernel.

forms: the same workload with Pij/Eij stored in bf16 and in scaled fp16.
batched: 512 independent networks in lock-step (batched_bf16: the same with
Pij/Eij in bf16).  frozen_recall: 8192 cues of
the trained network, exact and bf16-MFMA support.  hbm_copy: the library's
copy kernel on 2 GiB, the practical bandwidth ceiling of the box.

cpu_baseline: this project's CPU restatement of the same procedure (the
oracle, plain C, -O2, one thread pinned to one core) with the reference's
float closed forms and libm exp/log per update (Proj.cpp:241-277), on the
same timed steps (`same_window`) and on a bounded sample of the training
phase after them; the oracle's table arithmetic is timed beside it.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("synapse-trace updates/sec + sim steps/sec at 100×100 HCU×MCU, plain vs "
          "BF16/batched")
# algorithmic bytes per synapse update: FP32 form reads pij, eij, zi2, zj2, tij
# and writes them and wij (row) or reads/writes the five and reads ii
# (column); 16-bit forms move 2-byte pij/eij
BYTES_PER_SYN_UPDATE = {"fp32": 44, "bf16": 36, "fp16": 36}
# population kernel (population update with the postsynaptic traces fused):
# reads dsup, lginp, counter, epsc, bj, pj, ej, zj (32 B), writes dsup, act,
# the two draws, counter, epsc, bj, pj, ej, zj (40 B) and the spike (1 B);
# plus 4 B per support contribution of the previous step's row updates
POP_BYTES_PER_UNIT = 73
POP_BYTES_PER_CONTRIB = 4
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense bf16


def workload(name: str):
    from gsbn_amd import configs
    if name == "100x100":
        return configs.config1(), ("lazy BCPNN update, 100 HCU x 100 MCU, dense recurrent "
                                   "projection (9.9e7 synapses), plain FP32 state")
    if name == "batch512":
        return configs.batched_nets(512), ("lazy BCPNN update, 512 independent networks "
                                           "(HCU 8-64, MCU 16-128, sparse recurrent p in "
                                           "{0.1,0.5,0.9}), plain FP32 state, lock-step")
    if name == "10x10":
        return configs.config0(), ("lazy BCPNN update, 10 HCU x 10 MCU, dense recurrent "
                                   "projection (9e3 synapses), plain FP32 state")
    raise SystemExit("unknown workload " + name)


DEFERRED = None  # --deferred K, applied with the plan to every measured spec
POP_WAVE = None  # --pop-block: 0 (the block population kernel for batches)
OVF_KEEP = None  # --ovf-keep K: the one-wave row/column plan's row overflow queue


def apply_plan(spec, plan: str):
    """--plan: auto leaves the choice to the online tuner; --deferred K
    defers the column updates."""
    spec.deferred_columns = DEFERRED
    spec.pop_wave = POP_WAVE
    spec.ovf_keep = OVF_KEEP
    if plan == "auto":
        return
    if plan == "fused":
        spec.fused_step = 1
    elif plan == "chained":
        spec.chained_step = 1
    elif plan == "wave":
        spec.rowcol_wave = 1
        spec.chained_step = -1
    elif plan.startswith("graded-wave:"):
        spec.rowcol_wave = 1 + int(plan.split(":", 1)[1])
        spec.chained_step = -1
    elif plan.startswith("graded:"):
        spec.rowcol_graded = int(plan.split(":", 1)[1])
        spec.chained_step = -1
    elif plan.startswith("split:"):
        spec.rowcol_parts = int(plan.split(":", 1)[1])
        spec.chained_step = -1
    else:
        raise SystemExit("unknown plan " + plan)


def cpu_baseline(args) -> dict:
    """Run the oracle in child processes pinned to one core: first with the
    reference's own float closed-form arithmetic (Proj.cpp:241-277, libm
    exp per synapse update), the baseline, then with this project's
    coefficient-table arithmetic, reported beside it."""
    def child(arith, seconds):
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-only", "--workload",
               args.workload, "--warmup", str(args.warmup), "--steps", str(args.steps),
               "--cpu-seconds", str(seconds), "--cpu-arith", arith]
        out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                             timeout=900)
        if out.returncode != 0:
            return {"value": None, "unit": "synapse-trace updates/s", "cores": 1, "kind": "port",
                    "sample": "failed: " + (out.stderr.strip().splitlines() or ["?"])[-1][:200]}
        return json.loads(out.stdout.strip().splitlines()[-1])
    base = child("ref32", args.cpu_seconds)
    if args.cpu_seconds > 0:
        base["table_arith"] = child("table", args.cpu_seconds / 2)
    return base


def sweep_steps(spec, upto: int):
    """Executed steps at which the learning rate changes (a full catch-up
    sweep), from the schedule alone."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    eg = O.ExtGen(spec)
    out = []
    for _ in range(upto):
        eg.update()
        if eg.cycle_flag < 0:
            break
        if eg.cycle_flag == 1 and eg.prn != eg.old_prn:
            out.append(eg.simstep)
    return out


def cpu_only(args):
    """The oracle on one core: `warmup` steps, then the GPU's timed window
    (steps warmup+1 .. warmup+steps) and on, without crossing a full sweep,
    until about --cpu-seconds have passed."""
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    spec, _ = workload(args.workload)
    arith = {"ref32": O.ARITH_REF32, "table": O.ARITH_TABLE}[args.cpu_arith]
    net = O.OracleNet(spec, arith)
    for _ in range(args.warmup):
        net.step()
    sweeps = [t for t in sweep_steps(spec, args.warmup + 200000) if t > args.warmup]
    stop = sweeps[0] if sweeps else None
    w0 = net.work()
    t0 = time.perf_counter()
    n, win = 0, None
    while True:
        step = args.warmup + n + 1
        if stop is not None and step >= stop:
            break
        if n >= args.steps and time.perf_counter() - t0 >= args.cpu_seconds:
            break
        if net.step() < 0:
            break
        n += 1
        if n == args.steps:
            tw = time.perf_counter() - t0
            ww = net.work()
            win = {"steps": n, "seconds": tw,
                   "value": float(np.sum(ww[:3] - w0[:3])) / tw, "steps_per_sec": n / tw}
    dt = time.perf_counter() - t0
    w1 = net.work()
    upd = int(np.sum(w1[:3] - w0[:3]))
    desc = {"ref32": "the reference's float closed forms with libm exp/log per update "
                     "(Proj.cpp:241-277)",
            "table": "this project's FP64-built coefficient table, float FMA-free updates"}
    print(json.dumps({"value": upd / dt, "unit": "synapse-trace updates/s", "cores": 1,
                      "kind": "port", "steps_per_sec": n / dt, "arith": desc[args.cpu_arith],
                      "same_window": win,
                      "sample": "steps %d..%d of the same workload (the GPU's timed steps "
                                "%d..%d and the training steps after them, no full sweep), "
                                "%.1f s, 1 thread pinned to one core"
                                % (args.warmup + 1, args.warmup + n, args.warmup + 1,
                                   args.warmup + args.steps, dt)}))


PLAN_NAMES = {-1: "fused", -2: "chained", -20: "wave"}


def plan_name(parts: int) -> str:
    if -20 < parts <= -10:
        return "graded-%d" % (-10 - parts)
    if -24 < parts <= -21:
        return "graded-wave-%d" % (-20 - parts)
    return PLAN_NAMES.get(parts, "split:%d" % parts)


def fixed_plan(spec) -> bool:
    return bool(spec.fused_step or spec.chained_step > 0 or spec.rowcol_parts or
                spec.rowcol_graded or spec.rowcol_wave)


def fix_plan(spec, parts: int):
    """Fix the step plan named by a GlobalVar rowcol-parts value."""
    if parts == -1:
        spec.fused_step = 1
    elif parts == -2:
        spec.chained_step = 1
    elif parts <= -20:
        spec.rowcol_wave = -19 - parts
        spec.chained_step = -1
    elif parts <= -10:
        spec.rowcol_graded = -10 - parts
        spec.chained_step = -1
    else:
        spec.rowcol_parts = parts
        spec.chained_step = -1


def settle_plan(spec, max_steps: int = 5000):
    """Run the online plan tuner to completion on a solver of its own (same
    network, same schedule) and return (rowcol-parts it chose, steps it
    took), so that the measured solvers run that plan from their first step
    and no tuning sample falls inside a timed window."""
    from gsbn_amd import Solver, configs
    with tempfile.TemporaryDirectory() as d:
        s = Solver(configs.write(spec, d))
        configs.apply_conn_arrays(s, spec)
        n = 0
        while n < max_steps:
            if s.step(1) == 0:
                break
            n += 1
            if s.glob("plan-tuned"):
                break
        parts = s.glob("rowcol-parts")
        tuned = s.glob("plan-tuned")
        s.close()
    if not tuned:
        raise SystemExit("bench: the plan tuner did not finish within %d steps" % n)
    return parts, n


def _steps(s, n):
    if n > 0:
        s.step(n)


def _solver(spec, d):
    from gsbn_amd import Solver, configs
    t = time.perf_counter()
    s = Solver(configs.write(spec, d))
    configs.apply_conn_arrays(s, spec)
    s.sync()
    return s, time.perf_counter() - t


def measure(spec, steps, warmup, step_times=0):
    """Run `warmup` steps, then time `steps` steps with a fixed step plan
    (the one the online tuner chooses for this workload, found beforehand on
    a separate solver, unless the spec fixes one).

    Kernel times come from a second solver of the same spec and plan that
    replays the same steps (every result is deterministic) with HIP events
    recorded by each launch of the row/column kernel, so the timed steps run
    uninstrumented and the kernel figures cover exactly the launches of the
    timed window.  In the chained plan the launch of step x runs the row and
    column updates of step x-1; the work of a window is counted as what its
    launches execute, in both solvers' figures."""
    import copy
    spec = copy.deepcopy(spec)
    tune_steps = 0
    if not fixed_plan(spec):
        parts, tune_steps = settle_plan(spec)
        fix_plan(spec, parts)
    with tempfile.TemporaryDirectory() as d:
        s, t_init = _solver(spec, d)
        lag = 1 if spec.chained_step > 0 else 0
        if warmup < 1 or steps < 1:
            raise SystemExit("bench: needs --warmup >= 1 and --steps >= 1")
        _steps(s, warmup)
        s.sync()
        parts = s.glob("rowcol-parts")
        deferred = s.glob("deferred-columns")
        pop_wave = s.glob("pop-wave")
        t0 = time.perf_counter()
        done = s.step(steps)
        s.sync()
        t1 = time.perf_counter()
        step_ms = None
        if step_times:
            # per-step wall times (each step synchronised) for the p50/p99 report
            ts = []
            for _ in range(step_times):
                a0 = time.perf_counter()
                s.step(1)
                s.sync()
                ts.append(1e3 * (time.perf_counter() - a0))
            step_ms = {"p50": float(np.percentile(ts, 50)), "p99": float(np.percentile(ts, 99)),
                       "n": len(ts)}
        s.close()
    with tempfile.TemporaryDirectory() as d:
        r, _ = _solver(spec, d)
        _steps(r, warmup - lag)
        k0 = r.work()
        _steps(r, lag)
        r.timing(-1, collect=False)
        r.timing((1 << 0) | (1 << 1) | (1 << 6), collect=False)
        _steps(r, steps - lag)
        k1 = r.work()
        _steps(r, lag)
        r.sync()
        launches, ms = r.timing(0)
        r.timing(-1, collect=False)
        r.close()
    el = t1 - t0
    rows = int(k1[0] - k0[0])
    cols = int(k1[1] - k0[1])
    sweep = int(k1[2] - k0[2])
    bpu = BYTES_PER_SYN_UPDATE[spec.trace_precision]
    # one launch per step (the fused plan, -1, or the chained plan, -2) is
    # timed as kind 6; split plans time the population launches (kind 0) and
    # the row/column launches (kind 1)
    one_launch = parts in (-1, -2)
    batch = len(spec.pops) > 1
    rc_name = (("chained_kernel" if parts == -2 else "step_kernel_batch" if batch else "step_kernel")
               if one_launch else "rowcol_kernel_batch_wave" if parts == -20
               else "rowcol_kernel_graded_wave" if parts <= -21
               else "rowcol_kernel_graded" if parts <= -10
               else "rowcol_kernel_batch" if batch else "rowcol_kernel")
    kind = 6 if one_launch else 1
    other = int(launches[1]) if one_launch else int(launches[6])
    if other:
        raise SystemExit("bench: %d launches of another plan in the kernel window" % other)
    kernels = {}
    n_rc = int(launches[kind])
    rc_bytes = bpu * (rows + cols) / max(n_rc, 1)
    # the one-launch plans sum the support contributions on chip
    pop_bytes = POP_BYTES_PER_UNIT * spec.total_units
    if one_launch:
        rc_bytes += pop_bytes
    elif int(launches[0]):
        pop_bytes += POP_BYTES_PER_CONTRIB * rows / int(launches[0])
    kernels[rc_name] = (n_rc, float(ms[kind]), rc_bytes)
    if not one_launch and int(launches[0]):
        kernels[("pop_kernel_batch_wave" if pop_wave > 0 else "pop_kernel_batch") if batch
                else "pop_kernel"] = (
            int(launches[0]), float(ms[0]), float(pop_bytes))
    total_ms = sum(v[1] for v in kernels.values())
    table = {}
    for name, (n, tot, b) in kernels.items():
        avg = tot / max(n, 1)
        gbs = b / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
        table[name] = {"launches": n, "avg_launch_ms": avg, "bytes_per_launch": b,
                       "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
                       "share_of_kernel_time": tot / total_ms if total_ms > 0 else None}
    dominant = max(kernels, key=lambda k: kernels[k][1])
    d = table[dominant]
    return {"elapsed": el, "done": int(done), "rows": rows, "cols": cols, "sweep": sweep,
            "updates_per_s": (rows + cols + sweep) / el, "steps_per_s": done / el,
            "ms_per_step": 1e3 * el / max(done, 1), "init_s": t_init, "step_ms": step_ms,
            "rc_launches": table[rc_name]["launches"], "rc_avg_ms": table[rc_name]["avg_launch_ms"],
            "rc_bytes_per_launch": rc_bytes, "rc_achieved_gbs": table[rc_name]["achieved_gbs"],
            "rc_parts": parts if parts > 0 else None, "rc_plan": plan_name(parts),
            "tune_steps": tune_steps, "rc_kernel": rc_name, "kernels": table,
            "dominant": dominant, "dom_avg_ms": d["avg_launch_ms"],
            "dom_bytes_per_launch": d["bytes_per_launch"], "dom_achieved_gbs": d["achieved_gbs"],
            "dom_launches": d["launches"], "deferred": deferred,
            "kernel_window": "steps %d..%d, replayed with events on every launch"
                             % (warmup + 1, warmup + steps)}


def traffic_of(workload_name: str, m: dict, kernel: str):
    """HBM bytes per launch of `kernel` in the run `m`: the ratio of counter
    bytes to algorithmic bytes that the counter runs of this workload, kernel
    and plan measured (profiles/traffic_<workload>.json, written by
    gsbn_amd.traffic), applied to this run's algorithmic bytes per launch.
    (None, None) when no such measurement exists."""
    tp = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload_name)
    if not os.path.exists(tp):
        return None, None
    with open(tp) as f:
        rec = json.load(f).get("kernels", {}).get(kernel)
    plan = m["rc_plan"] if not kernel.startswith("pop_kernel") else "any"
    if not (rec and rec.get("plan") == plan and rec.get("deferred_columns") == m["deferred"] and
            rec.get("algorithmic_bytes_per_launch")):
        return None, None
    ratio = rec["hbm_bytes_per_launch"] / rec["algorithmic_bytes_per_launch"]
    return int(round(ratio * m["kernels"][kernel]["bytes_per_launch"])), ratio


def copy_calibration(mib: int):
    """Five copies of `mib` MiB by gsbn_device_copy (buffers from torch)."""
    import ctypes
    import torch
    from gsbn_amd import lib
    L = lib()
    L.gsbn_device_copy.restype = ctypes.c_int
    L.gsbn_device_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.c_void_p]
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(5):
        if L.gsbn_device_copy(dst.data_ptr(), src.data_ptr(), n, stream) != 0:
            raise SystemExit("gsbn_device_copy failed")
    torch.cuda.synchronize()
    del src, dst


def _hip_runtime():
    """The HIP runtime that libgsbn_amd.so itself loaded (ctypes)."""
    import ctypes
    from gsbn_amd import lib
    lib()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    if path is None:
        raise SystemExit("libamdhip64 not loaded")
    return ctypes.CDLL(path)


def hbm_copy(gib: int = 2, reps: int = 10) -> dict:
    """Device-to-device copy bandwidth with this library's copy kernel
    (gsbn_device_copy, 16-byte loads and stores): 2 x bytes per copy over
    the event-timed copies, beside the 8 TB/s the roofline divides by."""
    import ctypes
    from gsbn_amd import lib
    L = lib()
    H = _hip_runtime()
    vp = ctypes.c_void_p
    L.gsbn_device_copy.restype = ctypes.c_int
    L.gsbn_device_copy.argtypes = [vp, vp, ctypes.c_int64, vp]
    L.gsbn_i32_fill.restype = ctypes.c_int
    L.gsbn_i32_fill.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, vp]
    H.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    H.hipFree.argtypes = [vp]
    H.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    H.hipEventRecord.argtypes = [vp, vp]
    H.hipEventSynchronize.argtypes = [vp]
    H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    H.hipEventDestroy.argtypes = [vp]
    H.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    n = gib << 30
    src, dst = vp(), vp()
    if H.hipMalloc(ctypes.byref(src), n) or H.hipMalloc(ctypes.byref(dst), n):
        raise SystemExit("hipMalloc failed")
    try:
        if L.gsbn_i32_fill(src, n // 4, 0x5a5a1234, None) != 0:
            raise SystemExit("gsbn_i32_fill failed")
        for _ in range(2):
            L.gsbn_device_copy(dst, src, n, None)
        e0, e1 = vp(), vp()
        H.hipEventCreate(ctypes.byref(e0))
        H.hipEventCreate(ctypes.byref(e1))
        H.hipEventRecord(e0, None)
        for _ in range(reps):
            if L.gsbn_device_copy(dst, src, n, None) != 0:
                raise SystemExit("gsbn_device_copy failed")
        H.hipEventRecord(e1, None)
        H.hipEventSynchronize(e1)
        ms = ctypes.c_float(0.0)
        H.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        H.hipEventDestroy(e0)
        H.hipEventDestroy(e1)
        check = np.zeros(1 << 18, np.int32)
        H.hipMemcpy(check.ctypes.data, dst, check.nbytes, 2)  # device to host
        ok = bool((check == 0x5a5a1234).all())
    finally:
        H.hipDeviceSynchronize()
        H.hipFree(src)
        H.hipFree(dst)
    t = ms.value / reps
    gbs = 2.0 * n / (t * 1e-3) / 1e9
    return {"bytes_per_copy": n, "ms_per_copy": t, "gbs": gbs, "frac_of_8tbs": gbs / HBM_PEAK_GBS,
            "copy_checked": ok}


def batched(steps: int, warmup: int, form: str = "fp32") -> dict:
    """The 512-network form: aggregate and per-network step rates."""
    spec, desc = workload("batch512")
    spec.deferred_columns = DEFERRED
    spec.pop_wave = POP_WAVE
    spec.ovf_keep = OVF_KEEP
    spec.trace_precision = form
    if form != "fp32":
        desc = desc.replace("plain FP32 state", "Pij/Eij stored in " + form)
    m = measure(spec, steps, warmup, step_times=100)
    traffic, ratio = traffic_of("batch512", m, m["dominant"])
    n = len(spec.pops)
    return {"workload": desc, "networks": n, "steps": m["done"],
            "steps_per_sec": m["steps_per_s"], "network_steps_per_sec": n * m["steps_per_s"],
            "synapse_updates_per_sec": m["updates_per_s"],
            "step_ms_p50": m["step_ms"]["p50"], "step_ms_p99": m["step_ms"]["p99"],
            "rowcol_avg_launch_ms": m["rc_avg_ms"], "rowcol_achieved_gbs": m["rc_achieved_gbs"],
            "rowcol_workgroups_per_hcu": m["rc_parts"], "step_plan": m["rc_plan"],
            "timed_kernel": m["rc_kernel"], "deferred_columns": m["deferred"],
            "kernels": m["kernels"], "dominant_kernel": m["dominant"],
            "dominant_frac": m["dom_achieved_gbs"] / HBM_PEAK_GBS,
            "dominant_traffic": traffic, "dominant_traffic_ratio": ratio}


def vendor_gemm_tflops(m: int, n: int, k: int, reps: int = 10) -> float:
    """The vendor library's rate on the support product's shape: torch.matmul
    (hipBLASLt) of bf16 (m x k) and (k x n), bf16 out, no fused decay, timed
    with HIP events; a scale for the hand-written product, not a path."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(5)
    a = (torch.rand(m, k, device="cuda", generator=g) < 0.001).to(torch.bfloat16)
    b = torch.randn(k, n, device="cuda", generator=g).to(torch.bfloat16)
    for _ in range(3):
        c = torch.matmul(a, b)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        c = torch.matmul(a, b)
    e1.record()
    torch.cuda.synchronize()
    del a, b, c
    return 2.0 * m * n * k * reps / (e0.elapsed_time(e1) * 1e-3) / 1e12


def frozen_recall(batch: int = 8192, steps: int = 50, clamp_frac: float = 0.4) -> dict:
    """Train the 100x100 network (its training phase through the sweep of the
    first recall step), freeze W and bj, run `batch` cues for `steps` steps
    with the exact float support and with the bf16 matrix-core support."""
    from gsbn_amd import Solver, configs
    from gsbn_amd import recall as R
    spec = configs.config1()
    with tempfile.TemporaryDirectory() as d:
        s = Solver(configs.write(spec, d))
        configs.apply_conn_arrays(s, spec)
        trained = 0
        while True:
            s.step(1)
            trained += 1
            if s.glob("cycle-flag") == 1 and s.glob("prn", "f") == 0.0:
                break
        pats = np.asarray(spec.stim[:100], np.int64)
        clamp, wmask, want = R.cues(pats, batch, clamp_frac, seed=11)
        out = {}
        for name, mode in (("exact", R.EXACT), ("mfma", R.MFMA)):
            rc = R.Recall.from_solver(s, spec, batch, mode, seed=77)
            rc.run(clamp, wmask, 4)  # first launch of every kernel, untimed
            t0 = time.perf_counter()
            idx, cnt, _, _ = rc.run(clamp, wmask, steps, count_from=steps // 2)
            el = time.perf_counter() - t0
            ok = R.recall_correct(idx, want)
            out[name] = {"cues_per_sec": batch / el, "seconds": el, "recall_correct": float(ok.mean()),
                         "hcu_winner_accuracy": float(np.mean(idx == want)), "idx": idx}
            if mode == R.MFMA:
                # support products run from step delay+3 on
                _, _, _, sup_ms = rc.run(clamp, wmask, steps, count_from=steps // 2,
                                         time_support=True)
                N = spec.total_units
                nprod = steps - 3
                flops = 2.0 * batch * N * N * nprod
                out[name]["support_ms"] = sup_ms
                out[name]["mfma_tflops"] = flops / (sup_ms * 1e-3) / 1e12 if sup_ms > 0 else None
            rc.close()
        s.close()
    agree = float(np.mean(R.recall_correct(out["exact"]["idx"], want) ==
                          R.recall_correct(out["mfma"]["idx"], want)))
    winners = float(np.mean(out["exact"]["idx"] == out["mfma"]["idx"]))
    for v in out.values():
        v.pop("idx")
    mf = out["mfma"]
    vendor = vendor_gemm_tflops(batch, spec.total_units, spec.total_units)
    return {"workload": "100x100 network trained %d steps, W and bj frozen (bf16 for mfma), "
                        "%d cues (%.0f%% of HCUs clamped) x %d lock-step steps"
                        % (trained, batch, 100 * clamp_frac, steps),
            "exact": out["exact"], "mfma": mf, "outcome_agreement": agree,
            "winner_agreement": winners,
            "mfma_roofline": {"bound": "mfma", "achieved": mf.get("mfma_tflops"),
                              "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": (mf["mfma_tflops"] / MFMA_BF16_PEAK_TFLOPS
                                       if mf.get("mfma_tflops") else None),
                              "vendor_library_tflops": vendor,
                              "vs_vendor_library": (mf["mfma_tflops"] / vendor
                                                    if mf.get("mfma_tflops") else None)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="100x100")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the bf16 / fp16 trace-storage runs of the same workload")
    ap.add_argument("--no-batched", action="store_true", help="skip the 512-network form")
    ap.add_argument("--no-recall", action="store_true", help="skip the frozen-weight recall form")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB five times with the library's copy kernel "
                         "(reference point for rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--plan", default="auto",
                    help="step plan: auto (tuned online), split:N (N row/column workgroups "
                         "per HCU), fused, chained, or for batches wave (one wave per HCU), "
                         "graded:G (parts per size bucket, G = 1-3) or graded-wave:G (the "
                         "same with one wave per part)")
    ap.add_argument("--deferred", type=int, default=None, metavar="K",
                    help="deferred column updates with rings of K spike steps per unit "
                         "(ProcUpdLazy deferred-columns; 0: eager; default: the library's, 8 for one network, 16 for a batch)")
    ap.add_argument("--pop-block", action="store_true",
                    help="batches: the population update with one workgroup per hypercolumn "
                         "(ProcUpdLazy pop-wave 0) instead of one wave")
    ap.add_argument("--ovf-keep", type=int, default=None, metavar="K",
                    help="batches, one-wave row/column plan: active rows a hypercolumn keeps "
                         "for its own wave, the rest queued for helper waves (ProcUpdLazy "
                         "ovf-keep; 0: no queue; default: the library's, 2)")
    ap.add_argument("--cpu-only", action="store_true")
    ap.add_argument("--cpu-arith", default="ref32", choices=("ref32", "table"))
    ap.add_argument("--step-times", type=int, default=0,
                    help="after the timed region, time this many synchronised single steps")
    args = ap.parse_args()
    global DEFERRED, POP_WAVE, OVF_KEEP
    DEFERRED = args.deferred
    POP_WAVE = 0 if args.pop_block else None
    OVF_KEEP = args.ovf_keep
    if args.cpu_only:
        return cpu_only(args)
    if args.gpus != 1:
        print("bench.py: runs on one GPU only (sharding across GPUs by destination "
              "hypercolumn is a design, not built). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # torch before the library: it brings its own HIP runtime, and
    # libgsbn_amd.so then binds to that one (same soname); the other way round
    # the process would hold two runtimes and torch would find no device
    import torch  # noqa: F401
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    spec, desc = workload(args.workload)
    apply_plan(spec, args.plan)
    m = measure(spec, args.steps, args.warmup, args.step_times)
    traffic, traffic_ratio = traffic_of(args.workload, m, m["dominant"])
    achieved = m["dom_achieved_gbs"]
    line = {
        "metric": METRIC,
        "value": m["updates_per_s"],
        "unit": "synapse-trace updates/s",
        "n_gpus": 1,
        "steps": m["done"],
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "networks": len(spec.pops),
                   "hcu": spec.total_hcu, "units": spec.total_units,
                   "synapses": sum(p.hcu * int(q.slot_num) * p.mcu
                                   for p, q in zip(spec.pops, spec.projs)),
                   "phase": "training steps %d..%d" % (args.warmup + 1, args.warmup + m["done"])},
        "steps_per_sec": m["steps_per_s"],
        "work": {"row_synapse_updates": m["rows"], "col_synapse_updates": m["cols"],
                 "sweep_synapse_updates": m["sweep"]},
        "init_s": m["init_s"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_ratio": traffic_ratio,
                     "kernel": m["dominant"], "avg_launch_ms": m["dom_avg_ms"],
                     "launches": m["dom_launches"], "bytes_per_launch": m["dom_bytes_per_launch"],
                     "window": m["kernel_window"],
                     "workgroups_per_hcu": m["rc_parts"], "step_plan": m["rc_plan"],
                     "deferred_columns": m["deferred"],
                     "plan_tuning_steps": m["tune_steps"], "kernels": m["kernels"]},
    }
    if m["step_ms"] is not None:
        line["step_ms"] = m["step_ms"]
    if len(spec.pops) > 1:
        line["network_steps_per_sec"] = len(spec.pops) * m["done"] / m["elapsed"]
    if not args.no_forms and args.workload == "100x100":
        forms = {}
        for form in ("bf16", "fp16"):
            fspec, _ = workload(args.workload)
            fspec.trace_precision = form
            apply_plan(fspec, args.plan)
            f = measure(fspec, args.steps, args.warmup)
            forms[form] = {"value": f["updates_per_s"], "steps_per_sec": f["steps_per_s"],
                           "ms_per_step": f["ms_per_step"],
                           "speedup_steps_per_sec": f["steps_per_s"] / m["steps_per_s"],
                           "rowcol_avg_launch_ms": f["rc_avg_ms"],
                           "rowcol_achieved_gbs": f["rc_achieved_gbs"],
                           "timed_kernel": f["rc_kernel"], "step_plan": f["rc_plan"],
                           "bytes_per_update": BYTES_PER_SYN_UPDATE[form],
                           "kernels": f["kernels"]}
        line["forms"] = forms
    if args.workload == "100x100" and not args.no_batched:
        line["batched"] = batched(300, 100)
        line["batched"]["speedup_network_steps_per_sec"] = (
            line["batched"]["network_steps_per_sec"] / m["steps_per_s"])
        if not args.no_forms:
            # the BF16-state batched form against the plain single network
            line["batched_bf16"] = batched(300, 100, "bf16")
            line["batched_bf16"]["speedup_network_steps_per_sec"] = (
                line["batched_bf16"]["network_steps_per_sec"] / m["steps_per_s"])
    if args.workload == "100x100" and not args.no_recall:
        line["frozen_recall"] = frozen_recall()
    if args.workload == "100x100" and not args.no_forms:
        line["hbm_copy"] = hbm_copy()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
