"""Shared helpers for the parity tests: run the product solver and the oracle
on the same NetSpec and compare what they produce."""

import os

import numpy as np

import oracle as O
from gsbn_amd import Solver, configs

POP_VECTORS = ["dsup", "act", "epsc", "bj", "spike", ".rnd_uniform01", ".rnd_normal", ".counter"]
PROJ_VECTORS = ["ii", "qi", "di", ".siq", "pi", "ei", "zi", "ti", "pj", "ej", "zj", "pij", "eij",
                "zi2", "zj2", "tij", "wij", "si", ".ssi", ".ssj"]


def write_spec(spec, tmp_path):
    return configs.write(spec, str(tmp_path))


def state_names(spec):
    names = []
    for i in range(len(spec.pops)):
        names += ["%s_%d" % (n, i) for n in POP_VECTORS]
    for i in range(len(spec.projs)):
        names += ["%s_%d" % (n, i) for n in PROJ_VECTORS]
    return names


def compare_state(gpu, ora, spec, names=None):
    """Return a list of (name, n_mismatch, detail) for vectors that differ bitwise."""
    bad = []
    for name in names or state_names(spec):
        a = gpu.vector(name)
        b = ora.vector(name)
        if a.shape != b.shape:
            bad.append((name, -1, "shape %s vs %s" % (a.shape, b.shape)))
            continue
        if a.dtype.kind == "f":
            ne = a.view(np.uint32 if a.dtype == np.float32 else np.uint64) != \
                b.view(np.uint32 if b.dtype == np.float32 else np.uint64)
        else:
            ne = a != b
        if ne.any():
            i = int(np.nonzero(ne)[0][0])
            bad.append((name, int(ne.sum()), "first at %d: gpu %r oracle %r" % (i, a[i], b[i])))
    return bad


def run_lockstep(spec, tmp_path, max_steps=None, check_every=0):
    """Step product and oracle together; compare spikes every executed step.
    Returns (solver, oracle_net, first_divergent_step or None, steps)."""
    path = write_spec(spec, tmp_path)
    gpu = Solver(path)
    configs.apply_conn_arrays(gpu, spec)
    ora = O.OracleNet(spec)
    diverged = None
    steps = 0
    while True:
        gpu.step(1)
        f = ora.step()
        steps += 1
        if f == 1:
            sg = [gpu.vector("spike_%d" % i) for i in range(len(spec.pops))]
            so = [ora.vector("spike_%d" % i) for i in range(len(spec.pops))]
            if diverged is None and any((a != b).any() for a, b in zip(sg, so)):
                diverged = ora.extgen.simstep
                break
            if check_every and steps % check_every == 0:
                bad = compare_state(gpu, ora, spec)
                if bad:
                    diverged = ("state", ora.extgen.simstep, bad)
                    break
        if f < 0 or (max_steps is not None and steps >= max_steps):
            break
    return gpu, ora, diverged, steps


def subset(spec, nets):
    """The networks `nets` of a batch (network i = population i with its
    recurrent projection i) as a NetSpec of their own, in that order; the
    oracle keys their draws by their ids in the batch (pop_ids), so they run
    as they do inside the batch."""
    import copy
    import dataclasses
    starts = np.cumsum([0] + [p.hcu for p in spec.pops])
    cols = np.concatenate([np.arange(starts[i], starts[i + 1]) for i in nets])
    pops = [copy.deepcopy(spec.pops[i]) for i in nets]
    projs = [dataclasses.replace(spec.projs[i], src=k, dest=k) for k, i in enumerate(nets)]
    assert all(spec.projs[i].src == i and spec.projs[i].dest == i for i in nets)
    sub = dataclasses.replace(spec, pops=pops, projs=projs, stim=spec.stim[:, cols],
                              mask=spec.mask[:, cols], conn=[[] for _ in nets],
                              conn_arrays=[spec.conn_arrays[i] for i in nets])
    sub.pop_ids = list(nets)
    return sub


def compare_nets(gpu, ora, spec, nets, names=None):
    """compare_state for networks `nets` of a batch on the product side
    against networks 0.. of their subset on the oracle side."""
    bad = []
    for k, i in enumerate(nets):
        for base in (names or POP_VECTORS + PROJ_VECTORS):
            a = gpu.vector("%s_%d" % (base, i))
            b = ora.vector("%s_%d" % (base, k))
            if a.shape != b.shape:
                bad.append((i, base, "shape %s vs %s" % (a.shape, b.shape)))
                continue
            if a.dtype.kind == "f":
                ne = a.view(np.uint32) != b.view(np.uint32)
            else:
                ne = a != b
            if ne.any():
                j = int(np.nonzero(ne)[0][0])
                bad.append((i, base, int(ne.sum()), "first at %d: gpu %r oracle %r" % (j, a[j], b[j])))
    return bad
