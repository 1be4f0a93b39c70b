"""The chained step launch (gsbn_step_chained) on the GPU (-m gpu): forced
with ProcUpdLazy argi chained-step 1, each launch runs the previous step's
row updates (in the population workgroups, their weights added to the
support sums in ascending row order) and column updates (further
workgroups) together with this step's population update.  It must give
the oracle's spike rasters and state bit for bit.

The lockstep runs read only spike vectors between steps, which leaves the
row/column work pending from launch to launch; the state comparisons read
the synapse views, which complete it first (the flush), so both paths run.
Cases: the 10x10 workload with its recall results, periodic full-state
comparisons, the bf16 and fp16 forms across the learning-rate switches,
external spikes, 120 steps of the 100x100 network through its first sweep,
the sparse map with delays 1..4, a resume from a snapshot taken with work
pending, host writes of weights and traces between steps, structural
plasticity, the online choice among the plans, and the configurations the
plan refuses."""

import os

import numpy as np
import pytest

import oracle as O
from gsbn_amd import GsbnError, Solver, configs
from gsbn_amd import snapshot as S

import helpers

pytestmark = pytest.mark.gpu


def _check(spec, tmp_path, max_steps=None, recall=False, check_every=0):
    spec.chained_step = 1
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=max_steps,
                                                check_every=check_every)
    assert div is None, "diverged at %s" % (div,)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    if recall:
        c, t, flags = gpu.check()
        assert (c, t) == (ora.check.correct, ora.check.total)
    assert gpu.glob("rowcol-parts") == -2
    gpu.close()


def test_chained_config0_full_run(tmp_path):
    _check(configs.config0(), tmp_path, recall=True)


def test_chained_with_periodic_state_reads(tmp_path):
    _check(configs.config0(), tmp_path, max_steps=700, check_every=20)


@pytest.mark.parametrize("form", ["bf16", "fp16"])
def test_chained_16bit_forms(tmp_path, form):
    spec = configs.config0()
    spec.trace_precision = form
    _check(spec, tmp_path, max_steps=1300)


def test_chained_external_spikes(tmp_path):
    rng = np.random.default_rng(9)
    spec = configs.config0()
    spec.ext_spikes = {0: {s: sorted(rng.choice(100, size=rng.integers(0, 6), replace=False).tolist())
                           for s in range(1, 400)}}
    _check(spec, tmp_path, max_steps=400)


def test_chained_config1_size(tmp_path):
    _check(configs.config1(), tmp_path, max_steps=120)


def test_chained_config1_state_every_100_steps(tmp_path):
    _check(configs.config1(), tmp_path, max_steps=300, check_every=100)


def test_chained_sparse_delay_map(tmp_path):
    text = open(os.path.join(os.path.dirname(__file__), "golden", "init_conn_0.1.map")).read()
    spec = configs.recall_net(400, 10, 4, 60, 10, 60, 0.5, reset=20, name="sparse")
    spec.projs[0].slot_num = 60.0
    spec.conn_text = text
    spec.conn = configs.parse_conn_map(text)
    _check(spec, tmp_path, max_steps=320, check_every=40)


def test_chained_host_writes_between_steps(tmp_path):
    # weights (a synapse view) and pj (a population-side vector) written by
    # the host with row/column work pending: the work runs first, on the
    # state it was computed from, then the writes
    spec = configs.config0()
    spec.chained_step = 1
    path = configs.write(spec, str(tmp_path))
    gpu = Solver(path)
    ora = O.OracleNet(spec)
    rng = np.random.default_rng(3)
    for step in range(1, 501):
        gpu.step(1)
        f = ora.step()
        if f != 1:
            continue
        assert np.array_equal(gpu.vector("spike_0"), ora.vector("spike_0")), step
        if step in (150, 333):
            pj = ora.vector("pj_0") * np.float32(1.0 + 0.01 * rng.random())
            gpu.write("pj_0", pj)
            ora.write("pj_0", pj)
        if step in (200, 401):
            w = ora.vector("wij_0")
            w = (w + np.float32(0.05) * rng.standard_normal(w.shape).astype(np.float32))
            gpu.write("wij_0", w.astype(np.float32))
            ora.write("wij_0", w.astype(np.float32))
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    gpu.close()


def test_chained_resume_with_pending_work(tmp_path):
    spec = configs.config0()
    spec.chained_step = 1
    path = configs.write(spec, str(tmp_path))
    snap = str(tmp_path / "state.bin")
    a = Solver(path)
    while a.glob("simstep") < 52:
        a.step(1)
    a.snapshot(snap)  # reads every vector: the pending updates run first
    # vectors whose names start with '.' are not saved (the reference's rule)
    names = [n for n in helpers.state_names(spec) if not n.startswith(".")]
    spikes_a = []
    while a.glob("simstep") < 400:
        a.step(1)
        spikes_a.append(a.vector("spike_0").copy())
    want = {n: a.vector(n).copy() for n in names}
    a.close()
    b = Solver(path, snap)
    spikes_b = []
    while b.glob("simstep") < 400:
        b.step(1)
        spikes_b.append(b.vector("spike_0").copy())
    assert len(spikes_a) == len(spikes_b)
    for i, (x, y) in enumerate(zip(spikes_a, spikes_b)):
        assert np.array_equal(x, y), i
    for n in names:
        g = b.vector(n)
        assert np.array_equal(g.view(np.uint8), want[n].view(np.uint8)), n
    b.close()


def test_chained_structural_plasticity(tmp_path):
    spec = configs.config0()
    for m in spec.modes:
        m.plasticity = 1
    spec.structural = configs.StructuralSpec(period=20, pruning_period=60, t_th=30, seed=7)
    _check(spec, tmp_path, max_steps=400, check_every=20)


def test_online_choice_includes_chained(tmp_path):
    spec = configs.config0()
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=400)
    assert div is None, "diverged at %s" % (div,)
    assert not helpers.compare_state(gpu, ora, spec)
    assert gpu.glob("step-plan-ms-2", "f") > 0.0  # the chained plan was timed
    gpu.close()


@pytest.mark.parametrize("what", ["wta", "batched", "queue"])
def test_chained_refused(tmp_path, what):
    if what == "wta":
        spec = configs.config0()
        spec.wta_pop_max = 1
    elif what == "queue":
        # the population kernel does not push its own spikes
        spec = configs.config0()
        spec.exchange_direct = False
    else:
        spec = configs.batched_nets(3, hcu_range=(3, 5), mcu_range=(3, 6), n_pat=2, train=20,
                                    quiet=5, recall=10, reset=10)
    spec.chained_step = 1
    path = configs.write(spec, str(tmp_path))
    s = Solver(path)
    configs.apply_conn_arrays(s, spec)
    with pytest.raises(GsbnError):
        s.step(1)
    s.close()
