"""Deferred column updates on the GPU against the oracle, bit for bit (-m gpu).

With ProcUpdLazy argi deferred-columns K the column update of a
postsynaptic spike (Proj.cpp:481-536) is not done when the unit spikes: the
spike step goes into the unit's ring of K steps, and the update is replayed,
in the reference's order, by the next row update of each synapse, by a
column walk once K - 1 spikes of the unit are pending, by a learning-rate
sweep, or before the synapse state is read or written from the host.  The
oracle runs the reference's eager order; every case here must give the same
spikes every step and the same state, ProcCheck results and raster.  K = 2
walks at every spike (the eager order through the deferred code), larger K
defers up to K - 1 spikes.
"""

import numpy as np
import pytest

import oracle as O
from gsbn_amd import Solver, configs

import helpers

pytestmark = pytest.mark.gpu


def _run(spec, tmp_path, max_steps=None, check_every=0):
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=max_steps,
                                                check_every=check_every)
    assert div is None, "diverged at %s" % (div,)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    if spec.deferred_columns is not None:
        assert gpu.glob("deferred-columns") == spec.deferred_columns
    return gpu, ora


@pytest.mark.parametrize("K", [0, 2, 4, 16])
def test_config0_full_run(tmp_path, K):
    spec = configs.config0()
    spec.deferred_columns = K
    gpu, ora = _run(spec, tmp_path)
    c, t, flags = gpu.check()
    assert (c, t) == (ora.check.correct, ora.check.total)
    assert list(flags) == [int(r["ok"]) for r in ora.check.results]
    gpu.close()


@pytest.mark.parametrize("plan,K", [("chained", 4), ("split", 8), ("chained", 32)])
def test_config0_fixed_plans(tmp_path, plan, K):
    spec = configs.config0()
    spec.deferred_columns = K
    if plan == "chained":
        spec.chained_step = 1
    else:
        spec.chained_step = -1
        spec.rowcol_parts = 2
    gpu, ora = _run(spec, tmp_path, max_steps=1500, check_every=250)
    gpu.close()


def test_state_reads_between_steps(tmp_path):
    # every 7 steps the whole state is read (pending kicks materialised,
    # chained work flushed), then the run goes on deferred
    spec = configs.config0()
    spec.deferred_columns = 16
    spec.chained_step = 1
    gpu, ora = _run(spec, tmp_path, max_steps=700, check_every=7)
    gpu.close()


def test_queue_kernel_and_external_spikes(tmp_path):
    rng = np.random.default_rng(8)
    spec = configs.config0()
    spec.deferred_columns = 8
    spec.exchange_direct = False
    spec.ext_spikes = {0: {s: sorted(rng.choice(100, size=rng.integers(0, 9), replace=False).tolist())
                           for s in range(1, 500)}}
    gpu, ora = _run(spec, tmp_path, max_steps=500)
    gpu.close()


def test_reference_sparse_delay_map(tmp_path):
    # empty slots (ii < 0) never receive column kicks, delays 1..4
    import os
    text = open(os.path.join(os.path.dirname(__file__), "golden", "init_conn_0.1.map")).read()
    spec = configs.recall_net(400, 10, 4, 60, 10, 60, 0.5, reset=20, name="sparse")
    spec.projs[0].slot_num = 60.0
    spec.conn_text = text
    spec.conn = configs.parse_conn_map(text)
    spec.deferred_columns = 8
    gpu, ora = _run(spec, tmp_path, max_steps=450)
    gpu.close()


@pytest.mark.parametrize("wave,K", [(0, 4), (1, 4), (0, 0), (1, 8), (1, None), (0, None)])
def test_batched_networks(tmp_path, wave, K):
    # K None: the library's default for a batch, rings of 16
    spec = configs.batched_nets(7, hcu_range=(3, 9), mcu_range=(3, 140), n_pat=4, train=40,
                                quiet=5, recall=30, reset=20)
    spec.deferred_columns = K
    spec.rowcol_wave = wave
    gpu, ora = _run(spec, tmp_path)
    assert gpu.glob("deferred-columns") == (16 if K is None else K)
    c, t, flags = gpu.check()
    assert (c, t) == (ora.check.correct, ora.check.total)
    gpu.close()


@pytest.mark.parametrize("form,K", [("fp32", 8), ("bf16", 8), ("fp16", 4), ("fp32", 0)])
def test_config1_through_first_sweep(tmp_path, form, K):
    # 100 x 100, 300 steps while the tuner alternates split and chained
    # plans, state every 100 steps
    spec = configs.config1()
    spec.trace_precision = form
    spec.deferred_columns = K
    gpu, ora = _run(spec, tmp_path, max_steps=300, check_every=100)
    gpu.close()


def test_config1_sweeps_of_recall_windows(tmp_path):
    # sweeps replay the pending kicks with the old learning rate: 100 x 100
    # through the switches at 6001, 6051 and 6101, full state after each
    spec = configs.config1()
    spec.deferred_columns = 16
    path = helpers.write_spec(spec, tmp_path)
    gpu = Solver(path)
    ora = O.OracleNet(spec)
    while ora.extgen.simstep < 6105:
        gpu.step(1)
        assert ora.step() == 1
        t = ora.extgen.simstep
        assert np.array_equal(gpu.vector("spike_0"), ora.vector("spike_0")), t
        if t in (6001, 6051, 6101):
            bad = helpers.compare_state(gpu, ora, spec)
            assert not bad, (t, bad)
    gpu.close()
    ora.close()


def test_snapshot_resume_with_pending_kicks(tmp_path):
    # the snapshot holds the materialised state; the resumed run starts with
    # empty rings and continues bit for bit (one solver at a time: the
    # global variables are process-wide, as in the reference)
    spec = configs.config0()
    spec.deferred_columns = 16
    path = configs.write(spec, str(tmp_path))
    snap = str(tmp_path / "state.bin")
    a = Solver(path)
    ora = O.OracleNet(spec)
    for _ in range(333):
        a.step(1)
        ora.step()
    a.snapshot(snap)
    want = []
    for _ in range(400):
        a.step(1)
        ora.step()
        s = ora.vector("spike_0")
        assert np.array_equal(a.vector("spike_0"), s)
        want.append(s)
    bad = helpers.compare_state(a, ora, spec)
    assert not bad, bad
    a.close()
    b = Solver(path, snap)
    assert b.glob("simstep") == 333
    for k in range(400):
        b.step(1)
        assert np.array_equal(b.vector("spike_0"), want[k]), "resumed run differs at %d" % (334 + k)
    # the spike counters (.counter_0) are not in a snapshot, as in the reference
    names = [n for n in helpers.state_names(spec) if not n.startswith(".counter")]
    bad = helpers.compare_state(b, ora, spec, names)
    assert not bad, bad
    b.close()


def test_structural_plasticity_ends_deferral(tmp_path):
    # row structure changes are not replayed: the first pruning or growth
    # applies the pending kicks and the run continues eager
    spec = configs.config0()
    for m in spec.modes:
        m.plasticity = 1
    spec.deferred_columns = 8
    spec.structural = configs.StructuralSpec(period=100, pruning_period=300, t_th=150, seed=7)
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=50)
    assert div is None and gpu.glob("deferred-columns") == 8
    gpu.close()
    ora.close()
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=700, check_every=50)
    assert div is None, "diverged at %s" % (div,)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    assert gpu.glob("deferred-columns") == 0
    gpu.close()


def test_fused_step_runs_eager(tmp_path):
    # the fused launch does eager column updates: forcing it turns deferral off
    spec = configs.config0()
    spec.deferred_columns = 4
    spec.fused_step = 1
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=300)
    assert div is None
    assert gpu.glob("deferred-columns") == 0 and gpu.glob("rowcol-parts") == -1
    gpu.close()


def test_default_is_deferred(tmp_path):
    # without the argument ProcUpdLazy defers with rings of 8 (16 for a batch
    # of networks, test_batched_networks); 0 asks for eager
    spec = configs.config0()
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=30)
    assert div is None and gpu.glob("deferred-columns") == 8
    gpu.close()
    spec.deferred_columns = 0
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=30)
    assert div is None and gpu.glob("deferred-columns") == 0
    gpu.close()
