"""Snapshot, resume and the spike recorder on the GPU (-m gpu).

* resume: a run snapshotted at step k and resumed from that file continues
  bit for bit as the uninterrupted run: the spikes of every following step
  and, at the end, every saved vector.  Cases: the direct and the
  copy-and-queue spike exchange, the bf16 form with the snapshot on the last
  step before a learning-rate switch (the resumed run starts with the
  sweep), the reference's sparse map with delays 1..4 (rows in flight in the
  delay queue; ProcInitConn must not add the map again), and seven batched
  networks (several populations and projections);
* ProcSnapshot's files hold, under the reference's file names, exactly the
  vectors the solver reads back by name at those steps;
* the resumed step is exact at t = 1 s, where the reference's float
  truncation resumes one step early;
* ProcSpkRec's raster and active-row/unit files equal the oracle's spikes
  and active lists of every step, in the reference's CSV layout.
"""

import os

import numpy as np
import pytest

import oracle as O
from gsbn_amd import Solver, configs
from gsbn_amd import snapshot as S

import helpers

pytestmark = pytest.mark.gpu


def _saved_names(spec):
    names = []
    for i in range(len(spec.pops)):
        names += ["%s_%d" % (n, i) for n in ("dsup", "act", "ada", "epsc", "bj", "spike")]
    for i in range(len(spec.projs)):
        names += ["%s_%d" % (n, i) for n in ("ii", "qi", "di", "siq", "pi", "ei", "zi", "ti", "pj",
                                             "ej", "zj", "pij", "eij", "zi2", "zj2", "tij", "wij",
                                             "si")]
    return names


def _bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _run_to(s, step):
    while s.glob("simstep") < step:
        assert s.step(1) == 1


def _tail(s, spec, until):
    """Spikes of every step up to `until`, then every saved vector."""
    spikes = []
    while s.glob("simstep") < until:
        assert s.step(1) == 1
        spikes.append((s.glob("simstep"),
                       [s.vector("spike_%d" % i).copy() for i in range(len(spec.pops))]))
    return spikes, {n: s.vector(n).copy() for n in _saved_names(spec)}


def _resume_case(spec, tmp_path, k, n, pending=True):
    """`pending`: the snapshot step has rows in the delay queue (checked),
    so that the resume has queue state to restore"""
    path = configs.write(spec, str(tmp_path))
    snap = str(tmp_path / "state.bin")
    a = Solver(path)
    configs.apply_conn_arrays(a, spec)
    _run_to(a, k)
    a.snapshot(snap)
    if pending:
        assert any(a.vector("siq_%d" % i).any() for i in range(len(spec.projs)))
    want_spikes, want_state = _tail(a, spec, n)
    a.close()
    b = Solver(path, snap)
    assert b.glob("simstep") == k
    got_spikes, got_state = _tail(b, spec, n)
    b.close()
    assert len(got_spikes) == len(want_spikes)
    for (ta, sa), (tb, sb) in zip(want_spikes, got_spikes):
        assert ta == tb
        for i, (x, y) in enumerate(zip(sa, sb)):
            assert np.array_equal(x, y), "pop %d spikes differ at step %d" % (i, ta)
    for name in want_state:
        assert np.array_equal(_bits(want_state[name]), _bits(got_state[name])), name


# resume points are steps with rows in the delay queue (the 10x10 network
# spikes sparsely: about one step in four has a pending row)
def test_resume_direct_exchange(tmp_path):
    _resume_case(configs.config0(), tmp_path, 52, 400)


def test_resume_queue_kernel_exchange(tmp_path):
    spec = configs.config0()
    spec.exchange_direct = False
    _resume_case(spec, tmp_path, 1209, 1400)


def test_resume_bf16_across_learning_rate_switch(tmp_path):
    spec = configs.config0()
    spec.trace_precision = "bf16"
    # 10 patterns x (100 + 10) steps: step 1101 opens the first prn change
    _resume_case(spec, tmp_path, 1100, 1300, pending=False)


def test_resume_sparse_delay_map(tmp_path):
    text = open(os.path.join(os.path.dirname(__file__), "golden", "init_conn_0.1.map")).read()
    spec = configs.recall_net(400, 10, 4, 60, 10, 60, 0.5, reset=20, name="sparse")
    spec.projs[0].slot_num = 60.0
    spec.conn_text = text
    spec.conn = configs.parse_conn_map(text)
    _resume_case(spec, tmp_path, 45, 320)


def test_resume_batched_networks(tmp_path):
    spec = configs.batched_nets(7, hcu_range=(3, 9), mcu_range=(3, 17), n_pat=4, train=40,
                                quiet=5, recall=30, reset=20)
    _resume_case(spec, tmp_path, 37, 260)


def test_resume_sensitivity_without_saved_siq(tmp_path):
    # a snapshot without siq_ (as the reference writes them: its .siq_ is not
    # saved) resumes with no row active at the first step; the comparison
    # above must see that difference
    spec = configs.config0()
    path = configs.write(spec, str(tmp_path))
    snap = str(tmp_path / "state.bin")
    a = Solver(path)
    _run_to(a, 52)
    a.snapshot(snap)
    assert a.vector("siq_0").any()
    _, want = _tail(a, spec, 400)
    a.close()
    ts, prn, vecs = S.read_state(snap)
    cut = str(tmp_path / "cut.bin")
    S.write_state(cut, ts, prn, {n: v for n, (v, ld) in vecs.items() if n != "siq_0"},
                  {n: ld for n, (v, ld) in vecs.items()})
    b = Solver(path, cut)
    _, got = _tail(b, spec, 400)
    b.close()
    assert any(not np.array_equal(_bits(want[n]), _bits(got[n])) for n in want)


def test_procsnapshot_files_hold_the_vectors(tmp_path):
    spec = configs.config0()
    spec.snapshot = (200, 200)
    path = configs.write(spec, str(tmp_path))
    d = os.path.join(str(tmp_path), spec.name + "_rec", "ProcSnapshot")
    s = Solver(path)
    for k in (200, 400):
        _run_to(s, k)
        f = os.path.join(d, S.snapshot_name(k, spec.dt))
        ts, prn, got = S.read_state(f)
        assert ts == np.float32(np.float32(k) * np.float32(spec.dt)) and prn == 1.0
        assert set(got) == set(_saved_names(spec))
        for name, (arr, ld) in got.items():
            assert np.array_equal(_bits(arr), _bits(s.vector(name))), (k, name)
    assert sorted(os.listdir(d)) == [S.snapshot_name(200, spec.dt), S.snapshot_name(400, spec.dt)]
    s.close()


def test_resume_step_exact_at_one_second(tmp_path):
    spec = configs.config0()
    spec.snapshot = (1000, 1000)
    path = configs.write(spec, str(tmp_path))
    s = Solver(path)
    _run_to(s, 1000)
    s.close()
    f = os.path.join(str(tmp_path), spec.name + "_rec", "ProcSnapshot", "Snapshot_1.000000.bin")
    # the reference resumes at int(t / dt) in float
    t = np.float32(np.float32(1000) * np.float32(spec.dt))
    assert int(np.float32(t / np.float32(spec.dt))) == 999
    r = Solver(path, f)
    assert r.glob("simstep") == 1000
    r.step(1)
    assert r.glob("simstep") == 1001
    r.close()


def test_spike_recorder_matches_oracle(tmp_path):
    spec = configs.config0()
    spec.spike_record = (3, 2)
    # unit 9 and unit 19 (the source of row 9) spike now and then, so that
    # both active-list files hold ones
    spec.ext_spikes = {0: {s: [9, 19] for s in range(10, 400, 7)}}
    path = configs.write(spec, str(tmp_path))
    gpu = Solver(path)
    ora = O.OracleNet(spec)
    raster, ssi, ssj = [], [], []
    for _ in range(400):
        gpu.step(1)
        f = ora.step()
        t = ora.extgen.simstep
        if f != 1 or t < 3 or t % 2:
            continue
        sp = np.nonzero(ora.vector("spike_0") > 0)[0]
        if sp.size:
            raster.append(",".join([str(t)] + [str(int(i)) for i in sp]))
        ssi.append("%d,%d" % (t, int(9 in set(ora.vector(".ssi_0").tolist()))))
        ssj.append("%d,%d" % (t, int(9 in set(ora.vector(".ssj_0").tolist()))))
    gpu.close()
    d = os.path.join(str(tmp_path), spec.name + "_rec", "ProcSpkRec")
    lines = open(os.path.join(d, "spk_pop_0.csv")).read().splitlines()
    assert lines[0] == "10,10,0.001,100"
    assert lines[1:] == raster
    assert open(os.path.join(d, "ssi_pop_0.csv")).read().splitlines() == ssi
    assert open(os.path.join(d, "ssj_pop_0.csv")).read().splitlines() == ssj
    assert any(l.endswith(",1") for l in ssi) and any(l.endswith(",1") for l in ssj)
