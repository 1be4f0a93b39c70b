"""Structural plasticity on the GPU (-m gpu): ProcStructuralPlasticity
(pruning on the device, growth on the host) in lockstep with the oracle's
restatement, on the 10x10 network and on three batched networks.  Every step's spikes and, every 20 steps, every Database vector
must be bit-identical, and at the end the fan-out counters.  Cases: the
10x10 network with pruning every 60 and growth every 20 steps (direct and
copy-and-queue spike exchange, bf16 form), random initial connectivity grown
by the procedure with geometric delays, and a resume across a growth step.
"""

import numpy as np
import pytest

from gsbn_amd import Solver, configs
from gsbn_amd import snapshot as S

import helpers

pytestmark = pytest.mark.gpu


def _spec(**kw):
    spec = configs.config0()
    for m in spec.modes:
        m.plasticity = 1
    spec.structural = configs.StructuralSpec(period=20, pruning_period=60, t_th=30, seed=7, **kw)
    return spec


def _lockstep(spec, tmp_path, steps):
    gpu, ora, div, n = helpers.run_lockstep(spec, tmp_path, max_steps=steps, check_every=20)
    assert div is None, "diverged: %s" % (div,)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    for p in range(len(spec.pops)):
        assert np.array_equal(gpu.vector("fanout_%d" % p), ora.sp.pops[p]["fanout"])
    return gpu, ora


def test_structural_lockstep_direct(tmp_path):
    spec = _spec()
    gpu, ora = _lockstep(spec, tmp_path, 600)
    ii = gpu.vector("ii_0")
    assert (ii < 0).any() and (ii >= 0).any()
    # rows were both removed and grown on the way
    assert gpu.vector("fanout_0").max() > 10
    gpu.close()


def test_structural_lockstep_queue_kernel(tmp_path):
    spec = _spec()
    spec.exchange_direct = False
    gpu, ora = _lockstep(spec, tmp_path, 600)
    gpu.close()


def test_structural_lockstep_bf16(tmp_path):
    spec = _spec()
    spec.trace_precision = "bf16"
    gpu, ora = _lockstep(spec, tmp_path, 400)
    gpu.close()


def test_structural_init_conn_and_geometric_delays(tmp_path):
    spec = _spec(enable_init_conn=1, enable_geometry=1, d_norm=0.75, v_cond=2.0)
    spec.conn = [[]]  # every row grown by the procedure
    spec.pops[0].shape = (5, 2)
    gpu, ora = _lockstep(spec, tmp_path, 400)
    di = gpu.vector("di_0")
    assert len(set(di[gpu.vector("ii_0") >= 0].tolist())) > 1
    gpu.close()


def test_structural_batched_networks(tmp_path):
    # three networks of one batch, each pruned and grown on its own
    spec = configs.batched_nets(3, hcu_range=(3, 6), mcu_range=(3, 9), n_pat=4, train=40,
                                quiet=5, recall=30, reset=20)
    for m in spec.modes:
        m.plasticity = 1
    spec.structural = configs.StructuralSpec(period=20, pruning_period=60, t_th=30, seed=7)
    assert spec.conn_arrays
    ii0 = [np.asarray(ii).copy() for ii, di in spec.conn_arrays]
    gpu, ora = _lockstep(spec, tmp_path, 300)
    # rows were removed or grown in at least one network
    assert any(not np.array_equal(gpu.vector("ii_%d" % i), ii0[i]) for i in range(len(ii0)))
    gpu.close()


def test_structural_resume(tmp_path):
    spec = _spec()
    path = configs.write(spec, str(tmp_path))
    snap = str(tmp_path / "s.bin")
    a = Solver(path)
    while a.glob("simstep") < 130:
        a.step(1)
    a.snapshot(snap)
    names = ["ii_0", "di_0", "pij_0", "wij_0", "tij_0", "fanout_0", "spike_0"]
    while a.glob("simstep") < 260:
        a.step(1)
    want = {n: a.vector(n) for n in names}
    a.close()
    _, _, vecs = S.read_state(snap)
    assert "fanout_0" in vecs
    b = Solver(path, snap)
    while b.glob("simstep") < 260:
        b.step(1)
    for n in names:
        got = b.vector(n)
        assert np.array_equal(got.view(np.uint8), want[n].view(np.uint8)), n
    b.close()
