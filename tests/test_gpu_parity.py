"""GPU parity: the HIP path against the CPU oracle, bit for bit (-m gpu)."""

import numpy as np
import pytest

import oracle as O
from gsbn_amd import configs, device_math, host_expf, host_logf

import helpers

pytestmark = pytest.mark.gpu


def test_device_math_bitwise():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-120, 90, 200000).astype(np.float32),
                        np.array([0, -0.0, 1, -1, 88.7, -103.9, 1e-40, np.inf, -np.inf, np.nan],
                                 np.float32)])
    ge = device_math(0, x.view(np.uint32)).view(np.float32)
    he = np.array([host_expf(v) for v in x[:20000]], np.float32)
    assert (ge[:20000].view(np.uint32) == he.view(np.uint32)).all()
    y = np.abs(x[np.isfinite(x)]) + np.float32(1e-30)
    gl = device_math(1, y.view(np.uint32)).view(np.float32)
    hl = np.array([host_logf(v) for v in y[:20000]], np.float32)
    assert (gl[:20000].view(np.uint32) == hl.view(np.uint32)).all()
    # IEEE division must round identically (the WTA normalisation and Wij)
    ab = rng.uniform(1e-6, 10, (100000, 2)).astype(np.float32)
    q = device_math(3, ab.reshape(-1).view(np.uint32)).view(np.float32)
    assert (q.view(np.uint32) == (ab[:, 0] / ab[:, 1]).view(np.uint32)).all()


def test_device_draw_matches_oracle():
    rows = []
    for seed in (0, 12345678901):
        for step in (0, 1, 77, 100000):
            for unit in (0, 1, 63, 9999):
                rows.append([seed & 0xffffffff, seed >> 32, step, unit])
    out = device_math(2, np.array(rows, np.uint32).reshape(-1)).view(np.float32).reshape(-1, 2)
    for (lo, hi, step, unit), (u, z) in zip(rows, out):
        ou, oz = O.draw(lo | (hi << 32), 0, step, unit)
        assert np.float32(ou) == u and np.float32(oz) == z


def test_config0_full_run_bitwise(tmp_path):
    spec = configs.config0()
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path)
    assert div is None, "spike rasters diverge at step %s" % (div,)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    c, t, flags = gpu.check()
    assert (c, t) == (ora.check.correct, ora.check.total)
    assert list(flags) == [int(r["ok"]) for r in ora.check.results]
    gpu.close()


def _lockstep_and_compare(spec, tmp_path, max_steps=None):
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=max_steps)
    assert div is None, "spike rasters diverge at step %s" % (div,)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    return gpu, ora


@pytest.mark.parametrize("chained", [-1, 1])
def test_config0_per_mode_stimulus_rows(tmp_path, chained):
    # ProcExtGen's one-row .lginp, rebuilt at each mode change (the form it
    # takes beyond 2^28 resident floats), on the split and the chained plan
    spec = configs.config0()
    spec.resident_stimulus = False
    spec.chained_step = chained
    gpu, ora = _lockstep_and_compare(spec, tmp_path, max_steps=900)
    assert gpu.vector(".lginp").size == spec.total_units
    assert gpu.glob("lginp-idx") == 0
    gpu.close()


def test_config0_queue_kernel_path(tmp_path):
    # ProcExchangeSpike copies spikes into si_<proj> and the queue push runs
    # as its own kernel (gsbn_proj_que); same results as the direct path
    spec = configs.config0()
    spec.exchange_direct = False
    gpu, ora = _lockstep_and_compare(spec, tmp_path, max_steps=600)
    gpu.close()


def test_config0_wta_population_maximum(tmp_path):
    spec = configs.config0()
    spec.wta_pop_max = 1
    gpu, ora = _lockstep_and_compare(spec, tmp_path, max_steps=700)
    gpu.close()


def test_external_spikes(tmp_path):
    rng = np.random.default_rng(5)
    spec = configs.config0()
    spec.ext_spikes = {0: {s: sorted(rng.choice(100, size=rng.integers(0, 6), replace=False).tolist())
                           for s in range(1, 400)}}
    gpu, ora = _lockstep_and_compare(spec, tmp_path, max_steps=400)
    gpu.close()


def test_reference_sparse_delay_map(tmp_path):
    # 400 HCU x 10 MCU with the reference's data/init_conn_0.1.map: up to 60
    # slots per HCU, delays 1..4, many empty slots (ii < 0)
    import os
    text = open(os.path.join(os.path.dirname(__file__), "golden", "init_conn_0.1.map")).read()
    spec = configs.recall_net(400, 10, 4, 60, 10, 60, 0.5, reset=20, name="sparse")
    spec.projs[0].slot_num = 60.0
    spec.conn_text = text
    spec.conn = configs.parse_conn_map(text)
    gpu, ora = _lockstep_and_compare(spec, tmp_path, max_steps=450)
    gpu.close()


def test_config1_size_lockstep(tmp_path):
    # 100 x 100 (9.9e7 synapses): first 120 steps including the full sweep of
    # the first learning-rate switch, compared bit for bit at the end
    spec = configs.config1()
    gpu, ora = _lockstep_and_compare(spec, tmp_path, max_steps=120)
    w = gpu.work()
    assert w[2] == 100 * 9900 * 100  # one sweep over every synapse
    gpu.close()


def test_config1_tuned_plans_state_every_100_steps(tmp_path):
    # 100 x 100 for 300 steps while the online tuner alternates its plans
    # step by step from step 20 (splits of 1-32 parts, the chained launch):
    # spikes every step, every Database vector every 100 steps
    spec = configs.config1()
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=300, check_every=100)
    assert div is None, "diverged at %s" % (div,)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    gpu.close()


@pytest.mark.parametrize("pop_wave,mcu", [(None, (3, 17)), (0, (3, 17)), (None, (3, 128))])
def test_batched_networks_bitwise(tmp_path, pop_wave, mcu):
    # independent networks of different sizes and connection densities in one
    # batched population launch and one batched row/column launch; the
    # population launch one wave per hypercolumn (one or two units per lane)
    # or, pop_wave 0, one workgroup per hypercolumn
    spec = configs.batched_nets(7, hcu_range=(3, 9), mcu_range=mcu, n_pat=4, train=40,
                                quiet=5, recall=30, reset=20)
    spec.pop_wave = pop_wave
    gpu, ora = _lockstep_and_compare(spec, tmp_path)
    c, t, flags = gpu.check()
    assert (c, t) == (ora.check.correct, ora.check.total)
    gpu.close()


@pytest.mark.parametrize("graded", [1, 3])
def test_batched_networks_graded_split(tmp_path, graded):
    # the batch's row/column work split per size bucket of hypercolumns
    # (the largest quarter into 2 resp. 8 parts, the smallest into one)
    spec = configs.batched_nets(7, hcu_range=(3, 9), mcu_range=(3, 17), n_pat=4, train=40,
                                quiet=5, recall=30, reset=20)
    spec.rowcol_graded = graded
    gpu, ora = _lockstep_and_compare(spec, tmp_path)
    c, t, flags = gpu.check()
    assert (c, t) == (ora.check.correct, ora.check.total)
    assert gpu.glob("rowcol-parts") == -10 - graded
    gpu.close()


def test_batched_networks_tuned_plans(tmp_path):
    # 24 networks while the tuner alternates the batch plans step by step
    # (uniform splits, the fused launch, the graded splits) from step 20
    spec = configs.batched_nets(24, hcu_range=(3, 12), mcu_range=(3, 33), n_pat=4, train=40,
                                quiet=5, recall=30, reset=20)
    gpu, ora = _lockstep_and_compare(spec, tmp_path)
    c, t, flags = gpu.check()
    assert (c, t) == (ora.check.correct, ora.check.total)
    assert gpu.glob("step-plan-ms-11", "f") > 0.0  # a graded split was timed
    gpu.close()


def test_batched_networks_mixed_block_sizes(tmp_path):
    # populations of 40-140 MCUs in one batched population launch of
    # 192-thread blocks (most threads idle for the small ones), the
    # row/column workgroups largest hypercolumn first
    spec = configs.batched_nets(6, hcu_range=(2, 4), mcu_range=(40, 140), n_pat=3, train=30,
                                quiet=5, recall=20, reset=10)
    assert len({(p.mcu + 63) // 64 for p in spec.pops}) > 1
    gpu, ora = _lockstep_and_compare(spec, tmp_path)
    c, t, flags = gpu.check()
    assert (c, t) == (ora.check.correct, ora.check.total)
    gpu.close()


def _sweep_steps(spec, upto):
    eg = O.ExtGen(spec)
    out = []
    for _ in range(upto):
        eg.update()
        if eg.cycle_flag == 1 and eg.prn != eg.old_prn:
            out.append(eg.simstep)
    return out


def test_config1_through_first_recall_windows(tmp_path):
    # 100 x 100 through the end of training (6000 steps) and the first two
    # reset and recall windows: the learning rate switches at 6001 (1 ->
    # 1e-30), 6051 (-> 0), 6101 (-> 1e-30) and 6151 (-> 0), each with a full
    # sweep of 9.9e7 synapses over gaps of up to 6000 steps; spikes every
    # step, every Database vector after each of these sweeps
    spec = configs.config1()
    sweeps = [t for t in _sweep_steps(spec, 6160) if t > 1]
    assert sweeps == [6001, 6051, 6101, 6151]
    path = helpers.write_spec(spec, tmp_path)
    from gsbn_amd import Solver
    gpu = Solver(path)
    ora = O.OracleNet(spec)
    compared = 0
    while ora.extgen.simstep < 6160:
        gpu.step(1)
        assert ora.step() == 1
        t = ora.extgen.simstep
        assert np.array_equal(gpu.vector("spike_0"), ora.vector("spike_0")), \
            "spikes differ at step %d" % t
        if t in sweeps:
            bad = helpers.compare_state(gpu, ora, spec)
            assert not bad, (t, bad)
            compared += 1
    assert compared == 4
    w = gpu.work()
    assert w[2] == 5 * 100 * 9900 * 100  # five full sweeps, the first one at step 1
    gpu.close()
    ora.close()


def test_batch512_sampled_networks_bitwise(tmp_path):
    # the 512-network workload (HCU 8-64, MCU 16-128, p in {0.1, 0.5, 0.9},
    # 2.66e9 synapses) in one solver; the smallest network, the largest one
    # and a dense (p = 0.9) one of middle size run in the oracle as a network
    # of their own, with their draws keyed by their ids in the batch.  Through
    # the end of training (1200 steps) and the learning-rate switches at 1201
    # and 1251: spikes every step, every Database vector of the three after
    # each sweep
    spec = configs.batched_nets(512)
    syn = [p.hcu * int(q.slot_num) * p.mcu for p, q in zip(spec.pops, spec.projs)]
    dens = [(ii >= 0).sum() / (p.hcu * (p.hcu - 1) * p.mcu)
            for (ii, _), p in zip(spec.conn_arrays, spec.pops)]
    small, large = int(np.argmin(syn)), int(np.argmax(syn))
    dense = [int(i) for i in np.argsort(syn) if dens[i] > 0.8 and i not in (small, large)]
    nets = [small, large, dense[len(dense) // 2]]
    assert dens[nets[2]] > 0.8 and syn[large] > 5e7
    sub = helpers.subset(spec, nets)
    sweeps = _sweep_steps(spec, 1260)
    assert sweeps == [1, 1201, 1251]
    path = helpers.write_spec(spec, tmp_path)
    from gsbn_amd import Solver
    gpu = Solver(path)
    configs.apply_conn_arrays(gpu, spec)
    ora = O.OracleNet(sub)
    while ora.extgen.simstep < 1260:
        gpu.step(1)
        assert ora.step() == 1
        t = ora.extgen.simstep
        for k, i in enumerate(nets):
            assert np.array_equal(gpu.vector("spike_%d" % i), ora.vector("spike_%d" % k)), \
                "network %d: spikes differ at step %d" % (i, t)
        if t in sweeps[1:]:
            bad = helpers.compare_nets(gpu, ora, spec, nets)
            assert not bad, (t, bad[:4])
    gpu.close()
    ora.close()


@pytest.mark.parametrize("wave,keep,cap,form", [(1, None, None, "fp32"), (1, 0, None, "fp32"),
                                                (1, 1, None, "fp32"), (1, 1, 3, "fp32"),
                                                (1, 1, 128, "fp32"), (1, 1, None, "bf16"),
                                                (1, 1, None, "fp16"), (2, None, None, "fp32"),
                                                (4, None, None, "fp32")])
def test_batched_networks_wave_plan(tmp_path, wave, keep, cap, form):
    # the batch's row/column update with one wave per hypercolumn (plan -20)
    # and its row overflow queue (rows past the first `keep` of a
    # hypercolumn run in helper waves; cap 3: no room in any of the 64
    # sub-queues, every row runs in its hypercolumn's own wave; cap 128: two
    # rows per sub-queue, some overflow and some not; keep 0: no queue), or
    # one wave per part of the graded splits g = 1 and 3 (plans -21, -23: the
    # largest quarter of the hypercolumns in 2 resp. 8 waves)
    spec = configs.batched_nets(7, hcu_range=(3, 9), mcu_range=(3, 140), n_pat=4, train=40,
                                quiet=5, recall=30, reset=20)
    spec.rowcol_wave = wave
    spec.ovf_keep = keep
    spec.ovf_cap = cap
    spec.trace_precision = form
    gpu, ora = _lockstep_and_compare(spec, tmp_path)
    c, t, flags = gpu.check()
    assert (c, t) == (ora.check.correct, ora.check.total)
    assert gpu.glob("rowcol-parts") == -19 - wave
    gpu.close()
