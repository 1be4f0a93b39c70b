"""The fused step launch (gsbn_step_fused / gsbn_step_fused_batch) on the GPU
(-m gpu): forced with ProcUpdLazy argi fused-step 1 (one workgroup per
hypercolumn runs its population update and then its row/column update from
LDS), it must give the oracle's spike rasters and final state bit for bit,
like the split launches (sweep steps fall back to the split launches).
Cases: the full 10x10 workload with its recall results, the queue-kernel
exchange, the WTA population-maximum quirk, external spikes, the bf16 and
fp16 forms, 120 steps of the 100x100 network through its first sweep,
seven batched networks, and the online choice between fused and split
plans."""

import numpy as np
import pytest

from gsbn_amd import configs

import helpers

pytestmark = pytest.mark.gpu


def _check(spec, tmp_path, max_steps=None, recall=False):
    spec.fused_step = 1
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=max_steps)
    assert div is None, "spike rasters diverge at step %s" % (div,)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    if recall:
        c, t, flags = gpu.check()
        assert (c, t) == (ora.check.correct, ora.check.total)
    assert gpu.glob("rowcol-parts") == -1
    gpu.close()


def test_fused_config0_full_run(tmp_path):
    _check(configs.config0(), tmp_path, recall=True)


def test_fused_queue_kernel_exchange(tmp_path):
    spec = configs.config0()
    spec.exchange_direct = False
    _check(spec, tmp_path, max_steps=600)


def test_fused_wta_population_maximum(tmp_path):
    spec = configs.config0()
    spec.wta_pop_max = 1
    _check(spec, tmp_path, max_steps=700)


def test_fused_external_spikes(tmp_path):
    rng = np.random.default_rng(5)
    spec = configs.config0()
    spec.ext_spikes = {0: {s: sorted(rng.choice(100, size=rng.integers(0, 6), replace=False).tolist())
                           for s in range(1, 400)}}
    _check(spec, tmp_path, max_steps=400)


@pytest.mark.parametrize("form", ["bf16", "fp16"])
def test_fused_16bit_forms(tmp_path, form):
    spec = configs.config0()
    spec.trace_precision = form
    _check(spec, tmp_path, max_steps=1300)


def test_fused_config1_size(tmp_path):
    _check(configs.config1(), tmp_path, max_steps=120)


def test_fused_batched_networks(tmp_path):
    spec = configs.batched_nets(7, hcu_range=(3, 9), mcu_range=(3, 17), n_pat=4, train=40,
                                quiet=5, recall=30, reset=20)
    _check(spec, tmp_path, recall=True)


def test_online_plan_choice(tmp_path):
    # the tuner tries the split plans and the fused one from step 20 on
    # (results do not depend on the plan) and keeps one; eager column
    # updates, which the fused launch needs
    spec = configs.config0()
    spec.deferred_columns = 0
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=300)
    assert gpu.glob("step-plan-ms-1", "f") > 0.0  # the fused plan was timed
    assert div is None, "spike rasters diverge at step %s" % (div,)
    assert not helpers.compare_state(gpu, ora, spec)
    assert gpu.glob("rowcol-parts") != 0
    gpu.close()
