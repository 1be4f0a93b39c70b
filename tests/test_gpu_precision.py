"""GPU parity of the 16-bit trace forms (-m gpu).

The product's bf16 / fp16 storage of Pij and Eij against the oracle's
storage mode, bit for bit: the conversions themselves, whole runs of the
10x10 workload with their recall results, the 100x100 workload through its
first learning-rate sweep, and batched networks.
"""

import numpy as np
import pytest
import torch

import oracle as O
from gsbn_amd import configs, device_math

import helpers
from test_trace_format import _bf16_values, _edges, _f16_values, _random_floats

pytestmark = pytest.mark.gpu


def test_device_trace_conversions_bitwise():
    x = np.concatenate([_edges(_f16_values()), _edges(_bf16_values()), _random_floats(100000, 4),
                        np.array([65504, 65520, 1e30, 2.0 ** -25, 0.0, -0.0, 3.4e38], np.float32)])
    x = x.astype(np.float32)
    hbf = device_math(5, x.view(np.uint32)).astype(np.uint16)
    h16 = device_math(6, x.view(np.uint32)).astype(np.uint16)
    tb = torch.from_numpy(x.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(hbf, tb)
    with np.errstate(over="ignore"):
        assert np.array_equal(h16, x.astype(np.float16).view(np.uint16))
    codes = np.arange(0, 0x10000, dtype=np.uint32)
    dec = device_math(7, codes).view(np.float32)
    want = codes.astype(np.uint16).view(np.float16).astype(np.float32)
    fin = np.isfinite(want)
    assert np.array_equal(dec[fin].view(np.uint32), want[fin].view(np.uint32))
    assert np.all(np.isnan(dec[np.isnan(want)]))


def _check_shifts(gpu, ora, spec, form):
    for i in range(len(spec.projs)):
        assert gpu.glob("trace-precision-%d" % i) == O.TRACE_FORMS[form]
        if form == "fp16":
            assert gpu.glob("trace-p-shift-%d" % i) == ora.trace_shifts[i]
            assert gpu.glob("trace-e-shift-%d" % i) == ora.trace_shifts[i]


@pytest.mark.parametrize("form", ["bf16", "fp16"])
def test_config0_trace_form_bitwise(tmp_path, form):
    spec = configs.config0()
    spec.trace_precision = form
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path)
    assert div is None, "spike rasters diverge at step %s" % (div,)
    _check_shifts(gpu, ora, spec, form)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    c, t, flags = gpu.check()
    assert (c, t) == (ora.check.correct, ora.check.total)
    assert list(flags) == [int(r["ok"]) for r in ora.check.results]
    gpu.close()


@pytest.mark.parametrize("form", ["bf16", "fp16"])
def test_config1_trace_form_lockstep(tmp_path, form):
    # 100 x 100: the first 120 steps, including the full sweep of the first
    # learning-rate switch, compared bit for bit at the end
    spec = configs.config1()
    spec.trace_precision = form
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=120)
    assert div is None, "spike rasters diverge at step %s" % (div,)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    gpu.close()


def test_batched_networks_bf16(tmp_path):
    spec = configs.batched_nets(7, hcu_range=(3, 9), mcu_range=(3, 17), n_pat=4, train=40,
                                quiet=5, recall=30, reset=20)
    spec.trace_precision = "bf16"
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path)
    assert div is None, "spike rasters diverge at step %s" % (div,)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    c, t, flags = gpu.check()
    assert (c, t) == (ora.check.correct, ora.check.total)
    gpu.close()
