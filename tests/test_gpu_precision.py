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


@pytest.mark.parametrize("point", ["taup5", "close"])
@pytest.mark.parametrize("form", ["bf16", "fp16"])
def test_config2_points_lockstep(tmp_path, point, form):
    # the precision workload's two extra parameter points at 100 x 100: taup
    # 5 s, and taup 5 s with taue = taup - 1 ms (the cancellation case); the
    # first 120 steps through the first sweep, bit for bit
    import precision_study as PS
    spec = PS.point_spec(point, 100, 100)
    spec.trace_precision = form
    gpu, ora, div, steps = helpers.run_lockstep(spec, tmp_path, max_steps=120)
    assert div is None, "spike rasters diverge at step %s" % (div,)
    _check_shifts(gpu, ora, spec, form)
    bad = helpers.compare_state(gpu, ora, spec)
    assert not bad, bad
    gpu.close()
    ora.close()


# error bounds against the exact double-precision replay at a short horizon
# (300 steps of the 100 x 100 workload, product on the GPU); measured values
# with margin: fp32 Pij max ~6e-7; bf16 ~2.2e-2 (its first rounding of the
# initial 1/(C M) alone is 2.2e-3); fp16 ~2.6e-3
SHORT_BOUNDS = {"fp32": (2e-6, 1e-5), "bf16": (5e-2, 5e-2), "fp16": (6e-3, 6e-3)}


@pytest.mark.parametrize("point", ["default", "close"])
@pytest.mark.parametrize("form", ["fp32", "bf16", "fp16"])
def test_config2_error_against_exact_short(point, form):
    import precision_study as PS
    r = PS.run(form, point, 300, "gpu", 100, 100, log=lambda m: None)
    pmax, wmax = SHORT_BOUNDS[form]
    print("%s %s: Pij max %.2e mean %.2e, Wij abs max %.2e, bj max %.2e"
          % (form, point, r["pij"]["max"], r["pij"]["mean"], r["wij"]["max"], r["bj"]["max"]))
    assert r["pij"]["max"] <= pmax, r["pij"]
    assert r["wij"]["max"] <= wmax, r["wij"]
    assert r["bj"]["max"] <= 1e-5, r["bj"]
