"""The deterministic exp/log used by both paths, pinned against glibc (double
precision libm rounded to float, i.e. the correctly rounded value except in
double-rounding corner cases)."""

import math

import numpy as np

import oracle as O


def ulp_diff(a, b):
    ai = np.array(a, np.float32).view(np.int32).astype(np.int64)
    bi = np.array(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ai - bi)


def test_expf_within_one_ulp():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-103, 88, 30000), rng.uniform(-1, 1, 10000),
                         rng.uniform(-1e-6, 1e-6, 2000)]).astype(np.float32)
    got = np.array([O.det_expf(x) for x in xs], np.float32)
    ref = np.array([np.float32(math.exp(float(x))) for x in xs], np.float32)
    d = ulp_diff(got, ref)
    assert d.max() <= 1
    assert (d == 0).mean() > 0.999


def test_logf_within_one_ulp():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-100, 80, 30000)), rng.uniform(0.5, 2, 10000),
                         np.array([1e-45, 1e-40, 1e-38, 1.0, 2.0, 0.001, 1.001])]).astype(np.float32)
    xs = xs[xs > 0]
    got = np.array([O.det_logf(x) for x in xs], np.float32)
    ref = np.array([np.float32(math.log(float(x))) for x in xs], np.float32)
    d = ulp_diff(got, ref)
    assert d.max() <= 1
    assert (d == 0).mean() > 0.999


def test_special_values():
    assert O.det_expf(0.0) == 1.0
    assert O.det_expf(-200.0) == 0.0
    assert np.isinf(O.det_expf(100.0))
    assert O.det_logf(1.0) == 0.0
    assert np.isneginf(O.det_logf(0.0))
    assert np.isnan(O.det_logf(-1.0))


def test_product_host_math_matches_oracle():
    from gsbn_amd import host_expf, host_logf
    rng = np.random.default_rng(2)
    for x in rng.uniform(-50, 50, 500).astype(np.float32):
        assert host_expf(x) == O.det_expf(x)
        assert host_logf(abs(x) + 1e-3) == O.det_logf(abs(x) + 1e-3)
