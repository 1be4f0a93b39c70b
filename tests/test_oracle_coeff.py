"""Closed-form catch-up coefficients.

The table entry for an elapsed gap D maps (p, e, z) at t to t + D for the
cascade dz/dt = -kz z, de/dt = ke (z - e), dp/dt = kp (e - p).  Pins:
  * 50-digit decimal evaluation of the divided differences (textbook formula
    with enough digits that cancellation is harmless),
  * the matrix exponential of the 3x3 system (scipy.linalg.expm),
  * RK4 integration with a fine step,
  * the reference's own float expression (Proj.cpp:241-247, 269-277) where it
    is well conditioned,
and the product's host table builder must equal the oracle's bit for bit.
"""

from decimal import Decimal, getcontext

import numpy as np
import pytest
from scipy.linalg import expm

import oracle as O

getcontext().prec = 60

RATES = [
    (0.1, 0.1, 0.05, 0.001),        # dt=1ms, tauz=10ms, taue=20ms, taup=1s
    (0.1, 0.1, 0.05, 0.0),          # recall (kp = 0)
    (0.2, 0.19996, 0.199992, 1 / 60000),  # close time constants of data/10x10_rnd.prototxt
    (0.1, 0.1, 2.0004e-4, 2e-4),    # taue = taup - 1 ms at taup = 5 s
    (0.1, 0.1, 0.05, 1e-33),        # reset windows with a tiny non-zero rate
]
GAPS = [1, 2, 3, 7, 20, 100, 1000, 12345]


def dexp(x):
    return (-x).exp()


def dd2(f, a, b):
    return (f(a) - f(b)) / (a - b)


def dd3(f, a, b, c):
    return (dd2(f, a, b) - dd2(f, b, c)) / (a - c)


def decimal_entry(kzi, kzj, ke, kp, D):
    kzi, kzj, ke, kp = (Decimal(float(np.float64(v))) for v in (kzi, kzj, ke, kp))
    D = Decimal(D)
    F = lambda x: (-(x * D)).exp()
    a = kzi + kzj
    def phi(x, y):
        if x == y:
            return D * F(x)
        return -dd2(F, x, y)
    def psi(x, y, z):
        pts = sorted([x, y, z])
        eps = Decimal("1e-40")
        pts[1] += eps if pts[1] == pts[0] else 0
        pts[2] += 2 * eps if pts[2] == pts[1] else 0
        return dd3(F, *pts)
    return [F(kzi), F(kzj), F(ke), F(kp), ke * phi(a, ke), kp * phi(ke, kp),
            kp * ke * psi(a, ke, kp), ke * phi(kzi, ke), kp * ke * psi(kzi, ke, kp)]


@pytest.mark.parametrize("rates", RATES)
def test_table_vs_decimal(rates):
    for D in GAPS:
        got = O.coeff_entry(*rates, D)[:9].astype(np.float64)
        want = np.array([float(v) for v in decimal_entry(*rates, D)])
        # each entry is the double result rounded once to float
        tol = np.maximum(np.abs(want) * 1.2e-7, 1e-44)
        assert np.all(np.abs(got - want) <= tol), (rates, D, got, want)


@pytest.mark.parametrize("rates", RATES[:3])
def test_table_vs_expm_and_rk4(rates):
    kzi, kzj, ke, kp = rates
    for kz, cez, cpz in ((kzi + kzj, 4, 6), (kzi, 7, 8)):
        A = np.array([[-kp, kp, 0.0], [0.0, -ke, ke], [0.0, 0.0, -kz]])  # state (p, e, z)
        for D in (1, 5, 40, 300):
            E = expm(A * D)
            c = O.coeff_entry(*rates, D).astype(np.float64)
            np.testing.assert_allclose(c[3], E[0, 0], rtol=2e-7)
            np.testing.assert_allclose(c[5], E[0, 1], rtol=2e-7, atol=1e-30)
            np.testing.assert_allclose(c[cpz], E[0, 2], rtol=2e-7, atol=1e-30)
            np.testing.assert_allclose(c[2], E[1, 1], rtol=2e-7)
            np.testing.assert_allclose(c[cez], E[1, 2], rtol=2e-7, atol=1e-30)
    # RK4 of the cascade from (p, e, z) = (0.01, 0.2, 1.0) over 37 steps
    y = np.array([0.01, 0.2, 1.0])
    A = np.array([[-kp, kp, 0.0], [0.0, -ke, ke], [0.0, 0.0, -(kzi + kzj)]])
    h = 1e-3
    for _ in range(int(37 / h)):
        k1 = A @ y
        k2 = A @ (y + h / 2 * k1)
        k3 = A @ (y + h / 2 * k2)
        k4 = A @ (y + h * k3)
        y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    c = O.coeff_entry(*rates, 37).astype(np.float64)
    p = c[3] * 0.01 + c[5] * 0.2 + c[6] * 1.0
    e = c[2] * 0.2 + c[4] * 1.0
    np.testing.assert_allclose([p, e], y[:2], rtol=1e-6)


def test_f64_closed_form_vs_expm():
    for kz, ke, kp in ((0.2, 0.05, 0.001), (0.1, 0.05, 0.0), (0.39996, 0.199992, 1 / 60000),
                       (0.1, 2.0004e-4, 2e-4), (0.05, 0.05, 0.05)):
        A = np.array([[-kp, kp, 0.0], [0.0, -ke, ke], [0.0, 0.0, -kz]])
        s = np.array([0.3, 0.7, 1.9])
        for D in (1, 13, 500):
            got = O.closed_form_f64(kz, ke, kp, D, s)
            want = expm(A * D) @ s
            np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-300)


def test_reference_float_expression_agrees_when_well_conditioned():
    # the reference's float closed form and the table agree to float accuracy
    # for the test workload's rates (no near-equal rates)
    kzi, kzj, ke, kp = np.float32(0.1), np.float32(0.1), np.float32(0.05), np.float32(0.001)
    for D in (1, 3, 10, 50, 200):
        r = O.ref32_syn(kp, ke, kzi, kzj, D, 0.002, 0.3, 0.8, 0.6)
        c = O.coeff_entry(kzi, kzj, ke, kp, D)
        z0 = np.float32(0.8) * np.float32(0.6)
        p = (c[3] * np.float32(0.002) + c[5] * np.float32(0.3)) + c[6] * z0
        e = c[2] * np.float32(0.3) + c[4] * z0
        np.testing.assert_allclose([r[0], r[1], r[2], r[3]], [p, e, c[0] * np.float32(0.8),
                                   c[1] * np.float32(0.6)], rtol=2e-5, atol=1e-9)
        ri = O.ref32_row(kp, ke, kzi, D, 0.01, 0.2, 0.9)
        pi = (c[3] * np.float32(0.01) + c[5] * np.float32(0.2)) + c[8] * np.float32(0.9)
        ei = c[2] * np.float32(0.2) + c[7] * np.float32(0.9)
        np.testing.assert_allclose(ri, [pi, ei, c[0] * np.float32(0.9)], rtol=2e-5, atol=1e-9)


def test_reference_float_expression_loses_digits_with_close_rates():
    # documents why the table exists: with the shipped close time constants the
    # reference's float expression is far from the exact value
    kzi, kzj, ke, kp = np.float32(0.2), np.float32(0.19996), np.float32(0.199992), np.float32(1 / 60000)
    D = 5
    r = O.ref32_row(kp, ke, kzi, D, 0.01, 0.2, 0.9)
    exact = O.closed_form_f64(float(kzi), float(ke), float(kp), D, [0.01, 0.2, 0.9])
    c = O.coeff_entry(kzi, kzj, ke, kp, D)
    ei_tab = c[2] * np.float32(0.2) + c[7] * np.float32(0.9)
    assert abs(ei_tab - exact[1]) / exact[1] < 1e-6
    assert abs(r[1] - exact[1]) / exact[1] > 1e-4


@pytest.mark.parametrize("rates", RATES)
def test_product_table_equals_oracle_table(rates):
    import gsbn_amd
    for D in GAPS + [0, 4095, 70001]:
        a = gsbn_amd.coeff_entry(*[float(np.float32(r)) for r in rates], D)
        b = O.coeff_entry(*[float(np.float32(r)) for r in rates], D)
        assert a.tobytes() == b.tobytes(), (rates, D)


def test_saturation_is_exact():
    import gsbn_amd
    for rates in RATES[:2]:
        r = [float(np.float32(x)) for x in rates]
        sat = gsbn_amd.coeff_saturation(*r)
        assert sat > 0
        last = O.coeff_entry(*r, sat)
        for D in (sat, sat + 1, sat + 17, 2 * sat, 10 * sat):
            assert O.coeff_entry(*r, D).tobytes() == last.tobytes()
        assert O.coeff_entry(*r, sat - 1).tobytes() != last.tobytes()
