"""The transcendental bodies: the oracle's against the platform libm (the
mathematics they approximate), and the engine's host build against the
oracle's bit for bit (the engine runs the same algorithm on the device)."""
import math

import numpy as np
import pytest


def ulps(a, b):
    a, b = np.float64(a), np.float64(b)
    if a == b:
        return 0
    return abs(int(np.array(a).view(np.int64)) - int(np.array(b).view(np.int64)))


def test_exp_within_2_ulp(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-745, 709.7, 20000), rng.uniform(-1, 1, 20000), rng.uniform(-40, 40, 20000)])
    worst = max(ulps(oracle.om_exp(float(x)), math.exp(x)) for x in xs)
    assert worst <= 2


def test_log_within_2_ulp(oracle):
    rng = np.random.default_rng(2)
    xs = np.concatenate([10 ** rng.uniform(-300, 300, 20000), rng.uniform(0.5, 2.0, 20000), [5e-324, 2.2e-308, 1.0]])
    worst = max(ulps(oracle.om_log(float(x)), math.log(x)) for x in xs)
    assert worst <= 2


def test_special_values(oracle):
    inf, nan = math.inf, math.nan
    assert oracle.om_exp(0.0) == 1.0 and oracle.om_log(1.0) == 0.0
    assert oracle.om_exp(710.0) == inf and oracle.om_exp(-746.0) == 0.0
    assert math.isnan(oracle.om_exp(nan)) and math.isnan(oracle.om_log(-1.0))
    assert oracle.om_log(0.0) == -inf and oracle.om_log(inf) == inf
    assert oracle.om_pow(0.0, 2.0) == 0.0 and oracle.om_pow(2.0, 0.0) == 1.0
    assert oracle.om_pow(-2.0, 3.0) == pytest.approx(-8.0, rel=1e-15)
    assert math.isnan(oracle.om_pow(-2.0, 0.5))
    # pow within a few ulp of libm on moderate exponents
    rng = np.random.default_rng(3)
    for x, y in zip(rng.uniform(0.01, 3, 2000), rng.uniform(-3, 3, 2000)):
        assert ulps(oracle.om_pow(x, y), x ** y) <= 8


def test_engine_math_is_the_oracle_bitwise(oracle, hostlane):
    rng = np.random.default_rng(4)
    for x in np.concatenate([rng.uniform(-700, 700, 5000), rng.uniform(-2, 2, 5000)]):
        assert ulps(hostlane.hl_exp(float(x)), oracle.om_exp(float(x))) == 0
    for x in 10 ** rng.uniform(-300, 300, 5000):
        assert ulps(hostlane.hl_log(float(x)), oracle.om_log(float(x))) == 0
    for x, y in zip(rng.uniform(0.01, 3, 2000), rng.uniform(-3, 3, 2000)):
        assert ulps(hostlane.hl_pow(float(x), float(y)), oracle.om_pow(float(x), float(y))) == 0


def test_sin_within_1_ulp(oracle, hostlane):
    """fdlibm's sin against the platform libm (itself within 1 ulp), over the
    range a source's phase reaches and beyond; the engine's host build is
    the oracle's bit for bit."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-1, 1, 20000), rng.uniform(-50, 50, 20000), rng.uniform(-1e5, 1e5, 20000),
                         [0.0, -0.0, 1e-300, 5e-324, math.pi / 4, math.pi / 2, math.pi, 2 * math.pi, 710.0,
                          2 * math.pi * 50 * 0.06]])
    worst = max(ulps(oracle.om_sin(float(x)), math.sin(x)) for x in xs)
    assert worst <= 1
    assert all(ulps(hostlane.hl_sin(float(x)), oracle.om_sin(float(x))) == 0 for x in xs[::7])
    assert math.isnan(oracle.om_sin(math.inf)) and math.isnan(oracle.om_sin(math.nan))
    assert math.copysign(1, oracle.om_sin(-0.0)) == -1
