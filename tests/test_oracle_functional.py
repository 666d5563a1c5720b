"""Oracle pins for mean(f, result) / var(f, result) (src/statistics.jl:17-31,50-64,170-199):
the reference's known answers and its zero-weight and overflow rules."""

import math
from fractions import Fraction

import numpy as np
import pytest

import oracle


def test_known_answers_scalar():
    # test/results.jl:9-26: samples [1, 3, 10], log-weights 1000 + log [1, 2, 1] (1000 + log 2
    # rounds at 1.1e-13, so the weights are 1/4, 1/2, 1/4 to about 1e-13)
    x = np.array([1.0, 3.0, 10.0])
    lw = 1000 + np.log([1.0, 2.0, 1.0])
    m, v = oracle.functional_moments(lw, x)
    assert m == pytest.approx(4.25, rel=1e-12) and v == pytest.approx(11.6875, rel=1e-12)
    m, v = oracle.functional_moments(lw, x * x)
    assert m == pytest.approx(29.75, rel=1e-12) and v == pytest.approx(1655.6875, rel=1e-12)


def test_known_answers_structured():
    # test/results.jl:28-38: samples [[1, 3, 10], [2, 4, 8]], f(x) = x1 + 2 x2
    s = np.array([[1.0, 3.0, 10.0], [2.0, 4.0, 8.0]])
    lw = 1000 + np.log([1.0, 2.0, 1.0])
    m, v = oracle.functional_moments(lw, s[0] + 2 * s[1])
    assert m == pytest.approx(13.25, rel=1e-12) and v == pytest.approx(60.1875, rel=1e-12)


@pytest.mark.parametrize("logmass", [math.log(1e-20), -math.inf])
def test_variance_range(logmass):
    # test/cuda_statistics.jl:27-38: samples [0, 1e20], log-weights [0, logmass]:
    # var = mass 1e40 / (1 + mass)^2, exactly in rationals
    mass = Fraction(0) if logmass == -math.inf else Fraction(math.exp(logmass))
    expected = float(mass * Fraction(1e20) ** 2 / (1 + mass) ** 2)
    _, v = oracle.functional_moments([0.0, logmass], [0.0, 1e20])
    assert v == pytest.approx(expected, rel=1e-14, abs=0.0)


def test_zero_weight_rules():
    # a zero weight times a non-finite value poisons the mean (dot), the variance skips it
    m, _ = oracle.functional_moments([0.0, -math.inf], [1.0, math.inf])
    assert math.isnan(m)
    m, v = oracle.functional_moments([0.0, -math.inf], [2.0, 5.0])
    assert m == 2.0 and v == 0.0


def test_all_zero_weights_raise():
    with pytest.raises(ZeroDivisionError):
        oracle.functional_moments([-math.inf, -math.inf], [1.0, 2.0])


def test_equal_weights_reduce_to_sample_moments():
    rng = np.random.default_rng(3)
    f = rng.normal(size=1000)
    m, v = oracle.functional_moments(np.full(1000, -7.0), f)
    assert m == pytest.approx(f.mean(), rel=1e-13)
    assert v == pytest.approx(f.var(), rel=1e-12)
