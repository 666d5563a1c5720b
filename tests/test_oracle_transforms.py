"""Oracle pins for the scalar sample transforms (src/transforms.jl:160-235): the log
Jacobian against a central finite difference of the forward map (as test/transforms.jl:3-8
checks it), supports, and closed forms."""

import math

import numpy as np
import pytest

import oracle

CASES = [(0, -math.inf, math.inf), (1, -math.inf, math.inf), (2, -math.inf, math.inf), (3, 2.0, math.inf),
         (3, -math.inf, -1.0), (3, 0.0, 1.0), (3, -2.0, 5.0), (3, 3.0, 4.0)]


@pytest.mark.parametrize("kind,lo,hi", CASES)
def test_log_jacobian_matches_finite_difference(kind, lo, hi):
    z = np.linspace(-4.0, 4.0, 41)
    h = 1e-5
    x, lj = oracle.transform_forward(kind, lo, hi, z)
    xp, _ = oracle.transform_forward(kind, lo, hi, z + h)
    xm, _ = oracle.transform_forward(kind, lo, hi, z - h)
    fd = np.log(np.abs((xp - xm) / (2 * h)))
    assert np.max(np.abs(lj - fd)) < 1e-6
    if kind == 3:
        assert np.all(x > lo) and np.all(x < hi)
    if kind in (1, 2):
        assert np.all(x > 0)


def test_closed_forms():
    z = np.array([-3.0, 0.0, 2.5])
    x, lj = oracle.transform_forward(1, -math.inf, math.inf, z)
    assert np.array_equal(x, np.exp(z)) and np.array_equal(lj, z)
    x, lj = oracle.transform_forward(3, 0.0, 1.0, np.array([0.0]))
    assert x[0] == 0.5 and lj[0] == pytest.approx(math.log(0.25), rel=1e-15)
    x, _ = oracle.transform_forward(2, -math.inf, math.inf, np.array([0.0]))
    assert x[0] == pytest.approx(math.log(2.0), rel=1e-15)


NONFINITE_INPUT, NONFINITE_OUTPUT, OUTSIDE_SUPPORT = 0x0001, 0x0002, 0x0008


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_failure_reasons_known_answers(dtype):
    """test/transforms.jl:236-267,284-292: underflowing positive and softplus draws and an
    interval at +-floatmax fall outside the support, non-finite input, and an
    overflowing positive transform."""
    fi = np.finfo(dtype)
    underflow_z = float(np.log(dtype(np.nextafter(dtype(0), dtype(1))))) - 1.0
    for kind in (1, 2):
        assert oracle.transform_forward_checked(kind, -math.inf, math.inf, underflow_z, dtype)[2] == OUTSIDE_SUPPORT
    for z in (-float(fi.max), float(fi.max)):
        assert oracle.transform_forward_checked(3, -3.0, 5.0, z, dtype)[2] == OUTSIDE_SUPPORT
    for bad in (math.nan, math.inf, -math.inf):
        assert oracle.transform_forward_checked(0, -math.inf, math.inf, bad, dtype)[2] == NONFINITE_INPUT
    if dtype == np.float64:
        rounding_z = math.log(float(fi.eps) / 4)
        assert oracle.transform_forward_checked(3, -3.0, math.inf, rounding_z)[2] == OUTSIDE_SUPPORT
        assert oracle.transform_forward_checked(3, -math.inf, 5.0, rounding_z)[2] == OUTSIDE_SUPPORT
        assert oracle.transform_forward_checked(1, -math.inf, math.inf, float(fi.max))[2] == NONFINITE_OUTPUT
        for kind in (1, 2):
            assert oracle.transform_forward_c(kind, -math.inf, math.inf, underflow_z)[2] == OUTSIDE_SUPPORT
        assert oracle.transform_forward_c(3, -3.0, math.inf, rounding_z)[2] == OUTSIDE_SUPPORT
        assert oracle.transform_forward_c(3, -math.inf, 5.0, rounding_z)[2] == OUTSIDE_SUPPORT
        assert oracle.transform_forward_c(1, -math.inf, math.inf, float(fi.max))[2] == NONFINITE_OUTPUT
        assert oracle.transform_forward_c(0, -math.inf, math.inf, math.nan)[2] == NONFINITE_INPUT


@pytest.mark.parametrize("kind,lo,hi", CASES)
def test_checked_forms_agree_with_the_elementwise_map(kind, lo, hi):
    """The scalar checked restatements (numpy and C) give the elementwise map's values
    on regular inputs, with no failure bits."""
    z = np.linspace(-6.0, 6.0, 25)
    x, lj = oracle.transform_forward(kind, lo, hi, z)
    for k in range(z.size):
        xc, ljc, rc = oracle.transform_forward_c(kind, lo, hi, z[k])
        xn, ljn, rn = oracle.transform_forward_checked(kind, lo, hi, z[k])
        assert rc == 0 and rn == 0
        assert xc == pytest.approx(x[k], rel=1e-15, abs=1e-300) and ljc == pytest.approx(lj[k], rel=1e-14, abs=1e-15)
        assert xn == xc and ljn == ljc


def test_fp32_interval_promotes_to_the_fp64_bounds():
    """An FP32 draw through IntervalTransform(0.1, 0.3): the logistic is FP32, the
    interval value, log Jacobian and support check FP64 (the bounds' type).  Rounded to
    FP32 once, the value differs by an ulp from the all-FP32 form with the bounds rounded
    first for about a third of the draws, which is what the device test relies on."""
    rng = np.random.default_rng(0)
    lo32, hi32 = np.float32(0.1), np.float32(0.3)
    differ = 0
    for z in rng.normal(0, 3, 2000).astype(np.float32):
        x, lj, r = oracle.transform_forward_checked(3, 0.1, 0.3, float(z), np.float32)
        assert r == 0 and 0.1 < x < 0.3 and isinstance(x, float)
        p = np.float32(1) / (np.float32(1) + np.exp(-z)) if z >= 0 else np.exp(z) / (np.float32(1) + np.exp(z))
        x32 = lo32 + p * (hi32 - lo32) if p <= 0.5 else hi32 - (np.float32(1) - p) * (hi32 - lo32)
        differ += np.float32(x) != x32
    assert 300 < differ < 1500
