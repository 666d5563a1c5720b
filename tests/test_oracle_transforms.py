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
