"""GPU tests of per-coordinate sample transforms (prepare_sampler's transform=,
src/transforms.jl:160-235) through the C ABI: sampling stays in unconstrained
coordinates z, the target sees x = T(z) plus the forward log Jacobian, results report x.
Checked against the oracle's transform applied to the recovered z (log-weights within
1e-12 in FP64), known answers (the half-normal mass of exp(-x^2/2) on (0, inf) and on
(0, 1)), AMIS under a positive transform, the moments of the logical samples, and the
unsupported regression-target case."""

import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng(pkg):
    pkg.context(0)
    return pkg


def test_positive_transform_plain_is(eng, oracle_mod):
    _say("case: plain IS under PositiveTransform, d = 1, standard target")
    s = 1.5
    res = eng.importance_sample(5, eng.StandardGaussianTarget(),
                                eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(1), s), 200_000),
                                transform=eng.PositiveTransform())
    x = res.samples.reshape(-1)
    assert np.all(x > 0)
    z = np.log(x)
    xo, lj = oracle_mod.transform_forward(1, -math.inf, math.inf, z)
    lq = -0.5 * (z / s) ** 2 - math.log(s) - 0.5 * math.log(2 * math.pi)
    expected = -0.5 * xo ** 2 + lj - lq
    assert np.max(np.abs(res.logweights - expected)) < 1e-10
    # integral of exp(-x^2/2) over (0, inf) = sqrt(pi / 2)
    assert abs(eng.lognormalizer(res) - 0.5 * math.log(math.pi / 2)) < 5e-3
    # mean of the logical samples under the half-normal: sqrt(2 / pi)
    assert abs(eng.mean(res) - math.sqrt(2 / math.pi)) < 5e-3


def test_interval_transform_known_mass(eng, oracle_mod):
    _say("case: plain IS under IntervalTransform(0, 1), d = 1")
    s = 1.7
    res = eng.importance_sample(6, eng.StandardGaussianTarget(),
                                eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(1), s), 200_000),
                                transform=eng.IntervalTransform(0.0, 1.0))
    x = res.samples.reshape(-1)
    assert np.all((x > 0) & (x < 1))
    z = np.log(x) - np.log1p(-x)
    xo, lj = oracle_mod.transform_forward(3, 0.0, 1.0, z)
    lq = -0.5 * (z / s) ** 2 - math.log(s) - 0.5 * math.log(2 * math.pi)
    expected = -0.5 * xo ** 2 + lj - lq
    assert np.max(np.abs(res.logweights - expected)) < 1e-8
    mass = math.sqrt(math.pi / 2) * math.erf(1 / math.sqrt(2))
    assert abs(eng.lognormalizer(res) - math.log(mass)) < 2e-3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_amis_mixed_transforms(eng, precision):
    _say(f"case: AMIS under (Positive, Softplus, Interval(-1, 2)), d = 3, {precision}")
    tf = [eng.PositiveTransform(), eng.SoftplusTransform(), eng.IntervalTransform(-1.0, 2.0)]
    alg = eng.AMIS(eng.SphericalGaussian(np.zeros(3), 1.5), rounds=6, round_size=20000)
    res = eng.importance_sample(8, eng.StandardGaussianTarget(), alg, precision=precision, transform=tf)
    x = res.samples.reshape(len(res), 3)
    assert np.all(x[:, 0] > 0) and np.all(x[:, 1] > 0) and np.all((x[:, 2] > -1) & (x[:, 2] < 2))
    # exp(-|x|^2/2) over (0, inf)^2 x (-1, 2): (pi/2) * sqrt(pi/2) (erf(2/sqrt2) + erf(1/sqrt2))
    mass = (math.pi / 2) * math.sqrt(math.pi / 2) * (math.erf(2 / math.sqrt(2)) + math.erf(1 / math.sqrt(2)))
    lz = res.diagnostics["round_lognormalizers"][-1]
    assert abs(lz - math.log(mass)) < 1e-2


def test_transform_with_regression_target_is_unsupported(eng):
    _say("case: transform with a regression target is refused")
    X = np.ones((10, 2))
    y = np.zeros(10)
    with pytest.raises(Exception):
        eng.prepare_sampler(1, eng.LogisticRegressionTarget(X, y), eng.ImportanceSampling(
            eng.SphericalGaussian(np.zeros(2), 1.0), 100), transform=eng.PositiveTransform())
