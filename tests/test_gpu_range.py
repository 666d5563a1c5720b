"""GPU tests of proposal log-densities at the edges of the floating-point range
(src/proposals.jl:24-55,603-669 and the extended-range repair of
src/student_t_range.jl:55-153): the reference's known answers of
test/native_proposals.jl:205-306, with the expected values recomputed here in mpmath
(the BigFloat formulas of those tests), evaluated on the device by the sampling
kernels' own density code; plus an AMIS run whose Student-t factor has entries far
outside the ordinary range, against the same run with the factor rescaled."""

import math
import sys

import mpmath
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
mpmath.mp.prec = 256


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng(pkg):
    pkg.context(0)
    return pkg


def _close(got, expected, rtol):
    assert abs(got - expected) <= rtol * max(abs(expected), 1e-300), (got, expected)


def _gauss_expected(loc, L, point):
    """log N(point; loc, L L^T) in mpmath from the lower factor L (forward solve)."""
    d = len(point)
    y = []
    for r in range(d):
        v = mpmath.mpf(point[r]) - mpmath.mpf(loc[r])
        for c in range(r):
            v -= mpmath.mpf(L[r][c]) * y[c]
        y.append(v / mpmath.mpf(L[r][r]))
    r2 = sum(v * v for v in y)
    logdet = sum(mpmath.log(mpmath.mpf(L[r][r])) for r in range(d))
    return -logdet - d * mpmath.log(2 * mpmath.pi) / 2 - r2 / 2, r2


def _t_expected(nu, loc, L, point):
    d = len(point)
    _, r2 = _gauss_expected(loc, L, point)
    nu = mpmath.mpf(nu)
    logdet = sum(mpmath.log(mpmath.mpf(L[r][r])) for r in range(d))
    const = mpmath.loggamma((nu + d) / 2) - mpmath.loggamma(nu / 2) - d * mpmath.log(nu * mpmath.pi) / 2
    return const - logdet - (nu + d) / 2 * mpmath.log1p(r2 / nu)


@pytest.mark.parametrize("prec,magnitude,tail", [("fp32", 2e38, 2e19), ("fp64", 1e308, 1.5e154)])
def test_gaussian_representable_density_range(eng, prec, magnitude, tail):
    """test/native_proposals.jl:205-234."""
    _say(f"case: Gaussian density range, {prec}")
    rtol = 8 * (2.0 ** -23 if prec == "fp32" else 2.0 ** -52)
    T = np.float32 if prec == "fp32" else np.float64
    for location, scale, point in ((-magnitude, magnitude, magnitude), (0.0, 1.0, tail)):
        location, scale, point = float(T(location)), float(T(scale)), float(T(point))
        for p in (eng.SphericalGaussian(np.array([location, 0.0]), scale),
                  eng.DiagonalGaussian(np.array([location, 0.0]), [scale, scale]),
                  eng.FactorGaussian(np.array([location, 0.0]), np.diag([scale, scale]))):
            exp, _ = _gauss_expected([location, 0.0], [[scale, 0.0], [0.0, scale]], [point, 0.0])
            _close(eng.logdensityof(p, [point, 0.0], prec), float(exp), rtol)
    largest = float(np.finfo(T).max)
    p = eng.FactorGaussian(np.zeros(2), np.array([[1.0, 0.0], [largest, largest]]))
    exp, _ = _gauss_expected([0, 0], [[1, 0], [largest, largest]], [2.0, 0.0])
    _close(eng.logdensityof(p, [2.0, 0.0], prec), float(exp), rtol)
    pt = float(T(0.75 * math.sqrt(largest)))
    exp, _ = _gauss_expected([0, 0], [[1, 0], [0, 1]], [pt, pt])
    _close(eng.logdensityof(eng.SphericalGaussian(np.zeros(2), 1.0), [pt, pt], prec), float(exp), rtol)


def test_student_t_analytic_range(eng):
    """test/native_proposals.jl:236-262 (FP64): spherical scales and points far apart, and
    factors with entries from 1e-300 to 1e300 that only the extended-range log radius can
    weigh."""
    _say("case: Student-t density range, FP64")
    for location, scale, point in ((0.0, 1e-200, 1e200), (-1e308, 1.0, 1e308)):
        exp = _t_expected(1.0, [location], [[scale]], [point])
        _close(eng.logdensityof(eng.SphericalStudentT(1.0, location, scale), point), float(exp), 1e-12)
    factor = np.diag([1.0, 1e-200, 1e-300, 1e-300])
    factor[2, 0] = factor[3, 2] = 1e300
    eta = float(np.nextafter(0.0, 1.0))
    for L, point in ((np.array([[1e-200, 0.0], [1e-200, 1.0]]), [1e200, 1e200]),
                     (factor, [1e-300, 1e200, 0.0, 0.0]),
                     (np.array([[1e200, 0.0], [1e300, 1e-300]]), [1e-200, 0.0]),
                     (np.array([[1.0, 0.0], [2e-124, eta]]), [1e-200, eta])):
        d = len(point)
        exp = _t_expected(1.0, [0.0] * d, L.tolist(), point)
        got = eng.logdensityof(eng.FactorStudentT(1.0, np.zeros(d), L), point)
        _close(got, float(exp), 1e-12)
    for nu, point, prec in ((1.0, 1e200, "fp64"), (0.1, 8e18, "fp32"), (1e308, 2e154, "fp64")):
        if prec == "fp32":
            nu, point = float(np.float32(nu)), float(np.float32(point))
        exp = _t_expected(nu, [0.0], [[1.0]], [point])
        got = eng.logdensityof(eng.SphericalStudentT(nu, np.zeros(1), 1.0), [point], prec)
        _close(got, float(exp), 1e-5 if prec == "fp32" else 1e-12)
    pt = [1e200, -1e200]
    exp = _t_expected(1.0, [0, 0], [[1, 0], [0, 1]], pt)
    for p in (eng.SphericalStudentT(1.0, np.zeros(2), 1.0), eng.DiagonalStudentT(1.0, np.zeros(2), np.ones(2)),
              eng.FactorStudentT(1.0, np.zeros(2), np.eye(2))):
        _close(eng.logdensityof(p, pt), float(exp), 1e-12)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_student_t_subnormal_radial_density(eng, prec):
    """test/native_proposals.jl:292-306: nu = nextfloat(0), points at sqrt(nu)/2 and
    1.25 sqrt(nu)."""
    _say(f"case: Student-t subnormal radius, {prec}")
    T = np.float32 if prec == "fp32" else np.float64
    nu = float(np.nextafter(T(0), T(1)))
    rtol = 8 * (2.0 ** -23 if prec == "fp32" else 2.0 ** -52)
    for point in (float(T(math.sqrt(nu)) / T(2)), float(T(1.25) * T(math.sqrt(nu)))):
        exp = (-mpmath.log(2 * mpmath.pi) - (mpmath.mpf(nu) + 2) / 2 *
               mpmath.log1p(mpmath.mpf(point) ** 2 / mpmath.mpf(nu)))
        for p in (eng.SphericalStudentT(nu, np.zeros(2), 1.0), eng.DiagonalStudentT(nu, np.zeros(2), np.ones(2)),
                  eng.FactorStudentT(nu, np.zeros(2), np.eye(2))):
            _close(eng.logdensityof(p, [point, 0.0], prec), float(exp), max(rtol, 1e-12))


def test_ordinary_densities_match_the_oracle(eng, oracle_mod):
    """Away from the range edges the device density agrees with the oracle's
    (src/proposals.jl:24-55) for every family and scale kind."""
    _say("case: ordinary proposal densities vs oracle")
    rng = np.random.default_rng(3)
    d = 5
    loc = rng.normal(0, 1, d)
    F = np.tril(rng.normal(0, 0.3, (d, d)), -1) + np.diag(rng.uniform(0.5, 2.0, d))
    pts = rng.normal(0, 2, (200, d))
    for p in (eng.SphericalGaussian(loc, 1.7), eng.DiagonalGaussian(loc, np.diag(F)), eng.FactorGaussian(loc, F),
              eng.SphericalStudentT(3.0, loc, 1.7), eng.FactorStudentT(0.7, loc, F)):
        got = eng.logdensityof(p, pts)
        exp = np.array([oracle_mod.proposal_logdensity(p, x) for x in pts])
        assert np.max(np.abs(got - exp) / np.maximum(1.0, np.abs(exp))) < 1e-12


def test_sampling_with_subnormal_factors(eng):
    """Plain IS whose Student-t proposal and Gaussian target factors are subnormal
    (2^-1040 scale): the reciprocal diagonal is not representable, so the regular kernels
    give NaN densities and the engine replays the run on the extended-range kernels;
    every log-weight matches mpmath evaluated at the returned samples."""
    _say("case: sampling with subnormal Student-t and Gaussian factors")
    d, k = 2, 2.0 ** -1040
    L = np.array([[1.0, 0.0], [0.25, 0.75]])
    q = eng.FactorStudentT(5.0, np.zeros(d), 1.5 * k * L)
    # a diagonal target factor: the plain forward substitution of the Gaussian target
    # would otherwise round subnormal products L_rc y_c (2^-34 relative here), which the
    # reference's plain arithmetic does as well; the Student-t proposal's log radius is
    # formed in extended range and keeps full precision despite its off-diagonal entry
    Ltd = np.diag(np.diag(L))
    tgt = eng.GaussianTarget(np.zeros(d), k * Ltd)
    res = eng.importance_sample(3, tgt, eng.ImportanceSampling(q, 512))
    assert np.all(np.isfinite(res.logweights))
    Lq = (1.5 * k * L).tolist()
    Lt = (k * Ltd).tolist()
    for i in range(0, 512, 37):
        x = [float(v) for v in res.samples[i]]
        lt, _ = _gauss_expected([0.0, 0.0], Lt, x)
        lq = _t_expected(5.0, [0.0, 0.0], Lq, x)
        _close(res.logweights[i], float(lt - lq), 1e-10)
