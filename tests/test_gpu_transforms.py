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


TF_MIX = [(1, -math.inf, math.inf), (2, -math.inf, math.inf), (3, -1.0, 2.0)]


def _tf(eng):
    return [eng.PositiveTransform(), eng.SoftplusTransform(), eng.IntervalTransform(-1.0, 2.0)]


def _tf_target(oracle_mod, tgt, d):
    """The oracle's transformed target (kind 5) around the engine target `tgt`."""
    kinds = [k for k, _, _ in TF_MIX][:d]
    bounds = [(lo, hi) for _, lo, hi in TF_MIX][:d]
    return oracle_mod.transformed_target(tgt.kind, tgt.params(d), kinds, bounds)


def _relerr(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def test_transform_failures_carry_their_reason(eng):
    """Transform failures fail the draw (src/native_execution.jl:833-840) with
    InvalidTransformError(reason, location): nonfinite_output for an overflowing
    positive transform, outside_support for an underflowing softplus."""
    _say("case: transform failure reasons (plain IS, AMIS; FP64 and FP32)")
    alg = eng.ImportanceSampling(eng.SphericalGaussian(np.array([0.0, 800.0]), 1.0), 256)
    with pytest.raises(eng.SamplerExecutionError) as ei:
        eng.importance_sample(1, eng.StandardGaussianTarget(), alg,
                              transform=[eng.IdentityTransform(), eng.PositiveTransform()])
    assert ei.value.phase == "proposal_draw" and ei.value.index == 1
    assert isinstance(ei.value.cause, eng.InvalidTransformError)
    assert ei.value.cause.reason == "nonfinite_output" and ei.value.cause.location == 2
    alg = eng.ImportanceSampling(eng.SphericalGaussian(np.array([-800.0]), 1.0), 64)
    with pytest.raises(eng.SamplerExecutionError) as ei:
        eng.importance_sample(2, eng.StandardGaussianTarget(), alg, transform=eng.SoftplusTransform())
    assert ei.value.cause.reason == "outside_support" and ei.value.cause.location == 1
    alg = eng.ImportanceSampling(eng.SphericalGaussian(np.array([100.0]), 1.0), 64)
    with pytest.raises(eng.SamplerExecutionError) as ei:  # exp(100) overflows FP32
        eng.importance_sample(3, eng.StandardGaussianTarget(), alg, precision="fp32", transform=eng.PositiveTransform())
    assert ei.value.cause.reason == "nonfinite_output"
    s = eng.prepare_sampler(4, eng.StandardGaussianTarget(),
                            eng.AMIS(eng.SphericalGaussian(np.array([800.0, 0.0]), 1.0), rounds=2, round_size=128),
                            transform=eng.PositiveTransform())
    with pytest.raises(eng.AMISRoundError) as ei:
        eng.importance_sample_(s)
    assert ei.value.round == 1 and ei.value.phase == "sampling"
    assert ei.value.cause.cause.reason == "nonfinite_output" and ei.value.cause.cause.location == 1


def test_fp32_interval_values_use_fp64_bounds(eng, oracle_mod):
    """FP32 samples through IntervalTransform(0.1, 0.3): the value is formed in FP64 from
    the FP32 logistic (the bounds' type, src/transforms.jl:217-235) and rounded once; the
    all-FP32 form with rounded bounds differs by an ulp for about a third of the draws."""
    _say("case: FP32 interval transform against the FP64-bounds restatement")
    n = 4096
    s = eng.prepare_sampler(9, eng.StandardGaussianTarget(),
                            eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(1), 3.0), n), precision="fp32",
                            transform=eng.IntervalTransform(0.1, 0.3))
    z = (np.float32(3.0) * s.export_randomness(0)["normals"].astype(np.float32)).astype(np.float32)
    res = eng.importance_sample_(s)
    x = res.samples.reshape(-1).astype(np.float32)
    exp_x = np.array([np.float32(oracle_mod.transform_forward_checked(3, 0.1, 0.3, float(v), np.float32)[0])
                      for v in z])
    ulp = np.abs(x.view(np.int32).astype(np.int64) - exp_x.view(np.int32).astype(np.int64))
    # the device's FP32 exponential may differ by an ulp from numpy's, which 1 - p near 1
    # amplifies: compare match rates against both forms instead of demanding equality
    assert np.max(ulp) <= 8
    lo32, hi32 = np.float32(0.1), np.float32(0.3)
    fp32_form = []
    for v in z:
        p = np.float32(1) / (np.float32(1) + np.exp(-v)) if v >= 0 else np.exp(v) / (np.float32(1) + np.exp(v))
        fp32_form.append(lo32 + p * (hi32 - lo32) if p <= 0.5 else hi32 - (np.float32(1) - p) * (hi32 - lo32))
    promoted, rounded_bounds = np.mean(ulp == 0), np.mean(x == np.array(fp32_form, np.float32))
    assert promoted > 0.9 and promoted - rounded_bounds > 0.2, (promoted, rounded_bounds)
    assert np.all((x.astype(np.float64) > 0.1) & (x <= np.float32(0.3)))


@pytest.mark.parametrize("method", ["npmc", "dmpmc", "mis", "amis-logistic", "plain-logistic"])
def test_transforms_under_every_method_match_the_oracle(eng, oracle_mod, method):
    """(Positive, Softplus, Interval(-1, 2)) under NPMC, DM-PMC, static MIS, and with the
    batched logistic target, against the oracle's transformed target on the same
    counter-RNG draws (FP64, log-weights within 1e-10)."""
    _say(f"case: transforms under {method} vs oracle")
    d = 3
    rng = np.random.default_rng(31)
    A = rng.standard_normal((d, d))
    gauss = eng.GaussianTarget(rng.normal(0.0, 0.5, d), np.linalg.cholesky(A @ A.T / d + np.eye(d)))
    if method == "npmc":
        sched = [3000, 3000, 3000]
        q0 = eng.FactorGaussian(np.zeros(d), np.diag([1.2, 1.5, 1.1]))
        s = eng.prepare_sampler(5, gauss, eng.NPMC(q0, rounds=3, round_size=sched), transform=_tf(eng))
        normals = np.array([s.export_randomness(t)["normals"] for t in range(3)])
        res = eng.importance_sample_(s)
        means, facs, lns = s.history()
        tk, tp = _tf_target(oracle_mod, gauss, d)
        o = oracle_mod.adaptive(1, q0, sched, normals, tk, tp, replay=(means, facs, lns))
        assert o["rc"] == 0
        assert _relerr(res.logweights, o["logweights"]) < 1e-10
        assert _relerr(res.diagnostics["round_lognormalizers"], o["round_lognormalizers"]) < 1e-10
    elif method == "dmpmc":
        K, sched = 6, [1200, 1200, 1200]
        props = [eng.SphericalGaussian(rng.normal(0, 0.7, d), 1.0 + 0.1 * k) for k in range(K)]
        bank = eng.ProposalBank(props)
        s = eng.prepare_sampler(6, gauss, eng.DeterministicMixturePMC(bank, rounds=3, round_size=sched),
                                transform=_tf(eng))
        normals = np.array([s.export_randomness(t)["normals"] for t in range(3)])
        uniforms = np.array([s.export_randomness(t)["uniforms"] for t in range(3)])
        res = eng.importance_sample_(s)
        order = oracle_mod.active_order(list(bank.masses))
        counts, logc = oracle_mod.allocation_plan([bank.masses[i] for i in order], sched)
        P = [props[i] for i in order]
        tk, tp = _tf_target(oracle_mod, gauss, d)
        o = oracle_mod.dmpmc(0, 0.0, 0, d, np.array([p.location for p in P]), np.array([p.scale[0] for p in P]),
                             [p.lognormalizer for p in P], [i + 1 for i in order], sched, counts, logc, normals,
                             uniforms, tk, tp, resampling=2)
        assert o["rc"] == 0
        assert _relerr(res.logweights, o["logweights"]) < 1e-10
        assert np.array_equal(s.ancestors(), o["ancestors"])
    elif method == "mis":
        props = [eng.SphericalGaussian(rng.normal(0, 0.5, d), 1.3) for _ in range(4)]
        masses = np.array([1.0, 2.0, 0.5, 1.0])
        N = 4000
        s = eng.prepare_sampler(7, gauss, eng.ImportanceSampling(eng.ProposalBank(props, masses), N),
                                transform=_tf(eng))
        r = s.export_randomness(0)
        res = eng.importance_sample_(s)
        tk, tp = _tf_target(oracle_mod, gauss, d)
        o = oracle_mod.static_mis(props, masses, 0, None, r["normals"].reshape(N, d), r["uniforms"], tk, tp)
        assert np.array_equal(res.provenance["proposal_id"], o["proposal_ids"])
        assert _relerr(res.logweights, o["logweights"]) < 1e-10
    else:
        X = rng.standard_normal((500, d))
        y = (rng.random(500) < 0.5).astype(float)
        logi = eng.LogisticRegressionTarget(X, y, 10.0)
        tk, tp = _tf_target(oracle_mod, logi, d)
        q0 = eng.SphericalGaussian(np.zeros(d), 0.6)
        if method == "plain-logistic":
            s = eng.prepare_sampler(8, logi, eng.ImportanceSampling(q0, 8192), transform=_tf(eng))
            z = s.export_randomness(0)["normals"]
            res = eng.importance_sample_(s)
            xo, lwo, fail = oracle_mod.plain_is(q0, tk, tp, z)
            assert fail[0] == 0
            assert _relerr(res.logweights, lwo) < 1e-10
            # the reported samples are the logical ones
            xl = np.array([[oracle_mod.transform_forward_c(k, lo, hi, v)[0] for (k, lo, hi), v in zip(TF_MIX, row)]
                           for row in xo])
            assert np.max(np.abs(res.samples - xl)) < 1e-12
        else:
            sched = [2048, 2048, 2048]
            s = eng.prepare_sampler(9, logi, eng.AMIS(q0, rounds=3, round_size=sched), transform=_tf(eng))
            normals = np.array([s.export_randomness(t)["normals"] for t in range(3)])
            res = eng.importance_sample_(s)
            means, facs, lns = s.history()
            o = oracle_mod.adaptive(0, q0, sched, normals, tk, tp, replay=(means, facs, lns))
            assert o["rc"] == 0
            assert _relerr(res.logweights, o["logweights"]) < 1e-10
