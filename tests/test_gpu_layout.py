"""GPU tests of named flat selector layouts and SimplexTransform (prepare_sampler's
transform=(name = selector => transform, ...), src/proposal_composition.jl:179-240,
src/native_execution.jl:556-712, src/named_parameters.jl) through the C ABI: the chart
measure known answer (a target equal to the proposal's pushforward gives log-weights of
zero, test/product_proposal.jl:283-320), parity with the oracle's layout target on the
same counter-RNG draws under every method (FP64, log-weights within 1e-10), the exactly
uniform zero draw, the failing block as the error's location (test/kernel_execution.jl:
615-640), identity layouts equal to the plain run (test/product_proposal.jl:140-170),
named results and their statistics, named leaves in host batch targets
(test/batched_targets.jl:120-133), and a regression target over the logical dimension."""

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


FACTOR = np.array([[1.0, 0.0, 0.0, 0.0],
                   [0.3, 1.2, 0.0, 0.0],
                   [-0.2, 0.4, 0.8, 0.0],
                   [0.1, -0.3, 0.25, 1.5]])
ORACLE_LAYOUT = [(0, 2, 4), (2, 1, 1), (3, 1, 0)]


def _layout(eng):
    return {"weights": (range(0, 2), eng.SimplexTransform(3)), "rate": (2, eng.PositiveTransform()), "offset": 3}


def _flat(res):
    s = res.samples
    return np.column_stack([s["weights"], s["rate"], s["offset"]])


def _relerr(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def _structured(x, factor):
    """_dependent_structured_logdensity (test/product_proposal.jl:54-74) over named leaves."""
    lw = np.log(x["weights"])
    mean_lw = lw.sum(axis=1) / 3
    inv_root = 1 / math.sqrt(3)
    coef = inv_root / (1 - inv_root)
    last = lw[:, 2] - mean_lw
    coords = np.column_stack([lw[:, 0] - mean_lw + coef * last, lw[:, 1] - mean_lw + coef * last,
                              np.log(x["rate"]), x["offset"]])
    std = np.linalg.solve(factor, coords.T)
    base = -2 * math.log(2 * math.pi) - np.sum(np.log(np.diag(factor))) - 0.5 * np.sum(std ** 2, axis=0)
    return base - (0.5 * math.log(3) + lw.sum(axis=1) + np.log(x["rate"]))


def test_named_layout_preserves_the_chart_measure(eng):
    _say("case: named layout (Simplex(3), Positive, Identity), pushforward target, plain IS FP64")

    def batch(out, xs):
        out[:] = _structured(xs, FACTOR)

    n = 20_000
    res = eng.importance_sample(11, eng.HostBatchTarget(batch),
                                eng.ImportanceSampling(eng.FactorGaussian(np.zeros(4), FACTOR), n),
                                transform=_layout(eng))
    assert set(res.samples) == {"weights", "rate", "offset"}
    assert res.samples["weights"].shape == (n, 3) and res.samples["rate"].shape == (n,)
    assert res.samples["offset"].shape == (n,)
    assert np.all(res.samples["weights"] > 0) and np.all(res.samples["rate"] > 0)
    assert np.max(np.abs(res.samples["weights"].sum(axis=1) - 1.0)) <= 4 * np.finfo(float).eps
    assert np.max(np.abs(res.logweights)) <= 256 * np.finfo(float).eps
    assert abs(eng.lognormalizer(res)) <= 256 * np.finfo(float).eps


def _gauss(dl, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((dl, dl))
    return rng.normal(0.0, 0.3, dl), np.linalg.cholesky(A @ A.T / dl + np.eye(dl))


@pytest.mark.parametrize("method", ["plain", "amis", "npmc", "dmpmc", "mis"])
def test_layout_matches_the_oracle(eng, oracle_mod, method):
    _say(f"case: named layout under {method}, Gaussian target over dl = 5, FP64 vs oracle")
    d, dl = 4, 5
    m, L = _gauss(dl, 3)
    gauss = eng.GaussianTarget(m, L)
    tk, tp = oracle_mod.layout_target(gauss.kind, gauss.params(dl), ORACLE_LAYOUT)
    rng = np.random.default_rng(17)
    q0 = eng.FactorGaussian(np.zeros(d), FACTOR)
    if method == "plain":
        s = eng.prepare_sampler(5, gauss, eng.ImportanceSampling(q0, 6000), transform=_layout(eng))
        z = s.export_randomness(0)["normals"]
        res = eng.importance_sample_(s)
        xo, lwo, fail = oracle_mod.plain_is(q0, tk, tp, z)
        assert fail[0] == 0
        assert _relerr(res.logweights, lwo) < 1e-10
        xl, _, reason, _ = oracle_mod.layout_forward(ORACLE_LAYOUT, xo)
        assert np.all(reason == 0)
        assert np.max(np.abs(_flat(res) - xl)) < 1e-13
    elif method in ("amis", "npmc"):
        sched = [3000, 3000, 3000]
        alg = (eng.AMIS if method == "amis" else eng.NPMC)(q0, rounds=3, round_size=sched)
        s = eng.prepare_sampler(6, gauss, alg, transform=_layout(eng))
        normals = np.array([s.export_randomness(t)["normals"] for t in range(3)])
        res = eng.importance_sample_(s)
        means, facs, lns = s.history()
        o = oracle_mod.adaptive(0 if method == "amis" else 1, q0, sched, normals, tk, tp, replay=(means, facs, lns))
        assert o["rc"] == 0
        assert _relerr(res.logweights, o["logweights"]) < 1e-10
        assert _relerr(res.diagnostics["round_lognormalizers"], o["round_lognormalizers"]) < 1e-10
    elif method == "dmpmc":
        K, sched = 5, [1000, 1000, 1000]
        props = [eng.SphericalGaussian(rng.normal(0, 0.5, d), 1.0 + 0.1 * k) for k in range(K)]
        bank = eng.ProposalBank(props)
        s = eng.prepare_sampler(7, gauss, eng.DeterministicMixturePMC(bank, rounds=3, round_size=sched),
                                transform=_layout(eng))
        normals = np.array([s.export_randomness(t)["normals"] for t in range(3)])
        uniforms = np.array([s.export_randomness(t)["uniforms"] for t in range(3)])
        res = eng.importance_sample_(s)
        order = oracle_mod.active_order(list(bank.masses))
        counts, logc = oracle_mod.allocation_plan([bank.masses[i] for i in order], sched)
        P = [props[i] for i in order]
        o = oracle_mod.dmpmc(0, 0.0, 0, d, np.array([p.location for p in P]), np.array([p.scale[0] for p in P]),
                             [p.lognormalizer for p in P], [i + 1 for i in order], sched, counts, logc, normals,
                             uniforms, tk, tp)
        assert o["rc"] == 0
        assert _relerr(res.logweights, o["logweights"]) < 1e-10
        assert np.array_equal(s.ancestors(), o["ancestors"])
    else:
        props = [eng.SphericalGaussian(rng.normal(0, 0.5, d), 1.3) for _ in range(3)]
        masses = np.array([1.0, 2.0, 0.5])
        N = 3000
        s = eng.prepare_sampler(8, gauss, eng.ImportanceSampling(eng.ProposalBank(props, masses), N),
                                transform=_layout(eng))
        r = s.export_randomness(0)
        res = eng.importance_sample_(s)
        o = oracle_mod.static_mis(props, masses, 0, None, r["normals"].reshape(N, d), r["uniforms"], tk, tp)
        assert np.array_equal(res.provenance["proposal_id"], o["proposal_ids"])
        assert _relerr(res.logweights, o["logweights"]) < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_zero_draws_are_exactly_uniform(eng, precision):
    # test/transforms.jl:346-352 on the device: z = 0 gives x = 1/K exactly
    _say(f"case: SimplexTransform(4) at z = 0, {precision}")
    n = 64
    rng = eng.PrefilledRNG(normals=[np.zeros(3 * n)])
    res = eng.importance_sample(rng, eng.StandardGaussianTarget(),
                                eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(3), 1.0), n),
                                precision=precision, transform=eng.SimplexTransform(4))
    assert res.samples.shape == (n, 4)
    assert np.all(res.samples == 0.25)


def test_fp32_simplex_against_the_fp64_run(eng):
    _say("case: named layout FP32 vs FP64, Gaussian target over dl = 5")
    m, L = _gauss(5, 4)
    alg = eng.ImportanceSampling(eng.FactorGaussian(np.zeros(4), FACTOR), 8192)
    a = eng.importance_sample(9, eng.GaussianTarget(m, L), alg, transform=_layout(eng))
    b = eng.importance_sample(9, eng.GaussianTarget(m, L), alg, precision="fp32", transform=_layout(eng))
    fa, fb = _flat(a), _flat(b)
    assert np.max(np.abs(fb - fa) / np.maximum(1.0, np.abs(fa))) < 1e-5
    e32 = float(np.finfo(np.float32).eps)
    assert np.max(np.abs(b.samples["weights"].astype(np.float64).sum(axis=1) - 1.0)) <= 8 * e32
    assert np.max(np.abs(b.logweights - a.logweights)) < 2e-4
    assert abs(eng.lognormalizer(b) - eng.lognormalizer(a)) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_failing_blocks_name_their_selector(eng, precision):
    _say(f"case: simplex and layout transform failures, {precision}")
    big = float(np.finfo(np.float32 if precision == "fp32" else np.float64).max)
    # test/kernel_execution.jl:623-640: SphericalGaussian(zeros(2), floatmax), normals
    # (0, 0), (1, 1): the second draw's coordinate sum overflows
    rng = eng.PrefilledRNG(normals=[np.array([0.0, 0.0, 1.0, 1.0])])
    alg = eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(2), big), 2)
    with pytest.raises(eng.SamplerExecutionError) as ei:
        eng.importance_sample(rng, eng.StandardGaussianTarget(), alg, precision=precision,
                              transform=eng.SimplexTransform(3))
    assert ei.value.phase == "proposal_draw" and ei.value.index == 2
    assert isinstance(ei.value.cause, eng.InvalidTransformError)
    assert ei.value.cause.reason == "nonfinite_output" and ei.value.cause.location == range(0, 2)
    # the rate block (an Int selector) overflows in the third draw
    z = np.zeros(12)
    z[4 * 2 + 2] = 1.0
    rng = eng.PrefilledRNG(normals=[z])
    alg = eng.ImportanceSampling(eng.DiagonalGaussian(np.zeros(4), [1.0, 1.0, 900.0, 1.0]), 3)
    with pytest.raises(eng.SamplerExecutionError) as ei:
        eng.importance_sample(rng, eng.StandardGaussianTarget(), alg, precision=precision, transform=_layout(eng))
    assert ei.value.index == 3 and ei.value.cause.reason == "nonfinite_output" and ei.value.cause.location == 2


def test_identity_layout_equals_the_plain_run(eng):
    # test/product_proposal.jl:140-170: (scalar=1, vector=2:3, singleton=4:4)
    _say("case: identity layout vs no transform, plain IS and AMIS FP64")
    q = eng.DiagonalGaussian([0.5, -1.0, 2.0, 1.0], [1.25, 1.0, 2.0, 0.75])
    spec = {"scalar": 0, "vector": range(1, 3), "singleton": range(3, 4)}
    for alg in (eng.ImportanceSampling(q, 4096), eng.AMIS(q, rounds=3, round_size=2048)):
        a = eng.importance_sample(12, eng.StandardGaussianTarget(), alg, transform=spec)
        b = eng.importance_sample(12, eng.StandardGaussianTarget(), alg)
        assert a.samples["scalar"].shape == (len(b),) and a.samples["vector"].shape == (len(b), 2)
        assert a.samples["singleton"].shape == (len(b), 1)
        flat = np.column_stack([a.samples["scalar"], a.samples["vector"], a.samples["singleton"]])
        if isinstance(alg, eng.ImportanceSampling):
            assert np.array_equal(flat, b.samples)  # the same draws (the fits below round apart)
        else:
            assert _relerr(flat, b.samples) < 1e-11
        assert _relerr(a.logweights, b.logweights) < 1e-11


def test_named_statistics_and_resampling(eng):
    _say("case: mean / var / cov / resample of named results")
    m, L = _gauss(5, 5)
    res = eng.importance_sample(13, eng.GaussianTarget(m, L),
                                eng.ImportanceSampling(eng.FactorGaussian(np.zeros(4), FACTOR), 10_000),
                                transform=_layout(eng))
    w = np.exp(res.logweights - res.logweights.max())
    w /= w.sum()
    flat = _flat(res)
    mu = w @ flat
    va = w @ (flat - mu) ** 2
    mean, var = eng.mean(res), eng.var(res)
    assert set(mean) == {"weights", "rate", "offset"} and mean["weights"].shape == (3,)
    assert isinstance(mean["rate"], float) and isinstance(var["offset"], float)
    assert np.max(np.abs(np.concatenate([mean["weights"], [mean["rate"], mean["offset"]]]) - mu)) < 1e-12
    assert np.max(np.abs(np.concatenate([var["weights"], [var["rate"], var["offset"]]]) - va)) < 1e-12
    with pytest.raises(ValueError):
        eng.cov(res)
    u = np.random.default_rng(2).random(500)
    un = eng.resample(u, res, 500)
    got = un.samples
    anc = un.ancestors - 1
    assert np.array_equal(got["weights"], res.samples["weights"][anc])
    assert np.array_equal(got["rate"], res.samples["rate"][anc])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_host_batch_target_receives_named_leaves(eng, precision):
    # test/batched_targets.jl:120-133: the batch callback sees the logical named leaves
    _say(f"case: host batch target over named leaves vs the built-in target, {precision}")
    m, L = _gauss(5, 6)
    logn = -2.5 * math.log(2 * math.pi) - float(np.sum(np.log(np.diag(L))))
    seen = []

    def batch(out, xs):
        seen.append(sorted(xs))
        x = np.column_stack([xs["weights"], xs["rate"], xs["offset"]]).astype(np.float64)
        r = np.linalg.solve(L, (x - m).T)
        out[:] = logn - 0.5 * np.sum(r * r, axis=0)

    alg = eng.ImportanceSampling(eng.FactorGaussian(np.zeros(4), FACTOR), 8192)
    a = eng.importance_sample(14, eng.HostBatchTarget(batch), alg, precision=precision, transform=_layout(eng))
    b = eng.importance_sample(14, eng.GaussianTarget(m, L), alg, precision=precision, transform=_layout(eng))
    assert seen and seen[0] == ["offset", "rate", "weights"]
    assert np.array_equal(_flat(a), _flat(b))
    tol = 1e-11 if precision == "fp64" else 2e-4
    assert np.max(np.abs(a.logweights - b.logweights)) < tol


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_regression_target_over_the_logical_dimension(eng, oracle_mod, precision):
    _say(f"case: logistic target with p = dl = 5 under a layout, AMIS {precision}")
    rng = np.random.default_rng(21)
    X = rng.standard_normal((700, 5))
    y = (rng.random(700) < 0.5).astype(float)
    logi = eng.LogisticRegressionTarget(X, y, 10.0)
    q0 = eng.SphericalGaussian(np.zeros(4), 0.5)
    sched = [2048, 2048, 2048]
    s = eng.prepare_sampler(15, logi, eng.AMIS(q0, rounds=3, round_size=sched), precision=precision,
                            transform=_layout(eng))
    normals = np.array([s.export_randomness(t)["normals"] for t in range(3)])
    res = eng.importance_sample_(s)
    means, facs, lns = s.history()
    tk, tp = oracle_mod.layout_target(logi.kind, logi.params(5), ORACLE_LAYOUT)
    o = oracle_mod.adaptive(0, q0, sched, normals, tk, tp, replay=(means, facs, lns))
    assert o["rc"] == 0
    if precision == "fp64":
        assert _relerr(res.logweights, o["logweights"]) < 1e-10
    else:
        assert abs(res.diagnostics["round_lognormalizers"][-1] - o["round_lognormalizers"][-1]) < 2e-3
