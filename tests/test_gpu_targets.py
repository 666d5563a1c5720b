"""GPU parity of the batched regression targets (logistic and linear regression,
eta = X beta on MFMA) through the C ABI, against the oracle fed the same
counter-RNG normals.  Tolerances: log-weights 1e-10 relative with a 1.0 floor
(FP64); FP32 mode |delta log Z| < 1e-3 against FP64."""

import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b, floor=1.0):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(floor, np.abs(b)))) if a.size else 0.0


@pytest.fixture(scope="module")
def eng(pkg):
    pkg.context(0)
    return pkg


def _logistic_data(rng, n, p, scale=1.0):
    X = np.column_stack([rng.standard_normal((n, p - 1)) * scale, np.ones(n)])
    beta = rng.normal(0, 1 / math.sqrt(p), p)
    y = (rng.uniform(size=n) < 1 / (1 + np.exp(-X @ beta))).astype(float)
    return X, y, beta


def _linreg_data(rng, n, p, sigma=0.5):
    X = rng.standard_normal((n, p))
    beta = rng.normal(0, 1, p)
    y = X @ beta + sigma * rng.standard_normal(n)
    return X, y, beta


def _linreg_posterior(X, y, sigma, tau):
    p = X.shape[1]
    prec = X.T @ X / sigma**2 + np.eye(p) / tau**2
    S = np.linalg.inv(prec)
    m = S @ X.T @ y / sigma**2
    n = X.shape[0]
    C = sigma**2 * np.eye(n) + tau**2 * X @ X.T
    sign, logdet = np.linalg.slogdet(C)
    logev = -0.5 * (n * math.log(2 * math.pi) + logdet + y @ np.linalg.solve(C, y))
    return m, S, logev


@pytest.mark.parametrize("kind,n,p,N", [
    ("logistic", 777, 5, 3000),
    ("logistic", 64, 3, 100),
    ("logistic", 5000, 12, 700),     # several row slices
    ("logistic", 300, 33, 1000),     # 32 < p <= 64 bucket
    ("linreg", 513, 4, 2500),
    ("linreg", 2000, 20, 129),
])
def test_plain_is_regression_parity(eng, oracle_mod, kind, n, p, N):
    _say(f"case: plain IS {kind} n={n} p={p} N={N}")
    rng = np.random.default_rng(n + p)
    if kind == "logistic":
        X, y, beta = _logistic_data(rng, n, p)
        tgt = eng.LogisticRegressionTarget(X, y, prior_var=10.0)
    else:
        X, y, beta = _linreg_data(rng, n, p)
        tgt = eng.LinearRegressionTarget(X, y, sigma=0.5, tau=2.0)
    q = eng.DiagonalGaussian(beta + 0.05, np.full(p, 0.1))
    s = eng.prepare_sampler(99, tgt, eng.ImportanceSampling(q, N))
    rnd = s.export_randomness(0)
    res = eng.importance_sample_(s)
    x, lw, fail = oracle_mod.plain_is(q, tgt.kind, tgt.params(p), rnd["normals"], nthreads=8)
    assert fail[0] == 0
    assert relerr(res.samples.reshape(N, -1), x, 1.0) < 1e-13
    assert relerr(res.logweights, lw, 1.0) < RTOL
    assert abs(eng.lognormalizer(res) - oracle_mod.lognormalizer(lw)) < 1e-9 * max(1.0, abs(oracle_mod.lognormalizer(lw)))


def test_linreg_exact_posterior_gives_log_evidence(eng):
    """Proposal = the conjugate posterior, so every log-weight is log p(y)."""
    _say("case: linreg exact posterior")
    rng = np.random.default_rng(5)
    n, p, sigma, tau = 400, 6, 0.5, 2.0
    X, y, _ = _linreg_data(rng, n, p, sigma)
    m, S, logev = _linreg_posterior(X, y, sigma, tau)
    q = eng.FactorGaussian(m, np.linalg.cholesky(S))
    tgt = eng.LinearRegressionTarget(X, y, sigma=sigma, tau=tau)
    res = eng.importance_sample(3, tgt, eng.ImportanceSampling(q, 2048))
    assert np.max(np.abs(res.logweights - logev)) < 1e-8 * abs(logev)
    assert abs(eng.lognormalizer(res) - logev) < 1e-8 * abs(logev)


def test_amis_logistic_replay_parity(eng, oracle_mod):
    _say("case: AMIS logistic replay")
    rng = np.random.default_rng(11)
    n, p = 500, 4
    X, y, beta = _logistic_data(rng, n, p)
    tgt = eng.LogisticRegressionTarget(X, y, prior_var=10.0)
    q = eng.FactorGaussian(np.zeros(p), 1.5 * np.eye(p))
    schedule = [1000, 1200, 800]
    s = eng.prepare_sampler(8, tgt, eng.AMIS(q, rounds=3, round_size=schedule))
    maxn = max(schedule)
    normals = np.zeros((3, p * maxn))
    for t, nt in enumerate(schedule):
        normals[t, : p * nt] = s.export_randomness(t)["normals"]
    res = eng.importance_sample_(s)
    means, facs, lns = s.history()
    o = oracle_mod.adaptive(0, q, schedule, normals, tgt.kind, tgt.params(p), nthreads=8, replay=(means, facs, lns))
    assert o["rc"] == 0
    assert relerr(res.logweights, o["logweights"], 1.0) < RTOL
    assert relerr(res.diagnostics["round_lognormalizers"], o["round_lognormalizers"], 1.0) < RTOL
    assert relerr(means[1:], o["fit_means"][1:], 1.0) < 1e-9


@pytest.mark.parametrize("method", ["multinomial", "systematic"])
def test_dmpmc_logistic_parity(eng, oracle_mod, method):
    _say(f"case: DM-PMC logistic {method}")
    rng = np.random.default_rng(21)
    n, p, K = 300, 3, 8
    X, y, beta = _logistic_data(rng, n, p)
    tgt = eng.LogisticRegressionTarget(X, y, prior_var=10.0)
    props = [eng.SphericalGaussian(beta + rng.normal(0, 0.3, p), 0.4) for _ in range(K)]
    bank = eng.ProposalBank(props, np.ones(K))
    meth = eng.SystematicResampling() if method == "systematic" else eng.MultinomialResampling()
    schedule = [512, 512, 512]
    alg = eng.DeterministicMixturePMC(bank, rounds=3, round_size=schedule, method=meth)
    s = eng.prepare_sampler(4, tgt, alg)
    normals = np.zeros((3, p * 512))
    for t in range(3):
        normals[t] = s.export_randomness(t)["normals"]
    uniforms = np.array([s.export_randomness(t)["uniforms"] for t in range(3)])
    res = eng.importance_sample_(s)
    order = oracle_mod.active_order(list(bank.masses))
    counts, logc = oracle_mod.allocation_plan([bank.masses[i] for i in order], schedule)
    P = [bank.proposals[i] for i in order]
    o = oracle_mod.dmpmc(0, 0.0, 0, p, np.array([q.location for q in P]), np.array([q.scale[0] for q in P]),
                         [q.lognormalizer for q in P], [i + 1 for i in order], schedule, counts, logc, normals,
                         uniforms, tgt.kind, tgt.params(p), resampling=3 if method == "systematic" else 2)
    assert o["rc"] == 0
    assert relerr(res.logweights, o["logweights"], 1.0) < RTOL
    assert np.array_equal(s.ancestors(), o["ancestors"])


@pytest.mark.parametrize("method", ["plain", "amis"])
def test_logistic_fp32_matches_fp64_log_z(eng, method):
    """FP32 storage and FP32-MFMA eta against the FP64 run on the same normals, for a
    proposal on the posterior's scale (posterior sd ~ 0.03 at n = 4000)."""
    _say(f"case: logistic FP32 vs FP64 {method}")
    rng = np.random.default_rng(31)
    n, p = 4000, 8
    X, y, beta = _logistic_data(rng, n, p)
    tgt = eng.LogisticRegressionTarget(X, y, prior_var=10.0)
    q = eng.DiagonalGaussian(beta, np.full(p, 0.08))
    if method == "plain":
        alg = eng.ImportanceSampling(q, 20000)
    else:
        alg = eng.AMIS(eng.FactorGaussian(beta, 0.08 * np.eye(p)), rounds=4, round_size=5000)
    r64 = eng.importance_sample(2, tgt, alg, precision="fp64")
    r32 = eng.importance_sample(2, tgt, alg, precision="fp32")
    dlz = abs(eng.lognormalizer(r32) - eng.lognormalizer(r64))
    dm = np.max(np.abs(eng.mean(r32) - eng.mean(r64)))
    _say(f"  |dlogZ| = {dlz:.3g}, |dmean| = {dm:.3g}")
    assert dlz < 1e-3
    assert dm < 1e-3


def test_logistic_config3_shape_sampled_parity(eng, oracle_mod):
    """The config-3 data shape (100,000 x 51, one DM-PMC round of 64 x 512 samples),
    checked on a sample of 48 log-weights against the oracle's target and proposal
    densities at the device's own draws."""
    _say("case: logistic config-3 shape sampled parity")
    rng = np.random.default_rng(3)
    n, p, K = 100_000, 51, 64
    X, y, beta = _logistic_data(rng, n, p)
    tgt = eng.LogisticRegressionTarget(X, y, prior_var=10.0)
    props = [eng.SphericalGaussian(beta + rng.normal(0, 0.01, p), 0.01) for _ in range(K)]
    bank = eng.ProposalBank(props, np.ones(K))
    alg = eng.DeterministicMixturePMC(bank, rounds=1, round_size=[K * 512], method=eng.SystematicResampling())
    res = eng.importance_sample(1, tgt, alg)
    xs = res.samples.reshape(len(res), -1)
    lw = res.logweights
    assert np.all(np.isfinite(lw))
    params = tgt.params(p)
    idx = rng.choice(len(res), 48, replace=False)
    for i in idx:
        tv = oracle_mod.target(2, params, xs[i])
        dens = [oracle_mod.proposal_logdensity(q, xs[i]) for q in props]
        mix = np.logaddexp.reduce(dens) - math.log(K)
        assert abs(lw[i] - (tv - mix)) < 1e-9 * max(1.0, abs(tv))


def test_logistic_fp32_log_targets_unbiased_at_config3_shape(eng):
    """FP32 log-weights of 100,000-row logistic targets against FP64 at the same draws.
    Each log target sums 1e5 FP32 terms: a per-row bias of 1e-9 (a tenth of an FP32 ulp,
    the size of the bias of a float log1p or the hardware logarithm) would shift every
    log-weight by 1e-4, so the mean shift is held to 3e-5 (the per-row rounding noise of
    a sum of 1e5 terms is ~1e-5 per sample and averages out over the samples)."""
    _say("case: logistic FP32 vs FP64 log-weights, 100,000 x 51, mean shift")
    rng = np.random.default_rng(8)
    n, p = 100_000, 51
    X, y, beta = _logistic_data(rng, n, p)
    tgt = eng.LogisticRegressionTarget(X, y, prior_var=10.0)
    alg = eng.ImportanceSampling(eng.SphericalGaussian(beta, 0.003), 4096)
    r64 = eng.importance_sample(5, tgt, alg, precision="fp64")
    r32 = eng.importance_sample(5, tgt, alg, precision="fp32")
    x64 = r64.samples.reshape(len(r64), -1)
    x32 = r32.samples.reshape(len(r32), -1)
    assert np.max(np.abs(x32 - x64)) < 1e-6  # the same counter-RNG draws
    d = r32.logweights - r64.logweights
    assert abs(float(np.mean(d))) < 3e-5, float(np.mean(d))
    assert float(np.max(np.abs(d))) < 2e-3, float(np.max(np.abs(d)))
    assert abs(eng.lognormalizer(r32) - eng.lognormalizer(r64)) < 1e-4


def test_logistic_fp32_centred_bf16_far_from_the_centre(eng):
    """FP32 logistic log-weights (the centred BF16 form, p = 51) at draws about ten
    posterior standard deviations from the centre, against the FP64 run on the same
    draws.  The three-product split drops terms of about 2^-16 |x_i| |beta - x0| per
    row; over 1e5 rows those stay zero-mean, so the mean shift is held to 1e-4 and each
    log-weight to 1e-2."""
    _say("case: logistic FP32 centred BF16, 100,000 x 51, draws far from the centre")
    rng = np.random.default_rng(12)
    n, p = 100_000, 51
    X, y, beta = _logistic_data(rng, n, p)
    tgt = eng.LogisticRegressionTarget(X, y, prior_var=10.0)
    alg = eng.ImportanceSampling(eng.SphericalGaussian(beta, 0.05), 4096)
    r64 = eng.importance_sample(6, tgt, alg, precision="fp64")
    r32 = eng.importance_sample(6, tgt, alg, precision="fp32")
    d = r32.logweights - r64.logweights
    _say(f"  mean shift {float(np.mean(d)):.3g}, max {float(np.max(np.abs(d))):.3g}")
    assert np.all(np.isfinite(d))
    assert abs(float(np.mean(d))) < 1e-4, float(np.mean(d))
    assert float(np.max(np.abs(d))) < 1e-2, float(np.max(np.abs(d)))



@pytest.mark.parametrize("p,K,kind", [(20, 16, "gaussian"), (51, 64, "gaussian"), (33, 12, "student")])
def test_fp32_bank_mixture_on_matrix_cores(eng, oracle_mod, p, K, kind):
    """FP32 DM-PMC with a logistic target (samples drawn before the mixture pass): the
    bank's mixture density runs on the FP32 matrix cores (inverse factors, whitened
    coordinates as MFMA products).  Against the FP64 oracle's proposal densities and
    target at the device's own draws, on 64 sampled log-weights: the FP32 log target
    over 2,000 rows and the FP32 whitening stay within 2e-3."""
    _say(f"case: FP32 DM-PMC mixture on the matrix cores, p = {p}, K = {K}, {kind}")
    rng = np.random.default_rng(40 + p)
    n = 2000
    X, y, beta = _logistic_data(rng, n, p)
    tgt = eng.LogisticRegressionTarget(X, y, prior_var=10.0)
    if kind == "student":
        props = [eng.FactorStudentT(5.0, beta + rng.normal(0, 0.03, p),
                                    np.tril(rng.normal(0, 0.004, (p, p)), -1) + np.diag(rng.uniform(0.02, 0.05, p)))
                 for _ in range(K)]
    else:
        props = [eng.FactorGaussian(beta + rng.normal(0, 0.03, p),
                                    np.tril(rng.normal(0, 0.004, (p, p)), -1) + np.diag(rng.uniform(0.02, 0.05, p)))
                 for _ in range(K)]
    bank = eng.ProposalBank(props, np.ones(K))  # 64 draws each: mixture weights exactly 1/K
    alg = eng.DeterministicMixturePMC(bank, rounds=1, round_size=[K * 64], method=eng.SystematicResampling())
    res = eng.importance_sample(9, tgt, alg, precision="fp32")
    xs = res.samples.reshape(len(res), -1).astype(np.float64)
    lw = res.logweights
    assert np.all(np.isfinite(lw))
    params = tgt.params(p)
    logm = -math.log(K)
    idx = rng.choice(len(res), 64, replace=False)
    for i in idx:
        tv = oracle_mod.target(2, params, xs[i])
        dens = [oracle_mod.proposal_logdensity(q, xs[i]) for q in props]
        mix = np.logaddexp.reduce(np.array(dens) + logm)
        assert abs(lw[i] - (tv - mix)) < 2e-3, (i, lw[i], tv - mix)


def test_config3_fp32_gate_over_30_replayed_rounds(eng):
    """Config 3 at its full shape (X 100,000 x 51, 64 Laplace copies x 512 samples, 30
    rounds, systematic resampling): the FP32 form (centred BF16 target, MFMA mixture,
    compensated sums) against an FP64 run that replays its ancestors, so both weigh
    samples of the same banks: every round's log Z within 1e-3, and the pooled log Z."""
    _say("case: config 3, FP32 vs FP64 replay, 30 rounds")
    rng = np.random.default_rng(51)
    n, p = 100_000, 51
    X = np.hstack([rng.standard_normal((n, p - 1)), np.ones((n, 1))])
    beta = rng.normal(0.0, math.sqrt(1.0 / 50.0), p)
    y = (rng.random(n) < 1.0 / (1.0 + np.exp(-X @ beta))).astype(float)
    b = np.zeros(p)
    for _ in range(25):  # MAP by Newton
        mu = 1.0 / (1.0 + np.exp(-X @ b))
        H = (X * (mu * (1 - mu))[:, None]).T @ X + np.eye(p) / 10.0
        step = np.linalg.solve(H, X.T @ (y - mu) - b / 10.0)
        b = b + step
        if np.max(np.abs(step)) < 1e-10:
            break
    L = np.linalg.cholesky(np.linalg.inv(H))
    bank = eng.ProposalBank([eng.FactorGaussian(b, L) for _ in range(64)], np.ones(64))
    alg = eng.DeterministicMixturePMC(bank, rounds=30, round_size=64 * 512, method=eng.SystematicResampling())
    tgt = eng.LogisticRegressionTarget(X, y, 10.0)
    s32 = eng.prepare_sampler(3, tgt, alg, precision="fp32")
    r32 = eng.importance_sample_(s32)
    s64 = eng.prepare_sampler(3, tgt, alg, precision="fp64")
    s64.set_ancestors(s32.ancestors())
    r64 = eng.importance_sample_(s64)
    assert np.array_equal(s64.ancestors(), s32.ancestors())
    dlz = np.abs(r32.diagnostics["round_lognormalizers"] - r64.diagnostics["round_lognormalizers"])
    assert len(dlz) == 30 and np.all(np.isfinite(dlz))
    assert np.max(dlz) < 1e-3, np.max(dlz)
    assert abs(eng.lognormalizer(r32) - eng.lognormalizer(r64)) < 1e-3
