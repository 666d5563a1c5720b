"""GPU parity of the batched AMIS engine (one persistent launch over many
posteriors) through the C ABI: against the CPU oracle fed the same counter-RNG
normals (replay of each job's own fitted proposals, log-weights within 1e-10
relative in FP64), against the same samplers run alone through the single-job
engine, FP32 against FP64 (|delta log Z| < 1e-3), launch-shape and job-order
invariance, the conjugate linear-regression evidence, transactional commits and
failure isolation."""

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


def linreg_problem(rng, d, n, sigma=0.5, tau=2.0):
    """Config-4 style posterior: X ~ N(0,1), beta ~ N(0, tau^2 I), y = X beta + sigma eps.
    Returns X, y and the exact conjugate posterior (mean, covariance, log evidence)."""
    X = rng.standard_normal((n, d))
    beta = rng.normal(0.0, tau, d)
    y = X @ beta + sigma * rng.standard_normal(n)
    prec = X.T @ X / sigma ** 2 + np.eye(d) / tau ** 2
    S = np.linalg.inv(prec)
    S = 0.5 * (S + S.T)
    m = S @ (X.T @ y) / sigma ** 2
    # log p(y) = log N(y; 0, sigma^2 I + tau^2 X X^T)
    C = sigma ** 2 * np.eye(n) + tau ** 2 * X @ X.T
    sign, logdet = np.linalg.slogdet(C)
    logev = -0.5 * (n * math.log(2 * math.pi) + logdet + y @ np.linalg.solve(C, y))
    return X, y, m, S, logev


def make_jobs(eng, specs, seed=5):
    """specs: (d, n, N, rounds, family) with family "t5" or "gauss"."""
    rng = np.random.default_rng(seed)
    rngs, targets, algs, exact = [], [], [], []
    for k, (d, n, N, T, fam) in enumerate(specs):
        X, y, m, S, logev = linreg_problem(rng, d, n)
        L = np.linalg.cholesky(S)
        loc = m + 0.3 * (L @ rng.standard_normal(d))
        if fam == "t5":
            q0 = eng.FactorStudentT(5.0, loc, 1.6 * L)
        else:
            q0 = eng.FactorGaussian(loc, 1.4 * L)
        rngs.append(1000 + 17 * k)
        targets.append(eng.LinearRegressionTarget(X, y, 0.5, 2.0))
        algs.append(eng.AMIS(q0, rounds=T, round_size=N))
        exact.append((m, S, logev))
    return rngs, targets, algs, exact


SPECS = [(2, 20, 300, 3, "t5"), (5, 37, 256, 4, "t5"), (8, 100, 333, 3, "gauss"), (13, 150, 200, 3, "t5"),
         (17, 64, 150, 3, "t5"), (32, 300, 128, 3, "t5"), (3, 2000, 130, 2, "t5")]


def _export(batch, job):
    sched = batch.schedules[job]
    d, N = batch.dims[job], sched[0]
    normals = np.zeros((len(sched), d * N))
    rn = ru = None
    for t in range(len(sched)):
        r = batch.export_randomness(job, t)
        normals[t] = r["normals"]
        if "radial_normals" in r:
            if rn is None:
                rn = np.zeros((len(sched), r["radial_normals"].size))
                ru = np.zeros((len(sched), r["radial_uniforms"].size))
            rn[t] = r["radial_normals"]
            ru[t] = r["radial_uniforms"]
    return normals, rn, ru


def test_batch_fp64_matches_oracle_replay(eng, oracle_mod):
    _say("case: batch FP64 vs oracle replay, 7 jobs d in {2,3,5,8,13,17,32}")
    rngs, targets, algs, _ = make_jobs(eng, SPECS)
    b = eng.prepare_batch(rngs, targets, algs, precision="fp64")
    exported = [_export(b, j) for j in range(b.njobs)]
    st = b.run()
    assert not st.any(), st
    for j, alg in enumerate(algs):
        normals, rn, ru = exported[j]
        res = b.result(j)
        means, facs, lns = b.history(j)
        o = oracle_mod.adaptive(0, alg.proposal, alg.schedule, normals, targets[j].kind, targets[j].params(b.dims[j]),
                                rn, ru, nthreads=4, replay=(means, facs, lns))
        assert o["rc"] == 0
        assert relerr(res.samples, o["samples"]) < 1e-12, j
        assert relerr(res.logweights, o["logweights"]) < RTOL, j
        assert relerr(res.diagnostics["round_ess"], o["round_ess"]) < 1e-10, j
        assert relerr(res.diagnostics["round_lognormalizers"], o["round_lognormalizers"]) < RTOL, j
        # each round's FP64 fit (mean, factor, log normalizer) from the same weighted samples
        assert relerr(means[1:], o["fit_means"][1:]) < 1e-9, j
        assert relerr(facs[1:], o["fit_factors"][1:]) < 1e-9, j
        assert relerr(lns[1:], o["fit_lognormalizers"][1:]) < 1e-9, j
        assert eng.lognormalizer(res) == res.diagnostics["round_lognormalizers"][-1]


def test_batch_matches_single_job_engine(eng):
    _say("case: batch FP64 vs the same samplers run alone")
    rngs, targets, algs, _ = make_jobs(eng, SPECS[:5], seed=9)
    b = eng.prepare_batch(rngs, targets, algs, precision="fp64")
    assert not b.run().any()
    for j in range(b.njobs):
        s = eng.prepare_sampler(rngs[j], targets[j], algs[j], precision="fp64")
        alone = eng.importance_sample_(s)
        res = b.result(j)
        assert relerr(res.samples, alone.samples) < 1e-10, j
        assert relerr(res.logweights, alone.logweights) < 1e-9, j
        assert relerr(res.diagnostics["round_lognormalizers"], alone.diagnostics["round_lognormalizers"]) < 1e-9, j


def test_batch_fp32_matches_fp64(eng):
    _say("case: batch FP32 vs FP64 log Z and moments")
    specs = [(d, n, 1024, 4, "t5") for d, n in ((2, 40), (7, 200), (16, 500), (24, 900), (32, 1500))]
    rngs, targets, algs, exact = make_jobs(eng, specs, seed=21)
    b64 = eng.prepare_batch(rngs, targets, algs, precision="fp64")
    b32 = eng.prepare_batch(rngs, targets, algs, precision="fp32")
    assert not b64.run().any()
    assert not b32.run().any()
    lz64, lz32 = b64.lognormalizers(), b32.lognormalizers()
    assert np.max(np.abs(lz32 - lz64)) < 1e-3, (lz32, lz64)
    for j in range(b64.njobs):
        r64, r32 = b64.result(j), b32.result(j)
        w64 = np.exp(r64.logweights - r64.logweights.max())
        w32 = np.exp(r32.logweights - r32.logweights.max())
        m64 = (w64[:, None] * r64.samples).sum(0) / w64.sum()
        m32 = (w32[:, None] * r32.samples).sum(0) / w32.sum()
        assert np.max(np.abs(m64 - m32)) < 1e-3, j


def test_batch_launch_shape_and_order_invariance(eng):
    _say("case: batch results under waves 2/4/8, workgroups per CU and job order")
    rngs, targets, algs, _ = make_jobs(eng, SPECS[:6], seed=33)
    b = eng.prepare_batch(rngs, targets, algs, precision="fp64")
    ref = {}
    rev = np.arange(b.njobs)[::-1]
    # FP64 shapes whose LDS fits: (waves, workgroups per CU, order)
    for w, bpc, order in ((2, 1, None), (2, 3, rev), (4, 1, None), (4, 1, rev), (8, 1, None), (8, 1, rev)):
        b.set_order(order)
        b.set_tuning(w, bpc)
        b.trial(w, bpc)  # no commit: every configuration draws the same randomness
        lw = [b.result(j).logweights for j in range(b.njobs)]
        if w not in ref:
            ref[w] = lw
        else:  # same waves: the same reduction order, bit-identical
            for j in range(b.njobs):
                assert np.array_equal(lw[j], ref[w][j]), (w, bpc, j)
    for j in range(b.njobs):
        assert relerr(ref[2][j], ref[8][j]) < 1e-10


def test_batch_exact_posterior_proposal_gives_log_evidence(eng):
    _say("case: batch linear regression, exact posterior as proposal")
    rng = np.random.default_rng(4)
    rngs, targets, algs, evid = [], [], [], []
    for k, (d, n) in enumerate(((3, 50), (10, 400), (29, 1200))):
        X, y, m, S, logev = linreg_problem(rng, d, n)
        rngs.append(77 + k)
        targets.append(eng.LinearRegressionTarget(X, y, 0.5, 2.0))
        algs.append(eng.AMIS(eng.FactorGaussian(m, np.linalg.cholesky(S)), rounds=1, round_size=256))
        evid.append(logev)
    b = eng.prepare_batch(rngs, targets, algs, precision="fp64")
    assert not b.run().any()
    for j in range(b.njobs):
        lw = b.result(j).logweights  # one round: pi(x) / q0(x) = p(y) for every sample
        assert np.max(np.abs(lw - evid[j])) < 1e-7 * max(1.0, abs(evid[j])), (j, lw[:3], evid[j])
        assert abs(b.round_summary(j)[1][0] - evid[j]) < 1e-7 * max(1.0, abs(evid[j]))


def test_batch_repeated_runs_commit_and_advance(eng):
    _say("case: batch repeated runs start from the committed proposal")
    rngs, targets, algs, _ = make_jobs(eng, SPECS[:3], seed=41)
    b = eng.prepare_batch(rngs, targets, algs, precision="fp64")
    assert not b.run().any()
    last = [b.history(j) for j in range(b.njobs)]
    lw1 = b.result(0).logweights
    assert not b.run().any()
    for j in range(b.njobs):
        m, f, _ = b.history(j)
        assert np.array_equal(m[0], last[j][0][-1])
        assert np.allclose(f[0], last[j][1][-1], rtol=0, atol=0)
        cp = b.current_proposal(j)
        assert np.array_equal(cp.location, m[-1])
    assert not np.array_equal(b.result(0).logweights, lw1)  # new counter-RNG run index


def test_batch_failure_is_isolated(eng):
    _say("case: batch with one failing job (all weights zero)")
    rngs, targets, algs, _ = make_jobs(eng, SPECS[:4], seed=55)
    good = eng.prepare_batch(rngs, targets, algs, precision="fp64")
    assert not good.run().any()
    d = 4
    X = np.ones((10, d))
    y = np.zeros(10)
    bad_alg = eng.AMIS(eng.SphericalGaussian(np.zeros(d), 1e200), rounds=2, round_size=128)
    b = eng.prepare_batch(rngs[:2] + [5] + rngs[2:], targets[:2] + [eng.LinearRegressionTarget(X, y)] + targets[2:],
                          algs[:2] + [bad_alg] + algs[2:], precision="fp64")
    with pytest.raises(eng.BatchRoundErrors) as ei:
        eng.importance_sample_batch_(b)
    err = ei.value.errors
    assert list(err) == [2]
    assert err[2].round == 1 and err[2].phase == "moment"
    assert isinstance(err[2].cause, eng.AllZeroWeightsError)
    # the failed job keeps its initial proposal; the other jobs match the batch without it
    cp = b.current_proposal(2)
    assert np.array_equal(cp.location, np.zeros(d))
    for jb, jg in ((0, 0), (1, 1), (3, 2), (4, 3)):
        assert np.array_equal(b.result(jb).logweights, good.result(jg).logweights), jb


def _config4_jobs(eng, njobs, seed=4096, T=8, N=4096):
    """Config-4 posteriors: n ~ U{20..2000}, d ~ U{2..32}, X ~ N(0,1), sigma = 0.5, prior
    N(0, 4 I); AMIS 8 x 4,096 with Student-t (nu = 5) proposals from the conjugate
    posterior with the scale inflated 1.5x."""
    rng = np.random.default_rng(seed)
    rngs, targets, algs = [], [], []
    for j in range(njobs):
        n = int(rng.integers(20, 2001))
        d = int(rng.integers(2, 33))
        X = rng.standard_normal((n, d))
        y = X @ rng.normal(0.0, 2.0, d) + 0.5 * rng.standard_normal(n)
        S = np.linalg.inv(X.T @ X / 0.25 + np.eye(d) / 4.0)
        m = S @ (X.T @ y) / 0.25
        L = np.linalg.cholesky(0.5 * (S + S.T))
        rngs.append(1 + j)
        targets.append(eng.LinearRegressionTarget(X, y, 0.5, 2.0))
        algs.append(eng.AMIS(eng.FactorStudentT(5.0, m, 1.5 * L), rounds=T, round_size=N))
    return rngs, targets, algs


def test_batch_config4_512_posteriors_autotuned(eng, oracle_mod):
    """512 config-4 posteriors in one persistent launch whose shape the autotuner picks;
    two jobs of every padded dimension P = 4..32 replayed by the oracle (FP64,
    log-weights within 1e-10), and the tuned result bit-identical to a run of the same
    shape in another job order."""
    _say("case: batch config 4, 512 posteriors, autotuned, oracle replay per P bucket")
    rngs, targets, algs = _config4_jobs(eng, 512)
    b = eng.prepare_batch(rngs, targets, algs, precision="fp64")
    buckets = {}
    for j, d in enumerate(b.dims):
        buckets.setdefault((d + 3) // 4, [])
        if len(buckets[(d + 3) // 4]) < 2:
            buckets[(d + 3) // 4].append(j)
    assert sorted(buckets) == list(range(1, 9))
    picked = [j for js in buckets.values() for j in js]
    exported = {j: _export(b, j) for j in picked}
    choice = eng.autotune(b, candidates=[(4, 1, "lpt"), (8, 1, "lpt"), (4, 2, "input")], repeats=1)
    assert choice["waves"] in (4, 8) and not choice["cached"]
    st = b.run()
    assert not st.any(), np.nonzero(st)
    for j in picked:
        normals, rn, ru = exported[j]
        res = b.result(j)
        means, facs, lns = b.history(j)
        o = oracle_mod.adaptive(0, algs[j].proposal, algs[j].schedule, normals, targets[j].kind,
                                targets[j].params(b.dims[j]), rn, ru, nthreads=16, replay=(means, facs, lns))
        assert o["rc"] == 0, j
        assert relerr(res.logweights, o["logweights"]) < RTOL, (j, b.dims[j])
        assert relerr(res.diagnostics["round_lognormalizers"], o["round_lognormalizers"]) < RTOL, j
        assert relerr(means[1:], o["fit_means"][1:]) < 1e-9, j
    # the same batch in the reverse job order at the tuned shape: bit-identical (the
    # FP64 kernel fits one workgroup per CU at 4 and 8 waves)
    b2 = eng.prepare_batch(rngs, targets, algs, precision="fp64")
    b2.set_order(np.arange(b2.njobs)[::-1].copy())
    b2.set_tuning(choice["waves"], choice["blocks_per_cu"])
    assert not b2.run().any()
    for j in picked + [0, 511]:
        assert np.array_equal(b2.result(j).logweights, b.result(j).logweights), j


def test_batch_fp32_fits_rank_deficient_rounds(eng):
    """A wide initial proposal N(0, 9 I) leaves the first rounds with an effective sample
    size below d, so the weighted covariance is nearly rank-deficient and only its ridge
    keeps it positive definite.  The moment tiles are Gram matrices of the staged
    products with exact FP64 products, so FP32 jobs fit every job that FP64 fits."""
    _say("case: batch FP32 vs FP64 fits from a wide initial proposal, 64 jobs")
    rng = np.random.default_rng(5)
    rngs, targets, algs = [], [], []
    for j in range(64):
        n = int(rng.integers(50, 1000))
        d = int(rng.integers(4, 33))
        X = rng.standard_normal((n, d))
        y = X @ rng.normal(0, 2, d) + 0.5 * rng.standard_normal(n)
        rngs.append(100 + j)
        targets.append(eng.LinearRegressionTarget(X, y, 0.5, 2.0))
        algs.append(eng.AMIS(eng.FactorGaussian(np.zeros(d), 3.0 * np.eye(d)), rounds=6, round_size=2048))
    st = {p: eng.prepare_batch(rngs, targets, algs, precision=p).run() for p in ("fp64", "fp32")}
    only32 = np.nonzero((st["fp32"] != 0) & (st["fp64"] == 0))[0]
    assert only32.size == 0, only32
