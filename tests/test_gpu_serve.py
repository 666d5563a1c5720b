"""GPU tests of the serving loop (continuous batching): a concurrent mix of
plain IS, AMIS, NPMC, DM-PMC and static-MIS requests gives, for every request, the same
results as the request run alone (bit-identical log-weights and log Z);
admission keeps the admitted jobs within the memory budget; oversized
requests are refused; a failing request raises its own error while the rest of
the mix completes."""

import ctypes
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


def _gauss_target(eng, rng, d):
    A = rng.standard_normal((d, d))
    S = A @ A.T / d + np.eye(d)
    return eng.GaussianTarget(rng.normal(0, 1, d), np.linalg.cholesky(S))


def _logistic(eng, rng, d, n):
    X = rng.standard_normal((n, d))
    beta = rng.normal(0, 1 / np.sqrt(d), d)
    y = (rng.random(n) < 1 / (1 + np.exp(-X @ beta))).astype(float)
    return eng.LogisticRegressionTarget(X, y, 10.0)


def request_mix(eng, seed=3):
    rng = np.random.default_rng(seed)
    reqs = []
    reqs.append((11, eng.StandardGaussianTarget(), eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(3), 1.5), 20000),
                 "fp64"))
    reqs.append((12, _gauss_target(eng, rng, 10), eng.AMIS(eng.FactorGaussian(np.zeros(10), 4 * np.eye(10)), 8, 5000),
                 "fp32"))
    reqs.append((13, _gauss_target(eng, rng, 5), eng.NPMC(eng.DiagonalStudentT(5.0, np.zeros(5), np.full(5, 2.0)), 6, 4000),
                 "fp64"))
    reqs.append((14, _logistic(eng, rng, 6, 500), eng.AMIS(eng.SphericalGaussian(np.zeros(6), 1.0), 5, 3000), "fp32"))
    reqs.append((15, _logistic(eng, rng, 12, 1000),
                 eng.ImportanceSampling(eng.SphericalStudentT(5.0, np.zeros(12), 0.5), 8192), "fp64"))
    bank = eng.ProposalBank([eng.SphericalGaussian(rng.normal(0, 2, 4), 1.0) for _ in range(8)])
    reqs.append((16, _gauss_target(eng, rng, 4), eng.DeterministicMixturePMC(bank, 4, 2048), "fp64"))
    reqs.append((17, _gauss_target(eng, rng, 24), eng.AMIS(eng.SphericalStudentT(6.0, np.zeros(24), 3.0), 6, 4096),
                 "fp32"))
    reqs.append((18, eng.StandardGaussianTarget(), eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(40), 1.2), 50000),
                 "fp32"))
    mis_bank = eng.ProposalBank([eng.SphericalGaussian(rng.normal(0, 1, 6), 1.3) for _ in range(4)], [1, 2, 0, 3])
    reqs.append((19, _gauss_target(eng, rng, 6),
                 eng.ImportanceSampling(mis_bank, 6000, eng.PartialDeterministicMixture(((1, 3), (2, 4)))), "fp64"))
    return reqs


def test_serve_results_same_alone_and_in_mix(eng):
    _say("case: serving mix of 9 requests (plain IS, AMIS, NPMC, DM-PMC, static MIS; FP32 and FP64), 4 at once")
    reqs = request_mix(eng)
    srv = eng.Server(memory_budget=64 << 30, max_active=4)
    tickets = [srv.submit(r, t, a, p) for r, t, a, p in reqs]
    mixed = [srv.result(tk, timeout=120) for tk in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert 1 <= st["peak_active"] <= 4
    for (r, t, a, p), (res, timing) in zip(reqs, mixed):
        alone = eng.importance_sample(r, t, a, precision=p)
        assert np.array_equal(res.logweights, alone.logweights), type(a).__name__
        assert np.array_equal(res.samples, alone.samples)
        assert eng.lognormalizer(res) == eng.lognormalizer(alone)
        assert timing["latency_s"] >= timing["queue_s"] >= 0.0


def test_serve_admission_respects_memory_budget(eng):
    _say("case: serving admission against a small memory budget")
    reqs = request_mix(eng, seed=5)
    lib = eng._lib.load()
    from importancesamplers_jl_amd.api import _job_desc, FP32, FP64

    sizes = []
    for r, t, a, p in reqs:
        desc, keep, _, _ = _job_desc(r, t, a, FP32 if p == "fp32" else FP64)
        sizes.append(int(lib.is_job_device_bytes(ctypes.byref(desc))))
    assert min(sizes) > 0
    budget = int(1.2 * max(sizes))
    srv = eng.Server(memory_budget=budget, max_active=8)
    tickets = [srv.submit(r, t, a, p) for r, t, a, p in reqs]
    for tk in tickets:  # taking a result releases its memory for the queued requests
        srv.result(tk, timeout=120)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= budget
    assert st["bytes_in_use"] == 0


def test_serve_refuses_request_over_budget(eng):
    _say("case: serving refuses a request larger than the budget")
    srv = eng.Server(memory_budget=1 << 20, max_active=2)
    with pytest.raises(MemoryError):
        srv.submit(1, eng.StandardGaussianTarget(), eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(8), 1.0),
                                                                          1_000_000))
    srv.close()


def test_serve_failing_request_in_mix(eng):
    _say("case: serving mix with one failing AMIS request")
    reqs = request_mix(eng, seed=7)[:4]
    bad = (99, eng.GaussianTarget(np.zeros(3), np.eye(3)),
           eng.AMIS(eng.SphericalGaussian(np.zeros(3), 1e200), 3, 256), "fp64")
    reqs.insert(2, bad)
    srv = eng.Server(memory_budget=64 << 30, max_active=3)
    tickets = [srv.submit(r, t, a, p) for r, t, a, p in reqs]
    for i, tk in enumerate(tickets):
        if i == 2:
            with pytest.raises(eng.AMISRoundError) as ei:
                srv.result(tk, timeout=120)
            assert ei.value.round == 1
        else:
            res, _ = srv.result(tk, timeout=120)
            r, t, a, p = reqs[i]
            assert np.array_equal(res.logweights, eng.importance_sample(r, t, a, precision=p).logweights)
    st = srv.stats()
    srv.close()
    assert st["failed"] == 1 and st["completed"] == len(reqs) - 1


def test_serve_d64_million_sample_request_matches_oracle(eng, oracle_mod):
    """A served AMIS request at the largest dimension (d = 64, the k_finish fit) and
    N = 1e6 samples per round, beside two smaller requests, replayed by the FP64 oracle
    on the served job's own fitted proposals."""
    _say("case: serving an AMIS request d = 64, N = 1e6 x 2 rounds, FP64, vs oracle replay")
    d, N, T = 64, 1_000_000, 2
    rng = np.random.default_rng(64)
    F = np.tril(rng.normal(0, 0.02, (d, d)), -1) + np.diag(np.linspace(1.0, 1.3, d))
    q0 = eng.FactorGaussian(np.full(d, 0.05), F)
    tgt = eng.StandardGaussianTarget()
    alg = eng.AMIS(q0, rounds=T, round_size=N)
    ref = eng.prepare_sampler(640, tgt, alg)  # same descriptor and key: exports the normals it draws
    normals = np.zeros((T, d * N))
    for t in range(T):
        normals[t] = ref.export_randomness(t)["normals"]
    del ref
    others = request_mix(eng, seed=9)[:2]
    srv = eng.Server(memory_budget=64 << 30, max_active=3)
    tk = [srv.submit(r, t, a, p) for r, t, a, p in others[:1]]
    big = srv.submit(640, tgt, alg, "fp64")
    tk += [srv.submit(r, t, a, p) for r, t, a, p in others[1:]]
    res, timing = srv.result(big, timeout=110)
    for t in tk:
        srv.result(t, timeout=110)
    srv.close()
    means, facs, lns = res._sampler.history()
    o = oracle_mod.adaptive(0, q0, [N] * T, normals, tgt.kind, tgt.params(d), nthreads=16,
                            replay=(means, facs, lns))
    assert o["rc"] == 0
    rel = np.abs(res.logweights - o["logweights"]) / np.maximum(1.0, np.abs(o["logweights"]))
    assert np.max(rel) < 1e-10
    assert np.max(np.abs(res.diagnostics["round_lognormalizers"] - o["round_lognormalizers"])) < 1e-10
    assert np.max(np.abs(means[1:] - o["fit_means"][1:])) < 1e-9
    assert np.max(np.abs(facs[1:] - o["fit_factors"][1:])) < 1e-9
