"""GPU tests of the caller's batch target (IS_TARGET_CALLBACK, the batch!(values, samples)
contract of src/targets.jl:17-43 and src/batched_targets.jl:92-114) through the C ABI.
The callback here evaluates a Gaussian log density on the host (it synchronizes the
job's stream, copies the samples down through the library's own HIP runtime, and copies
the values up), so every result is checked against the same job with the built-in
GaussianTarget: identical draws, log-weights within rounding.  Also: NaN and untouched
values fail the round in the target phase, -Inf values are zero weights, a nonzero
return aborts the run without committing, and a callback job served beside others
gives its stand-alone result."""

import ctypes
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


class _HostGaussian:
    """Batch target log N(x; mean, L L^T) evaluated with numpy."""

    def __init__(self, mean, L, edit=None, ret=0):
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h, self.mean, self.L, self.edit, self.ret = h, np.asarray(mean, float), np.asarray(L, float), edit, ret
        d = self.mean.size
        self.lognorm = -0.5 * d * math.log(2 * math.pi) - float(np.sum(np.log(np.diag(self.L))))
        self.calls = 0

    def __call__(self, values, samples, stride, dim, count, is_float, stream):
        self.calls += 1
        assert self.h.hipStreamSynchronize(stream) == 0
        dt = np.float32 if is_float else np.float64
        sz = np.dtype(dt).itemsize
        X = np.empty((dim, count), dt)
        for c in range(dim):
            assert self.h.hipMemcpy(X[c].ctypes.data, samples + c * stride * sz, count * sz, 2) == 0
        x = X.T.astype(np.float64)
        z = np.linalg.solve(self.L, (x - self.mean).T)
        vals = (self.lognorm - 0.5 * np.sum(z * z, axis=0)).astype(dt)
        if self.edit is not None:
            self.edit(vals)
        if self.ret == 0 and self.edit is not None and getattr(self.edit, "skip", False):
            return 0  # leave the values untouched (they stay NaN)
        assert self.h.hipMemcpy(values, vals.ctypes.data, count * sz, 1) == 0
        return self.ret


def _problem(d, seed):
    rng = np.random.default_rng(seed)
    mean = rng.normal(0, 1, d)
    A = rng.normal(0, 0.4, (d, d))
    L = np.linalg.cholesky(A @ A.T + np.eye(d))
    return mean, L


def test_plain_is_callback_matches_builtin(eng):
    _say("case: batch-target callback, plain IS d = 3, FP64")
    mean, L = _problem(3, 1)
    cb = _HostGaussian(mean, L)
    alg = eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(3), 2.0), 20000)
    a = eng.importance_sample(7, eng.BatchTarget(cb), alg)
    b = eng.importance_sample(7, eng.GaussianTarget(mean, L), alg)
    assert cb.calls == 1
    assert np.array_equal(a.samples, b.samples)
    assert np.max(np.abs(a.logweights - b.logweights)) < 1e-11
    assert abs(eng.lognormalizer(a) - eng.lognormalizer(b)) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_amis_callback_matches_builtin(eng, precision):
    _say(f"case: batch-target callback, AMIS d = 4, 5 rounds, {precision}")
    mean, L = _problem(4, 2)
    cb = _HostGaussian(mean, L)
    alg = eng.AMIS(eng.SphericalGaussian(np.zeros(4), 3.0), rounds=5, round_size=4000)
    a = eng.importance_sample(3, eng.BatchTarget(cb), alg, precision=precision)
    b = eng.importance_sample(3, eng.GaussianTarget(mean, L), alg, precision=precision)
    assert cb.calls == 5
    tol = 1e-9 if precision == "fp64" else 2e-4
    assert np.max(np.abs(a.logweights - b.logweights) / np.maximum(1, np.abs(b.logweights))) < tol
    assert abs(eng.lognormalizer(a) - eng.lognormalizer(b)) < (1e-10 if precision == "fp64" else 1e-4)


def test_static_mis_callback_matches_builtin(eng):
    _say("case: batch-target callback, static MIS d = 3, K = 3")
    mean, L = _problem(3, 3)
    bank = eng.ProposalBank([eng.SphericalGaussian(np.full(3, v), 1.5) for v in (-1.0, 0.0, 1.0)], [1, 2, 1])
    alg = eng.ImportanceSampling(bank, 9000)
    a = eng.importance_sample(5, eng.BatchTarget(_HostGaussian(mean, L)), alg)
    b = eng.importance_sample(5, eng.GaussianTarget(mean, L), alg)
    assert np.array_equal(a.samples, b.samples)
    assert np.max(np.abs(a.logweights - b.logweights)) < 1e-11


def test_callback_nan_fails_in_target_phase(eng):
    _say("case: batch-target callback writes NaN at sample 17")
    mean, L = _problem(2, 4)

    def nan17(v):
        v[17] = np.nan

    alg = eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(2), 1.0), 100)
    with pytest.raises(eng.SamplerExecutionError) as ei:
        eng.importance_sample(1, eng.BatchTarget(_HostGaussian(mean, L, edit=nan17)), alg)
    assert ei.value.phase == "target" and ei.value.index == 18


def test_callback_untouched_values_fail(eng):
    _say("case: batch-target callback leaves the values untouched")
    mean, L = _problem(2, 5)

    def skip(v):
        pass

    skip.skip = True
    alg = eng.AMIS(eng.SphericalGaussian(np.zeros(2), 1.0), rounds=2, round_size=64)
    with pytest.raises(eng.AMISRoundError) as ei:
        eng.importance_sample(1, eng.BatchTarget(_HostGaussian(mean, L, edit=skip)), alg)
    assert ei.value.round == 1


def test_callback_neg_inf_is_zero_weight(eng):
    _say("case: batch-target callback returns -Inf for even samples")
    mean, L = _problem(2, 6)

    def neg_even(v):
        v[::2] = -np.inf

    alg = eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(2), 1.0), 1000)
    a = eng.importance_sample(2, eng.BatchTarget(_HostGaussian(mean, L, edit=neg_even)), alg)
    assert np.all(np.isneginf(a.logweights[::2])) and np.all(np.isfinite(a.logweights[1::2]))


def test_callback_error_aborts_without_commit(eng):
    _say("case: batch-target callback returns nonzero in an AMIS run")
    mean, L = _problem(2, 7)
    bad = _HostGaussian(mean, L, ret=3)
    s = eng.prepare_sampler(4, eng.BatchTarget(bad), eng.AMIS(eng.SphericalGaussian(np.zeros(2), 2.0), rounds=3,
                                                               round_size=500))
    before = eng.current_proposal(s).location.copy()
    with pytest.raises(eng.AMISRoundError) as ei:
        eng.importance_sample_(s)
    assert ei.value.round == 1 and bad.calls == 1
    assert np.array_equal(eng.current_proposal(s).location, before)
    bad.ret = 0  # the same prepared sampler runs once the callback succeeds
    res = eng.importance_sample_(s)
    assert np.all(np.isfinite(res.logweights))


def test_callback_request_served_beside_others(eng):
    _say("case: batch-target callback request in a serving mix")
    mean, L = _problem(3, 8)
    alg = eng.AMIS(eng.SphericalGaussian(np.zeros(3), 2.0), rounds=4, round_size=3000)
    srv = eng.Server(memory_budget=16 << 30, max_active=3)
    other = eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(5), 1.0), 50000)
    t0 = srv.submit(21, eng.StandardGaussianTarget(), other)
    t1 = srv.submit(9, eng.BatchTarget(_HostGaussian(mean, L)), alg, "fp64")
    t2 = srv.submit(22, eng.StandardGaussianTarget(), other)
    r1, _ = srv.result(t1, timeout=120)
    for t in (t0, t2):
        srv.result(t, timeout=120)
    srv.close()
    alone = eng.importance_sample(9, eng.BatchTarget(_HostGaussian(mean, L)), alg)
    assert np.array_equal(r1.logweights, alone.logweights)


def test_callback_jobs_match_the_oracle(eng, oracle_mod):
    """The callback path against the FP64 oracle fed the same counter-RNG normals, with the
    oracle evaluating the same Gaussian log density itself: plain IS (samples and
    log-weights), and AMIS replayed on the device's fitted proposals (log-weights, round
    log Z, fits)."""
    _say("case: batch-target callback vs oracle, plain IS d = 3 and AMIS d = 4")
    mean, L = _problem(3, 11)
    tgt = eng.GaussianTarget(mean, L)
    q = eng.SphericalGaussian(np.zeros(3), 2.0)
    s = eng.prepare_sampler(19, eng.BatchTarget(_HostGaussian(mean, L)), eng.ImportanceSampling(q, 12000))
    z = s.export_randomness(0)["normals"]
    res = eng.importance_sample_(s)
    xo, lwo, fail = oracle_mod.plain_is(q, tgt.kind, tgt.params(3), z)
    assert fail[0] == 0
    assert np.max(np.abs(res.samples - xo)) < 1e-13
    assert np.max(np.abs(res.logweights - lwo) / np.maximum(1.0, np.abs(lwo))) < 1e-10
    assert abs(eng.lognormalizer(res) - oracle_mod.lognormalizer(lwo)) < 1e-10

    mean, L = _problem(4, 12)
    tgt = eng.GaussianTarget(mean, L)
    sched = [3000, 3500, 4000, 2500]
    q0 = eng.SphericalGaussian(np.zeros(4), 3.0)
    s = eng.prepare_sampler(23, eng.BatchTarget(_HostGaussian(mean, L)), eng.AMIS(q0, rounds=4, round_size=sched))
    normals = np.zeros((4, 4 * max(sched)))
    for t, n in enumerate(sched):
        normals[t, : 4 * n] = s.export_randomness(t)["normals"]
    res = eng.importance_sample_(s)
    means, facs, lns = s.history()
    o = oracle_mod.adaptive(0, q0, sched, normals, tgt.kind, tgt.params(4), nthreads=4, replay=(means, facs, lns))
    assert o["rc"] == 0
    rel = np.abs(res.logweights - o["logweights"]) / np.maximum(1.0, np.abs(o["logweights"]))
    assert np.max(rel) < 1e-10
    assert np.max(np.abs(res.diagnostics["round_lognormalizers"] - o["round_lognormalizers"])) < 1e-10
    assert np.max(np.abs(means[1:] - o["fit_means"][1:])) < 1e-9
    assert np.max(np.abs(facs[1:] - o["fit_factors"][1:])) < 1e-9
