"""GPU parity of static MIS (ImportanceSampling with a ProposalBank and an MIS scheme)
through the C ABI: against the oracle fed the device's own counter-RNG values
(assignment uniforms, normals, Student-t radial draws) for all four schemes --
proposal ids bit-exact, samples within 1e-12 and log-weights within 1e-10
relative in FP64; the self-target identity (log-weights 0, zero-mass proposals
inert); the random-mixture assignment frequency; FP32 against FP64 log Z."""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def _schemes(eng, groups):
    return [(0, eng.StratifiedMixture(), None), (1, eng.RandomMixture(), None), (2, eng.StandardMIS(), None),
            (3, eng.PartialDeterministicMixture(groups), groups)]


def _bank(eng, rng, d, K, family):
    props = []
    for _ in range(K):
        loc = rng.normal(0.0, 1.0, d)
        A = 0.4 * rng.standard_normal((d, d))
        L = np.linalg.cholesky(A @ A.T + 0.6 * np.eye(d))
        props.append(eng.FactorGaussian(loc, L) if family == "gauss" else eng.FactorStudentT(5.0, loc, L))
    return props


@pytest.mark.parametrize("family,d,target_kind", [("gauss", 3, "gauss"), ("t", 5, "gauss"), ("gauss", 6, "logistic"),
                                                  ("gauss", 20, "gauss")])
def test_static_mis_matches_oracle_replay(eng, oracle_mod, family, d, target_kind):
    _say(f"case: static MIS vs oracle, family {family}, d {d}, target {target_kind}")
    rng = np.random.default_rng(11 + d)
    K = 5
    props = _bank(eng, rng, d, K, family)
    masses = np.array([1.0, 0.0, 2.5, 1.0, 0.5])
    groups = ((1, 3), (2, 4, 5))
    if target_kind == "gauss":
        A = rng.standard_normal((d, d))
        tgt = eng.GaussianTarget(rng.normal(0, 0.5, d), np.linalg.cholesky(A @ A.T / d + np.eye(d)))
    else:
        X = rng.standard_normal((300, d))
        y = (rng.random(300) < 0.5).astype(float)
        tgt = eng.LogisticRegressionTarget(X, y, 10.0)
    N = 1500
    bank = eng.ProposalBank(props, masses)
    for code, scheme, grp in _schemes(eng, groups):
        s = eng.prepare_sampler(40 + code, tgt, eng.ImportanceSampling(bank, N, scheme), precision="fp64")
        r = s.export_randomness(0)
        res = eng.importance_sample_(s)
        rn = r.get("radial_normals")
        ru = r.get("radial_uniforms")
        o = oracle_mod.static_mis(props, masses, code, grp, r["normals"].reshape(N, d), r["uniforms"], tgt.kind,
                                  tgt.params(d), None if rn is None else rn.reshape(N, -1),
                                  None if ru is None else ru.reshape(N, -1))
        assert np.array_equal(res.provenance["proposal_id"], o["proposal_ids"]), code
        assert relerr(res.samples, o["samples"]) < 1e-12, code
        assert relerr(res.logweights, o["logweights"]) < 1e-10, code
        assert res.diagnostics["mis_scheme"] == scheme.name
        assert 2 not in set(res.provenance["proposal_id"])  # zero mass: never drawn


def test_static_mis_self_target_identity(eng):
    _say("case: static MIS with every proposal equal to the target")
    rng = np.random.default_rng(5)
    d = 4
    A = rng.standard_normal((d, d))
    L = np.linalg.cholesky(A @ A.T / d + np.eye(d))
    mu = rng.normal(0, 1, d)
    props = [eng.FactorGaussian(mu, L) for _ in range(3)]
    bank = eng.ProposalBank(props, [1, 0, 3])
    tgt = eng.GaussianTarget(mu, L)
    for code, scheme, _ in _schemes(eng, ((1, 3), (2,))):
        res = eng.importance_sample(7, tgt, eng.ImportanceSampling(bank, 10_001, scheme))
        assert np.max(np.abs(res.logweights)) <= 4096 * np.finfo(float).eps, code
        assert abs(eng.lognormalizer(res)) <= 4096 * np.finfo(float).eps, code
        assert set(res.provenance["proposal_id"]) <= {1, 3}


def test_static_mis_assignment_frequencies(eng):
    _say("case: static MIS assignment frequencies (random and stratified)")
    d = 2
    props = [eng.SphericalGaussian(np.full(d, -1.0), 1.0), eng.SphericalGaussian(np.full(d, 50.0), 1.0),
             eng.SphericalGaussian(np.full(d, 1.0), 1.0)]
    bank = eng.ProposalBank(props, [1, 0, 3])
    n = 20_000
    rnd = eng.importance_sample(0x5301, eng.StandardGaussianTarget(),
                                eng.ImportanceSampling(bank, n, eng.RandomMixture()))
    assert abs(np.mean(rnd.provenance["proposal_id"] == 1) - 0.25) < 0.02
    st = eng.importance_sample(0x5302, eng.StandardGaussianTarget(), eng.ImportanceSampling(bank, n))
    assert abs(int(np.sum(st.provenance["proposal_id"] == 1)) - n // 4) <= 1  # stratified: within one


def test_static_mis_fp32_matches_fp64(eng):
    _say("case: static MIS FP32 vs FP64 log Z")
    rng = np.random.default_rng(9)
    d, K = 8, 6
    props = _bank(eng, rng, d, K, "gauss")
    bank = eng.ProposalBank(props, rng.integers(1, 4, K).astype(float))
    tgt = eng.GaussianTarget(np.zeros(d), np.eye(d))
    for code, scheme, _ in _schemes(eng, ((1, 2, 3), (4, 5, 6))):
        a = eng.ImportanceSampling(bank, 20_000, scheme)
        z64 = eng.lognormalizer(eng.importance_sample(3, tgt, a, precision="fp64"))
        z32 = eng.lognormalizer(eng.importance_sample(3, tgt, a, precision="fp32"))
        assert abs(z32 - z64) < 1e-4, (code, z32, z64)
