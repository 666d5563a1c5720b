"""GPU tests of job groups (is_group: same-shape samplers advanced together, one launch
per round over every member) through the C ABI: every member's samples, log-weights,
round summaries and fitted proposals are bit-identical to its stand-alone run on the same
descriptor, over two consecutive runs (the committed proposal carried between them), in
FP64 and FP32, for plain IS (the 3-D standard-Gaussian workload) and AMIS (the 10-D
correlated-Gaussian workload at reduced size) with members of different proposals,
targets and seeds; a failing member fails as it does alone while the others are
unaffected; and the shape checks refuse what a group cannot run."""

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


def _same(a, b):
    assert np.array_equal(a.samples, b.samples)
    assert np.array_equal(a.logweights, b.logweights)
    assert np.array_equal(a.diagnostics["round_ess"], b.diagnostics["round_ess"])
    assert np.array_equal(a.diagnostics["round_lognormalizers"], b.diagnostics["round_lognormalizers"])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_plain_group_members_equal_their_stand_alone_runs(eng, precision):
    _say(f"case: group of 48 plain-IS members (d = 3, N = 10,000), {precision}")
    alg = eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(3), 1.5), 10_000)
    specs = [(100 + m, eng.StandardGaussianTarget(), alg) for m in range(48)]
    g = eng.prepare_group(specs, precision)
    for _ in range(2):
        res = eng.importance_sample_group_(g)
        assert len(res) == 48
    for m in (0, 17, 47):
        s = eng.prepare_sampler(100 + m, eng.StandardGaussianTarget(), alg, precision)
        eng.importance_sample_(s)
        alone = eng.importance_sample_(s)  # the second run, as the group's
        _same(res[m], alone)
        assert eng.lognormalizer(res[m]) == eng.lognormalizer(alone)


def _amis_specs(eng, M, d=10, T=5, N=2000):
    specs = []
    for m in range(M):
        rng = np.random.default_rng(40 + m)
        Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
        lam = np.exp(rng.uniform(np.log(0.1), np.log(10.0), d))
        L = np.linalg.cholesky(Q @ np.diag(lam) @ Q.T)
        tgt = eng.GaussianTarget(rng.normal(0.0, 2.0, d), L)
        q0 = eng.FactorGaussian(rng.normal(0.0, 0.5, d), 5.0 * np.eye(d))
        specs.append((7 + m, tgt, eng.AMIS(q0, rounds=T, round_size=N)))
    return specs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_amis_group_members_equal_their_stand_alone_runs(eng, precision):
    _say(f"case: group of 12 AMIS members (d = 10, T = 5 x N = 2,000, distinct problems), {precision}")
    specs = _amis_specs(eng, 12)
    g = eng.prepare_group(specs, precision)
    first = eng.importance_sample_group_(g)
    second = eng.importance_sample_group_(g)
    for m in (0, 5, 11):
        rng, tgt, alg = specs[m]
        s = eng.prepare_sampler(rng, tgt, alg, precision)
        a1 = eng.importance_sample_(s)
        _same(first[m], a1)
        a2 = eng.importance_sample_(s)  # the second run starts from the first run's fit
        _same(second[m], a2)
        gm, gf, gl = g.members[m].history()
        sm, sf, sl = s.history()
        assert np.array_equal(gm, sm) and np.array_equal(gf, sf) and np.array_equal(gl, sl)
        assert np.array_equal(eng.mean(second[m]), eng.mean(a2))


def test_a_failing_member_fails_as_it_does_alone(eng):
    _say("case: AMIS group with one member whose target is unreachable (every weight zero)")
    specs = _amis_specs(eng, 4, T=3, N=1024)
    rng, tgt, alg = specs[2]
    bad = eng.GaussianTarget(np.full(10, 1e200), np.eye(10))
    specs[2] = (rng, bad, alg)
    g = eng.prepare_group(specs, "fp64")
    failed = g.run()
    assert failed == [2]
    with pytest.raises(eng.AMISRoundError) as ei_group:
        g.collect(2)
    s = eng.prepare_sampler(rng, bad, alg, "fp64")
    with pytest.raises(eng.AMISRoundError) as ei_alone:
        eng.importance_sample_(s)
    assert (ei_group.value.round, ei_group.value.phase) == (ei_alone.value.round, ei_alone.value.phase)
    for m in (0, 1, 3):
        r, t, a = specs[m]
        _same(g.collect(m), eng.importance_sample(r, t, a))


def test_group_shape_checks(eng):
    _say("case: group shape checks")
    alg = eng.AMIS(eng.SphericalGaussian(np.zeros(3), 1.0), rounds=2, round_size=256)
    with pytest.raises(ValueError):  # members of different dimensions
        eng.prepare_group([(1, eng.StandardGaussianTarget(), alg),
                           (2, eng.StandardGaussianTarget(), eng.AMIS(eng.SphericalGaussian(np.zeros(4), 1.0),
                                                                      rounds=2, round_size=256))])
    with pytest.raises(NotImplementedError):  # no fused round at d = 16
        eng.prepare_group([(1, eng.StandardGaussianTarget(),
                            eng.AMIS(eng.SphericalGaussian(np.zeros(16), 1.0), rounds=2, round_size=256))])
    with pytest.raises(NotImplementedError):  # groups run plain IS and AMIS
        eng.prepare_group([(1, eng.StandardGaussianTarget(),
                            eng.NPMC(eng.SphericalGaussian(np.zeros(3), 1.0), rounds=2, round_size=256))])
    one = eng.prepare_group([(5, eng.StandardGaussianTarget(), alg)])
    _same(eng.importance_sample_group_(one)[0], eng.importance_sample(5, eng.StandardGaussianTarget(), alg))


def test_group_reset_and_round_summary(eng):
    _say("case: group reset and gathered round summaries (4 AMIS members, d = 10, T = 4 x N = 1,500), fp32")
    specs = _amis_specs(eng, 4, T=4, N=1500)
    g = eng.prepare_group(specs, "fp32")
    first = eng.importance_sample_group_(g)
    for t in range(4):
        ess, lz = g.round_summary(t)
        for m in range(4):
            assert ess[m] == first[m].diagnostics["round_ess"][t]
            assert lz[m] == first[m].diagnostics["round_lognormalizers"][t]
    # the third run after a reset starts from the initial proposal again, as a stand-alone
    # sampler's run, run, reset, run does (the counter-RNG draws follow the run index)
    eng.importance_sample_group_(g)
    g.reset()
    third = eng.importance_sample_group_(g)
    for m in (0, 3):
        rng, tgt, alg = specs[m]
        s = eng.prepare_sampler(rng, tgt, alg, "fp32")
        eng.importance_sample_(s)
        eng.importance_sample_(s)
        s.reset()
        _same(third[m], eng.importance_sample_(s))
        assert g.round_summary(-1)[1][m] == third[m].diagnostics["round_lognormalizers"][-1]
