"""The static-MIS oracle (oracle.static_mis and helpers) against the reference's own
known answers: the assignment mapper truth table (test/static_mis.jl:31-67), the
active-bank ordering and zero-mass inertness (:69-183, :185-247), the compiled
partial groups (:380-432), and two identities fixed by the mathematics: a target
equal to the bank mixture gives log-weights 0 under the full mixture (:185-247),
and the partial denominator equals a brute-force mixture over the group
(:434-477)."""

import math

import numpy as np
import pytest

import oracle


def test_assignment_mapper_truth_table():
    cdf = [0.2, 0.5, 1.0]
    m = lambda c, u: oracle.mis_assignment(c, u, 0, 1, False) + 1  # noqa: E731 (1-based slots)
    assert m(cdf, 0.0) == 1
    assert m(cdf, 0.2) == 1
    assert m(cdf, 0.5) == 2
    assert m(cdf, np.nextafter(1.0, 0.0)) == 3
    assert m(cdf, 1.0) == 3
    for u in (0.0, 0.5, 1.0):
        assert m([1.0], u) == 1
    for b in (1e-8, np.nextafter(1.0, 0.0)):
        c = [b, 1.0]
        assert m(c, np.nextafter(b, 0.0)) == 1
        assert m(c, b) == 1
        assert m(c, np.nextafter(b, 2.0)) == 2
    # stratified: sample 2 of 3 (0-based 1) with u = 0.5 maps to 0.5
    assert oracle.mis_assignment(cdf, 0.5, 1, 3, True) + 1 == 2
    # 10,000 slots, the last one
    values = np.linspace(1 / 10_000, 1, 10_000)
    assert oracle.mis_assignment(values, np.nextafter(1.0, 0.0), 0, 1, False) + 1 == 10_000


def test_active_bank_order_and_log_masses():
    ids, cdf, lm = oracle.mis_active_bank([1, 0, 3])
    assert ids == [1, 3]
    assert np.allclose(cdf, [0.25, 1.0], rtol=0, atol=1e-15)
    assert np.allclose(np.exp(lm), [0.25, 0.75], rtol=1e-14)
    # stable sort by mass: equal masses keep the configured order
    ids2, _, _ = oracle.mis_active_bank([3, 1, 3, 1])
    assert ids2 == [2, 4, 1, 3]


def test_partial_groups_compile_like_the_reference():
    ids, cdf, lm = oracle.mis_active_bank([1, 2, 3, 4, 0])
    g, off, mem, coef = oracle.mis_partial_groups(ids, lm, ((1, 2), (5,), (3, 4)), 5)
    assert g == [0, 0, 1, 1]
    assert off == [0, 2, 4]
    assert mem == [0, 1, 2, 3]
    assert np.allclose(np.exp(coef), [1 / 3, 2 / 3, 3 / 7, 4 / 7], rtol=1e-14)
    for bad in ((), ((1, 2), (), (3, 4, 5)), ((1, 2), (2, 3, 4, 5)), ((1, 2), (3, 4)), ((1, 2), (3, 4, 5, 6))):
        with pytest.raises(ValueError):
            oracle.mis_partial_groups(ids, lm, bad, 5)


class _G:
    """Gaussian proposal record for the oracle (kind 0 spherical, 1 diagonal, 2 factor)."""

    def __init__(self, mu, sd):
        self.family, self.dof, self.kind = 0, 0.0, 0
        self.location = np.atleast_1d(np.asarray(mu, float))
        self.scale = np.array([float(sd)])
        d = self.location.size
        self.lognormalizer = -0.5 * d * math.log(2 * math.pi) - d * math.log(sd)


def _gauss_logpdf(x, mu, sd):
    return float(np.sum(-0.5 * ((x - mu) / sd) ** 2 - math.log(sd) - 0.5 * math.log(2 * math.pi)))


def test_full_mixture_identity_and_inert_zero_mass():
    rng = np.random.default_rng(1)
    props = [_G([-1.0, 0.0], 1.0), _G([50.0, 50.0], 1.0), _G([1.0, 0.5], 0.7)]
    masses = [1, 0, 3]
    n = 400
    # target = the mixture 0.25 N(-1, 1) + 0.75 N(1, 0.7) as a linear regression with
    # no data is not expressible; evaluate the identity on the denominator instead
    normals, uniforms = rng.standard_normal((n, 2)), rng.random(n)
    for scheme in (0, 1):
        o = oracle.static_mis(props, masses, scheme, None, normals, uniforms, 0, None)
        assert set(o["proposal_ids"]) <= {1, 3}
        for i in range(n):
            x = o["samples"][i]
            mix = math.log(0.25 * math.exp(_gauss_logpdf(x, props[0].location, 1.0))
                           + 0.75 * math.exp(_gauss_logpdf(x, props[2].location, 0.7)))
            assert abs((o["logtargets"][i] - o["logweights"][i]) - mix) < 1e-12 * max(1.0, abs(mix))
    # stratified counts are within one of n m_k; the random mixture draws near them
    o = oracle.static_mis(props, masses, 0, None, normals, uniforms, 0, None)
    assert abs(int(np.sum(o["proposal_ids"] == 1)) - 100) <= 1


def test_standard_mis_generating_denominator():
    rng = np.random.default_rng(2)
    props = [_G([0.3], 1.2), _G([9.0], 1.0), _G([-0.4], 0.8)]
    n = 300
    normals, uniforms = rng.standard_normal((n, 1)), rng.random(n)
    o = oracle.static_mis(props, [1, 0, 3], 2, None, normals, uniforms, 0, None)
    for i in range(n):
        k = o["proposal_ids"][i] - 1
        q = _gauss_logpdf(o["samples"][i], props[k].location, props[k].scale[0])
        assert abs((o["logtargets"][i] - o["logweights"][i]) - q) < 1e-12 * max(1.0, abs(q))


def test_partial_denominator_brute_force():
    rng = np.random.default_rng(3)
    props = [_G([float(k)], 0.6 + 0.1 * k) for k in range(4)]
    masses = [1.0, 2.0, 3.0, 4.0]
    groups = ((1, 2), (3, 4))
    n = 200
    normals, uniforms = rng.standard_normal((n, 1)), rng.random(n)
    o = oracle.static_mis(props, masses, 3, groups, normals, uniforms, 0, None)
    for i in range(n):
        pid = o["proposal_ids"][i]
        grp = groups[0] if pid <= 2 else groups[1]
        gm = np.array([masses[q - 1] for q in grp])
        c = gm / gm.sum()
        dens = sum(c[j] * math.exp(_gauss_logpdf(o["samples"][i], props[q - 1].location, props[q - 1].scale[0]))
                   for j, q in enumerate(grp))
        assert abs(-(o["logweights"][i] - o["logtargets"][i]) - math.log(dens)) < 1e-12
