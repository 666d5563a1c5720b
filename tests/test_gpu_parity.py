"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle fed the
same counter-RNG normals / uniforms, and against the reference's own known
answers.  Tolerances: log-weights 1e-10 relative with a 1e-10 absolute floor
(FP64); resampling ancestors bit-exact; FP32 mode |delta log Z| < 1e-3."""

import ctypes
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
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


# ---------------------------------------------------------------- RNG
def test_philox_matches_numpy(eng):
    _say("case: philox raw words vs numpy Philox4x64-10")
    lib = eng._lib.load()
    key = (0x1234_5678_9ABC_DEF0, 42)
    for c1, c2, c3 in ((0, 0, 0), (7, 1, 3), (2 ** 40, 5, 9)):
        n = 64
        out = np.zeros(4 * n, dtype=np.uint64)
        rc = lib.is_rng_raw(eng.context(), key[0], key[1], c1, c2, c3, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        assert rc == 0
        g = np.random.Philox(key=list(key), counter=[0, c1, c2, c3])
        ref = g.random_raw(4 * (n - 1))  # numpy emits counter+1 first
        assert np.array_equal(out[4:], ref.astype(np.uint64))


def test_exported_normals_are_box_muller(eng, oracle_mod):
    _say("case: exported normals vs host Box-Muller")
    q = eng.SphericalGaussian(np.zeros(5), 1.0)
    s = eng.prepare_sampler(99, eng.StandardGaussianTarget(), eng.ImportanceSampling(q, 300))
    z = s.export_randomness(0)["normals"].reshape(300, 5)
    assert abs(z.mean()) < 0.1 and abs(z.std() - 1) < 0.1
    # sample 0, coordinates 0..3 come from Philox block (0, round 0, purpose 0, 0)
    w = oracle_mod.philox4x64([0, 0, 0, 0], [99, 0x6D693335357873])
    n0, n1 = oracle_mod.box_muller(w[0], w[1])
    n2, n3 = oracle_mod.box_muller(w[2], w[3])
    assert np.allclose(z[0, :4], [n0, n1, n2, n3], rtol=1e-14, atol=1e-14)


# ---------------------------------------------------------------- plain IS
PLAIN_CASES = [
    ("spherical", lambda e: e.SphericalGaussian(np.zeros(3), 1.5), 0, None),
    ("diagonal", lambda e: e.DiagonalGaussian([0.5, -1.0, 2.0], [0.7, 1.3, 2.0]), 0, None),
    ("factor", lambda e: e.FactorGaussian([0.3, -0.2], [[1.2, 0], [0.4, 0.8]]), 1, "gauss2"),
    ("studentt-nu1", lambda e: e.SphericalStudentT(1.0, [0.0, 1.0], 2.0), 0, None),
    ("studentt-nu1.5", lambda e: e.DiagonalStudentT(1.5, [0.0, 1.0, -1.0], [1.0, 2.0, 0.5]), 0, None),
    ("studentt-nu5-factor", lambda e: e.FactorStudentT(5.0, [0.0, 1.0], [[1.1, 0], [-0.3, 0.9]]), 1, "gauss2"),
    ("scalar", lambda e: e.SphericalGaussian(0.25, 1.3), 0, None),
]


def _target(eng, name, d):
    if name == "gauss2":
        return eng.GaussianTarget([0.5, -0.5], [[1.0, 0.0], [0.6, 0.8]])
    return eng.StandardGaussianTarget()


@pytest.mark.parametrize("name,make,tk,tname", PLAIN_CASES)
def test_plain_is_parity(eng, oracle_mod, name, make, tk, tname):
    _say(f"case: plain IS parity {name}")
    q = make(eng)
    d = q.dim
    tgt = _target(eng, tname, d)
    n = 5000
    s = eng.prepare_sampler(1234, tgt, eng.ImportanceSampling(q, n))
    rnd = s.export_randomness(0)
    res = eng.importance_sample_(s)
    x, lw, fail = oracle_mod.plain_is(q, tk, tgt.params(d), rnd["normals"], rnd.get("radial_uniforms"))
    assert fail[0] == 0
    assert relerr(res.samples.reshape(n, -1), x, 1.0) < 1e-13
    assert relerr(res.logweights, lw, 1.0) < RTOL
    assert abs(eng.lognormalizer(res) - oracle_mod.lognormalizer(lw)) < 1e-10
    m, v, c = oracle_mod.moments(x, lw)
    assert relerr(eng.mean(res), m, 1.0) < 1e-10
    assert relerr(eng.var(res), v, 1.0) < 1e-10


@pytest.mark.parametrize("d", [16, 20, 33, 50])
def test_results_moments_multi_tile(eng, oracle_mod, d):
    # d + 1 > 16: the weighted moments span several 16x16 MFMA tiles
    _say(f"case: results moments d={d}")
    q = eng.DiagonalGaussian(np.linspace(-1.0, 1.0, d), np.linspace(0.8, 1.6, d))
    L = np.tril(np.full((d, d), 0.05), -1) + 1.3 * np.eye(d)
    tgt = eng.GaussianTarget(np.linspace(0.5, -0.5, d), L)
    s = eng.prepare_sampler(3, tgt, eng.ImportanceSampling(q, 3000))
    rnd = s.export_randomness(0)
    res = eng.importance_sample_(s)
    x, lw, fail = oracle_mod.plain_is(q, 1, tgt.params(d), rnd["normals"])
    assert fail[0] == 0
    assert relerr(res.logweights, lw, 1.0) < RTOL
    m, v, c = oracle_mod.moments(x, lw)
    assert relerr(eng.mean(res), m, 1.0) < 1e-10
    assert relerr(eng.var(res), v, 1.0) < 1e-10
    assert relerr(eng.cov(res), c, 1.0) < 1e-10


def test_plain_is_student_t_known_answer(eng):
    # test/kernel_execution.jl:490-505
    _say("case: plain IS Student-t buffered known answer")
    q = eng.SphericalStudentT(1.0, 0.0, 2.0)
    rng = eng.PrefilledRNG(normals=[[1.0, 2.0, -1.0, 4.0]])
    res = eng.importance_sample(rng, eng.GaussianTarget([0.0], [[1.0]]), eng.ImportanceSampling(q, 2))
    assert list(res.samples) == [1.0, -0.5]


def test_plain_is_exhausted_gamma_buffer(eng):
    # test/kernel_execution.jl:516-528: proposal_draw failure at index 1
    _say("case: plain IS exhausted gamma buffer")
    q = eng.SphericalStudentT(2.0, 0.0, 1.0)
    rng = eng.PrefilledRNG(normals=[np.concatenate([[0.0], np.full(8, -10.0)])], radial_uniforms=[np.ones(8)])
    with pytest.raises(eng.SamplerExecutionError) as ei:
        eng.importance_sample(rng, eng.StandardGaussianTarget(), eng.ImportanceSampling(q, 1))
    assert ei.value.phase == "proposal_draw" and ei.value.index == 1


def test_plain_is_self_target_zero_weights(eng):
    # test/native_proposals.jl:558-580: logw == 0 within 64 eps
    _say("case: plain IS self target")
    q = eng.FactorGaussian([1.0, -2.0], [[1.25, 0.0], [0.4, 0.8]])
    tgt = eng.GaussianTarget([1.0, -2.0], [[1.25, 0.0], [0.4, 0.8]])
    res = eng.importance_sample(5, tgt, eng.ImportanceSampling(q, 512))
    assert np.max(np.abs(res.logweights)) <= 64 * EPS
    assert abs(eng.lognormalizer(res)) <= 64 * EPS


# ---------------------------------------------------------------- AMIS
def _export_batches(s, schedule, d):
    maxn = max(schedule)
    normals = np.zeros((len(schedule), d * maxn))
    rn = ru = None
    for t, nt in enumerate(schedule):
        r = s.export_randomness(t)
        normals[t, : d * nt] = r["normals"]
        if "radial_normals" in r:
            if rn is None:
                rn = np.zeros((len(schedule), r["radial_normals"].size // nt * maxn))
            rn[t, : r["radial_normals"].size] = r["radial_normals"]
        if "radial_uniforms" in r:
            if ru is None:
                ru = np.zeros((len(schedule), r["radial_uniforms"].size // nt * maxn))
            ru[t, : r["radial_uniforms"].size] = r["radial_uniforms"]
    return normals, rn, ru


def test_amis_round_summary_known_answer(eng):
    # test/amis_kernel.jl:918-967 (Float64)
    _say("case: AMIS known answer 2 rounds x 2")
    q = eng.SphericalGaussian(0.0, 1.0)
    tgt = eng.GaussianTarget([0.0], [[1.0]])
    rng = eng.PrefilledRNG(normals=[[-2.0, 2.0], [0.0, 1.0]])
    s = eng.prepare_sampler(rng, tgt, eng.AMIS(q, rounds=2, round_size=2))
    res = eng.importance_sample_(s)
    ess = res.diagnostics["round_ess"]
    lz = res.diagnostics["round_lognormalizers"]
    assert abs(ess[0] - 2.0) < 1e-12 and abs(lz[0]) < 1e-12
    assert ess[1] == pytest.approx(3.473043499457537, rel=1e-10)
    assert lz[1] == pytest.approx(-0.2279354492780643, rel=1e-10)
    assert list(res.provenance["round"]) == [1, 1, 2, 2]


AMIS_CASES = [
    ("factor3-literal", 3, [257, 370, 483], "gauss", "factor", None),
    ("spherical2", 2, [17, 20, 23, 26, 29], "std", "spherical", None),
    ("scalar", 1, [64, 80, 96], "gauss", "scalar", None),
    ("d10-12rounds", 10, [2000] * 12, "gauss", "factor", None),
    ("d20-factor", 20, [3000] * 4, "gauss", "factor", None),
    ("studentt5-d4", 4, [300, 400, 500], "gauss", "diagonal", 5.0),
    # the fused round kernel's edges: d = 15 (largest fused d) with ragged round sizes,
    # d = 16 (first unfused exact d) and Student-t factor draws at d = 15
    # (rounds of several thousand samples keep the 15-d fits well conditioned)
    ("d15-ragged", 15, [4001, 5537, 4777, 6048, 4640, 5999], "std", "factor", None),
    ("d16-factor", 16, [1500] * 6, "gauss", "factor", None),
    ("studentt3-d15", 15, [4700, 4900, 4650, 4800, 4700, 4900], "gauss", "factor", 3.0),
    # d > 32: the wide buckets (48, 64) and the stand-alone fit kernel (k_finish);
    # near-target proposals keep the wide fits well conditioned
    ("d33-near", 33, [6000] * 4, "std", "near", None),
    ("d48-near", 48, [7000] * 4, "std", "near", None),
    ("d64-near", 64, [8000] * 4, "std", "near", None),
    ("studentt5-d48-near", 48, [7000] * 3, "std", "near", 5.0),
]


def _amis_problem(eng, d, tname, kind, dof, rng):
    A = rng.standard_normal((d, d))
    Q, _ = np.linalg.qr(A)
    lam = np.exp(rng.uniform(np.log(0.1), np.log(10), d))
    S = Q @ np.diag(lam) @ Q.T
    L = np.linalg.cholesky(S)
    mu = rng.normal(0, 2, d)
    tgt = eng.GaussianTarget(mu, L) if tname == "gauss" else eng.StandardGaussianTarget()
    loc = np.zeros(d) + 0.3
    if kind == "scalar":
        return eng.SphericalGaussian(0.3, 2.0), eng.GaussianTarget([0.75], [[1.1]])
    if kind == "spherical":
        p = eng.SphericalGaussian(loc, 2.5) if dof is None else eng.SphericalStudentT(dof, loc, 2.5)
    elif kind == "near":
        F = np.tril(rng.normal(0, 0.02, (d, d)), -1) + np.diag(np.linspace(1.0, 1.3, d))
        loc = np.full(d, 0.05)
        p = eng.FactorGaussian(loc, F) if dof is None else eng.FactorStudentT(dof, loc, F)
    elif kind == "diagonal":
        sc = np.linspace(1.5, 3.0, d)
        p = eng.DiagonalGaussian(loc, sc) if dof is None else eng.DiagonalStudentT(dof, loc, sc)
    else:
        F = np.tril(rng.normal(0, 0.2, (d, d)), -1) + np.diag(np.linspace(2.0, 3.0, d))
        p = eng.FactorGaussian(loc, F) if dof is None else eng.FactorStudentT(dof, loc, F)
    return p, tgt


@pytest.mark.parametrize("name,d,schedule,tname,kind,dof", AMIS_CASES)
def test_amis_parity(eng, oracle_mod, name, d, schedule, tname, kind, dof):
    _say(f"case: AMIS parity {name}")
    rng = np.random.default_rng(len(name) * 7 + d)
    p, tgt = _amis_problem(eng, d, tname, kind, dof, rng)
    alg = eng.AMIS(p, rounds=len(schedule), round_size=schedule)
    s = eng.prepare_sampler(77, tgt, alg)
    normals, rn, ru = _export_batches(s, schedule, d)
    res = eng.importance_sample_(s)
    x = res.samples.reshape(len(res), -1)
    # Replay (as validation/reproducers/cuda_amis.jl:298-313 does): the oracle samples
    # from the device's own proposal history, so every round is checked on its own and
    # rounding differences cannot compound through the adaptation.
    means, facs, lns = s.history()
    hist = (means, facs[:, 0, 0] if p.scalar else facs, lns)
    o = oracle_mod.adaptive(0, p, schedule, normals, tgt.kind, tgt.params(d), rn, ru, nthreads=4,
                            scalar=p.scalar, replay=hist)
    assert o["rc"] == 0
    assert relerr(x, o["samples"], 1.0) < 1e-13
    assert relerr(res.logweights, o["logweights"], 1.0) < RTOL
    assert relerr(res.diagnostics["round_ess"], o["round_ess"], 1.0) < 1e-10
    assert relerr(res.diagnostics["round_lognormalizers"], o["round_lognormalizers"], 1.0) < RTOL
    assert list(res.provenance["round"]) == list(o["round_ids"])
    # each round's FP64 fit (mean, covariance factor, log normalizer) from identical inputs
    assert relerr(means[1:], o["fit_means"][1:], 1.0) < 1e-9
    if p.scalar:
        assert relerr(facs[1:, 0, 0], o["fit_factors"][1:], 1.0) < 1e-9
    else:
        assert relerr(facs[1:], o["fit_factors"][1:], 1.0) < 1e-9
    assert relerr(lns[1:], o["fit_lognormalizers"][1:], 1.0) < 1e-9
    learned = eng.current_proposal(s)
    assert np.array_equal(learned.location, means[-1])
    if len(schedule) <= 5:
        # short, well-conditioned runs also match the oracle's own trajectory
        t = oracle_mod.adaptive(0, p, schedule, normals, tgt.kind, tgt.params(d), rn, ru, scalar=p.scalar)
        assert relerr(res.logweights, t["logweights"], 1.0) < RTOL
        assert relerr(learned.location, t["means"][-1], 1.0) < 1e-9


def test_amis_repeated_runs_start_from_committed_proposal(eng, oracle_mod):
    _say("case: AMIS repeated runs")
    rng = np.random.default_rng(3)
    p, tgt = _amis_problem(eng, 3, "gauss", "factor", None, rng)
    alg = eng.AMIS(p, rounds=3, round_size=300)
    s = eng.prepare_sampler(11, tgt, alg)
    eng.importance_sample_(s)
    first = eng.current_proposal(s)
    normals, _, _ = _export_batches(s, alg.schedule, 3)
    res2 = eng.importance_sample_(s)
    o = oracle_mod.adaptive(0, first, alg.schedule, normals, tgt.kind, tgt.params(3))
    assert relerr(res2.logweights, o["logweights"], 1.0) < RTOL


def _config2_problem(eng, seed=20260):
    """AMIS on a 10-D correlated Gaussian: mean ~ N(0, 4 I), covariance Q diag(lambda) Q^T
    with Q Haar and lambda log-uniform on [0.1, 10]; q0 = N(0, 25 I)."""
    d = 10
    rng = np.random.default_rng(seed)
    mu = rng.normal(0.0, 2.0, d)
    Q, R = np.linalg.qr(rng.standard_normal((d, d)))
    Q = Q * np.sign(np.diag(R))
    lam = np.exp(rng.uniform(np.log(0.1), np.log(10.0), d))
    L = np.linalg.cholesky(Q @ np.diag(lam) @ Q.T)
    return eng.GaussianTarget(mu, L), eng.FactorGaussian(np.zeros(d), 5.0 * np.eye(d))


def test_amis_config2_full_shape_fp64_replay(eng, oracle_mod):
    """The 50 x 20,000 workload at its own shape, FP64, each round replayed by the oracle
    on the device's fitted proposals: log-weights, fits and round log Z."""
    _say("case: AMIS config 2 full shape (T=50, N=20000, d=10) FP64 replay")
    T, N, d = 50, 20000, 10
    tgt, q0 = _config2_problem(eng)
    s = eng.prepare_sampler(1, tgt, eng.AMIS(q0, rounds=T, round_size=N))
    normals, _, _ = _export_batches(s, [N] * T, d)
    res = eng.importance_sample_(s)
    means, facs, lns = s.history()
    o = oracle_mod.adaptive(0, q0, [N] * T, normals, tgt.kind, tgt.params(d), nthreads=16,
                            replay=(means, facs, lns))
    assert o["rc"] == 0
    assert relerr(res.logweights, o["logweights"], 1.0) < RTOL
    assert relerr(res.diagnostics["round_lognormalizers"], o["round_lognormalizers"], 1.0) < 1e-10
    assert relerr(res.diagnostics["round_ess"], o["round_ess"], 1.0) < 1e-9
    assert relerr(means[1:], o["fit_means"][1:], 1.0) < 1e-9
    assert relerr(facs[1:], o["fit_factors"][1:], 1.0) < 1e-9
    assert relerr(lns[1:], o["fit_lognormalizers"][1:], 1.0) < 1e-9


def test_amis_config2_full_shape_fp32_against_fp64_oracle(eng, oracle_mod):
    """FP32 storage and arithmetic (FP64 fit and log Z) at the 50 x 20,000 shape against
    the oracle's own FP64 solve on the same normals: |delta log Z| < 1e-3 and moment error
    < 1e-3 (the accuracy gate of the effective-samples metric)."""
    _say("case: AMIS config 2 full shape FP32 vs FP64 oracle")
    T, N, d = 50, 20000, 10
    tgt, q0 = _config2_problem(eng)
    s = eng.prepare_sampler(1, tgt, eng.AMIS(q0, rounds=T, round_size=N), precision="fp32")
    normals, _, _ = _export_batches(s, [N] * T, d)
    res = eng.importance_sample_(s)
    o = oracle_mod.adaptive(0, q0, [N] * T, normals, tgt.kind, tgt.params(d), nthreads=16)
    assert o["rc"] == 0
    lz_o = oracle_mod.lognormalizer(o["logweights"])
    assert abs(eng.lognormalizer(res) - lz_o) < 1e-3
    om, _, oc = oracle_mod.moments(o["samples"].reshape(-1, d), o["logweights"])
    assert np.max(np.abs(eng.mean(res) - om)) < 1e-3
    assert np.max(np.abs(eng.cov(res) - oc)) < 1e-3


def test_amis_fp32_matches_fp64_log_z(eng):
    _say("case: AMIS fp32 vs fp64")
    rng = np.random.default_rng(5)
    p, tgt = _amis_problem(eng, 10, "gauss", "factor", None, rng)
    alg = eng.AMIS(eng.FactorGaussian(np.zeros(10), 5.0 * np.eye(10)), rounds=8, round_size=5000)
    r64 = eng.importance_sample(9, tgt, alg, precision="fp64")
    r32 = eng.importance_sample(9, tgt, alg, precision="fp32")
    lz64, lz32 = eng.lognormalizer(r64), eng.lognormalizer(r32)
    assert abs(lz32 - lz64) < 1e-3
    assert np.max(np.abs(eng.mean(r32) - eng.mean(r64))) < 1e-3


def test_amis_failure_is_transactional(eng):
    _say("case: AMIS nonfinite draw failure")
    p = eng.FactorGaussian([0.0, 0.0], [[1.0, 0.0], [0.0, 1.0]])
    tgt = eng.StandardGaussianTarget()
    bad = np.zeros(2 * 4)
    bad[5] = np.inf
    rng = eng.PrefilledRNG(normals=[np.zeros(8), bad])
    s = eng.prepare_sampler(rng, tgt, eng.AMIS(p, rounds=2, round_size=4))
    with pytest.raises(eng.AMISRoundError) as ei:
        eng.importance_sample_(s)
    assert ei.value.round == 2 and ei.value.phase == "sampling"
    assert ei.value.cause.index == 3
    kept = eng.current_proposal(s)
    assert np.array_equal(kept.location, [0.0, 0.0])


# ---------------------------------------------------------------- NPMC
def test_npmc_recurrence_known_answer(eng, oracle_mod):
    # test/npmc.jl:48-66
    _say("case: NPMC scalar recurrence")
    batches = [[-3, -2, -1, -0.5, 0, 0.5, 1, 2, 3], [3, 2, 1, 0.5, 0, -0.5, -1, -2, -3]]
    q = eng.SphericalGaussian(-1.0, 2.0)
    tgt = eng.GaussianTarget([0.75], [[math.sqrt(1.5)]], lognormalizer=0.0)
    s = eng.prepare_sampler(eng.PrefilledRNG(normals=batches), tgt, eng.NPMC(q, rounds=2, round_size=[9, 4]))
    res = eng.importance_sample_(s)
    o = oracle_mod.adaptive(1, q, [9, 4], np.array(batches, float), 1, tgt.params(1), scalar=True)
    assert relerr(res.logweights, o["logweights"], 1.0) < RTOL
    learned = eng.current_proposal(s)
    assert learned.location[0] == pytest.approx(o["means"][-1][0], rel=1e-10)
    assert learned.scale[0] == pytest.approx(o["factors"][-1], rel=1e-10)
    assert len(res) == 13 and res.diagnostics["proposal_evaluations"] == 13


@pytest.mark.parametrize("d,schedule", [(2, [16, 16]), (5, [1000, 1500, 700]), (10, [4000] * 6),
                                        (33, [6000] * 3), (48, [7000] * 3), (64, [9000] * 3),
                                        (4, [150000, 150000])])  # the clip's select over ~150 workgroups
def test_npmc_parity(eng, oracle_mod, d, schedule):
    _say(f"case: NPMC parity d={d}")
    rng = np.random.default_rng(d)
    if d > 32:  # the wide buckets and k_finish, on a near-target start
        p, tgt = _amis_problem(eng, d, "std", "near", None, rng)
    else:
        p, tgt = _amis_problem(eng, d, "gauss", "factor", None, rng)
    alg = eng.NPMC(p, rounds=len(schedule), round_size=schedule)
    s = eng.prepare_sampler(31, tgt, alg)
    normals, _, _ = _export_batches(s, schedule, d)
    res = eng.importance_sample_(s)
    means, facs, lns = s.history()
    o = oracle_mod.adaptive(1, p, schedule, normals, tgt.kind, tgt.params(d), replay=(means, facs, lns))
    assert o["rc"] == 0
    assert relerr(res.logweights, o["logweights"], 1.0) < RTOL
    assert relerr(res.diagnostics["round_lognormalizers"], o["round_lognormalizers"], 1.0) < RTOL
    assert relerr(res.diagnostics["adaptation_ess"], o["adaptation_ess"], 1.0) < 1e-9
    assert relerr(means[1:], o["fit_means"][1:], 1.0) < 1e-9
    assert relerr(facs[1:], o["fit_factors"][1:], 1.0) < 1e-9


# ---------------------------------------------------------------- DM-PMC and resampling
def _dm_problem(eng, d, K, kind, rng):
    if kind == "spherical":
        props = [eng.SphericalGaussian(rng.normal(0, 3, d), 1.0 + 0.1 * k) for k in range(K)]
    else:
        props = [eng.FactorGaussian(rng.normal(0, 3, d),
                                    np.tril(rng.normal(0, 0.1, (d, d)), -1) + np.diag(np.full(d, 1.2)))
                 for k in range(K)]
    masses = rng.uniform(0.5, 2.0, K)
    masses[rng.integers(K)] = 0.0  # one inert proposal
    return eng.ProposalBank(props, masses)


@pytest.mark.parametrize("d,K,schedule,kind,method", [
    (1, 3, [13, 14, 13], "spherical", "multinomial"),
    (3, 8, [512] * 5, "factor", "multinomial"),
    (6, 16, [2048, 1500, 3000], "spherical", "systematic"),
])
def test_dmpmc_parity(eng, oracle_mod, d, K, schedule, kind, method):
    _say(f"case: DM-PMC parity d={d} K={K} {method}")
    rng = np.random.default_rng(K + d)
    bank = _dm_problem(eng, d, K, kind, rng)
    meth = eng.SystematicResampling() if method == "systematic" else eng.MultinomialResampling()
    alg = eng.DeterministicMixturePMC(bank, rounds=len(schedule), round_size=schedule, method=meth)
    tgt = eng.GaussianTarget(np.zeros(d), 2.0 * np.eye(d))
    s = eng.prepare_sampler(5, tgt, alg)
    normals, _, _ = _export_batches(s, schedule, d)
    uniforms = np.array([s.export_randomness(t)["uniforms"] for t in range(len(schedule))])
    res = eng.importance_sample_(s)
    order = oracle_mod.active_order(list(bank.masses))
    counts, logc = oracle_mod.allocation_plan([bank.masses[i] for i in order], schedule)
    P = [bank.proposals[i] for i in order]
    locs = np.array([p.location for p in P])
    if kind == "spherical":
        scales = np.array([p.scale[0] for p in P]); ok = 0
    else:
        scales = np.array([p.factor() for p in P]); ok = 2
    o = oracle_mod.dmpmc(0, 0.0, ok, d, locs, scales, [p.lognormalizer for p in P], [i + 1 for i in order],
                         schedule, counts, logc, normals, uniforms, tgt.kind, tgt.params(d),
                         resampling=3 if method == "systematic" else 2)
    assert o["rc"] == 0
    assert relerr(res.logweights, o["logweights"], 1.0) < RTOL
    assert np.array_equal(res.provenance["proposal_id"], o["proposal_ids"])
    assert np.array_equal(s.ancestors(), o["ancestors"])  # bit-exact ancestors
    cur = eng.current_proposal(s)
    got = np.array([cur.proposals[i].location for i in order])
    assert relerr(got, o["final_locations"], 1.0) < 1e-12


def _host_cdf(eng, lw):
    lib = eng._lib.load()
    lw = np.ascontiguousarray(lw, dtype=np.float64)
    out = np.zeros(lw.size)
    F = eng._lib.F64P
    assert lib.is_cdf(eng.context(), lw.ctypes.data_as(F), lw.size, out.ctypes.data_as(F)) == 0
    return out


def _host_select(eng, cdf, lw, u, method):
    lib = eng._lib.load()
    F, I = eng._lib.F64P, eng._lib.I64P
    cdf, lw, u = (np.ascontiguousarray(a, dtype=np.float64) for a in (cdf, lw, u))
    count = u.size
    anc = np.zeros(count, dtype=np.int64)
    assert lib.is_select(eng.context(), cdf.ctypes.data_as(F), lw.ctypes.data_as(F), cdf.size,
                         u.ctypes.data_as(F), count, method, anc.ctypes.data_as(I)) == 0
    return anc


def test_device_cdf_bit_exact_and_selection(eng, oracle_mod):
    _say("case: device CDF + selection")
    rng = np.random.default_rng(2)
    for n in (1, 4095, 4096, 4097, 33000, 1048577, 3000000):  # the last two: more tiles than scan threads
        lw = rng.normal(0, 3, n)
        lw[rng.random(n) < 0.05] = -np.inf
        if np.all(lw == -np.inf):
            lw[0] = 0.0
        got = _host_cdf(eng, lw)
        ref = oracle_mod.cdf_blocked(lw)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
        u = rng.random(257)
        assert np.array_equal(_host_select(eng, ref, lw, u, 0), oracle_mod.select_ancestors(ref, lw, u))
        pts = oracle_mod.systematic_points(u[0], 257)
        assert np.array_equal(_host_select(eng, ref, lw, u, 1), oracle_mod.select_ancestors(ref, lw, pts))


def test_selection_known_answers_on_device(eng, oracle_mod):
    # test/dm_pmc_kernel.jl:205-272 on the device selection kernel
    _say("case: selection known answers")
    prev = np.nextafter
    cases = [
        ([0.2, 0.5, 0.9, 1.0], np.log([0.2, 0.3, 0.4, 0.1]), [0.1, 0.2, 0.2001, 0.95], [1, 2, 2, 4]),
        ([0, 0, 0.5, 1], None, [0, prev(0.5, 0), 0.5, prev(1.0, 0)], [3, 3, 4, 4]),
        ([0.25, 1, 1, 1], None, [0, 0.25, prev(0.25, 1), prev(1.0, 0)], [1, 2, 2, 2]),
        ([0.25, 0.5, 0.75, 1], None, [1.0], [4]),
    ]
    for cdf, lw, u, expected in cases:
        cdf = np.array(cdf, float)
        if lw is None:
            with np.errstate(divide="ignore"):
                lw = np.log(np.diff(np.concatenate([[0.0], cdf])))
        assert list(_host_select(eng, cdf, lw, np.array(u, float), 0)) == expected
    # test/resampling.jl:54-59: zero-mass support with the device CDF
    lw = np.concatenate([[-np.inf], np.zeros(6), [-np.inf], np.zeros(6), [-np.inf]])
    u = np.array([0.0, 0.5, np.nextafter(np.float32(1), np.float32(0))], dtype=np.float64)
    assert list(_host_select(eng, _host_cdf(eng, lw), lw, u, 0)) == [2, 9, 14]


def test_results_resample_matches_oracle(eng, oracle_mod):
    _say("case: results resample")
    q = eng.SphericalGaussian(np.zeros(2), 2.0)
    s = eng.prepare_sampler(8, eng.StandardGaussianTarget(), eng.ImportanceSampling(q, 10000))
    z = s.export_randomness(0)["normals"]
    res = eng.importance_sample_(s)
    x, lw, _ = oracle_mod.plain_is(q, 0, None, z)
    u = np.random.default_rng(1).random(500)
    us = eng.resample(u, res, 500)
    ref = oracle_mod.select_ancestors(oracle_mod.cdf_blocked(res.logweights), res.logweights, u)
    assert np.array_equal(us.ancestors, ref)
    # UnweightedSamples gathered on the device (src/resampling.jl:120-164,263-283)
    assert len(us) == 500 and np.array_equal(us.samples, res.samples[ref - 1])
    assert np.allclose(eng.mean(us), res.samples[ref - 1].mean(axis=0), rtol=1e-12, atol=1e-12)
    us_s = eng.resample([0.37], res, 1000, eng.SystematicResampling())
    ref_s = oracle_mod.select_ancestors(oracle_mod.cdf_blocked(res.logweights), res.logweights,
                                        oracle_mod.systematic_points(0.37, 1000))
    assert np.array_equal(us_s.ancestors, ref_s)
    assert np.array_equal(us_s.samples, res.samples[ref_s - 1])
    with pytest.raises(ValueError):
        eng.lognormalizer(us)  # unweighted samples carry no weights
    # FP32 results gather their FP32 samples
    s32 = eng.prepare_sampler(8, eng.StandardGaussianTarget(), eng.ImportanceSampling(q, 10000), precision="fp32")
    r32 = eng.importance_sample_(s32)
    us32 = eng.resample(u, r32, 500)
    ref32 = oracle_mod.select_ancestors(oracle_mod.cdf_blocked(r32.logweights), r32.logweights, u)
    assert np.array_equal(us32.ancestors, ref32)
    assert np.array_equal(us32.samples, r32.samples[ref32 - 1])


def test_results_keep_their_own_storage_across_runs(eng, oracle_mod):
    """A WeightedSamples does not change when the sampler runs again
    (src/methods/importance_sampling.jl:904-906)."""
    _say("case: results own their storage")
    rng = np.random.default_rng(5)
    p, tgt = _amis_problem(eng, 3, "gauss", "factor", None, rng)
    s = eng.prepare_sampler(12, tgt, eng.AMIS(p, rounds=3, round_size=400))
    r1 = eng.importance_sample_(s)
    lz1, m1, v1, c1 = eng.lognormalizer(r1), eng.mean(r1), eng.var(r1), eng.cov(r1)
    f1 = eng.mean(lambda x: x[:, 0] ** 2, r1)
    r2 = eng.importance_sample_(s)  # a second run (new counter-RNG draws, refitted proposals)
    assert eng.lognormalizer(r2) != lz1
    assert eng.lognormalizer(r1) == lz1
    assert np.array_equal(eng.mean(r1), m1) and np.array_equal(eng.var(r1), v1) and np.array_equal(eng.cov(r1), c1)
    assert eng.mean(lambda x: x[:, 0] ** 2, r1) == f1
    # and they are the oracle's statistics of r1's own arrays
    assert abs(lz1 - oracle_mod.lognormalizer(r1.logweights)) < 1e-12
    om, ov, oc = oracle_mod.moments(r1.samples, r1.logweights)
    assert np.max(np.abs(m1 - om)) < 1e-10 and np.max(np.abs(c1 - oc)) < 1e-10
    # the kept result still resamples, from its own weights
    us = eng.resample(rng.random(50), r1, 50)
    assert len(us) == 50


# ---------------------------------------------------------------- graph replay
def _graph_cases(eng):
    rng = np.random.default_rng(77)
    p, tgt = _amis_problem(eng, 4, "gauss", "factor", None, rng)
    yield "AMIS", tgt, eng.AMIS(p, rounds=4, round_size=[300, 200, 500, 100])
    yield "NPMC", tgt, eng.NPMC(p, rounds=3, round_size=400)
    bank = _dm_problem(eng, 3, 6, "factor", rng)
    yield "DM-PMC", eng.GaussianTarget(np.zeros(3), 2.0 * np.eye(3)), \
        eng.DeterministicMixturePMC(bank, rounds=3, round_size=[256, 300, 128],
                                    method=eng.SystematicResampling())


def test_graph_replays_match_eager_runs(eng):
    """Each run's launch sequence is captured once and replayed; replays after the
    first (with eager result kernels in between) must reproduce eager runs bit for bit."""
    for name, tgt, alg in _graph_cases(eng):
        _say(f"case: graph replay {name}")
        g = eng.prepare_sampler(21, tgt, alg)
        e = eng.prepare_sampler(21, tgt, alg)
        assert e._lib.is_job_set_graph(e._job, 0) == 0
        for rep in range(4):
            rg = eng.importance_sample_(g)
            lzg = eng.lognormalizer(rg)  # eager work between replays
            re_ = eng.importance_sample_(e)
            assert np.array_equal(rg.logweights, re_.logweights), (name, rep)
            assert np.array_equal(rg.samples, re_.samples), (name, rep)
            assert lzg == eng.lognormalizer(re_), (name, rep)
            if name == "DM-PMC":
                assert np.array_equal(g.ancestors(), e.ancestors()), (name, rep)
            if rep == 1:
                g.reset()
                e.reset()


def test_dmpmc_ancestor_replay_seam(eng, oracle_mod):
    """is_job_set_ancestors: a run given another run's ancestors follows its bank
    trajectory (FP64 with its own ancestors: bit-identical log-weights); an FP64 replay
    of an FP32 run's trajectory matches it per round within the FP32 gate."""
    _say("case: DM-PMC ancestor replay seam")
    rng = np.random.default_rng(44)
    d, K, sched = 4, 8, [1024] * 5
    bank = _dm_problem(eng, d, K, "factor", rng)
    tgt = eng.GaussianTarget(np.zeros(d), 2.0 * np.eye(d))
    alg = eng.DeterministicMixturePMC(bank, rounds=len(sched), round_size=sched)
    a = eng.prepare_sampler(3, tgt, alg)
    ra = eng.importance_sample_(a)
    b = eng.prepare_sampler(3, tgt, alg)
    b.set_ancestors(a.ancestors())
    rb = eng.importance_sample_(b)
    assert np.array_equal(ra.logweights, rb.logweights)
    assert np.array_equal(a.ancestors(), b.ancestors())
    # the FP32 run's banks, replayed in FP64
    s32 = eng.prepare_sampler(5, tgt, alg, precision="fp32")
    r32 = eng.importance_sample_(s32)
    s64 = eng.prepare_sampler(5, tgt, alg, precision="fp64")
    s64.set_ancestors(s32.ancestors())
    r64 = eng.importance_sample_(s64)
    assert np.array_equal(s64.ancestors(), s32.ancestors())
    dlz = np.abs(r32.diagnostics["round_lognormalizers"] - r64.diagnostics["round_lognormalizers"])
    assert np.max(dlz) < 1e-4
    with pytest.raises(ValueError):
        b.set_ancestors([[0] * K])  # 1-based


@pytest.mark.parametrize("d,K,schedule,precision", [(3, 6, [600, 900, 750], "fp64"), (5, 16, [4096] * 4, "fp64"),
                                                    (4, 8, [2000] * 3, "fp32")])
def test_dmpmc_local_resampling_parity(eng, oracle_mod, d, K, schedule, precision):
    """LocalResampling against the oracle fed the same draws and uniforms: bit-exact
    ancestors (each from its own proposal's group), log-weights."""
    _say(f"case: DM-PMC local resampling d={d} K={K} {precision}")
    rng = np.random.default_rng(100 + K)
    bank = _dm_problem(eng, d, K, "factor", rng)
    alg = eng.DeterministicMixturePMC(bank, rounds=len(schedule), round_size=schedule,
                                      resampling=eng.LocalResampling())
    tgt = eng.GaussianTarget(np.zeros(d), 2.0 * np.eye(d))
    s = eng.prepare_sampler(6, tgt, alg, precision=precision)
    normals, _, _ = _export_batches(s, schedule, d)
    uniforms = np.array([s.export_randomness(t)["uniforms"] for t in range(len(schedule))])
    res = eng.importance_sample_(s)
    assert res.diagnostics["resampling"] == "local"
    order = oracle_mod.active_order(list(bank.masses))
    counts, logc = oracle_mod.allocation_plan([bank.masses[i] for i in order], schedule)
    P = [bank.proposals[i] for i in order]
    o = oracle_mod.dmpmc(0, 0.0, 2, d, np.array([p.location for p in P]), np.array([p.factor() for p in P]),
                         [p.lognormalizer for p in P], [i + 1 for i in order], schedule, counts, logc, normals,
                         uniforms, tgt.kind, tgt.params(d), resampling=4)
    assert o["rc"] == 0
    if precision == "fp64":
        assert relerr(res.logweights, o["logweights"], 1.0) < RTOL
        assert np.array_equal(s.ancestors(), o["ancestors"])
    # every ancestor lies in its own proposal's group of the round
    Ka = len(order)  # active (positive-mass) proposals
    starts = np.concatenate([[0], np.cumsum(np.asarray(counts).reshape(Ka, -1)[:, 0])])
    anc = s.ancestors()
    for k in range(Ka):
        assert starts[k] < anc[0, k] <= starts[k + 1]


def test_dmpmc_local_resampling_known_answer_and_empty_group(eng):
    """test/dm_pmc.jl:798-817 and :885-911 on the device."""
    _say("case: DM-PMC local resampling known answer")
    bank = eng.ProposalBank([eng.SphericalGaussian(-2.0, 1.0), eng.SphericalGaussian(2.0, 1.0)])
    alg = eng.DeterministicMixturePMC(bank, rounds=1, round_size=4, resampling=eng.LocalResampling())
    rng = eng.PrefilledRNG(normals=[[-1.0, 1.0, -1.0, 1.0]], uniforms=[[0.5, 0.5]])
    s = eng.prepare_sampler(rng, eng.StandardGaussianTarget(), alg)
    res = eng.importance_sample_(s)
    assert list(res.samples.reshape(-1)) == [-3.0, -1.0, 1.0, 3.0]
    locs = [p.location[0] for p in eng.current_proposal(s).proposals]
    assert locs == [-1.0, 1.0]

    def left_empty(values, samples, stride, dim, count, is_float, stream):
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipStreamSynchronize(stream)
        x = np.empty(count)
        h.hipMemcpy(x.ctypes.data, samples, count * 8, 2)
        v = np.where(x < 0, -np.inf, -0.5 * x * x)
        h.hipMemcpy(values, v.ctypes.data, count * 8, 1)
        return 0

    rng = eng.PrefilledRNG(normals=[[-1.0, 1.0, -1.0, 1.0]], uniforms=[[0.5, 0.5]])
    s = eng.prepare_sampler(rng, eng.BatchTarget(left_empty), alg)
    initial = [p.location[0] for p in eng.current_proposal(s).proposals]
    with pytest.raises(eng.DMPMCRoundError) as ei:
        eng.importance_sample_(s)
    assert ei.value.phase == "resampling" and isinstance(ei.value.cause, eng.AllZeroWeightsError)
    assert [p.location[0] for p in eng.current_proposal(s).proposals] == initial
