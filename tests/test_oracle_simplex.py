"""Pins of the oracle's SimplexTransform and flat selector layout restatements against the
reference's own known answers (test/transforms.jl:305-470, test/product_proposal.jl:54-74,
283-320, test/kernel_execution.jl:615-640) and what the mathematics fixes: the exactly
uniform zero, sums to one, the inverse round trip, the orthonormal embedding, the log
Jacobian against a finite-difference determinant and the large-K closed form, the reason
bits and their order, and the chart measure of a flat layout (a target equal to the
proposal's pushforward gives log-weights of zero)."""

import math

import numpy as np
import pytest

from oracle import (TRANSFORM_NONFINITE_INPUT, TRANSFORM_NONFINITE_OUTPUT, TRANSFORM_OUTSIDE_SUPPORT,
                    layout_dim, layout_forward, layout_target, simplex_forward, simplex_forward_c, simplex_inverse,
                    target)

EPS = {np.float32: float(np.finfo(np.float32).eps), np.float64: float(np.finfo(np.float64).eps)}


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("K", [2, 4])
def test_zero_is_exactly_uniform(T, K):
    # test/transforms.jl:346-352
    x, _, r = simplex_forward(np.zeros(K - 1), T)
    assert r == 0
    assert np.all(x == T(1) / T(K)) and x.sum(dtype=T) == T(1)
    if T is np.float64:
        xc, _, rc = simplex_forward_c(np.zeros(K - 1))
        assert rc == 0 and np.all(xc == 1.0 / K)


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("K", [2, 3, 10])
def test_range_and_round_trip(T, K):
    # test/transforms.jl:325-344
    e = EPS[T]
    z = np.array([math.sin(i) / 3 for i in range(1, K)], dtype=T)
    x, lj, r = simplex_forward(z, T)
    assert r == 0 and x.size == K and np.all(np.isfinite(x)) and np.all(x > 0)
    assert abs(float(np.sum(x.astype(np.float64))) - 1.0) <= 8 * e
    zr, ljr, rr = simplex_inverse(x, T)
    assert rr == 0
    assert np.allclose(zr.astype(float), z.astype(float), rtol=256 * e, atol=256 * e)
    assert abs(float(ljr) - float(lj)) <= 64 * e * max(1.0, abs(float(lj)))


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_stable_finite_arithmetic(T):
    # test/transforms.jl:354-366
    e = EPS[T]
    z = np.array([20, -20, 15], dtype=T)
    x, lj, r = simplex_forward(z, T)
    assert r == 0 and np.all(np.isfinite(x)) and np.all(x > 0) and np.isfinite(lj)
    zr, ljr, rr = simplex_inverse(x, T)
    assert rr == 0
    assert np.allclose(zr.astype(float), z.astype(float), rtol=512 * e, atol=512 * e)
    assert abs(float(ljr) - float(lj)) <= 512 * e * abs(float(lj))


def test_float32_large_dimension_reductions():
    # test/transforms.jl:368-400
    x = [np.float32(v) for v in (0.2, 0.3, 0.5 + 32 * EPS[np.float32])]
    assert simplex_inverse(x, np.float32)[2] == TRANSFORM_OUTSIDE_SUPPORT
    for K in (1000, 10_000):
        _, lj, r = simplex_forward(np.zeros(K - 1, np.float32), np.float32)
        expected = 0.5 * math.log(K) - K * math.log(K)
        assert r == 0
        assert abs(float(lj) - expected) <= float(np.spacing(np.float32(abs(expected))))
    K = 10_000
    z = np.array([0.25 * math.sin(i) for i in range(1, K)], dtype=np.float32)
    x, _, r = simplex_forward(z, np.float32)
    assert r == 0 and abs(float(np.sum(x.astype(np.float64))) - 1.0) <= 2 * EPS[np.float32]


def test_orthonormal_permutation_neutral_geometry():
    # test/transforms.jl:402-432, on the C restatement
    K, scale = 5, 0.25
    cols = np.zeros((K, K - 1))
    for c in range(K - 1):
        z = np.zeros(K - 1)
        z[c] = scale
        x, _, r = simplex_forward_c(z)
        assert r == 0
        lx = np.log(x)
        cols[:, c] = (lx - lx.sum() / K) / scale
    assert np.max(np.abs(cols.T @ cols - np.eye(K - 1))) < 2e-14
    assert np.max(np.abs(cols @ cols.T - (np.eye(K) - 1.0 / K))) < 2e-14
    w = np.array([0.05, 0.15, 0.25, 0.20, 0.35])
    ref = float(np.sum(simplex_inverse(w)[0] ** 2))
    for perm in ([1, 0, 2, 4, 3], [4, 3, 2, 1, 0], [2, 4, 0, 3, 1]):
        assert abs(float(np.sum(simplex_inverse(w[perm])[0] ** 2)) - ref) < 2e-14


@pytest.mark.parametrize("K", [2, 3, 7])
def test_log_jacobian_closed_form_and_finite_difference(K):
    # test/transforms.jl:434-456: log|J| = log(K)/2 + sum(log, x), and for K = 3 the
    # determinant of the first K - 1 outputs' Jacobian by central differences
    for T in (np.float32, np.float64):
        x, lj, _ = simplex_forward(np.zeros(K - 1, T), T)
        expected = T(0.5) * np.log(T(K)) + np.sum(np.log(x), dtype=T)
        assert abs(float(lj) - float(expected)) <= 8 * EPS[T] * max(1.0, abs(float(expected)))
    z = np.array([0.2, -0.4])
    x, lj, _ = simplex_forward_c(z)
    h = np.cbrt(EPS[np.float64])
    J = np.zeros((2, 2))
    for c in range(2):
        off = np.zeros(2)
        off[c] = h
        J[:, c] = (simplex_forward_c(z + off)[0][:2] - simplex_forward_c(z - off)[0][:2]) / (2 * h)
    assert abs(math.log(abs(np.linalg.det(J))) - lj) < 2e-9


def test_reason_bits_in_the_reference_order():
    # test/transforms.jl:458-470 and the draw failure of test/kernel_execution.jl:623-640
    assert simplex_forward([0.0, math.inf])[2] == TRANSFORM_NONFINITE_INPUT
    assert simplex_forward_c([0.0, math.inf])[2] == TRANSFORM_NONFINITE_INPUT
    assert simplex_inverse([0.0, 0.5, 0.5])[2] == TRANSFORM_OUTSIDE_SUPPORT
    assert simplex_inverse([0.2, 0.3, 0.6])[2] == TRANSFORM_OUTSIDE_SUPPORT
    assert simplex_inverse([0.2, math.nan, 0.8])[2] == TRANSFORM_NONFINITE_INPUT
    big = np.finfo(np.float64).max
    assert simplex_forward_c([big, big])[2] == TRANSFORM_NONFINITE_OUTPUT
    for T in (np.float32, np.float64):
        m = np.finfo(T).max
        assert simplex_forward(np.array([m, m], T), T)[2] == TRANSFORM_NONFINITE_OUTPUT


def test_numpy_and_c_restatements_agree():
    rng = np.random.default_rng(3)
    for K in (2, 3, 5, 17, 64):
        for _ in range(20):
            z = rng.normal(0, 2, K - 1)
            x, lj, r = simplex_forward(z)
            xc, ljc, rc = simplex_forward_c(z)
            assert r == rc == 0
            assert np.max(np.abs(x - xc) / xc) < 8 * EPS[np.float64]
            assert abs(lj - ljc) <= 8 * EPS[np.float64] * max(1.0, abs(lj))


FACTOR = np.array([[1.0, 0.0, 0.0, 0.0],
                   [0.3, 1.2, 0.0, 0.0],
                   [-0.2, 0.4, 0.8, 0.0],
                   [0.1, -0.3, 0.25, 1.5]])
LAYOUT = [(0, 2, 4), (2, 1, 1), (3, 1, 0)]  # weights = 1:2 => Simplex(3), rate = 3 => Positive, offset = 4


def _structured_logdensity(x, factor):
    """_dependent_structured_logdensity (test/product_proposal.jl:54-74): the proposal's
    pushforward density at the logical sample (weights(3), rate, offset)."""
    lw = np.log(x[:3])
    mean_lw = lw.sum() / 3
    inv_root = 1 / math.sqrt(3)
    coef = inv_root / (1 - inv_root)
    last = lw[2] - mean_lw
    coords = np.array([lw[0] - mean_lw + coef * last, lw[1] - mean_lw + coef * last, math.log(x[3]), x[4]])
    std = np.linalg.solve(factor, coords)
    base = -2 * math.log(2 * math.pi) - np.sum(np.log(np.diag(factor))) - 0.5 * np.sum(std ** 2)
    return base - (0.5 * math.log(3) + lw.sum() + math.log(x[3]))


def test_layout_preserves_the_normalized_chart_measure():
    # test/product_proposal.jl:283-320: target = pushforward of the proposal => logw == 0
    rng = np.random.default_rng(5)
    z = (FACTOR @ rng.standard_normal((4, 256))).T
    x, lj, reason, block = layout_forward(LAYOUT, z)
    assert np.all(reason == 0) and np.all(block == 0)
    assert layout_dim(LAYOUT) == 5 and np.allclose(x[:, :3].sum(axis=1), 1.0, atol=4e-16)
    base = np.array([-2 * math.log(2 * math.pi) - np.sum(np.log(np.diag(FACTOR)))
                     - 0.5 * np.sum(np.linalg.solve(FACTOR, zi) ** 2) for zi in z])
    lw = np.array([_structured_logdensity(xi, FACTOR) for xi in x]) + lj - base
    assert np.max(np.abs(lw)) <= 256 * EPS[np.float64]
    # each block is the scalar or simplex map of its own coordinates, the log Jacobians
    # summed right to left (src/native_execution.jl:688-712)
    for i in range(8):
        xs, ljs, _ = simplex_forward_c(z[i, :2])
        assert np.array_equal(x[i, :3], xs) and x[i, 3] == math.exp(z[i, 2]) and x[i, 4] == z[i, 3]
        assert lj[i] == ljs + (z[i, 2] + 0.0)


def test_layout_target_and_failing_block():
    z = np.array([[0.1, -0.2, 0.3, 0.4], [0.0, 0.0, 800.0, 1.0], [np.inf, 0.0, 0.0, 0.0]])
    x, lj, reason, block = layout_forward(LAYOUT, z)
    assert list(block) == [0, 2, 1]
    assert reason[1] == TRANSFORM_NONFINITE_OUTPUT and reason[2] == TRANSFORM_NONFINITE_INPUT
    tk, tp = layout_target(0, None, LAYOUT)
    assert target(tk, tp, z[0]) == pytest.approx(-0.5 * np.sum(x[0] ** 2) + lj[0], rel=1e-15)
    assert math.isnan(target(tk, tp, z[1]))
