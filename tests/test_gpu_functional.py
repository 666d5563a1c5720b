"""GPU tests of mean(f, result) / var(f, result) (src/statistics.jl:17-31,50-64,144-199)
through the C ABI: the reference's known answers on device arrays (FP64 and FP32
log-weights and values), the variance range case (test/cuda_statistics.jl:27-38), the
zero-weight rules, and jobs' device-resident log-weights (AMIS in FP32 and FP64, plain
IS) with values passed from the host or from device memory, against the oracle fed the
same log-weights and values.  Device arrays are allocated through the engine's own HIP
runtime (libamdhip64.so.7, loaded by the library), so no second runtime is involved."""

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


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the engine library already loaded."""

    def __init__(self):
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        self.h = h
        self.live = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(1, a.nbytes)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device
        self.live.append(p)
        return p

    def close(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


@pytest.fixture
def hip(eng):
    h = _Hip()
    yield h
    h.close()


def _device_moments(eng, hip, lw, vals, want_var=True):
    """is_functional_moments_device over uploaded numpy arrays (their dtypes select FP32 /
    FP64); returns (rc, mean, var)."""
    lib = eng._lib.load()
    m, v = np.zeros(1), np.zeros(1)
    rc = lib.is_functional_moments_device(eng.context(0), hip.upload(lw), int(lw.dtype == np.float64),
                                          hip.upload(vals), int(vals.dtype == np.float64), lw.size,
                                          m.ctypes.data_as(eng._lib.F64P),
                                          v.ctypes.data_as(eng._lib.F64P) if want_var else None, None)
    return rc, float(m[0]), float(v[0])


@pytest.mark.parametrize("lw_dtype,v_dtype", [(np.float64, np.float64), (np.float32, np.float32),
                                              (np.float32, np.float64)])
def test_known_answers_on_device(eng, oracle_mod, hip, lw_dtype, v_dtype):
    _say(f"case: functional known answers, log-weights {lw_dtype.__name__}, values {v_dtype.__name__}")
    lw = (1000 + np.log([1.0, 2.0, 1.0])).astype(lw_dtype)  # FP32 rounds 1000 + log 2 at 6e-5
    s = np.array([[1.0, 3.0, 10.0], [2.0, 4.0, 8.0]])
    tol = 1e-12 if lw_dtype == np.float64 else 4e-4  # reference: isapprox at sqrt(eps(Float32))
    for f, m_ref, v_ref in [(s[0], 4.25, 11.6875), (s[0] ** 2, 29.75, 1655.6875), (s[0] + 2 * s[1], 13.25, 60.1875)]:
        vals = f.astype(v_dtype)
        rc, m, v = _device_moments(eng, hip, lw, vals)
        assert rc == 0
        assert m == pytest.approx(m_ref, rel=tol) and v == pytest.approx(v_ref, rel=tol)
        om, ov = oracle_mod.functional_moments(lw.astype(np.float64), vals.astype(np.float64))
        assert m == pytest.approx(om, rel=1e-13) and v == pytest.approx(ov, rel=1e-12)


@pytest.mark.parametrize("logmass", [math.log(1e-20), -math.inf])
def test_variance_range_on_device(eng, hip, logmass):
    _say(f"case: functional variance range, logmass {logmass}")
    # test/cuda_statistics.jl:27-38 (FP32 samples [0, 1e20]); the FP64 reduction needs no rescue
    lw = np.array([0.0, logmass], dtype=np.float32)
    vals = np.array([0.0, 1e20], dtype=np.float32)
    mass = math.exp(float(lw[1]))
    x = float(vals[1])
    expected = mass * x * x / (1 + mass) ** 2
    rc, _, v = _device_moments(eng, hip, lw, vals)
    assert rc == 0
    assert v == pytest.approx(expected, rel=8 * 1.1920929e-07, abs=0.0)


def test_zero_weight_rules_on_device(eng, oracle_mod, hip):
    _say("case: functional zero-weight rules")
    lw = np.array([0.0, -math.inf, 0.5])
    rc, m, _ = _device_moments(eng, hip, lw, np.array([1.0, math.inf, 2.0]), want_var=False)
    assert rc == 0 and math.isnan(m)
    rc, m, v = _device_moments(eng, hip, lw, np.array([1.0, 1e300, 2.0]))
    om, ov = oracle_mod.functional_moments(lw, [1.0, 1e300, 2.0])
    assert rc == 0 and m == pytest.approx(om, rel=1e-14) and v == pytest.approx(ov, rel=1e-13)
    rc, _, _ = _device_moments(eng, hip, np.full(5, -math.inf), np.ones(5))
    assert rc == eng._lib.IS_EFAILED


def test_many_blocks_on_device(eng, oracle_mod, hip):
    _say("case: functional over 3e6 values (many partial records)")
    rng = np.random.default_rng(5)
    n = 3_000_000
    lw = rng.normal(0, 3.0, n)
    vals = rng.normal(2.0, 1.0, n)
    rc, m, v = _device_moments(eng, hip, lw, vals)
    om, ov = oracle_mod.functional_moments(lw, vals)
    assert rc == 0 and m == pytest.approx(om, rel=1e-11) and v == pytest.approx(ov, rel=1e-11)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_job_functional_moments(eng, oracle_mod, hip, precision):
    _say(f"case: mean(f, result) / var(f, result) on an AMIS result, {precision}")
    d = 4
    rng = np.random.default_rng(9)
    mu = rng.normal(0, 1, d)
    A = rng.normal(0, 0.3, (d, d))
    S = A @ A.T + np.eye(d)
    tgt = eng.GaussianTarget(mu, S)
    alg = eng.AMIS(eng.SphericalGaussian(np.zeros(d), 3.0), rounds=4, round_size=5000)
    res = eng.importance_sample(11, tgt, alg, precision=precision)
    x = res.samples.reshape(len(res), d)

    def f(x):
        return x[:, 0] + 2.0 * x[:, 1] * x[:, 2]

    m = eng.mean(f, res)
    v = eng.var(f, res)
    om, ov = oracle_mod.functional_moments(res.logweights, f(x))
    assert m == pytest.approx(om, rel=1e-11, abs=1e-13)
    assert v == pytest.approx(ov, rel=1e-11)
    # the same values from device memory (the device functional path)
    dv = eng.DeviceValues(hip.upload(f(x)).value, fp64=True)
    assert eng.mean(dv, res) == m and eng.var(dv, res) == v
    dv32 = eng.DeviceValues(hip.upload(f(x).astype(np.float32)).value, fp64=False)
    om32, _ = oracle_mod.functional_moments(res.logweights, f(x).astype(np.float32))
    assert eng.mean(dv32, res) == pytest.approx(om32, rel=1e-11, abs=1e-13)
    # the identity functional agrees with the coordinate mean of the moments path (which, for
    # FP32 storage, forms each weight's factor sqrt(e) in FP32: 1e-7 relative)
    tol = 1e-10 if precision == "fp64" else 1e-6
    assert eng.mean(lambda x: x[:, 0], res) == pytest.approx(eng.mean(res)[0], rel=tol, abs=1e-12)
    with pytest.raises(ValueError):
        eng.mean(lambda x: x[:10, 0], res)


def test_plain_is_functional_d1(eng, oracle_mod):
    _say("case: mean(f, result) on plain IS, d = 1")
    res = eng.importance_sample(4, eng.StandardGaussianTarget(),
                                eng.ImportanceSampling(eng.SphericalGaussian(np.zeros(1), 1.5), 20000))
    m = eng.mean(lambda x: x * x, res)
    om, _ = oracle_mod.functional_moments(res.logweights, res.samples.reshape(-1) ** 2)
    assert m == pytest.approx(om, rel=1e-12)
    assert abs(m - 1.0) < 0.05  # E[x^2] = 1 under N(0, 1)
