"""GPU test of is_functional_moments_device_noalloc, the allocation-free form of
is_functional_moments_device (mean(f, result) / var(f, result) over device arrays,
src/statistics.jl:17-31,50-64): two calls in a row, the second on a smaller input, each
against the allocating entry point on the same device arrays (same kernels and
reduction order, so the results must be identical), plus the size limit."""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_SAMPLES = (131072 - 8) // 4 * 2048  # the 1 MiB scratch's limit, as the header states


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the engine library already loaded
    (test-side device arrays; the entry point under test allocates nothing)."""

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
        assert self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        self.live.append(p)
        return p

    def close(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


@pytest.fixture
def hip(pkg):
    pkg.context(0)
    h = _Hip()
    yield h
    h.close()


def _moments(pkg, fn, lwp, lw64, vp, v64, n):
    lib = pkg._lib.load()
    m, v = np.zeros(1), np.zeros(1)
    rc = getattr(lib, fn)(pkg.context(0), lwp, int(lw64), vp, int(v64), n, m.ctypes.data_as(pkg._lib.F64P),
                          v.ctypes.data_as(pkg._lib.F64P), None)
    return rc, float(m[0]), float(v[0])


@pytest.mark.parametrize("lw_dtype,v_dtype", [(np.float64, np.float64), (np.float32, np.float32)])
def test_noalloc_matches_allocating_entry_point(pkg, hip, lw_dtype, v_dtype):
    rng = np.random.default_rng(11)
    first = 3_000_001  # 1,465 partial records
    lw = (rng.normal(size=first) * 3.0 - 700.0).astype(lw_dtype)
    lw[::97] = -np.inf  # zero-weight samples
    vals = (rng.normal(size=first) * 2.0 + 1.5).astype(v_dtype)
    lwp, vp = hip.upload(lw), hip.upload(vals)
    lw64, v64 = lw_dtype == np.float64, v_dtype == np.float64
    for n in (first, 70_000):  # the second call is smaller: stale records past its end must not leak in
        _say(f"case: noalloc functional moments, n = {n}, {lw_dtype.__name__}/{v_dtype.__name__}")
        rc_a, m_a, v_a = _moments(pkg, "is_functional_moments_device", lwp, lw64, vp, v64, n)
        rc_n, m_n, v_n = _moments(pkg, "is_functional_moments_device_noalloc", lwp, lw64, vp, v64, n)
        assert rc_a == 0 and rc_n == 0, pkg._lib.last_error()
        assert m_n == m_a and v_n == v_a  # identical kernels and order
        # and the numbers themselves, against an FP64 evaluation of the same prefix
        w = np.exp(lw[:n].astype(np.float64) - np.max(lw[:n].astype(np.float64)))
        f = vals[:n].astype(np.float64)
        mu = np.sum(w * f) / np.sum(w)
        assert abs(m_n - mu) < 1e-10 * max(1.0, abs(mu))
        assert abs(v_n - np.sum(w * (f - mu) ** 2) / np.sum(w)) < 1e-9


def test_noalloc_rejects_sizes_beyond_its_scratch(pkg):
    lib = pkg._lib.load()
    m = np.zeros(1)
    # the size check comes before any device work, so the pointers are never read
    rc = lib.is_functional_moments_device_noalloc(pkg.context(0), ctypes.c_void_p(16), 1, ctypes.c_void_p(16), 1,
                                                  MAX_SAMPLES + 1, m.ctypes.data_as(pkg._lib.F64P), None, None)
    assert rc == pkg._lib.IS_EINVAL
    assert "allocation-free limit" in pkg._lib.last_error()
