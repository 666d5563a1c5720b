"""CPU FP64 oracle for the importance-sampling hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker (or the timed CPU baseline).  The product never imports it.

The arithmetic lives in ``oracle/oracle.cpp`` (a C++17 restatement of the
reference's CPU path, built by ``oracle/Makefile`` into ``liboracle.so``);
this module is its ctypes wrapper plus the exact-rational DM-PMC allocation
plan (``allocation_plan``), which restates
``src/population_preparation.jl:266-367`` with ``fractions.Fraction``.

Parity pins (what checks each oracle function; tests/test_oracle_golden.py):

* ``logaddexp`` / ``log1pexp`` -- LogExpFunctions 1.0.1 thresholds; checked
  against mpmath at the branch boundaries and against the append identity of
  ``test/log_mixture_accumulator.jl:7-16``.
* ``student_t_logconstant`` -- mpmath log-gamma closed form, all parities of d.
* ``proposal_logdensity`` -- the analytic densities of
  ``test/native_proposals.jl:179-306`` (Gaussian and Student-t, spherical,
  diagonal, factor).
* ``plain_is`` -- normalizer identity ``test/native_proposals.jl:558-597``
  (self-target logw == 0, log Z = log 2pi + log 3), Student-t buffered draws
  ``test/kernel_execution.jl:490-543`` (exact samples, exhaustion failure).
* ``adaptive`` (AMIS) -- the frozen round ESS / log Z of
  ``test/amis_kernel.jl:918-967`` (exact zero in round 1), the ridge
  bitstrings of ``test/amis_kernel.jl:282-329``, and the literal AMIS spec of
  ``validation/reproducers/amis.jl:190-309`` (independent numpy restatement,
  tolerance 4096*max(d,T)*eps).
* ``adaptive`` (NPMC) -- the ``npmc_oracle`` recurrence of
  ``test/npmc.jl:15-66`` (numpy restatement with clipping at the isqrt(n)-th
  largest weight) and the self-target case ``test/npmc.jl:96-111``.
* ``dmpmc`` -- realized denominator ``test/dm_pmc_kernel.jl:15-66``;
  ``allocation_plan`` -- ``test/dm_pmc.jl:71-155`` (exact integer counts).
* ``select_ancestors`` -- ``test/dm_pmc_kernel.jl:205-272`` and
  ``test/resampling.jl:54-66`` (bit-exact ancestors, plateau cases).
* ``systematic`` selection -- closed-form counts
  ceil(M*cdf_k - u) - ceil(M*cdf_{k-1} - u) in exact rationals (the reference
  has no systematic method; the mathematics fixes it).
* ``logweight_summary`` / ``lognormalizer`` / ``moments`` --
  ``test/dm_pmc_kernel.jl:274-296`` and ``test/results.jl:9-75,357-388``.
* ``cholesky`` -- numpy.linalg.cholesky on random SPD matrices.
* ``transform_forward`` -- the scalar transforms of ``src/transforms.jl:160-235``
  (numpy restatement): log Jacobian against a central finite difference of the
  forward map (the check of ``test/transforms.jl:3-8``), supports, and the
  closed forms of the positive and symmetric interval maps.
* ``functional_moments`` -- the known answers of ``test/results.jl:9-38``
  (mean/var of abs2 and of x1 + 2 x2) and the variance range case of
  ``test/cuda_statistics.jl:27-38`` (exact rationals), plus the zero-weight
  rules of ``src/statistics.jl:188-197``.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

F64P = ctypes.POINTER(ctypes.c_double)
I64P = ctypes.POINTER(ctypes.c_int64)
U64P = ctypes.POINTER(ctypes.c_uint64)
IP = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    d, i, i64 = ctypes.c_double, ctypes.c_int, ctypes.c_int64
    L.or_logaddexp.restype = d
    L.or_logaddexp.argtypes = [d, d]
    L.or_log1pexp.restype = d
    L.or_log1pexp.argtypes = [d]
    L.or_det_exp.restype = d
    L.or_det_exp.argtypes = [d]
    L.or_student_t_logconstant.restype = d
    L.or_student_t_logconstant.argtypes = [d, i]
    L.or_proposal_logdensity.restype = d
    L.or_proposal_logdensity.argtypes = [i, d, i, i, F64P, F64P, d, F64P, i]
    L.or_target.restype = d
    L.or_target.argtypes = [i, F64P, i, F64P]
    L.or_transform_forward.restype = i
    L.or_transform_forward.argtypes = [i, d, d, d, F64P, F64P]
    L.or_student_t_multiplier.restype = i
    L.or_student_t_multiplier.argtypes = [d, F64P, F64P, F64P]
    L.or_plain_is.restype = None
    L.or_plain_is.argtypes = [i, d, i, i, F64P, F64P, d, i, F64P, i64, F64P, F64P, i, F64P, F64P, U64P]
    L.or_logweight_summary.restype = None
    L.or_logweight_summary.argtypes = [F64P, i64, F64P, F64P]
    L.or_merge_logweight_moments.restype = None
    L.or_merge_logweight_moments.argtypes = [F64P, F64P, F64P]
    L.or_logsumexp_accumulator.restype = None
    L.or_logsumexp_accumulator.argtypes = [F64P, i64, F64P]
    L.or_lognormalizer.restype = d
    L.or_lognormalizer.argtypes = [F64P, i64]
    L.or_normalized_weights.restype = i
    L.or_normalized_weights.argtypes = [F64P, i64, F64P]
    L.or_moments.restype = i
    L.or_moments.argtypes = [F64P, F64P, i64, i, F64P, F64P, F64P]
    L.or_cholesky.restype = i
    L.or_cholesky.argtypes = [F64P, i, F64P]
    L.or_cdf_sequential.restype = i
    L.or_cdf_sequential.argtypes = [F64P, i64, F64P]
    L.or_cdf_blocked.restype = i
    L.or_cdf_blocked.argtypes = [F64P, i64, F64P, i, i]
    L.or_select_ancestors.restype = None
    L.or_select_ancestors.argtypes = [F64P, F64P, i64, F64P, i64, I64P]
    L.or_local_select.restype = i
    L.or_local_select.argtypes = [F64P, I64P, i, F64P, I64P, i]
    I32P = ctypes.POINTER(ctypes.c_int32)
    L.or_simplex_forward.restype = i
    L.or_simplex_forward.argtypes = [i, F64P, F64P, F64P]
    L.or_layout_forward.restype = None
    L.or_layout_forward.argtypes = [i, F64P, i, i64, F64P, F64P, F64P, I32P, I32P]
    L.or_systematic_points.restype = None
    L.or_systematic_points.argtypes = [d, i64, F64P]
    L.or_adaptive.restype = i
    L.or_adaptive.argtypes = [i, i, d, i, i, F64P, F64P, d, i, I64P, F64P, F64P, F64P, i, F64P, i,
                              F64P, F64P, F64P, F64P, I64P, F64P, F64P, F64P, F64P, F64P, F64P,
                              U64P, IP, i, F64P, F64P, F64P]
    L.or_dmpmc.restype = i
    L.or_dmpmc.argtypes = [i, d, i, i, i, F64P, F64P, F64P, I64P, i, I64P, I64P, F64P, i, i, i,
                           F64P, F64P, F64P, F64P, i, F64P, i, F64P, F64P, I64P, I64P, I64P,
                           F64P, F64P, F64P, U64P, IP]


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _p(a, t=F64P):
    if a is None:
        return None
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------- scalars
def logaddexp(a, b):
    return lib().or_logaddexp(float(a), float(b))


def log1pexp(x):
    return lib().or_log1pexp(float(x))


def det_exp(x):
    return lib().or_det_exp(float(x))


def student_t_logconstant(dof, d):
    return lib().or_student_t_logconstant(float(dof), int(d))


# ---------------------------------------------------------------- proposals
class Slot:
    """Flat proposal description: family (0 Gaussian, 1 Student-t), dof,
    kind (0 spherical, 1 diagonal, 2 factor), location, scale storage, lognormalizer."""

    def __init__(self, family, dof, kind, location, scale, lognormalizer):
        self.family = int(family)
        self.dof = float(dof)
        self.kind = int(kind)
        self.location = _f64(np.atleast_1d(location))
        self.d = self.location.size
        if kind == 2:
            self.scale = _f64(np.asarray(scale, dtype=np.float64).reshape(self.d, self.d).T.ravel())
        else:
            self.scale = _f64(np.atleast_1d(scale))
        self.lognormalizer = float(lognormalizer)


def as_slot(p):
    """Accept any object exposing family/dof/kind/location/scale/lognormalizer."""
    if isinstance(p, Slot):
        return p
    return Slot(p.family, p.dof, p.kind, p.location, p.scale, p.lognormalizer)


def proposal_logdensity(p, x, packed=False):
    s = as_slot(p)
    x = _f64(np.atleast_1d(x))
    return lib().or_proposal_logdensity(s.family, s.dof, s.kind, s.d, _p(s.location), _p(s.scale),
                                        s.lognormalizer, _p(x), int(bool(packed)))


TRANSFORMED_TARGET = 5


def transformed_target(inner_kind, inner_params, kinds, bounds):
    """(kind, params) of the target inner(T(z)) + sum log|dT_c/dz_c| for per-coordinate
    transforms (kinds: d codes, bounds: d (lower, upper) pairs), the engine's
    transform= contract; a failing transform gives NaN."""
    kinds = np.asarray(kinds, dtype=np.float64).ravel()
    bounds = _f64(bounds).ravel()
    inner = _f64(inner_params if inner_params is not None else [0.0]).ravel()
    return TRANSFORMED_TARGET, np.concatenate([[float(inner_kind)], kinds, bounds, inner])


LAYOUT_TARGET = 6


def layout_blocks(layout):
    """The (B, 5) block table (first, count, kind, lower, upper) of a flat selector layout
    given as a list of (first, count, kind[, lower, upper]) or as the engine's C blocks."""
    rows = []
    for b in layout:
        if hasattr(b, "first"):
            rows.append([b.first, b.count, b.kind, b.lower, b.upper])
        else:
            b = list(b) + [-math.inf, math.inf][len(b) - 3:]
            rows.append([float(v) for v in b[:5]])
    return _f64(rows).reshape(-1, 5)


def layout_dim(layout):
    t = layout_blocks(layout)
    return int(sum(int(c) + (1 if int(k) == 4 else 0) for c, k in zip(t[:, 1], t[:, 2])))


def layout_target(inner_kind, inner_params, layout):
    """(kind, params) of the target inner(T(z)) + log|J| for a flat selector layout (inner
    over the logical dimension); a failing block gives NaN."""
    t = layout_blocks(layout)
    inner = _f64(inner_params if inner_params is not None else [0.0]).ravel()
    return LAYOUT_TARGET, np.concatenate([[float(inner_kind), float(t.shape[0])], t.ravel(), inner])


def layout_forward(layout, z):
    """The C restatement of _native_transform_flat_blocks! over samples z (n, d): logical
    samples (n, dl), log Jacobians (n,), reason bits (n,) and failing blocks (n,)."""
    t = layout_blocks(layout)
    z = _f64(np.atleast_2d(z))
    n, d = z.shape
    dl = layout_dim(layout)
    x, lj = np.zeros((n, dl)), np.zeros(n)
    reason, block = np.zeros(n, np.int32), np.zeros(n, np.int32)
    I32P = ctypes.POINTER(ctypes.c_int32)
    lib().or_layout_forward(t.shape[0], _p(t), d, n, _p(z), _p(x), _p(lj), _p(reason, I32P), _p(block, I32P))
    return x, lj, reason, block


def simplex_forward_c(z):
    """The C restatement of the FP64 simplex forward map: (x, log|J|, reason bits)."""
    z = _f64(np.atleast_1d(z))
    x, lj = np.zeros(z.size + 1), ctypes.c_double()
    r = lib().or_simplex_forward(z.size, _p(z), _p(x), ctypes.byref(lj))
    return x, lj.value, int(r)


def _cadd(total, correction, value):
    # _compensated_add (src/transforms.jl:347-352)
    corrected = value - correction
    updated = total + corrected
    return updated, (updated - total) - corrected


def simplex_forward(z, dtype=np.float64):
    """_transform_with_logjac(SimplexTransform(K), z) (src/transforms.jl:400-449) in the
    scalar type `dtype` (numpy scalars round every operation to it): (x, log|J|, reason
    bits), the checks in the reference's order."""
    T = np.dtype(dtype).type
    z = [T(v) for v in np.atleast_1d(z)]
    K = len(z) + 1
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        total, corr = T(0), T(0)
        for v in z:
            if not np.isfinite(v):
                return None, T(0), TRANSFORM_NONFINITE_INPUT
            total, corr = _cadd(total, corr, v)
        if not np.isfinite(total):
            return None, T(0), TRANSFORM_NONFINITE_OUTPUT
        inv_root = T(1) / np.sqrt(T(K))
        shared = (T(1) + inv_root) / T(K - 1)
        maxl = inv_root * total
        x = [T(0)] * K
        x[K - 1] = maxl
        for i, v in enumerate(z):
            logit = v - shared * total
            if not np.isfinite(logit):
                return None, T(0), TRANSFORM_NONFINITE_OUTPUT
            x[i] = logit
            maxl = max(maxl, logit)
        es, ec = T(0), T(0)
        for i in range(K):
            x[i] = np.exp(x[i] - maxl)
            es, ec = _cadd(es, ec, x[i])
        if not np.isfinite(es):
            return None, T(0), TRANSFORM_NONFINITE_OUTPUT
        lj, lc = T(0.5) * np.log(T(K)), T(0)
        for i in range(K):
            w = x[i] / es
            if not w > 0:
                return None, lj, TRANSFORM_OUTSIDE_SUPPORT
            x[i] = w
            lj, lc = _cadd(lj, lc, np.log(w))
        if not np.isfinite(lj):
            return np.array(x, dtype=dtype), lj, TRANSFORM_NONFINITE_LOGJAC
    return np.array(x, dtype=dtype), lj, 0


def simplex_inverse(x, dtype=np.float64):
    """_inverse_with_logjac(SimplexTransform(K), x) via _native_simplex_inverse!
    (src/transforms.jl:354-394,451-456) in `dtype`: (z, log|J| of the forward map, reason
    bits); weights must be positive, finite and sum to one within 8 eps."""
    T = np.dtype(dtype).type
    x = [T(v) for v in np.atleast_1d(x)]
    K = len(x)
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        ws, wc, ls, lc = T(0), T(0), T(0), T(0)
        z = [T(0)] * (K - 1)
        last = T(0)
        for i, w in enumerate(x):
            if not np.isfinite(w):
                return None, T(0), TRANSFORM_NONFINITE_INPUT
            if not w > 0:
                return None, T(0), TRANSFORM_OUTSIDE_SUPPORT
            ws, wc = _cadd(ws, wc, w)
            lw = np.log(w)
            ls, lc = _cadd(ls, lc, lw)
            if i < K - 1:
                z[i] = lw
            else:
                last = lw
        if not np.isfinite(ws):
            return None, T(0), TRANSFORM_NONFINITE_INPUT
        if not abs(ws - T(1)) <= T(8) * np.finfo(dtype).eps:
            return None, T(0), TRANSFORM_OUTSIDE_SUPPORT
        mean_lw = ls / T(K)
        inv_root = T(1) / np.sqrt(T(K))
        coef = inv_root / (T(1) - inv_root)
        last_c = last - mean_lw
        z = [zi - mean_lw + coef * last_c for zi in z]
        lj = T(0.5) * np.log(T(K)) + ls
        if not np.isfinite(lj):
            return np.array(z, dtype=dtype), lj, TRANSFORM_NONFINITE_LOGJAC
    return np.array(z, dtype=dtype), lj, 0


def transform_forward_c(kind, lo, hi, z):
    """The C restatement of the scalar forward transforms: (x, log|J|, reason bits)."""
    x, lj = ctypes.c_double(), ctypes.c_double()
    r = lib().or_transform_forward(int(kind), float(lo), float(hi), float(z), ctypes.byref(x), ctypes.byref(lj))
    return x.value, lj.value, int(r)


def target(kind, params, x):
    x = _f64(np.atleast_1d(x))
    params = _f64(params if params is not None else [0.0])
    return lib().or_target(int(kind), _p(params), x.size, _p(x))


def student_t_multiplier(dof, radial_normals, radial_uniforms=None):
    out = ctypes.c_double(0.0)
    rn = _f64(radial_normals)
    ru = None if radial_uniforms is None else _f64(radial_uniforms)
    ok = lib().or_student_t_multiplier(float(dof), _p(rn), _p(ru), ctypes.byref(out))
    return (out.value if ok == 0 else None)


def normal_stride(family, dof, d):
    return d + (0 if family == 0 else (1 if dof == 1.0 else 8))


def uniform_stride(family, dof):
    return 0 if (family == 0 or dof == 1.0) else 8 + (1 if dof < 2.0 else 0)


def radial_normal_stride(family, dof):
    return 0 if family == 0 else (1 if dof == 1.0 else 8)


# ---------------------------------------------------------------- plain IS
def plain_is(p, tkind, tparams, normals, uniforms=None, nthreads=1):
    """Returns (samples (n, d), logweights (n,), failure uint64[3])."""
    s = as_slot(p)
    stride = normal_stride(s.family, s.dof, s.d)
    normals = _f64(normals)
    n = normals.size // stride
    uniforms = None if uniforms is None else _f64(uniforms)
    samples = np.zeros((n, s.d))
    logw = np.full(n, np.nan)
    fail = np.zeros(3, dtype=np.uint64)
    tp = _f64(tparams if tparams is not None else [0.0])
    lib().or_plain_is(s.family, s.dof, s.kind, s.d, _p(s.location), _p(s.scale), s.lognormalizer,
                      int(tkind), _p(tp), n, _p(normals), _p(uniforms), int(nthreads),
                      _p(samples), _p(logw), _p(fail, U64P))
    return samples, logw, fail


# ---------------------------------------------------------------- summaries / results
def logweight_summary(lw):
    lw = _f64(lw)
    ess, logz = ctypes.c_double(), ctypes.c_double()
    lib().or_logweight_summary(_p(lw), lw.size, ctypes.byref(ess), ctypes.byref(logz))
    return ess.value, logz.value


def merge_logweight_moments(a, b):
    a, b = _f64(a), _f64(b)
    out = np.zeros(3)
    lib().or_merge_logweight_moments(_p(a), _p(b), _p(out))
    return out


def logsumexp_accumulator(lw):
    lw = _f64(lw)
    out = np.zeros(2)
    lib().or_logsumexp_accumulator(_p(lw), lw.size, _p(out))
    return out


def lognormalizer(lw):
    lw = _f64(lw)
    return lib().or_lognormalizer(_p(lw), lw.size)


def normalized_weights(lw):
    lw = _f64(lw)
    w = np.zeros(lw.size)
    if lib().or_normalized_weights(_p(lw), lw.size, _p(w)):
        raise ZeroDivisionError("all weights are zero")
    return w


def moments(samples, lw):
    """samples (n, d) sample-major; returns mean (d,), var (d,), cov (d, d)."""
    x = _f64(samples)
    if x.ndim == 1:
        x = x.reshape(-1, 1)
    n, d = x.shape
    lw = _f64(lw)
    mean, var, cov = np.zeros(d), np.zeros(d), np.zeros(d * d)
    if lib().or_moments(_p(x), _p(lw), n, d, _p(mean), _p(var), _p(cov)):
        raise ZeroDivisionError("all weights are zero")
    return mean, var, cov.reshape(d, d).T.copy()


def _softplus(z):
    return np.where(z > 0, z + np.log1p(np.exp(-np.abs(z))), np.log1p(np.exp(np.minimum(z, 0))))


def transform_forward(kind, lo, hi, z):
    """src/transforms.jl:160-235 with the helpers of :295-333, elementwise over z:
    returns (x, log|dx/dz|).  kind 0 identity, 1 positive, 2 softplus, 3 interval."""
    z = _f64(z)
    with np.errstate(over="ignore", invalid="ignore"):
        if kind == 0:
            return z.copy(), np.zeros_like(z)
        if kind == 1:
            return np.exp(z), z.copy()
        if kind == 2:
            x = _softplus(z)
            return x, z - x
        if math.isfinite(lo) and not math.isfinite(hi):
            return lo + np.exp(z), z.copy()
        if not math.isfinite(lo) and math.isfinite(hi):
            return hi - np.exp(z), z.copy()
        p = np.where(z >= 0, 1.0 / (1.0 + np.exp(-np.abs(z))), np.exp(-np.abs(z)) / (1.0 + np.exp(-np.abs(z))))
        if lo < 0 < hi:
            x = (1.0 - p) * lo + p * hi
            span = math.log(max(-lo, hi)) + math.log1p(min(-lo, hi) / max(-lo, hi))
        else:
            x = np.where(p <= 0.5, lo + p * (hi - lo), hi - (1.0 - p) * (hi - lo))
            span = math.log(hi - lo)
        return x, span - _softplus(-z) - _softplus(z)


TRANSFORM_NONFINITE_INPUT, TRANSFORM_NONFINITE_OUTPUT = 0x0001, 0x0002
TRANSFORM_NONFINITE_LOGJAC, TRANSFORM_OUTSIDE_SUPPORT = 0x0004, 0x0008


def transform_forward_checked(kind, lo, hi, z, dtype=np.float64):
    """One scalar z through _native_transform_with_logjac (src/transforms.jl:160-235)
    with its checks in the reference's order: returns (x, log|dx/dz|, reason bits).  z is
    rounded to `dtype` (the storage type); the FP64 interval bounds promote the interval
    arithmetic to FP64, as Julia promotes T to the bounds' type."""
    T = np.dtype(dtype).type
    z = T(z)
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        if not np.isfinite(z):
            return z, (T(0) if kind == 0 else z), TRANSFORM_NONFINITE_INPUT
        if kind == 0:
            return z, T(0), 0
        if kind == 1:
            x = np.exp(z)
            if not np.isfinite(x):
                return x, z, TRANSFORM_NONFINITE_OUTPUT
            return x, z, (0 if x > 0 else TRANSFORM_OUTSIDE_SUPPORT)
        if kind == 2:
            x = z + np.log1p(np.exp(-z)) if z > 0 else np.log1p(np.exp(z))
            lj = z - x
            if not np.isfinite(x):
                return x, lj, TRANSFORM_NONFINITE_OUTPUT
            if not np.isfinite(lj):
                return x, lj, TRANSFORM_NONFINITE_LOGJAC
            return x, lj, (0 if x > 0 else TRANSFORM_OUTSIDE_SUPPORT)
        hl, hh = math.isfinite(lo), math.isfinite(hi)
        if hl != hh:
            dist = np.exp(z)
            if not np.isfinite(dist):
                return dist, z, TRANSFORM_NONFINITE_OUTPUT
            if not dist > 0:
                return dist, z, TRANSFORM_OUTSIDE_SUPPORT
            x = lo + float(dist) if hl else hi - float(dist)
            if not math.isfinite(x):
                return x, z, TRANSFORM_NONFINITE_OUTPUT
            return x, z, (0 if (x > lo if hl else x < hi) else TRANSFORM_OUTSIDE_SUPPORT)
        p = T(1) / (T(1) + np.exp(-z)) if z >= 0 else np.exp(z) / (T(1) + np.exp(z))
        if lo < 0 < hi:
            x = float(T(1) - p) * lo + float(p) * hi
            big, small = max(-lo, hi), min(-lo, hi)
            span = math.log(big) + math.log1p(small / big)
        else:
            x = lo + float(p) * (hi - lo) if p <= T(0.5) else hi - float(T(1) - p) * (hi - lo)
            span = math.log(hi - lo)

        def lsig(v):  # -softplus(-v) in T
            v = -v
            return -(v + np.log1p(np.exp(-v)) if v > 0 else np.log1p(np.exp(v)))

        lj = (span + float(lsig(z))) + float(lsig(-z))
        if not math.isfinite(x):
            return x, lj, TRANSFORM_NONFINITE_OUTPUT
        if not math.isfinite(lj):
            return x, lj, TRANSFORM_NONFINITE_LOGJAC
        return x, lj, (0 if lo < x < hi else TRANSFORM_OUTSIDE_SUPPORT)


def functional_moments(lw, values):
    """mean(f, result) and var(f, result) of src/statistics.jl:17-31,50-64,170-199 for
    per-sample values f_i: w = normalized_weights (src/results.jl:589-598),
    mean = dot(w, f) (a zero weight times a non-finite value gives NaN), and
    var = sum of _weighted_squared_deviation(f_i, mean, w_i): (f - c) (w (f - c)),
    recomputed as (sqrt(w) f - sqrt(w) c)^2 when f - c is not finite, 0 when w = 0."""
    w = normalized_weights(lw)
    f = _f64(values)
    with np.errstate(invalid="ignore", over="ignore"):
        mean = float(np.dot(w, f))
        dev = f - mean
        contrib = dev * (w * dev)
        bad = ~np.isfinite(dev)
        r = np.sqrt(w)
        contrib[bad] = (r[bad] * f[bad] - r[bad] * mean) ** 2
        contrib[w == 0.0] = 0.0
        var = float(np.sum(contrib))
    return mean, var


def cholesky(A):
    A = np.asarray(A, dtype=np.float64)
    d = A.shape[0]
    a = _f64(A.T.ravel())
    L = np.zeros(d * d)
    info = lib().or_cholesky(_p(a), d, _p(L))
    if info:
        raise np.linalg.LinAlgError(f"not positive definite at {info}")
    return L.reshape(d, d).T.copy()


# ---------------------------------------------------------------- resampling
def cdf_sequential(lw):
    lw = _f64(lw)
    cdf = np.zeros(lw.size)
    if lib().or_cdf_sequential(_p(lw), lw.size, _p(cdf)):
        raise ZeroDivisionError("all weights are zero")
    return cdf


def cdf_blocked(lw, threads=256, per=16):
    lw = _f64(lw)
    cdf = np.zeros(lw.size)
    if lib().or_cdf_blocked(_p(lw), lw.size, _p(cdf), int(threads), int(per)):
        raise ZeroDivisionError("all weights are zero")
    return cdf


def select_ancestors(cdf, lw, u):
    """1-based ancestors (binary search with plateau fix-up)."""
    cdf, lw, u = _f64(cdf), _f64(lw), _f64(np.atleast_1d(u))
    anc = np.zeros(u.size, dtype=np.int64)
    lib().or_select_ancestors(_p(cdf), _p(lw), cdf.size, _p(u), u.size, _p(anc, I64P))
    return anc


def local_select(lw, group_sizes, u, lanes=256):
    """LocalResampling: each group's 1-based round index (lanes 256: the engine's
    cooperative order; 0: the reference CPU kernel's); raises ZeroDivisionError when a
    group has no positive weight."""
    lw = _f64(lw)
    group = np.ascontiguousarray(np.concatenate([[0], np.cumsum(group_sizes)]).astype(np.int64))
    K = len(group_sizes)
    u = _f64(np.atleast_1d(u))
    anc = np.zeros(K, dtype=np.int64)
    if lib().or_local_select(_p(lw), _p(group, I64P), K, _p(u), _p(anc, I64P), int(lanes)):
        raise ZeroDivisionError("a local group has no positive weight")
    return anc


def systematic_points(u, count):
    pts = np.zeros(int(count))
    lib().or_systematic_points(float(u), int(count), _p(pts))
    return pts


def systematic_counts_exact(cdf, u):
    """Closed-form systematic counts ceil(M*cdf_k - u) - ceil(M*cdf_{k-1} - u),
    evaluated in exact rationals from the FP64 values of cdf and u (M = count)."""
    raise NotImplementedError  # tests compute it directly with a given M


# ---------------------------------------------------------------- AMIS / NPMC
def adaptive(method, p, schedule, normals, tkind, tparams, radial_normals=None,
             radial_uniforms=None, nthreads=1, scalar=None, replay=None):
    """method 0 AMIS, 1 NPMC.  normals: (T, d*maxN) per-round batches.

    replay=(means (T+1, d), factors (T+1, d, d) or (T+1,), lognormalizers (T+1,)):
    every round samples from the given slots; each round's own fit is returned
    in fit_means / fit_factors / fit_lognormalizers."""
    s = as_slot(p)
    schedule = np.ascontiguousarray(np.asarray(schedule, dtype=np.int64))
    T = schedule.size
    total = int(schedule.sum())
    maxN = int(schedule.max())
    d = s.d
    if scalar is None:
        scalar = getattr(p, "scalar", False)
    if scalar:
        scale0 = _f64([s.scale[0]])
    else:
        if s.kind == 2:
            scale0 = s.scale.copy()
        else:
            diag = np.full(d, s.scale[0]) if s.kind == 0 else s.scale
            scale0 = _f64(np.diag(diag).T.ravel())
    normals = _f64(normals).reshape(-1)
    assert normals.size == T * d * maxN, (normals.size, T, d, maxN)
    rn = None if radial_normals is None else _f64(radial_normals)
    ru = None if radial_uniforms is None else _f64(radial_uniforms)
    tp = _f64(tparams if tparams is not None else [0.0])
    samples = np.zeros((total, d))
    lt, ln, lw = np.zeros(total), np.zeros(total), np.zeros(total)
    rid = np.zeros(total, dtype=np.int64)
    ress, rlz, aess = np.zeros(T), np.zeros(T), np.zeros(T)
    fsz = 1 if scalar else d * d
    hm, hf, hl = np.zeros((T + 1) * d), np.zeros((T + 1) * fsz), np.zeros(T + 1)
    fm, ff, fl = np.zeros((T + 1) * d), np.zeros((T + 1) * fsz), np.zeros(T + 1)
    if replay is not None:
        rm, rf, rl = replay
        hm[:] = np.asarray(rm, float).ravel()
        rf = np.asarray(rf, float)
        hf[:] = rf.ravel() if scalar else rf.reshape(T + 1, d, d).transpose(0, 2, 1).ravel()
        hl[:] = np.asarray(rl, float).ravel()
    fail = np.zeros(3, dtype=np.uint64)
    info = np.zeros(2, dtype=np.int32)
    rc = lib().or_adaptive(int(method), s.family, s.dof, int(bool(scalar)), d, _p(s.location),
                           _p(scale0), s.lognormalizer, T, _p(schedule, I64P), _p(normals),
                           _p(rn), _p(ru), int(tkind), _p(tp), int(nthreads), _p(samples), _p(lt),
                           _p(ln), _p(lw), _p(rid, I64P), _p(ress), _p(rlz), _p(aess), _p(hm),
                           _p(hf), _p(hl), _p(fail, U64P), info.ctypes.data_as(IP),
                           int(replay is not None), _p(fm), _p(ff), _p(fl))
    factors = hf.reshape(T + 1, d, d).transpose(0, 2, 1).copy() if not scalar else hf.copy()
    fit_factors = ff.reshape(T + 1, d, d).transpose(0, 2, 1).copy() if not scalar else ff.copy()
    return dict(rc=rc, samples=samples, logtargets=lt, lognumerators=ln, logweights=lw,
                round_ids=rid, round_ess=ress, round_lognormalizers=rlz, adaptation_ess=aess,
                means=hm.reshape(T + 1, d), factors=factors, lognormalizers=hl, failure=fail,
                fail_round=int(info[0]), fail_phase=int(info[1]),
                fit_means=fm.reshape(T + 1, d), fit_factors=fit_factors, fit_lognormalizers=fl)


# ---------------------------------------------------------------- DM-PMC
def active_order(masses):
    """Active slots sorted by mass, stable (src/proposal_banks.jl:301-307).
    Returns 0-based original indices."""
    ids = [i for i, m in enumerate(masses) if m != 0]
    return sorted(ids, key=lambda i: masses[i])


def round_counts(exact_props, round_size, round_index):
    """_dm_pmc_round_counts (src/population_preparation.jl:283-312), exact rationals.
    round_index is 1-based."""
    K = len(exact_props)
    counts, rem = [], []
    for p in exact_props:
        q = p * round_size
        c = q.numerator // q.denominator
        counts.append(c)
        rem.append(q - c)
    remaining = round_size - sum(counts)
    assert 0 <= remaining < K
    first_tie = (round_index - 1) % K + 1
    order = sorted(range(K), key=lambda s: (-rem[s], (s + 1 - first_tie) % K))
    for idx in range(remaining):
        counts[order[idx]] += 1
    return counts


def allocation_plan(active_masses, schedule):
    """(counts K x T int64, logcoefficients K x T) for masses in active order."""
    exact = [Fraction(float(m)) for m in active_masses]
    total = sum(exact)
    props = [m / total for m in exact]
    K, T = len(props), len(schedule)
    counts = np.zeros((K, T), dtype=np.int64)
    for t, n in enumerate(schedule):
        counts[:, t] = round_counts(props, int(n), t + 1)
    if (counts <= 0).any():
        raise ValueError("a round cannot allocate at least one sample to every positive-mass proposal")
    logc = np.log(counts.astype(np.float64) / np.asarray(schedule, dtype=np.float64)[None, :])
    return counts, logc


def dmpmc(family, dof, kind, d, locs, scales, lognorms, proposal_ids, schedule, counts, logcoef,
          normals, uniforms, tkind, tparams, resampling=0, radial_normals=None,
          radial_uniforms=None, nthreads=1, cdf_threads=256, cdf_per=16):
    """locs (K, d); scales: spherical (K,), diagonal (K, d), factor (K, d, d) row-major per slot.
    normals (T, d*maxN); uniforms (T, K)."""
    locs = _f64(locs).reshape(-1)
    K = locs.size // d
    sc = np.asarray(scales, dtype=np.float64)
    if kind == 2:
        sc = _f64(sc.reshape(K, d, d).transpose(0, 2, 1).ravel())
    else:
        sc = _f64(sc.ravel())
    schedule = np.ascontiguousarray(np.asarray(schedule, dtype=np.int64))
    T = schedule.size
    total = int(schedule.sum())
    counts_c = np.ascontiguousarray(np.asarray(counts, dtype=np.int64).T.ravel())
    logc_c = _f64(np.asarray(logcoef, dtype=np.float64).T.ravel())
    pid = np.ascontiguousarray(np.asarray(proposal_ids, dtype=np.int64))
    ln = _f64(lognorms)
    normals = _f64(normals).reshape(-1)
    uniforms = _f64(uniforms).reshape(-1)
    rn = None if radial_normals is None else _f64(radial_normals)
    ru = None if radial_uniforms is None else _f64(radial_uniforms)
    tp = _f64(tparams if tparams is not None else [0.0])
    samples = np.zeros((total, d))
    lw = np.zeros(total)
    rid = np.zeros(total, dtype=np.int64)
    pids = np.zeros(total, dtype=np.int64)
    anc = np.zeros(T * K, dtype=np.int64)
    ress, rlz = np.zeros(T), np.zeros(T)
    final = np.zeros(K * d)
    fail = np.zeros(3, dtype=np.uint64)
    info = np.zeros(2, dtype=np.int32)
    rc = lib().or_dmpmc(int(family), float(dof), int(kind), int(d), K, _p(locs), _p(sc), _p(ln),
                        _p(pid, I64P), T, _p(schedule, I64P), _p(counts_c, I64P), _p(logc_c),
                        int(resampling), int(cdf_threads), int(cdf_per), _p(normals), _p(rn), _p(ru),
                        _p(uniforms), int(tkind), _p(tp), int(nthreads), _p(samples), _p(lw),
                        _p(rid, I64P), _p(pids, I64P), _p(anc, I64P), _p(ress), _p(rlz), _p(final),
                        _p(fail, U64P), info.ctypes.data_as(IP))
    return dict(rc=rc, samples=samples, logweights=lw, round_ids=rid, proposal_ids=pids,
                ancestors=anc.reshape(T, K), round_ess=ress, round_lognormalizers=rlz,
                final_locations=final.reshape(K, d), failure=fail, fail_round=int(info[0]),
                fail_phase=int(info[1]))


# ---------------------------------------------------------------- Philox4x64-10 (host restatement)
_M0 = 0xD2E7470EE14C6C93
_M1 = 0xCA5A826395121157
_W0 = 0x9E3779B97F4A7C15
_W1 = 0xBB67AE8584CAA73B
_MASK = (1 << 64) - 1


def philox4x64(counter, key, rounds=10):
    """Philox4x64-R (Salmon et al., SC'11), pure Python: 4 counter words, 2 key words."""
    c0, c1, c2, c3 = (int(v) & _MASK for v in counter)
    k0, k1 = (int(v) & _MASK for v in key)
    for r in range(rounds):
        if r:
            k0 = (k0 + _W0) & _MASK
            k1 = (k1 + _W1) & _MASK
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> 64, p0 & _MASK
        hi1, lo1 = p1 >> 64, p1 & _MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return [c0, c1, c2, c3]


def u53(word):
    """Uniform in (0, 1): ((w >> 11) + 0.5) * 2^-53."""
    return ((int(word) >> 11) + 0.5) * 2.0 ** -53


def box_muller(w0, w1):
    u1, u2 = u53(w0), u53(w1)
    r = math.sqrt(-2.0 * math.log(u1))
    return r * math.cos(2.0 * math.pi * u2), r * math.sin(2.0 * math.pi * u2)


def ridge_scalar(scale, family=0, dof=0.0):
    L = lib()
    L.or_ridge_scalar.restype = ctypes.c_double
    L.or_ridge_scalar.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
    return L.or_ridge_scalar(float(scale), int(family), float(dof))


def ridge_factor(factor, family=0, dof=0.0):
    L = lib()
    L.or_ridge_factor.restype = ctypes.c_double
    L.or_ridge_factor.argtypes = [F64P, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    f = np.asarray(factor, dtype=np.float64)
    d = f.shape[0]
    a = _f64(f.T.ravel())
    return L.or_ridge_factor(_p(a), d, int(family), float(dof))


# ---------------------------------------------------------------- static MIS (numpy restatement)
# src/proposal_banks.jl:301-326 (active bank), src/methods/static_mis.jl:172-218
# (assignment), :245-310 (partial groups), :383-468 (denominators), :526-592
# (log-weights).  Pinned in tests/test_oracle_static_mis.py by the reference's
# assignment truth table and partial-group compilation (test/static_mis.jl:31-67,
# 380-432), the zero-mass inertness and mixture identities (:185-247, :345-378)
# and a brute-force partial denominator (:434-477).
def mis_active_bank(masses):
    """Active slots (configured ids with nonzero mass, stably sorted by mass), cumulative
    masses with the last one exactly 1, and log interval masses less their logsumexp."""
    masses = _f64(masses)
    ids = [i + 1 for i in sorted((i for i in range(masses.size) if masses[i] != 0), key=lambda i: masses[i])]
    m = np.array([masses[i - 1] for i in ids])
    m = m / m.sum()
    cdf = np.cumsum(m)
    cdf[-1] = 1.0
    prev = np.concatenate([[0.0], cdf[:-1]])
    if np.any(cdf <= prev):
        raise ValueError("every positive proposal mass must retain a positive assignment interval")
    lm = np.log(cdf - prev)
    mx = lm.max()
    lm = lm - (mx + math.log(np.sum(np.exp(lm - mx))))
    return ids, cdf, lm


def mis_assignment(cdf, u, i, n, stratified):
    """0-based slot of sample i (0-based): the first slot whose cumulative mass is >= the
    mapped uniform, (i + u) / n when stratified, u otherwise."""
    v = ((i + u) / n) if stratified else u
    lo, hi = 0, len(cdf) - 1
    while lo < hi:
        mid = lo + (hi - lo) // 2
        if v <= cdf[mid]:
            hi = mid
        else:
            lo = mid + 1
    return lo


def mis_partial_groups(ids, logmass, groups, nproposals):
    """Per active slot its group, the group offsets, the members (0-based slots, in the
    listed order, inactive ids skipped, empty groups dropped) and log coefficients."""
    slot_of = {pid: s for s, pid in enumerate(ids)}
    seen = set()
    slot_group = [0] * len(ids)
    offsets, members, coef = [0], [], []
    for g in groups:
        if not g:
            raise ValueError("partial-MIS groups cannot be empty")
        first = len(members)
        for pid in g:
            if not 1 <= pid <= nproposals or pid in seen:
                raise ValueError("invalid partial-MIS proposal ID")
            seen.add(pid)
            if pid in slot_of:
                members.append(slot_of[pid])
        if len(members) == first:
            continue
        glm = -math.inf
        for s in members[first:]:
            glm = logaddexp(glm, logmass[s])
        for s in members[first:]:
            slot_group[s] = len(offsets) - 1
            coef.append(logmass[s] - glm)
        offsets.append(len(members))
    if len(seen) != nproposals:
        raise ValueError("partial-MIS groups must include every proposal ID exactly once")
    return slot_group, offsets, members, coef


def _draw_factor(p):
    s = as_slot(p)
    if s.kind == 0:
        return np.eye(s.d) * s.scale[0]
    if s.kind == 1:
        return np.diag(s.scale)
    return s.scale.reshape(s.d, s.d).T  # lower factor (stored column-major)


def static_mis(props, masses, scheme, groups, normals, uniforms, tkind, tparams,
               radial_normals=None, radial_uniforms=None):
    """scheme 0 stratified mixture, 1 random mixture, 2 standard MIS, 3 partial mixture.
    normals (n, d), uniforms (n,): the sample's standard normals and assignment uniform;
    Student-t banks take rs radial normals (and us uniforms) per sample.  Returns the
    samples, log-weights, log targets and 1-based proposal ids."""
    ids, cdf, lm = mis_active_bank(masses)
    n = len(uniforms)
    normals = _f64(normals).reshape(n, -1)
    slots = [props[pid - 1] for pid in ids]
    if scheme == 3:
        slot_group, offsets, members, coef = mis_partial_groups(ids, lm, groups, len(props))
    d = as_slot(slots[0]).d
    xs, lw, lt, pid = np.zeros((n, d)), np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int64)
    for i in range(n):
        g = mis_assignment(cdf, uniforms[i], i, n, scheme != 1)
        sl = as_slot(slots[g])
        mult = 1.0
        if sl.family == 1:
            mult = student_t_multiplier(sl.dof, radial_normals[i],
                                        None if radial_uniforms is None else radial_uniforms[i])
        x = sl.location + _draw_factor(sl) @ (normals[i] * mult)
        if scheme in (0, 1):
            terms = [lm[k] + proposal_logdensity(slots[k], x) for k in range(len(slots))]
        elif scheme == 2:
            terms = [proposal_logdensity(slots[g], x)]
        else:
            grp = slot_group[g]
            terms = [coef[m] + proposal_logdensity(slots[members[m]], x)
                     for m in range(offsets[grp], offsets[grp + 1])]
        den = -math.inf
        for t_ in terms:
            den = logaddexp(den, t_)
        tv = target(tkind, tparams, x)
        xs[i], lt[i], lw[i], pid[i] = x, tv, tv - den, ids[g]
    return {"samples": xs, "logweights": lw, "logtargets": lt, "proposal_ids": pid, "ids": ids, "cdf": cdf,
            "logmass": lm}
