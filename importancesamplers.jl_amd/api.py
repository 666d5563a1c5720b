"""Host mirror of the ImportanceSamplers.jl user API over the gfx950 engine.

Names follow the reference (src/ImportanceSamplers.jl:19-91); Julia's
mutating `importance_sample!` is `importance_sample_` here.  Everything that
computes runs in libissampler.so on the GPU; this module only builds job
descriptors, moves host buffers across the C ABI and raises the reference's
error types.
"""

from __future__ import annotations

import ctypes
import math
import threading
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import F64P, I32P, I64P, VP, JobDesc

GAUSSIAN, STUDENT_T = 0, 1
SPHERICAL, DIAGONAL, FACTOR = 0, 1, 2
PLAIN, AMIS_M, NPMC_M, DMPMC_M, MIS_M = 0, 1, 2, 3, 4
FP64, FP32 = 0, 1

PHASES = {0: None, 1: "sampling", 2: "target", 3: "denominator", 4: "weight", 5: "moment",
          6: "factorization", 7: "resampling"}
PLAIN_PHASES = {1: "proposal_draw", 2: "target", 3: "proposal_logdensity", 4: "logweight"}


# ---------------------------------------------------------------- errors (reference types)
class SamplerExecutionError(RuntimeError):
    """src/methods/importance_sampling.jl:210-225: phase, first logical index and the
    captured cause (an InvalidTransformError for transform failures)."""

    def __init__(self, phase, index, reason_bits=0, message="", cause=None):
        super().__init__(f"sampler failed during {phase} at sample {index}: {message or cause or ''}")
        self.phase = phase
        self.index = index
        self.reason_bits = reason_bits
        self.cause = cause


class _RoundError(RuntimeError):
    method = "sampler"

    def __init__(self, round, phase, cause=None, diagnostics=None):
        super().__init__(f"{self.method} failed in round {round} during {phase}: {cause}")
        self.round = round
        self.phase = phase
        self.cause = cause
        self.diagnostics = diagnostics or {}


class AMISRoundError(_RoundError):
    method = "AMIS"


class NPMCRoundError(_RoundError):
    method = "NPMC"


class DMPMCRoundError(_RoundError):
    method = "DM-PMC"


class InvalidTransformError(ValueError):
    """src/transforms.jl:79-95: a sample transform failed (`reason`: nonfinite_input,
    nonfinite_output, nonfinite_logabsjac or outside_support) at `location` (the
    1-based coordinate of a per-coordinate transform)."""

    def __init__(self, reason, location=None):
        super().__init__(f"invalid transform: {reason}" + ("" if location is None else f" at {location}"))
        self.reason = reason
        self.location = location


def _transform_reason(bits):
    """_native_transform_failure_reason (src/transforms.jl:282-287)."""
    if bits & 0x0001:
        return "nonfinite_input"
    if bits & 0x0002:
        return "nonfinite_output"
    if bits & 0x0004:
        return "nonfinite_logabsjac"
    return "outside_support"


class AllZeroWeightsError(ArithmeticError):
    """src/results.jl: every raw log weight is -Inf."""


class SamplerBusyError(RuntimeError):
    """src/methods/importance_sampling.jl:105-116."""


class EngineError(RuntimeError):
    pass


def _check(rc):
    if rc == _lib.IS_OK:
        return
    msg = _lib.last_error()
    if rc == _lib.IS_EBUSY:
        raise SamplerBusyError(msg)
    if rc == _lib.IS_EINVAL:
        raise ValueError(msg)
    if rc == _lib.IS_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise EngineError(f"engine error {rc}: {msg}")


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _p(a, t=F64P):
    return None if a is None else a.ctypes.data_as(t)


# ---------------------------------------------------------------- proposals (src/proposals.jl)
def _gaussian_lognormalizer(d, logabsdet):
    return -0.5 * d * math.log(2.0 * math.pi) - logabsdet


def student_t_logconstant(dof, d):
    """src/proposals.jl:410-437 (host evaluation for proposal constructors)."""
    value = -d * math.log(2.0 * math.pi) / 2.0
    if d % 2 == 1:
        shifted, correction = dof, 0.0
        while shifted < 32.0:
            if shifted < 1.0:
                correction += (math.log(shifted) + math.log(shifted + 2.0)) / 2.0 - math.log1p(shifted)
            else:
                correction += math.log1p(-((1.0 / (shifted + 1.0)) ** 2)) / 2.0
            shifted += 2.0
        r = 1.0 / shifted
        r2 = r * r
        poly = 691.0 / 88.0
        for coef in (-31.0 / 36.0, 17.0 / 112.0, -1.0 / 20.0, 1.0 / 24.0, -1.0 / 4.0):
            poly = poly * r2 + coef
        value += correction + r * poly
    acc = 0.0
    for offset in range(1 if d % 2 == 1 else 0, d - 1, 2):
        x = float(offset)
        acc += math.log1p(x / dof) if x <= dof else math.log(x) + math.log1p(dof / x) - math.log(dof)
    return value + acc


@dataclass
class NativeProposal:
    """A native radial proposal (`_GaussianProposal` / `_StudentTProposal`)."""

    family: int
    kind: int
    location: np.ndarray
    scale: np.ndarray
    lognormalizer: float
    dof: float = 0.0
    scalar: bool = False

    @property
    def dim(self):
        return self.location.size

    def factor(self):
        d = self.dim
        if self.kind == FACTOR:
            return self.scale.reshape(d, d).copy()
        if self.kind == DIAGONAL:
            return np.diag(self.scale)
        return np.eye(d) * self.scale[0]


def _loc(location):
    scalar = np.ndim(location) == 0
    loc = np.atleast_1d(np.asarray(location, dtype=np.float64)).copy()
    if loc.size == 0 or not np.all(np.isfinite(loc)):
        raise ValueError("native parameter must contain only finite values")
    return loc, scalar


def SphericalGaussian(location, scale):
    loc, scalar = _loc(location)
    s = float(scale)
    if not (math.isfinite(s) and s > 0):
        raise ValueError("scale must be finite and positive")
    ln = _gaussian_lognormalizer(loc.size, loc.size * math.log(s))
    return NativeProposal(GAUSSIAN, SPHERICAL, loc, np.array([s]), ln, scalar=scalar)


def DiagonalGaussian(location, scales):
    loc, scalar = _loc(location)
    if scalar:
        raise ValueError("DiagonalGaussian requires a vector location")
    sc = _f64(scales).ravel()
    if sc.size != loc.size or not np.all(np.isfinite(sc) & (sc > 0)):
        raise ValueError("diagonal scales must be finite, positive and match the location length")
    ln = _gaussian_lognormalizer(loc.size, float(np.sum(np.log(sc))))
    return NativeProposal(GAUSSIAN, DIAGONAL, loc, sc, ln)


def _factor(loc, factor):
    L = np.array(factor, dtype=np.float64)
    d = loc.size
    if L.shape != (d, d):
        raise ValueError("factor must be square and match the location length")
    if not np.all(np.isfinite(L)):
        raise ValueError("factor must contain only finite values")
    if np.any(np.triu(L, 1) != 0):
        raise ValueError("factor must be lower triangular")
    if np.any(np.diag(L) <= 0):
        raise ValueError("factor diagonal must be positive")
    return L


def FactorGaussian(location, factor):
    loc, scalar = _loc(location)
    if scalar:
        raise ValueError("FactorGaussian requires a vector location")
    L = _factor(loc, factor)
    ln = _gaussian_lognormalizer(loc.size, float(np.sum(np.log(np.diag(L)))))
    return NativeProposal(GAUSSIAN, FACTOR, loc, L.ravel(), ln)


def _dof(dof):
    v = float(dof)
    if not (math.isfinite(v) and v > 0):
        raise ValueError("degrees of freedom must be finite and positive")
    return v


def SphericalStudentT(dof, location, scale):
    nu = _dof(dof)
    p = SphericalGaussian(location, scale)
    p.family, p.dof = STUDENT_T, nu
    p.lognormalizer = student_t_logconstant(nu, p.dim) - p.dim * math.log(float(scale))
    return p


def DiagonalStudentT(dof, location, scales):
    nu = _dof(dof)
    p = DiagonalGaussian(location, scales)
    p.family, p.dof = STUDENT_T, nu
    p.lognormalizer = student_t_logconstant(nu, p.dim) - float(np.sum(np.log(p.scale)))
    return p


def FactorStudentT(dof, location, factor):
    nu = _dof(dof)
    p = FactorGaussian(location, factor)
    p.family, p.dof = STUDENT_T, nu
    p.lognormalizer = student_t_logconstant(nu, p.dim) - float(np.sum(np.log(np.diag(p.factor()))))
    return p


@dataclass
class ProposalBank:
    """src/proposal_banks.jl:12-97: proposals with nonnegative masses (zero = inert)."""

    proposals: list
    masses: np.ndarray = None

    def __post_init__(self):
        if self.masses is None:
            self.masses = np.ones(len(self.proposals))
        self.masses = _f64(self.masses).ravel()
        if len(self.proposals) == 0 or self.masses.size != len(self.proposals):
            raise ValueError("bank needs one mass per proposal")
        p0 = self.proposals[0]
        for p in self.proposals:
            if (p.family, p.kind, p.dim, p.dof) != (p0.family, p0.kind, p0.dim, p0.dof):
                raise ValueError("bank proposals must share family, scale kind, dimension and dof")


# ---------------------------------------------------------------- targets (engine built-ins)
@dataclass
class StandardGaussianTarget:
    """log pi(x) = -sum(abs2, x)/2 (the plain-IS known-answer target)."""

    kind: int = 0

    def params(self, d):
        return np.zeros(1)


@dataclass
class GaussianTarget:
    """Normalized Gaussian log density N(mean, L L^T) (or with `lognormalizer` overridden)."""

    mean: np.ndarray
    factor: np.ndarray
    lognormalizer: float = None
    kind: int = 1

    def __post_init__(self):
        self.mean = np.atleast_1d(_f64(self.mean))
        self.factor = np.atleast_2d(_f64(self.factor))
        if self.lognormalizer is None:
            self.lognormalizer = _gaussian_lognormalizer(self.mean.size, float(np.sum(np.log(np.diag(self.factor)))))

    def params(self, d):
        return np.concatenate([self.mean, self.factor.T.ravel(), [self.lognormalizer]])


class BatchTarget:
    """A caller's batch target (the `batch=batch!` contract, src/targets.jl:17-43,
    src/batched_targets.jl:92-114): batch(values, samples, coord_stride, dim, count,
    is_float, stream) sets values[i], i < count, to the log density of sample i, whose
    coordinate c is element c * coord_stride + i of `samples`.  values and samples are
    device pointers (FP32 when is_float, else FP64) ordered on the job's HIP stream
    `stream`; values start as NaN, -Inf is valid, NaN and +Inf fail the round.  Return 0,
    or raise / return nonzero to abort the run."""

    kind: int = 4

    def __init__(self, batch):
        self.batch = batch

        def _cb(user, samples, stride, dim, count, values, is_float, stream):
            try:
                rc = self.batch(values, samples, int(stride), int(dim), int(count), bool(is_float), stream)
                return int(rc or 0)
            except Exception:  # the callback's own failure aborts the run
                return -1

        self._fn = _lib.BATCH_TARGET_FN(_cb)

    def params(self, d):
        return np.zeros(1)


_hip_runtime = None


def _hip():
    global _hip_runtime
    if _hip_runtime is None:
        for name in ("libamdhip64.so.7", "libamdhip64.so"):
            try:
                h = ctypes.CDLL(name)
                break
            except OSError:
                h = None
        if h is None:
            raise EngineError("the HIP runtime library is not loadable")
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip_runtime = h
    return _hip_runtime


class HostBatchTarget(BatchTarget):
    """A batch target evaluated on the host (LogTarget(f; batch=batch!), src/targets.jl:17-43,
    src/batched_targets.jl:92-114): batch(out, xs) fills the float64 array out (one value
    per sample) from the logical samples xs, a (count, dim) array, or under a named layout
    a dict of the named leaves ((count,) for an int selector, (count, len) for a range), as
    the reference's batch callbacks receive them.  The job's stream is synchronised, the
    samples are copied down and the values up through the HIP runtime."""

    def __init__(self, batch):
        self.host_batch = batch
        self._layout = None

        def dev(values, samples, stride, dim, count, is_float, stream):
            h = _hip()
            if h.hipStreamSynchronize(stream) != 0:
                return -1
            dt = np.float32 if is_float else np.float64
            sz = np.dtype(dt).itemsize
            X = np.empty((dim, count), dt)
            for c in range(dim):
                if h.hipMemcpy(X[c].ctypes.data, samples + c * stride * sz, count * sz, 2) != 0:
                    return -1
            xs = np.ascontiguousarray(X.T)
            if self._layout is not None:
                xs = self._layout.split(xs)
            out = np.full(count, np.nan)
            rc = self.host_batch(out, xs)
            vals = np.ascontiguousarray(out, dtype=dt)
            if h.hipMemcpy(values, vals.ctypes.data, count * sz, 1) != 0:
                return -1
            return int(rc or 0)

        super().__init__(dev)

    def _bind_layout(self, layout):
        self._layout = layout


@dataclass
class LogisticRegressionTarget:
    """Bayesian logistic regression with prior N(0, prior_var I)."""

    X: np.ndarray
    y: np.ndarray
    prior_var: float = 10.0
    kind: int = 2

    def params(self, d):
        X = _f64(self.X)
        return np.concatenate([[X.shape[0], X.shape[1], self.prior_var], X.ravel(), _f64(self.y)])


@dataclass
class LinearRegressionTarget:
    """Bayesian linear regression y ~ N(X beta, sigma^2), prior N(0, tau^2 I)."""

    X: np.ndarray
    y: np.ndarray
    sigma: float = 0.5
    tau: float = 2.0
    kind: int = 3

    def params(self, d):
        X = _f64(self.X)
        return np.concatenate([[X.shape[0], X.shape[1], self.sigma, self.tau], X.ravel(), _f64(self.y)])


# ---------------------------------------------------------------- algorithms (src/methods/*.jl)
def _schedule(rounds, round_size):
    if int(rounds) < 1:
        raise ValueError("rounds must be positive")
    if np.ndim(round_size) == 0:
        sched = [int(round_size)] * int(rounds)
    else:
        sched = [int(v) for v in round_size]
        if len(sched) != int(rounds):
            raise ValueError("round_size must have one entry per round")
    if any(v <= 0 for v in sched):
        raise ValueError("round sizes must be positive")
    return sched


@dataclass
class ImportanceSampling:
    """`ImportanceSampling(proposal; nsamples)`, or static MIS with a `ProposalBank`
    and a `mis_scheme` (src/methods/static_mis.jl; default StratifiedMixture)."""

    proposal: object
    nsamples: int
    mis_scheme: object = None

    def __post_init__(self):
        if int(self.nsamples) <= 0:
            raise ValueError("nsamples must be positive")
        if isinstance(self.proposal, ProposalBank):
            if self.mis_scheme is None:
                self.mis_scheme = StratifiedMixture()
            if not isinstance(self.mis_scheme, (StratifiedMixture, RandomMixture, StandardMIS,
                                                PartialDeterministicMixture)):
                raise ValueError(f"{type(self.mis_scheme).__name__} static-MIS execution is not implemented")
            if isinstance(self.mis_scheme, PartialDeterministicMixture):
                self.mis_scheme.validate(len(self.proposal.proposals))
        elif self.mis_scheme is not None:
            raise ValueError("an MIS scheme needs a ProposalBank")


# ---------------------------------------------------------------- static MIS schemes
class StratifiedMixture:
    """Stratified slot assignment, full deterministic-mixture denominator."""

    code, name = 0, "stratified_mixture"


class RandomMixture:
    """Independent slot assignment, full deterministic-mixture denominator."""

    code, name = 1, "random_mixture"


class StandardMIS:
    """Stratified slot assignment, the generating proposal alone in the denominator."""

    code, name = 2, "standard_mis"


class PartialDeterministicMixture:
    """Stratified slot assignment; the denominator is the mixture of the generating
    proposal's group (`groups`: tuples of 1-based proposal ids, each id exactly once;
    src/methods/static_mis.jl:245-310)."""

    code, name = 3, "partial_deterministic_mixture"

    def __init__(self, groups):
        self.groups = tuple(tuple(g) for g in groups)

    def validate(self, nproposals):
        if not self.groups:
            raise ValueError("partial-MIS groups cannot be empty")
        seen = set()
        for g in self.groups:
            if not g:
                raise ValueError("partial-MIS groups cannot be empty")
            for pid in g:
                if isinstance(pid, bool) or not isinstance(pid, (int, np.integer)):
                    raise ValueError("partial-MIS proposal IDs must be integers")
                if not 1 <= pid <= nproposals:
                    raise ValueError(f"partial-MIS proposal ID {pid} is out of range")
                if pid in seen:
                    raise ValueError(f"partial-MIS proposal ID {pid} appears more than once")
                seen.add(int(pid))
        if len(seen) != nproposals:
            raise ValueError("partial-MIS groups must include every proposal ID exactly once")


@dataclass
class AMIS:
    proposal: NativeProposal
    rounds: int
    round_size: object

    def __post_init__(self):
        self.schedule = _schedule(self.rounds, self.round_size)
        if self.proposal.family == STUDENT_T and not self.proposal.dof > 2:
            raise ValueError("covariance adaptation requires Student-t degrees of freedom > 2")


@dataclass
class NPMC(AMIS):
    pass


class MultinomialResampling:
    code = 0


class SystematicResampling:
    """New method (the reference has only multinomial): one uniform, points (j + u)/M."""

    code = 1


class GlobalResampling:
    """Every next-round location from the complete weighted round (src/methods/dm_pmc.jl:12-17)."""

    name = "global"


class LocalResampling:
    """Each next-round location from its own proposal's weighted sample group
    (src/methods/dm_pmc.jl:19-24, src/dm_pmc_execution.jl:80-265); one uniform per proposal."""

    name = "local"
    code = 2


@dataclass
class DeterministicMixturePMC:
    bank: ProposalBank
    rounds: int
    round_size: object
    resampling: object = None
    method: object = None  # GlobalResampling's selection: MultinomialResampling (default) or SystematicResampling

    def __post_init__(self):
        self.schedule = _schedule(self.rounds, self.round_size)
        if self.resampling is None:
            self.resampling = GlobalResampling()
        if not isinstance(self.resampling, (GlobalResampling, LocalResampling)):
            raise ValueError("resampling must be GlobalResampling() or LocalResampling()")
        if self.method is None:
            self.method = MultinomialResampling()
        if isinstance(self.resampling, LocalResampling) and not isinstance(self.method, MultinomialResampling):
            raise ValueError("LocalResampling draws one multinomial ancestor per proposal")


# ---------------------------------------------------------------- RNG seams
@dataclass
class PrefilledRNG:
    """Feeds exact normals/uniforms per round (the reference tests' prefilled RNG seam).

    normals[t]: plain IS -> stride*n (per sample d normals then radial normals);
    adaptive -> d*N_t.  radial_normals/radial_uniforms[t]: Student-t gamma buffers.
    uniforms[t]: DM-PMC resampling uniforms (K per round)."""

    normals: list
    radial_normals: list = None
    radial_uniforms: list = None
    uniforms: list = None
    seed: int = 0


# ---------------------------------------------------------------- context
_ctx = None
_ctx_lock = threading.Lock()


def context(device=0):
    global _ctx
    with _ctx_lock:
        if _ctx is None:
            lib = _lib.load()
            h = VP()
            _check(lib.is_ctx_create(device, ctypes.byref(h)))
            _ctx = h
    return _ctx


@dataclass
class IdentityTransform:
    """x = z (src/transforms.jl:17)."""
    code = 0
    lower: float = -math.inf
    upper: float = math.inf


@dataclass
class PositiveTransform:
    """x = exp(z), log|J| = z (src/transforms.jl:30)."""
    code = 1
    lower: float = -math.inf
    upper: float = math.inf


@dataclass
class SoftplusTransform:
    """x = softplus(z), log|J| = z - softplus(z) (src/transforms.jl:43)."""
    code = 2
    lower: float = -math.inf
    upper: float = math.inf


@dataclass
class IntervalTransform:
    """x in (lower, upper), one side may be open (src/transforms.jl:45-48,249-270)."""
    lower: float = -math.inf
    upper: float = math.inf
    code = 3

    def __post_init__(self):
        lo = -math.inf if self.lower is None else float(self.lower)
        hi = math.inf if self.upper is None else float(self.upper)
        if not lo < hi or (math.isinf(lo) and math.isinf(hi)):
            raise ValueError("IntervalTransform needs lower < upper with a finite side")
        self.lower, self.upper = lo, hi


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch: a transform given the wrong number of coordinates."""


class SimplexTransform:
    """SimplexTransform(K) (src/transforms.jl:51-70): K - 1 unconstrained coordinates to K
    positive weights that sum to one, through the orthonormal embedding into the sum-zero
    logit subspace; log|J| = log(K)/2 + sum(log, x), measured against the first K - 1
    simplex coordinates.  Evaluated on the device in the sampler's precision."""

    code = 4
    lower = -math.inf
    upper = math.inf

    def __init__(self, dimension):
        if type(dimension) is not int or dimension < 2:
            raise ValueError("simplex dimension must be an Int of at least two")
        self.dimension = dimension

    def __eq__(self, other):
        return isinstance(other, SimplexTransform) and other.dimension == self.dimension

    def __hash__(self):
        return hash(("simplex", self.dimension))

    def __repr__(self):
        return f"SimplexTransform({self.dimension})"


_SCALAR_TRANSFORMS = (PositiveTransform, SoftplusTransform, IntervalTransform)
_TRANSFORMS = (IdentityTransform,) + _SCALAR_TRANSFORMS + (SimplexTransform,)


def _flat_selector(sel, d):
    """_validated_flat_selector (src/proposal_composition.jl:179-201) with Python's 0-based
    indices: an int, or a nonempty contiguous range (a range or slice with step 1)."""
    if isinstance(sel, slice):
        if sel.step not in (None, 1) or sel.start is None or sel.stop is None:
            raise ValueError("flat transform selectors must be ints or contiguous ranges")
        sel = range(sel.start, sel.stop)
    if isinstance(sel, (int, np.integer)) and not isinstance(sel, (bool, np.bool_)):
        sel = int(sel)
        if not 0 <= sel < d:
            raise ValueError(f"flat transform selector {sel} is out of bounds for dimension {d}")
        return sel
    if isinstance(sel, range):
        if sel.step != 1:
            raise ValueError("flat transform selectors must be ints or contiguous ranges")
        if len(sel) == 0:
            raise ValueError("flat transform ranges must be nonempty")
        if sel.start < 0 or sel.stop > d:
            raise ValueError(f"flat transform selector {sel} is out of bounds for dimension {d}")
        return sel
    raise ValueError("flat transform selectors must be ints or contiguous ranges")


class _Layout:
    """A flat selector layout (_FlatTransformLayout, src/proposal_composition.jl:11-14,
    203-240) of named blocks over the d sampling coordinates, or one SimplexTransform over
    all of them (`names` None: results stay one (n, K) array).  `blocks` is the C
    is_transform_block array in name order; `dl` the logical dimension."""

    def __init__(self, spec, d):
        if isinstance(spec, SimplexTransform):
            if d != spec.dimension - 1:
                raise DimensionMismatch(f"{spec!r} requires {spec.dimension - 1} coordinates")
            self.names = None
            items = [("", (range(0, d), spec))]
        else:
            if not spec:
                raise ValueError("flat selector transforms must contain blocks")
            self.names = [str(k) for k in spec]
            items = list(spec.items())
        selected = np.zeros(d, dtype=bool)
        self.selectors, self.transforms, self.lengths, self.scalars = [], [], [], []
        blocks = (_lib.TransformBlock * len(items))()
        row = 0
        for b, (name, value) in enumerate(items):
            sel, tf = value if isinstance(value, tuple) else (value, IdentityTransform())
            if not isinstance(tf, _TRANSFORMS):
                raise ValueError("flat transform selector values must be sample transforms")
            sel = _flat_selector(sel, d)
            scalar = isinstance(sel, int)
            idx = range(sel, sel + 1) if scalar else sel
            if isinstance(tf, _SCALAR_TRANSFORMS) and not scalar:
                raise DimensionMismatch(f"{type(tf).__name__} requires one unconstrained scalar")
            if isinstance(tf, SimplexTransform) and (scalar or len(idx) != tf.dimension - 1):
                raise DimensionMismatch(f"{tf!r} requires {tf.dimension - 1} coordinates")
            if selected[idx.start:idx.stop].any():
                raise ValueError("flat transform selectors must be disjoint")
            selected[idx.start:idx.stop] = True
            n_out = tf.dimension if isinstance(tf, SimplexTransform) else len(idx)
            blk = blocks[b]
            blk.first, blk.count, blk.kind, blk.scalar = idx.start, len(idx), tf.code, int(scalar)
            blk.lower, blk.upper = float(tf.lower), float(tf.upper)
            self.selectors.append(sel)
            self.transforms.append(tf)
            self.lengths.append(n_out)
            self.scalars.append(scalar)
            row += n_out
        if not selected.all():
            raise ValueError("flat transform selectors must collectively cover the base dimension")
        self.blocks = blocks
        self.dl = row
        self.offsets = np.concatenate([[0], np.cumsum(self.lengths)]).astype(int)

    def location(self, block):
        """_native_failure_location (src/native_execution.jl:826-830): the failing
        block's selector."""
        return self.selectors[block - 1] if 1 <= block <= len(self.selectors) else None

    def split(self, flat):
        """(n, dl) logical samples -> {name: (n,) for an int selector, (n, len) for a range}
        (_allocate_flat_samples, src/proposal_composition.jl:70-76)."""
        if self.names is None:
            return flat
        out = {}
        for b, name in enumerate(self.names):
            o, k = self.offsets[b], self.lengths[b]
            out[name] = flat[:, o].copy() if self.scalars[b] else flat[:, o:o + k].copy()
        return out


def _layout_of(transform, d):
    """A transform= argument that needs the layout path (a dict of named blocks or a
    SimplexTransform), else None."""
    if isinstance(transform, _Layout):
        return transform
    if isinstance(transform, (dict, SimplexTransform)):
        return _Layout(transform, d)
    return None


def _transform_arrays(transform, d):
    ts = list(transform) if isinstance(transform, (list, tuple)) else [transform] * d
    if len(ts) != d:
        raise ValueError("one transform per coordinate")
    if any(not isinstance(t, _TRANSFORMS[:-1]) for t in ts):
        raise ValueError("per-coordinate transforms are scalar sample transforms")
    kinds = np.asarray([t.code for t in ts], dtype=np.int32)
    bounds = np.asarray([[t.lower, t.upper] for t in ts], dtype=np.float64).ravel()
    return kinds, bounds


def _job_desc(rng, target, algorithm, precision, transform=None):
    """Builds the C descriptor of one prepared sampler; returns (desc, keep, proposal,
    schedule) where `keep` holds the numpy buffers the descriptor points into."""
    desc = JobDesc()
    keep = []

    def arr(a):
        a = _f64(a)
        keep.append(a)
        return _p(a)

    seed = rng.seed if isinstance(rng, PrefilledRNG) else int(rng)
    desc.seed = seed & 0xFFFFFFFFFFFFFFFF
    desc.stream = 0x6D693335357873  # fixed second key word
    desc.precision = precision
    if isinstance(algorithm, ImportanceSampling) and isinstance(algorithm.proposal, ProposalBank):
        p, desc.method, sched = algorithm.proposal.proposals[0], MIS_M, [int(algorithm.nsamples)]
    elif isinstance(algorithm, ImportanceSampling):
        p, desc.method, sched = algorithm.proposal, PLAIN, [int(algorithm.nsamples)]
    elif isinstance(algorithm, NPMC):
        p, desc.method, sched = algorithm.proposal, NPMC_M, algorithm.schedule
    elif isinstance(algorithm, AMIS):
        p, desc.method, sched = algorithm.proposal, AMIS_M, algorithm.schedule
    elif isinstance(algorithm, DeterministicMixturePMC):
        p, desc.method, sched = algorithm.bank.proposals[0], DMPMC_M, algorithm.schedule
    else:
        raise TypeError("unsupported algorithm")
    desc.family, desc.scale_kind, desc.scalar, desc.dim, desc.dof = p.family, p.kind, int(p.scalar), p.dim, p.dof
    desc.rounds = len(sched)
    rs = np.asarray(sched, dtype=np.int64)
    keep.append(rs)
    desc.round_size = rs.ctypes.data_as(I64P)
    if desc.method in (DMPMC_M, MIS_M):
        bank = algorithm.bank if desc.method == DMPMC_M else algorithm.proposal
        desc.bank_size = len(bank.proposals)
        desc.bank_locations = arr(np.concatenate([q.location for q in bank.proposals]))
        desc.bank_scales = arr(np.concatenate([q.scale if q.kind != FACTOR else q.factor().T.ravel()
                                               for q in bank.proposals]))
        desc.bank_masses = arr(bank.masses)
        if desc.method == DMPMC_M:
            local = isinstance(algorithm.resampling, LocalResampling)
            desc.resampling = LocalResampling.code if local else algorithm.method.code
        else:
            scheme = algorithm.mis_scheme
            desc.mis_scheme = scheme.code
            if isinstance(scheme, PartialDeterministicMixture):
                offs = np.cumsum([0] + [len(g) for g in scheme.groups]).astype(np.int32)
                mem = np.asarray([int(pid) for g in scheme.groups for pid in g], dtype=np.int32)
                keep.extend([offs, mem])
                desc.mis_group_count = len(scheme.groups)
                desc.mis_group_offsets = offs.ctypes.data_as(I32P)
                desc.mis_group_members = mem.ctypes.data_as(I32P)
    else:
        desc.location = arr(p.location)
        desc.scale = arr(p.scale if p.kind != FACTOR else p.factor().T.ravel())
    desc.target_kind = target.kind
    tp = _f64(target.params(p.dim))
    keep.append(tp)
    desc.target_params = _p(tp)
    desc.target_params_len = tp.size
    if isinstance(target, BatchTarget):
        desc.target_batch = target._fn
        keep.append(target)
    layout = _layout_of(transform, p.dim)
    if layout is not None:
        keep.append(layout)
        desc.transform_block_count = len(layout.blocks)
        desc.transform_blocks = ctypes.cast(layout.blocks, ctypes.POINTER(_lib.TransformBlock))
    elif transform is not None:
        kinds, bounds = _transform_arrays(transform, p.dim)
        keep.extend([kinds, bounds])
        desc.transform_kinds = kinds.ctypes.data_as(I32P)
        desc.transform_bounds = bounds.ctypes.data_as(F64P)
    return desc, keep, p, sched


# ---------------------------------------------------------------- prepared sampler
class PreparedSampler:
    """`_PreparedImportanceSampler` on the MI355X device (one engine job)."""

    def __init__(self, rng, target, algorithm, precision=FP64, transform=None):
        self.algorithm = algorithm
        self.target = target
        self.precision = precision
        lib = _lib.load()
        ctx = context()
        dim = (algorithm.bank.proposals[0] if isinstance(algorithm, DeterministicMixturePMC) else
               algorithm.proposal.proposals[0] if isinstance(getattr(algorithm, "proposal", None), ProposalBank)
               else algorithm.proposal).dim
        self.layout = _layout_of(transform, dim)
        if self.layout is not None:
            transform = self.layout
            if hasattr(target, "_bind_layout"):
                target._bind_layout(self.layout)
        desc, keep, p, sched = _job_desc(rng, target, algorithm, precision, transform)
        h = VP()
        _check(lib.is_job_create(ctx, ctypes.byref(desc), ctypes.byref(h)))
        self._adopt(h, desc, keep, p, sched)
        if isinstance(rng, PrefilledRNG):
            self.set_prefilled(rng)

    def _adopt(self, handle, desc, keep, p, sched):
        self._keep = keep
        self.proposal0 = p
        self.dim = p.dim
        self.logical_dim = int(_lib.load().is_job_logical_dim(handle))
        self.scalar = bool(p.scalar)
        self.schedule = list(sched)
        self._job = handle
        self._lib = _lib.load()
        self.total = int(self._lib.is_job_total(handle))
        if desc.method in (DMPMC_M, MIS_M):
            bank = self.algorithm.bank if desc.method == DMPMC_M else self.algorithm.proposal
            order = sorted([i for i, m in enumerate(bank.masses) if m != 0], key=lambda i: bank.masses[i])
            self.active_ids = [i + 1 for i in order]

    @classmethod
    def from_handle(cls, handle, rng, target, algorithm, precision, owned_elsewhere=False):
        """Wraps a job created elsewhere (the serving loop, a job group) with the same
        descriptor; a job owned elsewhere (a group's member) is destroyed by its owner."""
        self = cls.__new__(cls)
        self.algorithm, self.target, self.precision = algorithm, target, precision
        self.layout = None
        self._owned_elsewhere = owned_elsewhere
        desc, keep, p, sched = _job_desc(rng, target, algorithm, precision)
        self._adopt(handle, desc, keep, p, sched)
        return self

    def __del__(self):
        try:
            if getattr(self, "_job", None) and not getattr(self, "_owned_elsewhere", False):
                self._lib.is_job_destroy(self._job)
                self._job = None
        except Exception:
            pass

    @property
    def method(self):
        a = self.algorithm
        if isinstance(a, ImportanceSampling):
            return MIS_M if isinstance(a.proposal, ProposalBank) else PLAIN
        if isinstance(a, NPMC):
            return NPMC_M
        if isinstance(a, AMIS):
            return AMIS_M
        return DMPMC_M

    # ---- RNG seams
    def set_prefilled(self, rng: PrefilledRNG):
        lib = self._lib
        for t, batch in enumerate(rng.normals or []):
            if batch is None:
                continue
            b = _f64(batch).ravel()
            _check(lib.is_job_set_normals(self._job, t, _p(b), b.size))
        n_rad = max(len(rng.radial_normals or []), len(rng.radial_uniforms or []))
        for t in range(n_rad):
            rn = _f64(rng.radial_normals[t]).ravel() if rng.radial_normals and rng.radial_normals[t] is not None else None
            ru = _f64(rng.radial_uniforms[t]).ravel() if rng.radial_uniforms and rng.radial_uniforms[t] is not None else None
            _check(lib.is_job_set_radial(self._job, t, _p(rn), 0 if rn is None else rn.size, _p(ru),
                                         0 if ru is None else ru.size))
        for t, batch in enumerate(rng.uniforms or []):
            if batch is None:
                continue
            b = _f64(batch).ravel()
            _check(lib.is_job_set_uniforms(self._job, t, _p(b), b.size))

    def export_randomness(self, round):
        """The counter-RNG values the next run consumes in `round` (for the oracle)."""
        lib, d, n = self._lib, self.dim, self.schedule[round]
        p = self.proposal0
        rs = 0 if p.family == GAUSSIAN else (1 if p.dof == 1.0 else 8)
        us = 0 if (p.family == GAUSSIAN or p.dof == 1.0) else 8 + (1 if p.dof < 2 else 0)
        out = {}
        if self.method == PLAIN:
            z = np.zeros((d + rs) * n)
            _check(lib.is_job_export_normals(self._job, round, _p(z), z.size))
            out["normals"] = z
        else:
            z = np.zeros(d * n)
            _check(lib.is_job_export_normals(self._job, round, _p(z), z.size))
            out["normals"] = z
            if rs:
                rn = np.zeros(rs * n)
                _check(lib.is_job_export_radial(self._job, round, _p(rn), rn.size, None, 0))
                out["radial_normals"] = rn
        if us:
            ru = np.zeros(us * n)
            _check(lib.is_job_export_radial(self._job, round, None, 0, _p(ru), ru.size))
            out["radial_uniforms"] = ru
        if self.method == DMPMC_M:
            K = len(self.active_ids)
            u = np.zeros(K)
            _check(lib.is_job_export_uniforms(self._job, round, _p(u), K))
            out["uniforms"] = u
        if self.method == MIS_M:
            u = np.zeros(n)  # one assignment uniform per sample
            _check(lib.is_job_export_uniforms(self._job, round, _p(u), n))
            out["uniforms"] = u
        return out

    def set_ancestors(self, ancestors):
        """DM-PMC replay: round t resamples to ancestors[t] (K 1-based indices; a (T, K)
        array, or a list with None for rounds left to the sampler)."""
        for t, a in enumerate(ancestors):
            if a is None:
                continue
            a = np.ascontiguousarray(np.asarray(a, dtype=np.int64))
            _check(self._lib.is_job_set_ancestors(self._job, t, _p(a, I64P), a.size))

    def clear_prefilled(self):
        _check(self._lib.is_job_clear_prefilled(self._job))

    def reset(self):
        _check(self._lib.is_job_reset(self._job))

    # ---- execution
    def run(self):
        rc = self._lib.is_job_run(self._job)
        if rc == _lib.IS_EFAILED:
            self._raise_failure()
        _check(rc)

    def failure(self):
        f = _lib.Failure()
        _check(self._lib.is_job_failure(self._job, ctypes.byref(f)))
        return f

    def _raise_failure(self):
        f = self.failure()
        if f.draw_first_logical_index:
            index, bits = f.draw_first_logical_index, f.draw_reason_bits
        else:
            index, bits = f.first_logical_index, f.reason_bits
        phase = PHASES.get(f.phase)
        # transform reasons come with the failing coordinate (or layout block) in the
        # block field (src/native_execution.jl:826-840)
        block = int(f.draw_first_block if f.draw_first_logical_index else f.first_block)
        where = self.layout.location(block) if self.layout is not None else (block or None)
        tcause = InvalidTransformError(_transform_reason(bits), where) if bits & 0x000F else None
        if self.method == PLAIN:
            raise SamplerExecutionError(PLAIN_PHASES.get(f.phase, phase), index, bits, cause=tcause)
        if f.phase in (5, 7):
            cause = AllZeroWeightsError()
        elif f.phase == 6:
            cause = np.linalg.LinAlgError("fitted covariance is not positive definite")
        else:
            cause = SamplerExecutionError(phase, index, bits, cause=tcause)
        err = {AMIS_M: AMISRoundError, NPMC_M: NPMCRoundError, DMPMC_M: DMPMCRoundError}[self.method]
        raise err(f.round, phase, cause, {"failure_count": int(f.count)})

    # ---- results
    def _get(self, name, n, dtype=np.float64):
        out = np.zeros(n, dtype=dtype)
        _check(getattr(self._lib, name)(self._job, _p(out, I64P if dtype == np.int64 else F64P)))
        return out

    def collect(self):
        d, n, T = self.logical_dim, self.total, len(self.schedule)
        samples = self._get("is_job_get_samples", n * d).reshape(n, d)
        logw = self._get("is_job_get_logweights", n)
        prov = {"round": self._get("is_job_get_round_ids", n, np.int64)}
        ess, lz, aess = np.zeros(T), np.zeros(T), np.zeros(T)
        _check(self._lib.is_job_get_round_summary(self._job, _p(ess), _p(lz), _p(aess)))
        method = self.method
        diag = {"method": {PLAIN: "importance_sampling", MIS_M: "importance_sampling", AMIS_M: "amis",
                           NPMC_M: "npmc", DMPMC_M: "deterministic_mixture_pmc"}[method],
                "execution": "mi355x", "rounds": T, "round_sizes": list(self.schedule),
                "round_ess": ess, "round_lognormalizers": lz,
                "target_evaluations": n,
                "proposal_evaluations": (T * n if method == AMIS_M else n),
                "failures": 0, "precision": "fp32" if self.precision == FP32 else "fp64"}
        if method == NPMC_M:
            diag["adaptation_ess"] = aess
            diag["clipped_counts"] = [math.isqrt(v) for v in self.schedule]
        if method == DMPMC_M:
            prov["proposal_id"] = self._get("is_job_get_proposal_ids", n, np.int64)
            diag["resampling"] = self.algorithm.resampling.name
        if method == MIS_M:
            prov = {"proposal_id": self._get("is_job_get_proposal_ids", n, np.int64)}
            diag["mis_scheme"] = self.algorithm.mis_scheme.name
            if isinstance(self.algorithm.mis_scheme, (StratifiedMixture, RandomMixture)):
                diag["proposal_evaluations"] = n * len(self.active_ids)
        if self.layout is not None:
            samples = self.layout.split(samples)
        elif self.scalar:
            samples = samples[:, 0].copy()
        return WeightedSamples(samples, logw, prov, diag, _sampler=self, _result=_ResultHandle.of_job(self))

    def history(self):
        T, d = len(self.schedule), self.dim
        means, facs, lns = np.zeros((T + 1) * d), np.zeros((T + 1) * d * d), np.zeros(T + 1)
        _check(self._lib.is_job_get_history(self._job, _p(means), _p(facs), _p(lns)))
        return means.reshape(T + 1, d), facs.reshape(T + 1, d, d).transpose(0, 2, 1).copy(), lns

    def ancestors(self):
        K, T = len(self.active_ids), len(self.schedule)
        a = self._get("is_job_get_ancestors", K * T, np.int64)
        return a.reshape(T, K)


def prepare_sampler(rng, target, algorithm, precision="fp64", transform=None):
    """src/methods/importance_sampling.jl:357-404 (device-prepared); transform: one
    transform for every coordinate or one per coordinate (sampling stays in unconstrained
    coordinates, the target sees the logical sample, results report it)."""
    prec = {"fp64": FP64, "fp32": FP32, FP64: FP64, FP32: FP32}[precision]
    return PreparedSampler(rng, target, algorithm, prec, transform)


class PreparedGroup:
    """Same-shape samplers advanced together (is_group): plain IS or AMIS with a built-in
    standard or Gaussian target, d <= 15, one family, precision and round schedule;
    proposals, targets' parameters and seeds may differ.  Every round is one launch over
    all members; member m's results are those of its own prepared sampler run alone
    (`members[m]` is that sampler, for the whole results API)."""

    def __init__(self, specs, precision="fp64"):
        prec = {"fp64": FP64, "fp32": FP32, FP64: FP64, FP32: FP32}[precision]
        specs = list(specs)
        if not specs:
            raise ValueError("a group needs at least one sampler")
        lib = _lib.load()
        ctx = context()
        descs = (JobDesc * len(specs))()
        self._keep = []
        for m, (rng, target, alg) in enumerate(specs):
            desc, keep, _, _ = _job_desc(rng, target, alg, prec)
            descs[m] = desc
            self._keep.append(keep)
        h = VP()
        _check(lib.is_group_create(ctx, descs, len(specs), ctypes.byref(h)))
        self._group, self._lib = h, lib
        self.members = []
        for m, (rng, target, alg) in enumerate(specs):
            jh = VP()
            _check(lib.is_group_job(h, m, ctypes.byref(jh)))
            self.members.append(PreparedSampler.from_handle(jh, rng, target, alg, prec, owned_elsewhere=True))

    def __len__(self):
        return len(self.members)

    def run(self):
        """Runs every member once; returns the indices of the members that failed (each
        member's failure is raised by `collect`)."""
        rc = self._lib.is_group_run(self._group)
        if rc not in (0, _lib.IS_EFAILED):
            _check(rc)
        if rc == 0:
            return []
        return [m for m, s in enumerate(self.members) if s.failure().round != 0]

    def reset(self):
        """Every member's next run starts from its initial proposal (is_group_reset)."""
        _check(self._lib.is_group_reset(self._group))

    def round_summary(self, round_index=-1):
        """(ess, lognormalizers): every member's round ESS and log normaliser of the last
        run, member order (is_group_round_summary)."""
        T_ = len(self.members[0].schedule)
        t = round_index % T_
        ess = np.zeros(len(self.members))
        lz = np.zeros(len(self.members))
        _check(self._lib.is_group_round_summary(self._group, t, ess.ctypes.data_as(F64P), lz.ctypes.data_as(F64P)))
        return ess, lz

    def collect(self, m):
        """Member m's WeightedSamples of the last run (raising its failure as a stand-alone
        run would)."""
        s = self.members[m]
        f = s.failure()
        if f.round != 0:
            s._raise_failure()
        return s.collect()

    def __del__(self):
        try:
            if getattr(self, "_group", None):
                for s in self.members:
                    s._job = None
                self._lib.is_group_destroy(self._group)
                self._group = None
        except Exception:
            pass


def prepare_group(specs, precision="fp64"):
    """A PreparedGroup of (rng, target, algorithm) triples."""
    return PreparedGroup(specs, precision)


def importance_sample_group_(group: PreparedGroup):
    """Runs the group once and returns every member's WeightedSamples (in member order);
    a failed member's entry is its exception."""
    group.run()
    out = []
    for m in range(len(group)):
        try:
            out.append(group.collect(m))
        except Exception as exc:  # the member's own failure, as its stand-alone run raises it
            out.append(exc)
    return out


def importance_sample_(sampler: PreparedSampler):
    """`importance_sample!` (src/methods/importance_sampling.jl:912-930)."""
    sampler.run()
    return sampler.collect()


def importance_sample(rng, target, algorithm, precision="fp64", transform=None):
    """`importance_sample` (src/methods/importance_sampling.jl:857-874)."""
    return importance_sample_(prepare_sampler(rng, target, algorithm, precision, transform))


def current_proposal(sampler: PreparedSampler):
    """src/adaptive_gaussian.jl:382-412 (and DM-PMC current locations)."""
    lib, d = sampler._lib, sampler.dim
    if sampler.method == DMPMC_M:
        K = len(sampler.active_ids)
        out = np.zeros(K * d)
        _check(lib.is_job_get_locations(sampler._job, _p(out)))
        locs = out.reshape(K, d)
        bank = sampler.algorithm.bank
        props = [p for p in bank.proposals]
        for slot, pid in enumerate(sampler.active_ids):
            q = props[pid - 1]
            props[pid - 1] = NativeProposal(q.family, q.kind, locs[slot].copy(), q.scale.copy(),
                                            q.lognormalizer, q.dof, q.scalar)
        return ProposalBank(props, bank.masses.copy())
    loc, fac, ln = np.zeros(d), np.zeros(d * d), np.zeros(1)
    _check(lib.is_job_get_proposal(sampler._job, _p(loc), _p(fac), _p(ln)))
    p0 = sampler.proposal0
    if sampler.scalar:
        return NativeProposal(p0.family, SPHERICAL, loc, fac[:1].copy(), float(ln[0]), p0.dof, True)
    L = fac.reshape(d, d).T.copy()
    return NativeProposal(p0.family, FACTOR, loc, L.ravel(), float(ln[0]), p0.dof, False)


def logdensityof(proposal: NativeProposal, x, precision="fp64"):
    """DensityInterface.logdensityof(proposal, x) (src/proposals.jl:24-55,603-669, with the
    extended-range repair of src/student_t_range.jl:55-153), evaluated on the device: x is
    one point (d,) or scalar, or points (n, d); returns a float or an (n,) array."""
    lib = _lib.load()
    d = proposal.dim
    pts = np.atleast_1d(np.asarray(x, dtype=np.float64))
    single = pts.ndim == 1 and (pts.size == d)
    pts = _f64(pts.reshape(-1, d))
    scale = proposal.scale if proposal.kind != FACTOR else proposal.factor().T.ravel()
    scale = _f64(scale)
    out = np.zeros(pts.shape[0])
    prec = {"fp64": FP64, "fp32": FP32, FP64: FP64, FP32: FP32}[precision]
    _check(lib.is_proposal_logdensity(context(), proposal.family, proposal.dof, proposal.kind, d,
                                      _p(_f64(proposal.location)), _p(scale), _p(pts), pts.shape[0], prec, _p(out)))
    return float(out[0]) if single else out


# ---------------------------------------------------------------- results (src/results.jl, src/statistics.jl)
class _ResultHandle:
    """An engine result (is_result): the run's samples and log-weights copied on the
    device, which later runs of the sampler leave unchanged
    (src/methods/importance_sampling.jl:904-906); destroyed with its owner."""

    def __init__(self, handle, dim, scalar, layout=None):
        self.h = handle
        self.dim = dim
        self.scalar = scalar
        self.layout = layout
        self._lib = _lib.load()

    @classmethod
    def of_job(cls, sampler):
        h = VP()
        _check(sampler._lib.is_job_result(sampler._job, ctypes.byref(h)))
        return cls(h, sampler.logical_dim, sampler.scalar, sampler.layout)

    def __del__(self):
        try:
            if self.h:
                self._lib.is_result_destroy(self.h)
                self.h = None
        except Exception:
            pass


@dataclass
class WeightedSamples:
    samples: np.ndarray
    logweights: np.ndarray
    provenance: dict = field(default_factory=dict)
    diagnostics: dict = field(default_factory=dict)
    _sampler: object = None
    _result: object = None

    def __len__(self):
        return self.logweights.size


class UnweightedSamples:
    """`UnweightedSamples` (src/resampling.jl:24-48) of a device resample: the gathered
    samples stay on the device (`device_arrays()`); `samples` copies them to the host."""

    def __init__(self, handle: _ResultHandle, ancestors=None):
        self._result = handle
        self.ancestors = ancestors

    def __len__(self):
        return int(self._result._lib.is_result_count(self._result.h))

    @property
    def samples(self):
        r = self._result
        n, d = len(self), r.dim
        out = np.zeros(n * d)
        _check(r._lib.is_result_get_samples(r.h, _p(out)))
        out = out.reshape(n, d)
        if r.layout is not None:
            return r.layout.split(out)
        return out[:, 0].copy() if r.scalar else out

    def device_arrays(self):
        """(samples device pointer, coordinate stride, is_float)."""
        r = self._result
        sp, stride, isf = VP(), ctypes.c_int64(), ctypes.c_int32()
        _check(r._lib.is_result_device_arrays(r.h, ctypes.byref(sp), ctypes.byref(stride), None, ctypes.byref(isf)))
        return sp.value, int(stride.value), bool(isf.value)


def _handle(result):
    h = getattr(result, "_result", None)
    if h is None or not h.h:
        raise ValueError("result is detached from its device storage")
    return h


def lognormalizer(result: WeightedSamples):
    """LSE(logw) - log N (src/results.jl:560-567), reduced on the device over the
    result's own device copy."""
    if isinstance(result, UnweightedSamples):
        raise ValueError("unweighted samples carry no weights")
    if getattr(result, "_result", None) is None and getattr(result, "_lognormalizer", None) is not None:
        return result._lognormalizer  # batch jobs: the device's final-round reduction
    h = _handle(result)
    out = np.zeros(3)
    _check(h._lib.is_result_lognormalizer(h.h, _p(out)))
    return float(out[0])


def _device_moments(result):
    h = _handle(result)
    d = h.dim
    mean, var, cov = np.zeros(d), np.zeros(d), np.zeros(d * d)
    rc = h._lib.is_result_moments(h.h, _p(mean), _p(var), _p(cov))
    if rc == _lib.IS_EFAILED:
        raise AllZeroWeightsError()
    _check(rc)
    return mean, var, cov.reshape(d, d).T.copy()


@dataclass
class DeviceValues:
    """Per-sample values f(x_i) already in device memory (a pointer on the engine's HIP
    runtime, e.g. written by the caller's own kernel over is_job_device_results'
    samples), FP64 when fp64 else FP32, in sample order."""
    ptr: int
    fp64: bool = True


def _functional_moments(f, result, want_var):
    """mean(f, result) / var(f, result) (src/statistics.jl:17-31,50-64,144-199) against the
    device-resident log-weights, reduced in FP64 on the device.  f is a DeviceValues, or
    a callable mapping the samples (n,) or (n, d) to one number per sample, evaluated on
    the result's host samples and uploaded."""
    h = _handle(result)
    m, v = np.zeros(1), np.zeros(1)
    if isinstance(f, DeviceValues):
        rc = h._lib.is_result_functional_moments(h.h, VP(int(f.ptr)), int(bool(f.fp64)), 1, _p(m),
                                                 _p(v) if want_var else None)
    else:
        vals = np.asarray(f(result.samples), dtype=np.float64)
        if vals.size != len(result) or (vals.ndim > 1 and vals.shape[0] != len(result)):
            raise ValueError("functionals must return one number per sample")
        vals = np.ascontiguousarray(vals.reshape(-1))
        rc = h._lib.is_result_functional_moments(h.h, vals.ctypes.data_as(VP), 1, 0, _p(m),
                                                 _p(v) if want_var else None)
    if rc == _lib.IS_EFAILED:
        raise AllZeroWeightsError()
    _check(rc)
    return float(m[0]), float(v[0])


def _rowwise_moments(result, want_var):
    """Per-coordinate weighted mean and variance of the logical samples: each coordinate's
    device row through the functional reduction (src/statistics.jl:186-197,204-213)."""
    h = _handle(result)
    sp, stride, isf = VP(), ctypes.c_int64(), ctypes.c_int32()
    _check(h._lib.is_result_device_arrays(h.h, ctypes.byref(sp), ctypes.byref(stride), None, ctypes.byref(isf)))
    size = 4 if isf.value else 8
    m, v = np.zeros(h.dim), np.zeros(h.dim)
    mc, vc = np.zeros(1), np.zeros(1)
    for c in range(h.dim):
        row = VP(sp.value + c * stride.value * size)
        rc = h._lib.is_result_functional_moments(h.h, row, 0 if isf.value else 1, 1, _p(mc),
                                                 _p(vc) if want_var else None)
        if rc == _lib.IS_EFAILED:
            raise AllZeroWeightsError()
        _check(rc)
        m[c], v[c] = mc[0], vc[0]
    return m, v


def _named_summary(result, values):
    """Per-block summaries of a named layout (_weighted_sum / _weighted_variance over a
    NamedTuple, src/statistics.jl:191-194,215-223)."""
    lay = _handle(result).layout
    if lay.names is None:
        return values
    out = {}
    for b, name in enumerate(lay.names):
        o, k = lay.offsets[b], lay.lengths[b]
        out[name] = float(values[o]) if lay.scalars[b] else values[o:o + k].copy()
    return out


def _layout_result(result):
    h = getattr(result, "_result", None)
    return h is not None and h.layout is not None and (h.layout.names is not None or h.dim > 64)


def mean(f_or_result, result: WeightedSamples = None):
    """mean(result): self-normalized weighted mean (src/statistics.jl:11-15), per name for
    named samples; mean(f, result): weighted expectation of the scalar functional f,
    evaluated on the device samples (src/statistics.jl:17-31)."""
    if result is not None:
        return _functional_moments(f_or_result, result, False)[0]
    result = f_or_result
    if _layout_result(result):
        return _named_summary(result, _rowwise_moments(result, False)[0])
    m, _, _ = _device_moments(result)
    return float(m[0]) if result.samples.ndim == 1 else m


def var(f_or_result, result: WeightedSamples = None, corrected=False):
    """var(result): uncorrected weighted variance (src/statistics.jl:39-46); var(f, result):
    that of the scalar functional f (src/statistics.jl:50-64)."""
    if corrected:
        raise ValueError("corrected=true is unsupported")
    if result is not None:
        return _functional_moments(f_or_result, result, True)[1]
    result = f_or_result
    if _layout_result(result):
        return _named_summary(result, _rowwise_moments(result, True)[1])
    _, v, _ = _device_moments(result)
    return float(v[0]) if result.samples.ndim == 1 else v


def cov(result: WeightedSamples, corrected=False):
    """Uncorrected weighted covariance (src/statistics.jl:96-103); named samples have none
    (src/statistics.jl:238-240)."""
    if corrected:
        raise ValueError("corrected=true is unsupported")
    h = getattr(result, "_result", None)
    if h is not None and h.layout is not None and h.layout.names is not None:
        raise ValueError("covariance requires vector-valued samples")
    _, _, c = _device_moments(result)
    return c


def resample(uniforms, result: WeightedSamples, count=None, method=None):
    """resample (src/resampling.jl:245-283): multinomial (count uniforms) or systematic
    (one uniform) ancestors of the normalised weights, the samples gathered on the
    device into an UnweightedSamples (its `ancestors` are the 1-based indices)."""
    method = method or MultinomialResampling()
    h = _handle(result)
    count = len(result) if count is None else int(count)
    if count <= 0:
        raise ValueError("resampling count must be positive")
    u = _f64(np.atleast_1d(uniforms))
    if u.size < (1 if isinstance(method, SystematicResampling) else count):
        raise ValueError("one uniform per draw (one for systematic resampling)")
    anc = np.zeros(count, dtype=np.int64)
    out = VP()
    rc = h._lib.is_result_resample(h.h, _p(u), count, method.code, _p(anc, I64P), ctypes.byref(out))
    if rc == _lib.IS_EFAILED:
        raise AllZeroWeightsError()
    _check(rc)
    return UnweightedSamples(_ResultHandle(out, h.dim, h.scalar, h.layout), anc)
