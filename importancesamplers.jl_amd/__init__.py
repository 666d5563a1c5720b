"""MI355X (gfx950) engine for the ImportanceSamplers.jl sampling hot path.

The compute lives in ``libissampler.so`` (hand-written HIP kernels behind the
C ABI of ``include/issampler.h``).  This package is the host-side mirror of the
reference API used by the tests and the benchmark.
"""

from .api import (  # noqa: F401
    AMIS,
    NPMC,
    AllZeroWeightsError,
    AMISRoundError,
    DeterministicMixturePMC,
    DiagonalGaussian,
    DiagonalStudentT,
    DMPMCRoundError,
    FactorGaussian,
    FactorStudentT,
    GaussianTarget,
    GlobalResampling,
    ImportanceSampling,
    InvalidTransformError,
    LinearRegressionTarget,
    LocalResampling,
    LogisticRegressionTarget,
    MultinomialResampling,
    NPMCRoundError,
    PartialDeterministicMixture,
    PrefilledRNG,
    RandomMixture,
    ProposalBank,
    SamplerBusyError,
    SamplerExecutionError,
    SphericalGaussian,
    SphericalStudentT,
    StandardGaussianTarget,
    StandardMIS,
    StratifiedMixture,
    SystematicResampling,
    UnweightedSamples,
    WeightedSamples,
    context,
    cov,
    BatchTarget,
    DeviceValues,
    IdentityTransform,
    IntervalTransform,
    PositiveTransform,
    SoftplusTransform,
    current_proposal,
    importance_sample,
    importance_sample_,
    logdensityof,
    lognormalizer,
    mean,
    prepare_sampler,
    resample,
    student_t_logconstant,
    var,
)
from .batch import (  # noqa: F401
    BatchRoundErrors,
    PreparedBatch,
    autotune,
    importance_sample_batch_,
    prepare_batch,
)
from .serve import Server  # noqa: F401
from . import _lib  # noqa: F401
