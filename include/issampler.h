/*
 * issampler.h -- C ABI of the MI355X (gfx950) importance-sampling engine.
 *
 * The engine replaces the device-side execution of ImportanceSamplers.jl's
 * sampling hot path.  A host binding (Julia `ccall`, Python ctypes, ...) maps
 * the reference's backend hooks onto these entry points; every declaration
 * names the reference interface it stands in for (paths relative to the
 * reference repository root).
 *
 * Conventions
 *   - Plain C types only; host pointers unless a name says `device`.
 *   - Samples cross the boundary in the reference's layout: a d x n column-major
 *     matrix (each sample's d coordinates contiguous), src/storage.jl:74-85.
 *   - Return value: 0 = IS_OK, negative = API misuse or device error,
 *     positive = the run failed inside the sampler (IS_EFAILED); details via
 *     is_job_failure(), which mirrors the packed failure record of
 *     src/execution.jl:669-767.
 *   - A job handle is non-reentrant: a concurrent call on a busy handle
 *     returns IS_EBUSY (SamplerBusyError, src/methods/importance_sampling.jl:105-116).
 *     Distinct handles may be used from distinct host threads.
 */
#ifndef ISSAMPLER_H
#define ISSAMPLER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define IS_OK 0
#define IS_EINVAL (-1)        /* ArgumentError / DimensionMismatch at preparation */
#define IS_ENOMEM (-2)
#define IS_EBUSY (-3)         /* SamplerBusyError */
#define IS_EDEVICE (-4)       /* HIP runtime error */
#define IS_EUNSUPPORTED (-5)  /* configuration outside the engine's scope */
#define IS_ESTATE (-6)        /* call order violated (e.g. results before a run) */
#define IS_EFAILED 1          /* sampler execution failed: see is_job_failure */

/* ---- enumerations (reference types in parentheses) ------------------------------- */
enum is_method {
  IS_PLAIN = 0,   /* ImportanceSampling(proposal; nsamples), src/methods/importance_sampling.jl */
  IS_AMIS = 1,    /* AMIS(proposal; rounds, round_size), src/methods/amis.jl:42-47 */
  IS_NPMC = 2,    /* NPMC(proposal; rounds, round_size), src/methods/npmc.jl:21-32 */
  IS_DMPMC = 3,   /* DeterministicMixturePMC(bank; rounds, round_size), src/methods/dm_pmc.jl:72-120 */
  IS_MIS = 4      /* ImportanceSampling(bank::ProposalBank; nsamples, mis_scheme), src/methods/static_mis.jl */
};
enum is_family { IS_GAUSSIAN = 0, IS_STUDENT_T = 1 };             /* GaussianFamily / StudentTFamily */
enum is_scale_kind { IS_SPHERICAL = 0, IS_DIAGONAL = 1, IS_FACTOR = 2 }; /* src/proposals.jl:158-169 */
enum is_precision { IS_FP64 = 0, IS_FP32 = 1 };
enum is_resampling {        /* DM-PMC ancestor selection (src/methods/dm_pmc.jl:10-26, src/resampling.jl:13) */
  IS_MULTINOMIAL = 0,       /* GlobalResampling with MultinomialResampling                              */
  IS_SYSTEMATIC = 1,        /* GlobalResampling, systematic points (new)                                */
  IS_LOCAL = 2              /* LocalResampling: each proposal's ancestor from its own group             */
};
enum is_mis_scheme {       /* static MIS designs (src/methods/static_mis.jl:100-147)             */
  IS_MIS_STRATIFIED = 0,   /* StratifiedMixture: stratified assignment, full mixture             */
  IS_MIS_RANDOM = 1,       /* RandomMixture: independent assignment, full mixture                */
  IS_MIS_STANDARD = 2,     /* StandardMIS: stratified assignment, generating proposal only       */
  IS_MIS_PARTIAL = 3       /* PartialDeterministicMixture(groups): stratified, the slot's group  */
};
enum is_target_kind {
  IS_TARGET_STANDARD = 0,  /* -sum(abs2, x)/2                                              */
  IS_TARGET_GAUSSIAN = 1,  /* params: mean[d], L[d*d] column-major lower, lognormalizer     */
  IS_TARGET_LOGISTIC = 2,  /* params: n, p, prior_var, X[n*p] row-major, y[n]               */
  IS_TARGET_LINREG = 3,    /* params: n, p, sigma, tau, X[n*p] row-major, y[n]              */
  IS_TARGET_CALLBACK = 4   /* a caller's batch target: target_batch, target_user             */
};
/* Batch target callback (batch!(values, samples), src/targets.jl:17-43,
 * src/batched_targets.jl:92-114): set values[i], i < count, to the log density of
 * sample i, whose coordinate c is element c * coord_stride + i of `samples` (FP32
 * when is_float, else FP64, as `values`).  Both are device memory of the job and
 * stream-ordered on `stream` (a hipStream_t): enqueue the work there, or synchronize
 * it before reading and finish private work before returning.  values start as NaN;
 * -Inf is a valid value, NaN and +Inf fail the round in the target phase.  Inputs are
 * borrowed; a nonzero return aborts the run.  Jobs with a callback run eagerly (no
 * graph replay). */
typedef int (*is_batch_target_fn)(void* user, const void* samples, int64_t coord_stride, int32_t dim,
                                  int64_t count, void* values, int32_t is_float, void* stream);
enum is_transform_kind {   /* scalar sample transforms (src/transforms.jl:17-235)             */
  IS_TRANSFORM_IDENTITY = 0,
  IS_TRANSFORM_POSITIVE = 1, /* x = exp(z), log|J| = z                                         */
  IS_TRANSFORM_SOFTPLUS = 2, /* x = softplus(z), log|J| = z - softplus(z)                      */
  IS_TRANSFORM_INTERVAL = 3, /* bounds (l, u): l + exp(z), u - exp(z), or l + (u - l) logistic(z) */
  IS_TRANSFORM_SIMPLEX = 4   /* SimplexTransform(K) (src/transforms.jl:51-70,402-456): K - 1
                                coordinates to K positive weights summing to one, through the
                                orthonormal sum-zero logit embedding; log|J| = log(K)/2 + sum(log, x) */
};
/* One block of a named flat transform layout (prepare_sampler's transform=(name = selector
 * => transform, ...), src/proposal_composition.jl:179-240, src/native_execution.jl:556-712).
 * Blocks are listed in name order, their selectors disjoint and together covering the d
 * sampling coordinates; block b's logical values follow those of blocks 0..b-1 (a simplex
 * block gives K = count + 1 of them, every other block count). */
typedef struct is_transform_block {
  int32_t first;           /* first sampling coordinate of the selector, 0-based */
  int32_t count;           /* its sampling coordinates (a simplex: K - 1) */
  int32_t kind;            /* enum is_transform_kind */
  int32_t scalar;          /* 1: an Int selector (a scalar leaf, count 1), 0: a UnitRange */
  double lower, upper;     /* IS_TRANSFORM_INTERVAL bounds (-inf / +inf for an open side) */
} is_transform_block;
enum is_phase {            /* AMISRoundError / NPMCRoundError / DMPMCRoundError phases       */
  IS_PHASE_NONE = 0,
  IS_PHASE_SAMPLING = 1,   /* :sampling / :proposal_draw                                     */
  IS_PHASE_TARGET = 2,     /* :target                                                        */
  IS_PHASE_DENOMINATOR = 3,/* :denominator / :proposal_logdensity                            */
  IS_PHASE_WEIGHT = 4,     /* :weight / :logweight                                           */
  IS_PHASE_MOMENT = 5,     /* :moment (AllZeroWeightsError)                                  */
  IS_PHASE_FACTORIZATION = 6, /* :factorization (PosDefException / invalid candidate)        */
  IS_PHASE_RESAMPLING = 7  /* :resampling (AllZeroWeightsError)                              */
};

typedef struct is_ctx is_ctx;
typedef struct is_job is_job;

/* Job descriptor: the prepared-sampler state of _prepare_importance_sampler
 * (src/methods/importance_sampling.jl:357-440) and the method states
 * (src/adaptive_gaussian.jl:208-234, src/population_preparation.jl:370-400). */
typedef struct is_job_desc {
  int32_t method;          /* enum is_method */
  int32_t family;          /* enum is_family */
  int32_t scale_kind;      /* enum is_scale_kind */
  int32_t scalar;          /* 1: scalar location (d = 1, scalar history) */
  int32_t dim;             /* d */
  double dof;              /* Student-t degrees of freedom */
  int32_t rounds;          /* T (plain IS: 1) */
  const int64_t* round_size; /* T entries (plain IS: nsamples) */
  /* single proposal (plain, AMIS, NPMC) */
  const double* location;  /* d */
  const double* scale;     /* spherical: 1, diagonal: d, factor: d*d column-major lower */
  /* proposal bank (DM-PMC, static MIS): configured order, zero masses are inert */
  int32_t bank_size;       /* K */
  const double* bank_locations; /* d*K column-major */
  const double* bank_scales;    /* spherical: K, diagonal: d*K, factor: d*d*K */
  const double* bank_masses;    /* K */
  int32_t resampling;      /* enum is_resampling */
  /* target */
  int32_t target_kind;     /* enum is_target_kind */
  const double* target_params;
  int64_t target_params_len;
  /* execution */
  int32_t precision;       /* enum is_precision */
  uint64_t seed;           /* counter-RNG key word 0 */
  uint64_t stream;         /* counter-RNG key word 1 */
  /* static MIS (method IS_MIS, the bank fields above hold the ProposalBank) */
  int32_t mis_scheme;      /* enum is_mis_scheme */
  int32_t mis_group_count; /* partial mixture: number of groups G */
  const int32_t* mis_group_offsets; /* G + 1 offsets into the members, starting at 0 */
  const int32_t* mis_group_members; /* configured 1-based proposal ids, in group order, each id once */
  /* per-coordinate transforms (prepare_sampler's transform=, src/methods/importance_sampling.jl:308-313):
   * NULL, or d enum is_transform_kind values with 2d bounds (lower, upper, -inf / +inf for an
   * open side).  Sampling, fits and mixtures stay in unconstrained coordinates z, the target
   * sees x = T(z), the weights gain the forward log Jacobian and results report x (every
   * target kind, batched and callback targets receiving the logical samples).  A failing
   * transform fails the draw with its reason bits (0x000f) and the coordinate as block. */
  const int32_t* transform_kinds;
  const double* transform_bounds;
  /* caller's batch target (target_kind IS_TARGET_CALLBACK) */
  is_batch_target_fn target_batch;
  void* target_user;       /* passed back as the callback's first argument */
  /* named flat transform layout: 0, or that many blocks (transform_kinds then NULL).  The
   * target sees the logical sample of dl = sum of the blocks' logical lengths values
   * (is_job_logical_dim, the dimension of the target's parameters), results report it, and
   * a failing block fails the draw with its reason bits and its 1-based block index. */
  int32_t transform_block_count;
  const is_transform_block* transform_blocks;
} is_job_desc;

/* Failure snapshot: _device_failure_snapshot (src/execution.jl:719-734) plus
 * the round / phase of the adaptive error types. */
typedef struct is_failure {
  int32_t round;           /* 1-based failing round, 0 when none */
  int32_t phase;           /* enum is_phase */
  uint64_t count;
  int64_t first_logical_index;
  int32_t first_block;
  uint16_t reason_bits;
  int64_t draw_first_logical_index; /* 0 when no draw failure */
  uint16_t draw_reason_bits;
  uint64_t words[3];       /* raw packed record */
  int32_t draw_first_block; /* block of the first draw failure (a transform failure: the
                               1-based coordinate of a per-coordinate transform, or the
                               1-based block of a named layout) */
} is_failure;

/* ---- context ---------------------------------------------------------------------
 * _with_backend_device / _backend_functional (src/methods/importance_sampling.jl:507-508,
 * ext/ImportanceSamplersCUDAExt.jl:9-19). */
int is_ctx_create(int device, is_ctx** out);
int is_ctx_destroy(is_ctx* ctx);
const char* is_last_error(void);

/* ---- jobs ------------------------------------------------------------------------
 * is_job_create: prepare_sampler + device transfer + _prepare_backend_execution
 * (src/methods/importance_sampling.jl:357-404, 599-719). */
int is_job_create(is_ctx* ctx, const is_job_desc* desc, is_job** out);
int is_job_destroy(is_job* job);
/* Device memory (bytes) a job of this descriptor allocates, or -1 for an invalid
 * descriptor; no reference counterpart (the serving loop's admission estimate). */
int64_t is_job_device_bytes(const is_job_desc* desc);

/* importance_sample! -> _execute_prepared_sampler! (src/methods/importance_sampling.jl:912-953);
 * adaptive loops src/adaptive_gaussian_execution.jl:897-1112 and src/population_execution.jl:407-567.
 * Asynchronous on the job's stream up to the final failure snapshot. */
int is_job_run(is_job* job);

/* RNG seams mirroring the prefilled-RNG test pattern (test/amis_kernel.jl:15-34,
 * test/npmc.jl:5-13, test/support/dm_pmc.jl:10-42): the given round uses these
 * normals / uniforms instead of the counter RNG until is_job_clear_prefilled.
 * normals: plain IS stride*n per src/execution.jl:560-583; adaptive rounds d*N_t;
 * radial: Student-t gamma buffers (8 normals, 8 + (nu < 2) uniforms per sample; 1 normal for nu == 1). */
int is_job_set_normals(is_job* job, int32_t round, const double* normals, int64_t count);
int is_job_set_radial(is_job* job, int32_t round, const double* normals, int64_t ncount,
                      const double* uniforms, int64_t ucount);
int is_job_set_uniforms(is_job* job, int32_t round, const double* uniforms, int64_t count);
/* DM-PMC replay seam: the given round resamples to these K ancestors (1-based indices into
 * the round's samples, active-proposal order) instead of selecting them from its weights
 * and uniforms, until is_job_clear_prefilled.  It replays another run's bank trajectory
 * (is_job_get_ancestors), e.g. an FP64 run on an FP32 run's banks, so the per-round log Z
 * of the two precisions is compared on the same proposals (src/dm_pmc_execution.jl:267-295). */
int is_job_set_ancestors(is_job* job, int32_t round, const int64_t* ancestors, int64_t count);
int is_job_clear_prefilled(is_job* job);
/* The counter-RNG values the job consumes in `round` (for feeding the oracle). */
int is_job_export_normals(is_job* job, int32_t round, double* out, int64_t count);
int is_job_export_radial(is_job* job, int32_t round, double* normals, int64_t ncount,
                         double* uniforms, int64_t ucount);
int is_job_export_uniforms(is_job* job, int32_t round, double* out, int64_t count);

/* Event timing of the round kernel (draw / append / mixture / weight pass) of
 * each round, on the job's stream; no reference counterpart (measurement seam). */
int is_job_set_timing(is_job* job, int32_t enable);
int is_job_round_kernel_ms(is_job* job, double* out); /* T entries, last run */
/* Replay a run's launch sequence from one captured HIP graph (default on); a timed run
 * is always issued eagerly.  No reference counterpart (launch-overhead control). */
int is_job_set_graph(is_job* job, int32_t enable);

/* ---- results (WeightedSamples, src/results.jl:79-98) -------------------------------- */
int64_t is_job_total(const is_job* job);
/* Logical dimension of the samples results report: d, or dl under a named transform layout. */
int32_t is_job_logical_dim(const is_job* job);
int is_job_failure(const is_job* job, is_failure* out);
int is_job_get_samples(is_job* job, double* out);           /* logical dim x total */
int is_job_get_logweights(is_job* job, double* out);        /* total */
int is_job_get_round_ids(is_job* job, int64_t* out);        /* provenance.round */
int is_job_get_proposal_ids(is_job* job, int64_t* out);     /* provenance.proposal_id (DM-PMC) */
/* diagnostics.round_ess / round_lognormalizers / adaptation_ess (src/adaptive_gaussian_execution.jl:1068-1089) */
int is_job_get_round_summary(is_job* job, double* ess, double* lognormalizers, double* adaptation_ess);
/* current_proposal (src/adaptive_gaussian.jl:382-412): committed mean, factor (d*d column-major
 * lower; scalar: 1) and lognormalizer. */
int is_job_get_proposal(is_job* job, double* location, double* factor, double* lognormalizer);
/* Fitted proposal history slots 1..T+1 (means (T+1)*d, factors (T+1)*d*d, lognormalizers T+1). */
int is_job_get_history(is_job* job, double* means, double* factors, double* lognormalizers);
/* DM-PMC: current_proposal locations (d x K, active order) and per-round ancestors (K x T, 1-based). */
int is_job_get_locations(is_job* job, double* out);
int is_job_get_ancestors(is_job* job, int64_t* out);
/* Device-resident logweights / samples (SoA, coordinate stride = capacity) of the last run.
 * The device log-weights are the log-weights minus is_job_logweight_offset. */
int is_job_device_results(is_job* job, const void** samples, int64_t* coord_stride,
                          const void** logweights, int32_t* is_float);
/* Offset c kept off the device-resident log-targets and log-weights: in FP32 mode
 * with a regression target (n data rows) they are stored minus c = log pi(x0), the
 * FP64 log target at the initial location (mean of the bank locations for DM-PMC),
 * so their FP32 rounding stays at the scale of the differences between samples.
 * Every result getter adds c back; 0 in every other case.  No reference
 * counterpart (mixed-precision storage). */
int is_job_logweight_offset(const is_job* job, double* out);
/* Reset the committed proposal (AMIS/NPMC) or locations (DM-PMC) to the descriptor's. */
int is_job_reset(is_job* job);

/* ---- results API on the last run --------------------------------------------------
 * lognormalizer (src/results.jl:560-567): out = {log Z, max, scaled_sum};
 * mean / var / cov (src/statistics.jl:11-103): FP64 device reductions. */
int is_job_lognormalizer(is_job* job, double* out3);
int is_job_moments(is_job* job, double* mean, double* var, double* cov);
/* mean(f, result) / var(f, result) (src/statistics.jl:17-31,50-64,144-199) for a scalar
 * functional: values holds f(x_i) for the last run's samples in sample order (total
 * samples; FP64 when values_fp64, else FP32), in device memory when values_on_device
 * (the device functional path: the caller evaluates f on is_job_device_results'
 * samples), else in host memory (uploaded first).  mean = sum_i w_i f_i and
 * var = sum_i w_i (f_i - mean)^2 with w the normalised weights of the device-resident
 * log-weights, reduced in FP64; var may be NULL.  IS_EFAILED when every weight is zero. */
int is_job_functional_moments(is_job* job, const void* values, int32_t values_fp64, int32_t values_on_device,
                              double* mean, double* var);
/* resample (src/resampling.jl:245-283): multinomial or systematic ancestors of the
 * last run's logweights for host uniforms (count, or 1 for systematic). */
int is_job_resample(is_job* job, const double* uniforms, int64_t count, int32_t method,
                    int64_t* ancestors);

/* ---- results with their own device storage ---------------------------------------
 * A WeightedSamples owns its storage: later runs of the sampler do not change it
 * (src/methods/importance_sampling.jl:904-906, src/results.jl:79-98).  is_job_result
 * copies the last run's logical samples and log-weights into a result on the device;
 * the results API below works on that copy.  Destroy with is_result_destroy. */
typedef struct is_result is_result;
int is_job_result(is_job* job, is_result** out);
int is_result_destroy(is_result* result);
int64_t is_result_count(const is_result* result);
/* 1 for WeightedSamples, 0 for the UnweightedSamples of is_result_resample */
int is_result_is_weighted(const is_result* result);
int is_result_get_samples(is_result* result, double* out);     /* logical dim x count */
int is_result_get_logweights(is_result* result, double* out);  /* count */
/* device arrays (coordinate-major, stride = count; log-weights minus the offset, NULL
 * for unweighted samples) -- valid until is_result_destroy */
int is_result_device_arrays(is_result* result, const void** samples, int64_t* coord_stride,
                            const void** logweights, int32_t* is_float);
int is_result_logweight_offset(const is_result* result, double* out);
/* lognormalizer, mean / var / cov and mean(f, r) / var(f, r) as the is_job_ forms above
 * (src/results.jl:560-567, src/statistics.jl:11-103,144-199); unweighted samples have
 * equal weights in the moments and no lognormalizer */
int is_result_lognormalizer(is_result* result, double* out3);
int is_result_moments(is_result* result, double* mean, double* var, double* cov);
int is_result_functional_moments(is_result* result, const void* values, int32_t values_fp64,
                                 int32_t values_on_device, double* mean, double* var);
/* resample(rng, result, count; method) (src/resampling.jl:245-283): the CDF of the
 * result's weights in the engine's fixed wavefront-scan order, the ancestors of the host
 * uniforms (count of them, or one for systematic), optionally copied to `ancestors`
 * (1-based), and optionally the UnweightedSamples gathered on the device
 * (src/resampling.jl:120-164,263-283) as a new result in `*unweighted`. */
int is_result_resample(is_result* result, const double* uniforms, int64_t count, int32_t method,
                       int64_t* ancestors, is_result** unweighted);

/* ---- primitives on device buffers (stream = hipStream_t or NULL) -------------------
 * _resampling_cdf! (src/resampling.jl:199-225) with the engine's fixed wavefront-scan
 * order (per-thread runs of 16, shuffle scans within waves, wave totals in order, tile
 * carries by the same scheme),
 * and ancestor selection (src/resampling.jl:92-118). */
int is_cdf_device(is_ctx* ctx, const double* logweights_device, int64_t n, double* cdf_device,
                  void* stream);
int is_select_device(is_ctx* ctx, const double* cdf_device, const double* logweights_device, int64_t n,
                     const double* uniforms_device, int64_t count, int32_t method,
                     int32_t* ancestors_device, void* stream);
/* mean(f, result) / var(f, result) over device arrays (the device functional path of
 * src/statistics.jl:17-31,50-64,144-163 for a result adopted from device arrays,
 * src/results.jl:485-546): log-weights and per-sample values, each FP64 or FP32.
 * Zero-weight samples enter the mean as 0 * f_i (a non-finite f_i there gives NaN, as
 * dot(weights, values) does) and are skipped by the variance
 * (src/statistics.jl:188-197). */
int is_functional_moments_device(is_ctx* ctx, const void* logweights_device, int32_t logweights_fp64,
                                 const void* values_device, int32_t values_fp64, int64_t n, double* mean,
                                 double* var, void* stream);
/* Allocation-free form of is_functional_moments_device, with the same arguments and results.
 * No counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * calls no allocating routine: its partial records live in a fixed 1 MiB device array per GPU
 * (bound at is_ctx_create), which limits n to 67,104,768 samples (IS_EINVAL beyond).  Every
 * call on one GPU shares that array, so calls that overlap on one GPU (different streams or
 * host threads) race; serialise them. */
int is_functional_moments_device_noalloc(is_ctx* ctx, const void* logweights_device, int32_t logweights_fp64,
                                         const void* values_device, int32_t values_fp64, int64_t n,
                                         double* mean, double* var, void* stream);
/* Host-buffer forms of the two primitives above (copy in, run, copy out); ancestors 1-based. */
int is_cdf(is_ctx* ctx, const double* logweights, int64_t n, double* cdf);
int is_select(is_ctx* ctx, const double* cdf, const double* logweights, int64_t n, const double* uniforms,
              int64_t count, int32_t method, int64_t* ancestors);
/* logdensityof(proposal, x) (DensityInterface, src/proposals.jl:24-55,603-669 with the
 * extended-range repair of src/student_t_range.jl:55-153) of a native radial proposal at n
 * points (d x n column-major), evaluated on the device in FP64 or FP32 by the sampling
 * kernels' own density code; d <= 64. */
int is_proposal_logdensity(is_ctx* ctx, int32_t family, double dof, int32_t scale_kind, int32_t dim,
                           const double* location, const double* scale, const double* points, int64_t n,
                           int32_t precision, double* out);
/* Raw Philox4x64-10 words for counters (i, c1, c2, c3), i = 0..n-1 (RNG known-answer seam). */
int is_rng_raw(is_ctx* ctx, uint64_t key0, uint64_t key1, uint64_t c1, uint64_t c2, uint64_t c3,
               int64_t n, uint64_t* out);

/* ---- job groups: same-shape samplers advanced together ---------------------------
 * A group holds `count` prepared samplers of one shape -- plain IS or AMIS, one family,
 * dimension d <= 15, precision and round sizes; a built-in standard or Gaussian target;
 * no transforms -- whose proposals, target parameters and seeds are free.  is_group_run
 * runs every member once: each round is one launch over all members (the stand-alone
 * fused round with the member in blockIdx.y) and one grouped reduction, the whole run
 * one captured graph between one staging upload and one status download.  Member m's
 * outcome and results are those of importance_sample! on its own descriptor (the same
 * launch arguments and counter-RNG draws): is_group_job returns it for every is_job_
 * results call and its is_job_failure.  is_group_run returns IS_EFAILED when a member
 * failed.  The group owns its members (destroyed with it); a member may also run alone,
 * not while its group runs (IS_EBUSY).  No reference counterpart: the reference runs
 * such samplers one at a time (src/methods/importance_sampling.jl:912-953). */
typedef struct is_group is_group;
int is_group_create(is_ctx* ctx, const is_job_desc* descs, int32_t count, is_group** out);
int is_group_run(is_group* group);
int32_t is_group_size(const is_group* group);
int is_group_job(is_group* group, int32_t index, is_job** out);
/* is_job_reset on every member: each next run starts from its initial proposal. */
int is_group_reset(is_group* group);
/* Every member's ESS and log normaliser of round `round` (0-based) of the last run, in
 * member order (`count` = group size entries each; either pointer may be null): the
 * values is_job_get_round_summary returns member by member.  A failed member's entries
 * are those its own status holds (zero ESS, -inf log normaliser for the failed round). */
int is_group_round_summary(is_group* group, int32_t round, double* ess, double* lognormalizers);
int is_group_destroy(is_group* group);

/* ---- batched AMIS: many independent posteriors in one persistent launch ----------
 * A batch is a vector of prepared AMIS samplers run together; job j means
 * importance_sample!(prepare_sampler(rng_j, target_j, AMIS(q_j; rounds, round_size)))
 * (src/methods/amis.jl:42-47, src/adaptive_gaussian_execution.jl:897-1112) and its
 * results are those of the job run alone on the same counter-RNG key.  Jobs must be
 * AMIS with a built-in regression target (linear or logistic), d <= 32, equal round
 * sizes; precision is per batch (IS_FP32 or IS_FP64).  One workgroup runs one job for
 * all of its rounds; workgroups pull jobs from a device queue in `order`.
 * No reference counterpart exists for the batch itself: the reference runs such jobs
 * one prepared sampler at a time. */
typedef struct is_batch is_batch;
int is_batch_create(is_ctx* ctx, const is_job_desc* descs, int32_t njobs, int32_t precision, is_batch** out);
int is_batch_destroy(is_batch* batch);
/* importance_sample! over every job; IS_EFAILED when any job failed (per-job status
 * via is_batch_status / is_batch_job_failure; failed jobs keep their committed proposal). */
int is_batch_run(is_batch* batch);
/* Launch shape: waves per workgroup (2, 4, 8) and workgroups per CU; job order (a
 * permutation of 0..njobs-1, NULL = longest estimated job first). */
int is_batch_set_tuning(is_batch* batch, int32_t waves, int32_t blocks_per_cu);
int is_batch_get_tuning(const is_batch* batch, int32_t* waves, int32_t* blocks_per_cu);
int is_batch_set_order(is_batch* batch, const int32_t* order);
int is_batch_get_order(const is_batch* batch, int32_t* order);
/* Autotuning trial: one run with the given launch shape, event-timed, without commit
 * (the next run draws the same counter-RNG values); kernel_ms = persistent-kernel time. */
int is_batch_trial(is_batch* batch, int32_t waves, int32_t blocks_per_cu, double* kernel_ms);
int is_batch_set_timing(is_batch* batch, int32_t enable);
int is_batch_kernel_ms(const is_batch* batch, double* ms);
/* per-job results of the last run */
int is_batch_status(const is_batch* batch, int32_t* status);  /* njobs: 0 ok, else failing round */
int64_t is_batch_job_total(const is_batch* batch, int32_t job);
int is_batch_job_failure(is_batch* batch, int32_t job, is_failure* out);
int is_batch_job_samples(is_batch* batch, int32_t job, double* out);     /* d x total */
int is_batch_job_logweights(is_batch* batch, int32_t job, double* out);  /* total */
int is_batch_job_round_summary(is_batch* batch, int32_t job, double* ess, double* lognormalizers);
int is_batch_final_summary(is_batch* batch, double* ess, double* lognormalizers);  /* njobs each */
int is_batch_job_history(is_batch* batch, int32_t job, double* means, double* factors, double* lognormalizers);
int is_batch_job_proposal(is_batch* batch, int32_t job, double* location, double* factor, double* lognormalizer);
int is_batch_export_normals(is_batch* batch, int32_t job, int32_t round, double* out, int64_t count);
int is_batch_export_radial(is_batch* batch, int32_t job, int32_t round, double* normals, int64_t ncount,
                           double* uniforms, int64_t ucount);

/* ---- serving: a request loop with continuous batching -----------------------------
 * Requests are job descriptors (any method); each is copied at submission, admitted
 * in arrival order when one of `max_active` execution slots is free and its device
 * memory (is_job_device_bytes) fits the budget beside every admitted and untaken job,
 * created (its proposal slots, mixture sums and samples stay resident on the device)
 * and run on its own stream, so jobs in flight overlap on the GPU.  Results are those
 * of the same descriptor run alone.  No reference counterpart: the reference runs
 * importance_sample! (src/methods/importance_sampling.jl:912-930) one call at a time. */
typedef struct is_server is_server;
enum is_request_state { IS_REQ_QUEUED = 0, IS_REQ_RUNNING = 1, IS_REQ_DONE = 2, IS_REQ_FAILED = 3 };
typedef struct is_server_config {
  int64_t memory_budget;   /* bytes of device memory admitted jobs may hold at once */
  int32_t max_active;      /* jobs running at once (1..64) */
  int32_t use_graph;       /* replay each job's launches from a captured graph (0: eager launches) */
} is_server_config;
typedef struct is_server_timing {
  int64_t submit_ns, admit_ns, done_ns;  /* steady-clock times of the request */
  int64_t device_bytes;                  /* admission estimate */
} is_server_timing;
typedef struct is_server_stats {
  int64_t submitted, completed, failed, active, queued, peak_active, bytes_in_use, peak_bytes;
} is_server_stats;
int is_server_create(is_ctx* ctx, const is_server_config* cfg, is_server** out);
/* IS_ENOMEM when the job alone exceeds the budget */
int is_server_submit(is_server* server, const is_job_desc* desc, uint64_t* ticket);
int is_server_poll(is_server* server, uint64_t ticket, int32_t* state);
/* waits until the request finished (timeout_ms < 0: no limit); *state as of the return */
int is_server_wait(is_server* server, uint64_t ticket, int32_t timeout_ms, int32_t* state);
/* hands the finished job (results, failure record) to the caller, who destroys it with
 * is_job_destroy; returns the job's run status (IS_EFAILED: see is_job_failure) */
int is_server_take(is_server* server, uint64_t ticket, is_job** job, is_server_timing* timing);
int is_server_get_stats(is_server* server, is_server_stats* out);
/* finishes the admitted requests, drops the queued ones, frees untaken jobs */
int is_server_destroy(is_server* server);

#ifdef __cplusplus
}
#endif

#endif /* ISSAMPLER_H */
