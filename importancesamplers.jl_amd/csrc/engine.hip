// Host side of the gfx950 importance-sampling engine: job preparation, the
// per-round launch sequences of plain IS / AMIS / NPMC / DM-PMC, failure
// snapshots and the C ABI declared in include/issampler.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <memory>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/issampler.h"
#include "host.h"
#include "launch.h"

#include <type_traits>

using namespace issampler;
using namespace issampler::host;

namespace {

// ---------------------------------------------------------------- exact allocation plan
// _dm_pmc_round_counts / _deterministic_allocation_plan
// (src/population_preparation.jl:266-367): exact largest-remainder counts with
// ties rotated by round.  Masses are converted to integers over a common
// power of two so every comparison is exact.
typedef unsigned __int128 u128;

bool exact_integer_masses(const std::vector<double>& m, std::vector<u128>& out) {
  int emin = 1 << 30;
  std::vector<uint64_t> mant(m.size());
  std::vector<int> ex(m.size());
  for (size_t k = 0; k < m.size(); ++k) {
    int e;
    const double f = std::frexp(m[k], &e);
    uint64_t M = (uint64_t)std::ldexp(f, 53);
    e -= 53;
    while (M && !(M & 1)) {
      M >>= 1;
      ++e;
    }
    mant[k] = M;
    ex[k] = e;
    emin = std::min(emin, e);
  }
  out.resize(m.size());
  for (size_t k = 0; k < m.size(); ++k) {
    const int shift = ex[k] - emin;
    if (shift + 53 > 90) return false;
    out[k] = (u128)mant[k] << shift;
  }
  return true;
}

bool round_counts(const std::vector<u128>& I, int64_t n, int round1, std::vector<int64_t>& counts) {
  const int K = (int)I.size();
  u128 S = 0;
  for (auto v : I) S += v;
  counts.assign(K, 0);
  std::vector<u128> rem(K);
  int64_t used = 0;
  for (int k = 0; k < K; ++k) {
    const u128 q = I[k] * (u128)n;
    counts[k] = (int64_t)(q / S);
    rem[k] = q % S;
    used += counts[k];
  }
  const int64_t remaining = n - used;
  if (remaining < 0 || remaining >= K) return false;
  const int first_tie = (round1 - 1) % K + 1;
  std::vector<int> order(K);
  for (int k = 0; k < K; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (rem[a] != rem[b]) return rem[a] > rem[b];
    const int ra = ((a + 1 - first_tie) % K + K) % K, rb = ((b + 1 - first_tie) % K + K) % K;
    return ra < rb;
  });
  for (int64_t i = 0; i < remaining; ++i) counts[order[i]] += 1;
  return true;
}

}  // namespace

struct is_job {
  is_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  std::atomic<bool> busy{false};
  // configuration
  int method = 0, family = 0, kind = 0, scalar = 0, d = 0, T = 1, K = 0, resampling = 0;
  int tkind = 0, precision = 0;
  double dof = 0.0, logconstant = 0.0;
  std::vector<int64_t> schedule, offsets;
  int64_t total = 0, maxN = 0;
  RngKey key{0, 0};
  uint64_t run_index = 0;
  int stride = 0;   // slot stride
  int nslots = 0;   // slot table entries
  int D = 4;        // compile-time dimension bucket
  int64_t pcap = 1; // partial-record column length
  // committed state
  HostProp initial, committed;
  std::vector<double> bank_slots_init;   // K slots (FP64)
  std::vector<double> committed_locs;    // K*d
  std::vector<int32_t> bank_ids;         // 1-based configured ids of active slots
  std::vector<int64_t> plan_counts;      // K*T
  std::vector<double> plan_logcoef;      // K*T
  // static MIS: scheme, active-slot cumulative masses and log masses, partial groups
  int mis_scheme = 0;
  std::vector<double> mis_cdf, mis_logmass;
  std::vector<int32_t> mis_slot_group, mis_goff, mis_gmem;
  std::vector<double> mis_gcoef;
  DevBuf mis_cdf_d, mis_ints_d, mis_gcoefT;
  std::vector<double> tslot;             // Gaussian target slot
  // batched regression target (logistic / linear): data on the device
  int64_t tn = 0;                        // data rows
  double t_inv_two_var = 0.0;            // linear regression 1 / (2 sigma^2)
  double t_inv_two_prior = 0.0;          // 1 / (2 prior variance)
  double t_const = 0.0;                  // log normalising constants
  double t_shift = 0.0;                  // log-target offset kept off the device values (FP32 regression)
  std::vector<double> tdata;             // X (n x p row-major) then y (n), FP64, until upload
  DevBuf Xd, yd, tpre, tpart;
  // FP32 logistic target, p <= 64: the centred BF16 form (k_logistic_bf16): the rows'
  // BF16 split fragments, x_i . x0 and the responses (FP32, padded), the centre x0
  std::vector<double> t_x0;              // the centre: the log-target offset's anchor
  bool t_bf16 = false;
  int t_ks2 = 0;                          // 32-wide K steps
  DevBuf t_xfrag, t_base, t_x0d;
  // FP32 jobs with a two-pass (batch-target) round, 16 <= d <= 64: the mixture pass on
  // the matrix cores (k_mix_prep, k_mix_mfma): inverse-factor fragments, centres and
  // log normalisers of up to nslots components, and the samples' mixture values
  DevBuf mix_frag, mix_mu, mix_lognorm, mix_out;
  // device buffers
  DevBuf X, logtargets, lognums, logweights, round_ids, proposal_ids;
  DevBuf slots64, slotsT, logmassT, zerosT, tslotT, assign, bank_ids_d;
  DevBuf partials, partials2, reduced, reduced2, summary, clip, cdf, tile_total, uniforms, ancestors, scratch;
  // per-run control block: one device allocation holding the run word, the
  // finish counters, the failure records and the round summaries.  Each run
  // uploads [0, ctl_up_bytes) (run word + zeroed status) from a pinned image and
  // reads [ctl_down_off, ctl.bytes) (status + round summaries) back in one copy.
  DevBuf ctl;
  DevView run_dev, counter, fail_rec, fit_status, cdf_status, round_ess, round_logz, adapt_ess;
  size_t ctl_up_bytes = 0, ctl_down_off = 0;
  HostPinned ctl_up, ctl_down;
  // pinned staging for the committed slot upload and the commit read-back
  HostPinned slot_up, slot_down;
  bool slots_resident = false;  // plain IS / static MIS: slot table already holds the committed state
  DevBuf pre_normals, pre_rn, pre_ru, pre_uniforms;
  std::vector<int> prefilled, prefilled_radial, prefilled_uniforms, prefilled_ancestors;
  DevBuf pre_ancestors;  // DM-PMC: K x T given ancestors (1-based), replacing the resampling step
  DevBuf plan_group;     // DM-PMC: per round, the K + 1 group boundaries of the samples by proposal
  // last run
  bool have_results = false;
  int last_status = 0;
  is_failure failure{};
  std::vector<int64_t> ancestors_host;
  // the run word (run index << 32) lives in the control block: RNG counters read
  // it from the device, so the captured launch sequence stays valid across runs
  // the run's launch sequence captured as a HIP graph (replayed while valid)
  int use_graph = 1;
  // caller's batch target (IS_TARGET_CALLBACK): called between the draw and the
  // mixture pass of every round, on the job's stream, outside any graph capture
  is_batch_target_fn tbatch = nullptr;
  void* tuser = nullptr;
  // per-coordinate transforms or a named layout: kinds and bounds (or the layout's blocks)
  // on the device, logical samples XL (dl rows; dl = d without a layout)
  bool has_tf = false;
  DevBuf tf_kinds_d, tf_bounds_d, XL;
  bool has_layout = false;
  int dl = 0;
  std::vector<is_transform_block> layout;
  DevBuf layout_d;
  bool graph_dirty = true;
  // replay runs on the extended-range kernels (set after a run's densities left the
  // floating-point range; kept while runs succeed there)
  bool range_mode = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // optional per-round event timing of the round kernel
  int timing = 0;
  std::vector<hipEvent_t> events;  // 2 per round
  std::vector<double> round_ms;
};

namespace {

struct BusyGuard {
  is_job* job;
  bool ok;
  explicit BusyGuard(is_job* j) : job(j) {
    bool expected = false;
    ok = job->busy.compare_exchange_strong(expected, true);
  }
  ~BusyGuard() {
    if (ok) job->busy.store(false);
  }
};

void tick(is_job* j, int t, int which, hipStream_t s) {
  if (j->timing && (size_t)(2 * t + which) < j->events.size()) (void)hipEventRecord(j->events[2 * t + which], s);
}



// ---------------------------------------------------------------- launch dispatch
constexpr int64_t kSweepPerBlock = 512;  // samples per sweep block (one partial record)
// samples per sweep block of the fused plain-IS / static-MIS round (it has no sweep
// blocks; the AMIS rounds size theirs with round_sweep_per_block)
constexpr int64_t kRoundPerBlock = 1024;

// Calls f(std::integral_constant<int, D>) for the job's dimension instantiation.
template <class F>
void with_dim(int D, F&& f) {
  switch (D) {
#define IS_DIM_CASE(v)                     \
  case v:                                  \
    f(std::integral_constant<int, v>{});   \
    break;
    IS_DIMS(IS_DIM_CASE)
#undef IS_DIM_CASE
    default:
      break;
  }
}

// Waves per new-sample block: enough that each wave evaluates a few mixture
// components and the Box-Muller pairs of the block's 64 samples spread over
// its threads; at most 8, so that three blocks fit on a CU and a
// 20,000-sample round (313 blocks) starts in one pass.
int new_waves(int d, int ncomp) {
  const int np = (d + 1) / 2;
  return (ncomp >= 16 || np > 4) ? 8 : 4;
}

template <class T, int MODE>
void launch_mix(int D, int ncomp, const RoundArgs<T>& a, hipStream_t s) {
  const int W = new_waves(a.d, ncomp);
  const int blocks = (int)std::max<int64_t>(1, cdiv(a.n_new, 64));
  if (D == kRangeDims)
    RangeKernels::new_samples(MODE, a, ncomp, blocks, W, s);
  else
    with_dim(D, [&](auto k) { DimKernels<decltype(k)::value>::new_samples(MODE, a, ncomp, blocks, W, s); });
}

// Launches one sweep over [a.lo, a.hi); returns the number of partial records.
template <class T>
int launch_sweep(int D, int nt, SweepArgs<T> a, hipStream_t s) {
  const int nw = sweep_waves(nt);
  const int blocks = (int)std::max<int64_t>(1, cdiv(a.hi - a.lo, kSweepPerBlock));
  a.per_block = kSweepPerBlock;
  const size_t lds = sweep_lds(nt, nw);
  if (D == kRangeDims)
    RangeKernels::sweep(nt, a, blocks, nw, lds, s);
  else
    with_dim(D, [&](auto k) { DimKernels<decltype(k)::value>::sweep(nt, a, blocks, nw, lds, s); });
  return blocks;
}

// Launches one fused AMIS round (k_round); false when the job's dimension
// instantiation has none, and nothing was launched.
template <class T>
bool launch_round(int D, int nt, int ncomp, const RoundArgs<T>& a, const SweepArgs<T>& w, int nb_new, int nrec_old,
                  hipStream_t s) {
  bool done = false;
  with_dim(D, [&](auto k) { done = DimKernels<decltype(k)::value>::round(nt, a, w, ncomp, nb_new, nrec_old, s); });
  return done;
}

template <class T>
void launch_target(int D, const TargetArgs<T>& t, int64_t splits, hipStream_t s) {
  with_dim(D, [&](auto k) { DimKernels<decltype(k)::value>::target(t, splits, s); });
}

// Summary / fit on a reduced record (one wavefront, any d <= 64).
#define IS_EXACT_DIMS(X) \
  X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

void launch_finish(int d, const double* reduced, const FinishArgs& f, hipStream_t s) {
  switch (d) {
#define IS_FINISH_CASE(v)                                                                     \
  case v:                                                                                     \
    hipLaunchKernelGGL((k_finish_small<v, true>), dim3(1), dim3(64), 0, s, reduced, f); \
    return;
    IS_EXACT_DIMS(IS_FINISH_CASE)
#undef IS_FINISH_CASE
    default:
      break;
  }
  if (d <= 32)
    hipLaunchKernelGGL((k_finish_small<32, false>), dim3(1), dim3(64), 0, s, reduced, f);
  else
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, s, reduced, f);
}

// Reduction of `nrec` partial records into `reduced`, then summary / fit: one
// launch for d <= 32 (the last reduction block fits; f.counter is its ticket
// word; d <= 16 compiled for the exact d), a reduction and a finish launch above.
void launch_reduce_finish(int d, const double* partials, int64_t nrec, int64_t pcap, int nv, double* reduced,
                          const FinishArgs& f, hipStream_t s) {
  switch (d) {
#define IS_REDUCE_CASE(v)                                                                          \
  case v:                                                                                          \
    hipLaunchKernelGGL((k_reduce_finish<v, true>), dim3(nv), dim3(kBlock), 0, s, partials, nrec, pcap, nv, \
                       reduced, f);                                                                \
    return;
    IS_EXACT_DIMS(IS_REDUCE_CASE)
#undef IS_REDUCE_CASE
    default:
      break;
  }
  if (d <= 32) {
    hipLaunchKernelGGL((k_reduce_finish<32, false>), dim3(nv), dim3(kBlock), 0, s, partials, nrec, pcap, nv,
                       reduced, f);
  } else {
    hipLaunchKernelGGL((k_reduce_finish<0, false>), dim3(nv), dim3(kBlock), 0, s, partials, nrec, pcap, nv,
                       reduced, f);
    launch_finish(d, reduced, f, s);
  }
}

// Row slices for the batched target: enough blocks to fill the chip when a round
// has few sample tiles; each slice a whole number of 64-row chunks.
int64_t target_splits(int64_t n_rows, int64_t n_new) {
  const int64_t tiles = std::max<int64_t>(1, cdiv(n_new, 64));
  const int64_t want = std::max<int64_t>(1, cdiv(1024, tiles));
  return std::min(want, std::max<int64_t>(1, cdiv(n_rows, 64)));
}

// ---------------------------------------------------------------- job helpers
// log pi(beta) of the job's regression target at one point, FP64 on the host:
// logistic sum_i y_i eta_i - log(1 + e^eta_i), or linear -(y_i - eta_i)^2 / (2
// sigma^2), plus the Gaussian log prior and the normalising constants.
double regression_logtarget(const is_job* j, const double* beta) {
  const int64_t n = j->tn;
  const int p = j->dl;
  const double* X = j->tdata.data();
  const double* y = X + n * p;
  double ll = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double eta = 0.0;
    for (int c = 0; c < p; ++c) eta += X[i * p + c] * beta[c];
    if (j->tkind == IS_TARGET_LOGISTIC) {
      const double sp = eta > 0.0 ? eta + std::log1p(std::exp(-eta)) : std::log1p(std::exp(eta));
      ll += y[i] * eta - sp;
    } else {
      const double r = y[i] - eta;
      ll -= r * r * j->t_inv_two_var;
    }
  }
  double b2 = 0.0;
  for (int c = 0; c < p; ++c) b2 += beta[c] * beta[c];
  return ll - b2 * j->t_inv_two_prior + j->t_const;
}

double host_logaddexp(double a, double b) {
  if (a == -HUGE_VAL) return b;
  if (b == -HUGE_VAL) return a;
  const double m = std::max(a, b);
  return m + std::log1p(std::exp(-std::fabs(a - b)));
}

// Static MIS design over the active slots (configured order of the nonzero
// masses, stably sorted by mass): cumulative masses with the last exactly 1
// and the log interval masses (src/proposal_banks.jl:301-326); for the partial
// mixture, the active members of each group in the listed order, empty groups
// dropped, with log coefficients log m - log sum_group m
// (src/methods/static_mis.jl:245-310).  Uploads the cumulative masses, the
// group tables and allocates the assignment buffer.
int setup_mis(is_job* j, const is_job_desc* desc, const std::vector<double>& masses) {
  if (desc->mis_scheme < IS_MIS_STRATIFIED || desc->mis_scheme > IS_MIS_PARTIAL)
    return fail(IS_EINVAL, "unknown MIS scheme");
  j->mis_scheme = desc->mis_scheme;
  const int K = j->K;
  double tot = 0.0;
  for (double m : masses) tot += m;
  j->mis_cdf.assign(K, 0.0);
  double acc = 0.0;
  for (int s = 0; s < K; ++s) {
    acc += masses[s] / tot;
    j->mis_cdf[s] = acc;
  }
  j->mis_cdf[K - 1] = 1.0;
  j->mis_logmass.assign(K, 0.0);
  double prev = 0.0, mx = -HUGE_VAL;
  for (int s = 0; s < K; ++s) {
    if (!(j->mis_cdf[s] > prev))
      return fail(IS_EINVAL, "every positive proposal mass must retain a positive assignment interval");
    j->mis_logmass[s] = std::log(j->mis_cdf[s] - prev);
    prev = j->mis_cdf[s];
    mx = std::max(mx, j->mis_logmass[s]);
  }
  double se = 0.0;
  for (int s = 0; s < K; ++s) se += std::exp(j->mis_logmass[s] - mx);
  const double lse = mx + std::log(se);
  for (int s = 0; s < K; ++s) j->mis_logmass[s] -= lse;
  if (j->mis_scheme == IS_MIS_PARTIAL) {
    const int G = desc->mis_group_count, Kc = desc->bank_size;
    if (G < 1 || !desc->mis_group_offsets || !desc->mis_group_members)
      return fail(IS_EINVAL, "partial-MIS groups cannot be empty");
    if (desc->mis_group_offsets[0] != 0) return fail(IS_EINVAL, "partial-MIS group offsets must start at 0");
    std::vector<int> slot_of_id(Kc + 1, -1);
    for (int s = 0; s < K; ++s) slot_of_id[j->bank_ids[s]] = s;
    std::vector<char> seen(Kc + 1, 0);
    j->mis_slot_group.assign(K, 0);
    j->mis_goff.assign(1, 0);
    j->mis_gmem.clear();
    j->mis_gcoef.clear();
    for (int g = 0; g < G; ++g) {
      const int e0 = desc->mis_group_offsets[g], e1 = desc->mis_group_offsets[g + 1];
      if (e1 <= e0) return fail(IS_EINVAL, "partial-MIS groups cannot be empty");
      const size_t first = j->mis_gmem.size();
      for (int e = e0; e < e1; ++e) {
        const int id = desc->mis_group_members[e];
        if (id < 1 || id > Kc) return fail(IS_EINVAL, "partial-MIS proposal ID %d is out of range", id);
        if (seen[id]) return fail(IS_EINVAL, "partial-MIS proposal ID %d appears more than once", id);
        seen[id] = 1;
        if (slot_of_id[id] >= 0) j->mis_gmem.push_back(slot_of_id[id]);
      }
      if (j->mis_gmem.size() == first) continue;
      const int ag = (int)j->mis_goff.size() - 1;
      double glm = -HUGE_VAL;
      for (size_t m = first; m < j->mis_gmem.size(); ++m) glm = host_logaddexp(glm, j->mis_logmass[j->mis_gmem[m]]);
      for (size_t m = first; m < j->mis_gmem.size(); ++m) {
        j->mis_slot_group[j->mis_gmem[m]] = ag;
        j->mis_gcoef.push_back(j->mis_logmass[j->mis_gmem[m]] - glm);
      }
      j->mis_goff.push_back((int32_t)j->mis_gmem.size());
    }
    for (int id = 1; id <= Kc; ++id)
      if (!seen[id]) return fail(IS_EINVAL, "partial-MIS groups must include every proposal ID exactly once");
    std::vector<int32_t> ints(j->mis_slot_group);
    ints.insert(ints.end(), j->mis_goff.begin(), j->mis_goff.end());
    ints.insert(ints.end(), j->mis_gmem.begin(), j->mis_gmem.end());
    int rc;
    if ((rc = j->mis_ints_d.alloc(ints.size() * sizeof(int32_t)))) return rc;
    HIP_TRY(locked_memcpy(j->mis_ints_d.p, ints.data(), ints.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  int rc;
  if ((rc = j->mis_cdf_d.alloc((size_t)K * sizeof(double)))) return rc;
  HIP_TRY(locked_memcpy(j->mis_cdf_d.p, j->mis_cdf.data(), (size_t)K * sizeof(double), hipMemcpyHostToDevice));
  if ((rc = j->assign.alloc((size_t)j->schedule[0] * sizeof(int32_t)))) return rc;
  return IS_OK;
}

int64_t isqrt64(int64_t n) {
  int64_t r = (int64_t)std::sqrt((double)n);
  while (r * r > n) --r;
  while ((r + 1) * (r + 1) <= n) ++r;
  return r;
}

uint64_t round_word(const is_job* j, int t) { return ((uint64_t)j->run_index << 32) | (uint64_t)t; }

template <class T>
int upload_T(DevBuf& buf, const std::vector<double>& v) {
  std::vector<T> t(v.begin(), v.end());
  if (buf.alloc(std::max<size_t>(1, t.size()) * sizeof(T))) return IS_ENOMEM;
  if (!t.empty()) HIP_TRY(locked_memcpy(buf.p, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice));
  return IS_OK;
}

template <class T>
RoundArgs<T> base_args(is_job* j, int t) {
  RoundArgs<T> a{};
  a.d = j->d;
  a.round = t;
  a.run_hi = j->run_dev.as<unsigned long long>();
  a.slots = j->slotsT.as<T>();
  a.stride = j->stride;
  a.fam.student = j->family;
  a.fam.dof = (T)j->dof;
  a.fam.half_dof_plus_d = (T)((j->dof + (double)j->d) / 2.0);
  a.dof = j->dof;
  a.key = j->key;
  a.rs = radial_normal_stride(j->family, j->dof);
  a.us = radial_uniform_stride(j->family, j->dof);
  a.tkind = j->tkind;
  a.tslot = j->tslotT.as<T>();
  a.X = j->X.as<T>();
  // a named layout maps the draws in its own pass (k_layout_forward), so the draw kernels
  // see no per-coordinate transforms; both forms carry the log Jacobian to the pre-target pass
  const bool per_coordinate = j->has_tf && !j->has_layout;
  a.tf_kinds = per_coordinate ? j->tf_kinds_d.as<int32_t>() : nullptr;
  a.tf_bounds = per_coordinate ? j->tf_bounds_d.as<double>() : nullptr;
  a.XL = j->has_tf ? j->XL.as<T>() : nullptr;
  a.tf_carry = j->has_tf ? 1 : 0;
  a.cap = j->total;
  a.logtargets = j->logtargets.as<T>();
  a.lognums = j->lognums.as<T>();
  a.logweights = j->logweights.as<T>();
  a.round_ids = j->round_ids.as<int32_t>();
  a.proposal_ids = (j->method == IS_DMPMC || j->method == IS_MIS) ? j->proposal_ids.as<int32_t>() : nullptr;
  a.fail = j->fail_rec.as<unsigned long long>() + 3 * t;
  a.partials = j->partials.as<double>();
  return a;
}

// prefilled buffers for round t
template <class T>
void set_prefilled(is_job* j, int t, RoundArgs<T>& a) {
  const int d = j->d;
  const int rs = a.rs, us = a.us;
  if (j->method == IS_PLAIN) {
    const int stride = d + rs;
    if (j->prefilled[0]) {
      a.pre_normals = j->pre_normals.as<double>();
      a.nstride = stride;
      a.pre_rn = rs ? a.pre_normals + d : nullptr;
      a.rn_stride = stride;
    }
    if (us && j->prefilled_radial[0] && j->pre_ru.p) {
      a.pre_ru = j->pre_ru.as<double>();
      a.ru_stride = us;
    }
    return;
  }
  if (j->prefilled[t]) {
    a.pre_normals = j->pre_normals.as<double>() + (size_t)t * d * j->maxN;
    a.nstride = d;
  }
  if (j->family == IS_STUDENT_T && j->prefilled_radial[t]) {
    // a part that was never uploaded is drawn from the counter RNG
    a.pre_rn = (rs && j->pre_rn.p) ? j->pre_rn.as<double>() + (size_t)t * rs * j->maxN : nullptr;
    a.rn_stride = rs;
    a.pre_ru = (us && j->pre_ru.p) ? j->pre_ru.as<double>() + (size_t)t * us * j->maxN : nullptr;
    a.ru_stride = us;
  }
}

FinishArgs base_finish(is_job* j, int t, unsigned int* counter) {
  FinishArgs f{};
  f.d = j->d;
  f.round = t;
  f.student = j->family;
  f.dof = j->dof;
  f.logconstant = j->logconstant;
  f.round_ess = j->round_ess.as<double>();
  f.round_logz = j->round_logz.as<double>();
  f.adapt_ess = j->adapt_ess.as<double>();
  f.fit_status = j->fit_status.as<int>();
  f.counter = counter;
  f.t_is_float = j->precision == IS_FP32;
  return f;
}


template <class T>
SweepArgs<T> sweep_args(is_job* j, int t) {
  SweepArgs<T> w{};
  w.d = j->d;
  w.stride = j->stride;
  w.fam.student = j->family;
  w.fam.dof = (T)j->dof;
  w.fam.half_dof_plus_d = (T)((j->dof + (double)j->d) / 2.0);
  w.X = j->X.as<T>();
  w.cap = j->total;
  w.logtargets = j->logtargets.as<T>();
  w.lognums = j->lognums.as<T>();
  w.logweights = j->logweights.as<T>();
  w.round_ids = j->round_ids.as<int32_t>();
  w.fail = j->fail_rec.as<unsigned long long>() + 3 * t;
  w.partials = j->partials.as<double>();
  w.pcap = j->pcap;
  return w;
}

// FP32 jobs whose rounds draw their samples before the mixture pass (a batch target,
// or a named layout), at 16 <= d <= 64: that pass's component densities run on the
// matrix cores (k_mix_prep, k_mix_mfma) for rounds of at least kMixMfmaComponents
// components, except on the extended-range replay kernels.
constexpr int kMixMfmaComponents = 8;
inline bool mix_on_mfma(const is_job* j) {
  const bool two_pass = j->tkind == IS_TARGET_LOGISTIC || j->tkind == IS_TARGET_LINREG ||
                        j->tkind == IS_TARGET_CALLBACK || j->has_layout;
  return j->precision == IS_FP32 && two_pass && j->d >= 16 && j->d <= 64 && j->method != IS_NPMC &&
         j->method != IS_PLAIN;
}

// The mixture values of a round's new samples on the matrix cores: each component's
// inverse factor, then the samples' log-sum-exp over the components [a.k0, a.k1].
template <class T>
void launch_mix_mfma(is_job* j, const RoundArgs<T>& a, int ncomp, hipStream_t s) {
  const int d = j->d, kt = cdiv(d, 16);
  MixPrepArgs pa{};
  pa.d = d;
  pa.kp = 16 * kt;
  pa.k0 = a.k0;
  pa.ncomp = ncomp;
  pa.slots64 = j->slots64.as<double>();
  pa.stride = j->stride;
  pa.frag = j->mix_frag.as<float>();
  pa.mu = j->mix_mu.as<float>();
  pa.lognorm = j->mix_lognorm.as<float>();
  MixMfmaArgs ma{};
  ma.d = d;
  ma.kp = 16 * kt;
  ma.ncomp = ncomp;
  ma.k0 = a.k0;
  ma.frag = pa.frag;
  ma.mu = pa.mu;
  ma.lognorm = pa.lognorm;
  ma.logmass = reinterpret_cast<const float*>(a.logmass);
  ma.fam.student = a.fam.student;
  ma.fam.dof = (float)a.fam.dof;
  ma.fam.half_dof_plus_d = (float)a.fam.half_dof_plus_d;
  ma.X = reinterpret_cast<const float*>(a.X);
  ma.cap = a.cap;
  ma.first = a.first;
  ma.n_new = a.n_new;
  ma.assign = a.assign;
  ma.gen_slot = a.gen_slot;
  ma.out = j->mix_out.as<float>();
  const dim3 gp((unsigned)ncomp), gm((unsigned)cdiv(a.n_new, 64));
  switch (kt) {
#define IS_MIX_CASE(v)                                                                     \
  case v:                                                                                  \
    hipLaunchKernelGGL(k_mix_prep<v>, gp, dim3(64), 0, s, pa);                             \
    hipLaunchKernelGGL(k_mix_mfma<v>, gm, dim3(256), 0, s, ma);                            \
    break;
    IS_MIX_CASE(1) IS_MIX_CASE(2) IS_MIX_CASE(3) IS_MIX_CASE(4)
#undef IS_MIX_CASE
    default:
      break;
  }
}

// Samples per sweep block of a fused round over `first` retained samples: the
// fewest whole 64-sample steps per wave (4 waves per block) that keep the sweep
// blocks and the round's nb_new new-sample blocks within one pass of the chip
// (kRoundBlocksPerCu resident blocks per CU), so that small rounds spread over many
// short blocks; the records stay within the partial-record columns.
constexpr int kRoundBlocksPerCu = 4;
int64_t round_sweep_per_block(const is_job* j, int64_t first, int nb_new) {
  const int cus = j->ctx ? j->ctx->cus : 256;
  const int64_t budget = std::max<int64_t>(cus, (int64_t)kRoundBlocksPerCu * cus - nb_new);
  int64_t per_block = std::max<int64_t>(256, cdiv(cdiv(first, budget), 256) * 256);
  while (per_block < first && cdiv(first, per_block) + nb_new > j->pcap) per_block += 256;
  return per_block;
}

// The launch arguments of one round of plain IS (t = 0) or AMIS (round t) as the fused
// form issues them: the new samples' blocks beside the sweep share wo of the retained
// samples [0, first), then the reduction and finish f.  w is the sweep over every
// retained sample (the unfused form).  Stand-alone runs and job groups share it.
template <class T>
struct RoundPlan {
  RoundArgs<T> a;
  SweepArgs<T> w, wo;
  FinishArgs f;
  int nt = 0, ncomp = 1, nv = 2, nb_new = 0, nrec_old = 0;
};

template <class T>
RoundPlan<T> round_plan(is_job* j, int t, unsigned int* counter) {
  RoundPlan<T> p;
  const int d = j->d;
  const int64_t N = j->schedule[t];
  RoundArgs<T>& a = p.a;
  a = base_args<T>(j, t);
  set_prefilled(j, t, a);
  a.n_new = N;
  p.nb_new = (int)cdiv(N, 64);
  if (j->method == IS_PLAIN) {
    a.first = 0;
    a.k0 = a.k1 = 0;
    a.gen_slot = 0;
    a.logmass = j->zerosT.as<T>();
    a.logtotal = T(0);
    a.record_nonfinite_after_exhaust = 0;
    p.w = sweep_args<T>(j, 0);
    p.w.lo = 0;
    p.w.hi = N;
    p.w.nv = 2;
    p.wo = p.w;
    p.wo.hi = 0;
    p.wo.per_block = kRoundPerBlock;
    p.f = base_finish(j, 0, counter);
    p.f.flags = kFinishSummary;
    p.f.summary_count = N;
    p.f.summary_out = j->summary.as<double>();
    return p;
  }
  // AMIS: new samples against the (t + 1)-term temporal mixture, then one sweep over
  // every retained sample: append of the newest component, weights, summary and moments
  const int64_t first = j->offsets[t];
  const int64_t last = j->offsets[t + 1];
  a.first = first;
  a.record_nonfinite_after_exhaust = 1;
  a.k0 = 0;
  a.k1 = t;
  a.gen_slot = t;
  a.logmass = j->logmassT.as<T>();
  a.logtotal = (T)std::log((double)last);
  double* slot64_t = j->slots64.as<double>() + (size_t)t * j->stride;
  SweepArgs<T>& w = p.w;
  w = sweep_args<T>(j, t);
  w.lo = 0;
  w.hi = last;
  w.append = first > 0;
  w.append_hi = first;
  w.slot = j->slotsT.as<T>() + (size_t)t * j->stride;
  w.logmass = (T)std::log((double)N);
  w.logtotal = a.logtotal;
  w.center = slot64_t;
  w.nv = 2 + d + d * (d + 1) / 2;
  p.wo = w;
  p.wo.hi = first;
  p.wo.per_block = round_sweep_per_block(j, first, p.nb_new);
  p.nt = 1;
  p.ncomp = t + 1;
  p.nv = w.nv;
  p.nrec_old = first > 0 ? (int)cdiv(first, p.wo.per_block) : 0;
  FinishArgs& f = p.f;
  f = base_finish(j, t, counter);
  f.flags = kFinishSummary | kFinishFit;
  f.summary_count = last;
  f.center = slot64_t;
  f.prev_slot = slot64_t;
  f.next_slot64 = j->slots64.as<double>() + (size_t)(t + 1) * j->stride;
  f.next_slot_t = j->slotsT.as<T>() + (size_t)(t + 1) * j->stride;
  return p;
}

template <class T>
int run_rounds(is_job* j) {
  hipStream_t s = j->stream;
  const int d = j->d;
  // the extended-range kernels replay a run whose densities left the floating-point
  // range (no fused round among them: the round falls back to new-sample + sweep)
  const int D = j->range_mode ? kRangeDims : j->D;
  const int nv_mom = 2 + d + d * (d + 1) / 2;
  const int nt_mom = sweep_nt(d, true);
  unsigned int* counter = j->counter.as<unsigned int>();
  // a named layout runs every target on the logical samples after the draw pass
  const bool batch_target = j->tkind == IS_TARGET_LOGISTIC || j->tkind == IS_TARGET_LINREG ||
                            j->tkind == IS_TARGET_CALLBACK || j->has_layout;
  const int dl = j->dl;
  // the mixture pass after the target values: on the matrix cores where mix_on_mfma
  // allows, then the new-sample kernel in pre-target mode
  auto pretarget = [&](const RoundArgs<T>& a, int ncomp) {
    RoundArgs<T> b = a;
    if constexpr (sizeof(T) == 4) {
      if (mix_on_mfma(j) && !j->range_mode && b.mis_terms == 0 && ncomp >= kMixMfmaComponents && b.n_new > 0) {
        launch_mix_mfma<T>(j, b, ncomp, s);
        b.pre_mix = j->mix_out.as<T>();
      }
    }
    launch_mix<T, kPreTarget>(D, ncomp, b, s);
  };
  auto mix = [&](RoundArgs<T>& a, int ncomp) -> int {
    if (!batch_target) {
      launch_mix<T, kDrawFused>(D, ncomp, a, s);
      return IS_OK;
    }
    launch_mix<T, kDrawOnly>(D, ncomp, a, s);
    if (j->has_layout && a.n_new > 0)
      hipLaunchKernelGGL((k_layout_forward<T>), dim3((unsigned)cdiv(a.n_new, 256)), dim3(256), 0, s,
                         j->layout_d.as<LayoutBlock>(), (int)j->layout.size(), a.X, a.XL, a.cap, a.first, a.n_new,
                         a.lognums, a.fail);
    if (j->tkind == IS_TARGET_STANDARD || j->tkind == IS_TARGET_GAUSSIAN) {
      // the built-in targets over the layout's logical rows
      if (a.n_new > 0)
        hipLaunchKernelGGL((k_target_logical<T>), dim3((unsigned)cdiv(a.n_new, 256)), dim3(256), 0, s, j->tkind, dl,
                           j->tslotT.as<T>(), a.XL, a.cap, a.first, a.n_new, j->tpre.as<T>());
      a.pre_logtarget = j->tpre.as<T>();
      pretarget(a, ncomp);
      return IS_OK;
    }
    if (j->tkind == IS_TARGET_CALLBACK) {
      // draw, the caller's batch target over the round's samples (values start as
      // NaN, src/batched_targets.jl:96-98), then the mixture with those values; the
      // mixture pass checks them (NaN / +Inf fail in the target phase)
      HIP_TRY(hipMemsetAsync(j->tpre.p, 0xff, (size_t)a.n_new * sizeof(T), s));
      HIP_TRY(hipGetLastError());
      const int cb = j->tbatch(j->tuser, (j->has_tf ? j->XL : j->X).as<T>() + a.first, j->total, dl, a.n_new, j->tpre.p,
                               sizeof(T) == 4 ? 1 : 0, (void*)s);
      if (cb != 0) {
        // the run fails in the target phase of this round and commits nothing
        (void)hipStreamSynchronize(s);
        is_failure fr{};
        fr.round = a.round + 1;
        fr.phase = IS_PHASE_TARGET;
        j->failure = fr;
        j->have_results = false;
        j->last_status = IS_EFAILED;
        return fail(IS_EFAILED, "batch target callback returned %d in round %d", cb, a.round + 1);
      }
      a.pre_logtarget = j->tpre.as<T>();
      pretarget(a, ncomp);
      return IS_OK;
    }
    // draw, batched target on MFMA, then the mixture with the precomputed target values
    const int64_t splits = target_splits(j->tn, a.n_new);
    if (j->t_bf16) {
      // FP32 logistic: the centred BF16 form
      LogisticBf16Args ba{};
      ba.p = dl;
      ba.n = j->tn;
      ba.xfrag = j->t_xfrag.as<u32x4_t>();
      ba.base = j->t_base.as<float>();
      ba.x0 = j->t_x0d.as<double>();
      ba.beta = (j->has_tf ? j->XL : j->X).as<float>();
      ba.cap = j->total;
      ba.first = a.first;
      ba.n_new = a.n_new;
      ba.ntiles = cdiv(j->tn, 16);
      ba.tiles_per_split = cdiv(cdiv(ba.ntiles, splits), 4) * 4;
      ba.part = j->tpart.as<double>();
      const int64_t used = cdiv(ba.ntiles, ba.tiles_per_split);
      const dim3 grid((unsigned)cdiv(a.n_new, 64), (unsigned)used);
      if (j->t_ks2 == 1)
        hipLaunchKernelGGL(k_logistic_bf16<1>, grid, dim3(256), 0, s, ba);
      else
        hipLaunchKernelGGL(k_logistic_bf16<2>, grid, dim3(256), 0, s, ba);
      hipLaunchKernelGGL((k_target_combine<T>), dim3((unsigned)cdiv(a.n_new, 256)), dim3(256), 0, s, ba.part,
                         (int)used, a.n_new, (const T*)ba.beta, ba.cap, ba.first, dl, j->t_inv_two_prior,
                         j->t_const - j->t_shift, j->tpre.as<T>());
      a.pre_logtarget = j->tpre.as<T>();
      pretarget(a, ncomp);
      return IS_OK;
    }
    TargetArgs<T> ta{};
    ta.kind = j->tkind;
    ta.p = dl;
    ta.n = j->tn;
    ta.Xd = j->Xd.as<T>();
    ta.yd = j->yd.as<T>();
    ta.inv_two_var = j->t_inv_two_var;
    ta.beta = (j->has_tf ? j->XL : j->X).as<T>();  // the logical samples under a transform
    ta.cap = j->total;
    ta.first = a.first;
    ta.n_new = a.n_new;
    ta.rows_per_split = cdiv(cdiv(j->tn, splits), 64) * 64;
    ta.part = j->tpart.as<double>();
    const int64_t used = cdiv(j->tn, ta.rows_per_split);
    launch_target<T>(dim_bucket(dl), ta, used, s);
    hipLaunchKernelGGL((k_target_combine<T>), dim3((unsigned)cdiv(a.n_new, 256)), dim3(256), 0, s, ta.part, (int)used,
                       a.n_new, ta.beta, ta.cap, ta.first, dl, j->t_inv_two_prior, j->t_const - j->t_shift,
                       j->tpre.as<T>());
    a.pre_logtarget = j->tpre.as<T>();
    pretarget(a, ncomp);
    return IS_OK;
  };
  auto reduce = [&](const DevBuf& partials, int nrec, int nv, DevBuf& reduced, const FinishArgs& f) {
    launch_reduce_finish(d, partials.as<double>(), (int64_t)nrec, j->pcap, nv, reduced.as<double>(), f, s);
  };

  if (j->method == IS_MIS) {
    // static MIS (src/methods/static_mis.jl:310-340,553-592): slot assignment,
    // draws from the assigned slots, the design's denominator, one summary
    const int64_t N = j->schedule[0];
    hipLaunchKernelGGL(k_mis_assign, dim3((unsigned)cdiv(N, 256)), dim3(256), 0, s, j->key,
                       j->run_dev.as<unsigned long long>(), N, j->mis_scheme == IS_MIS_RANDOM ? 0 : 1,
                       j->mis_cdf_d.as<double>(), j->K,
                       j->prefilled_uniforms[0] ? j->pre_uniforms.as<double>() : nullptr, j->assign.as<int32_t>());
    RoundArgs<T> a = base_args<T>(j, 0);
    set_prefilled(j, 0, a);
    a.first = 0;
    a.n_new = N;
    a.assign = j->assign.as<int32_t>();
    a.bank_ids = j->bank_ids_d.as<int32_t>();
    a.gen_slot = 0;
    a.logmass = j->logmassT.as<T>();
    a.logtotal = T(0);
    a.record_nonfinite_after_exhaust = 0;
    int ncomp = j->K;
    if (j->mis_scheme == IS_MIS_STANDARD) {
      a.mis_terms = 1;
      ncomp = 1;
    } else if (j->mis_scheme == IS_MIS_PARTIAL) {
      const int32_t* ints = j->mis_ints_d.as<int32_t>();
      a.mis_terms = 2;
      a.mis_slot_group = ints;
      a.mis_goff = ints + j->K;
      a.mis_gmem = a.mis_goff + j->mis_goff.size();
      a.mis_gcoef = j->mis_gcoefT.as<T>();
      ncomp = 1;
    }
    a.k0 = 0;
    a.k1 = ncomp - 1;
    SweepArgs<T> w = sweep_args<T>(j, 0);
    w.lo = 0;
    w.hi = N;
    w.nv = 2;
    int nrec = 0;
    bool fused = false;
    if (!batch_target) {
      // one launch: the new-sample blocks write their own summary records
      SweepArgs<T> wo = w;
      wo.hi = 0;
      wo.per_block = kRoundPerBlock;
      const int nb_new = (int)cdiv(N, 64);
      tick(j, 0, 0, s);
      fused = launch_round<T>(D, 0, ncomp, a, wo, nb_new, 0, s);
      tick(j, 0, 1, s);
      nrec = nb_new;
    }
    if (!fused) {
      tick(j, 0, 0, s);
      if (const int mrc = mix(a, ncomp)) return mrc;
      tick(j, 0, 1, s);
      nrec = launch_sweep<T>(D, 0, w, s);
    }
    FinishArgs f = base_finish(j, 0, counter);
    f.flags = kFinishSummary;
    f.summary_count = N;
    f.summary_out = j->summary.as<double>();
    reduce(j->partials, nrec, 2, j->reduced, f);
    HIP_TRY(hipGetLastError());
    return IS_OK;
  }
  if (j->method == IS_PLAIN) {
    RoundPlan<T> p = round_plan<T>(j, 0, counter);
    int nrec = 0;
    bool fused = false;
    if (!batch_target) {
      // one launch: the new-sample blocks write their own summary records
      tick(j, 0, 0, s);
      fused = launch_round<T>(D, 0, 1, p.a, p.wo, p.nb_new, 0, s);
      tick(j, 0, 1, s);
      nrec = p.nb_new;
    }
    if (!fused) {
      tick(j, 0, 0, s);
      if (const int mrc = mix(p.a, 1)) return mrc;
      tick(j, 0, 1, s);
      nrec = launch_sweep<T>(D, 0, p.w, s);
    }
    reduce(j->partials, nrec, 2, j->reduced, p.f);
    HIP_TRY(hipGetLastError());
    return IS_OK;
  }

  for (int t = 0; t < j->T; ++t) {
    const int64_t N = j->schedule[t];
    const int64_t first = j->offsets[t];
    const int64_t last = j->offsets[t + 1];
    RoundArgs<T> a = base_args<T>(j, t);
    set_prefilled(j, t, a);
    a.first = first;
    a.n_new = N;
    a.record_nonfinite_after_exhaust = 1;
    double* slot64_t = j->slots64.as<double>() + (size_t)t * j->stride;
    if (j->method == IS_AMIS) {
      RoundPlan<T> p = round_plan<T>(j, t, counter);
      int nrec = 0;
      bool fused = false;
      if (!batch_target && !p.w.clip) {
        // one launch: the new samples' blocks (with their own moment records) beside
        // the sweep of the retained samples [0, first)
        tick(j, t, 0, s);
        fused = launch_round<T>(D, 1, t + 1, p.a, p.wo, p.nb_new, p.nrec_old, s);
        tick(j, t, 1, s);
        nrec = p.nrec_old + p.nb_new;
      }
      if (!fused) {
        if (const int mrc = mix(p.a, t + 1)) return mrc;
        tick(j, t, 0, s);
        nrec = launch_sweep<T>(D, nt_mom, p.w, s);
        tick(j, t, 1, s);
      }
      reduce(j->partials, nrec, nv_mom, j->reduced, p.f);
    } else if (j->method == IS_NPMC) {
      a.k0 = a.k1 = t;
      a.gen_slot = t;
      a.logmass = j->zerosT.as<T>();
      a.logtotal = T(0);
      tick(j, t, 0, s);
      if (const int mrc = mix(a, 1)) return mrc;
      tick(j, t, 1, s);
      SweepArgs<T> w = sweep_args<T>(j, t);
      w.lo = first;
      w.hi = last;
      w.nv = 2;
      const int nrec = launch_sweep<T>(D, 0, w, s);
      FinishArgs f = base_finish(j, t, counter);
      f.flags = kFinishSummary;
      f.summary_count = N;
      reduce(j->partials, nrec, 2, j->reduced, f);
      // clip at the isqrt(n)-th largest raw weight (src/npmc_execution.jl:21-33)
      const int64_t kth = N - isqrt64(N) + 1;
      {
        const int cus = j->ctx ? j->ctx->cus : 256;
        const int sb = (int)std::max<int64_t>(1, std::min<int64_t>(4 * cus, cdiv(N, 4 * kSelectBlock)));
        SelectState* st = reinterpret_cast<SelectState*>(j->clip.as<char>() + 64);
        for (int pass = 0; pass < 8; ++pass)
          hipLaunchKernelGGL((k_select_pass<T>), dim3(sb), dim3(kSelectBlock), 0, s, j->logweights.as<T>(), first, N,
                             pass, kth, st, j->clip.as<double>());
      }
      SweepArgs<T> w2 = sweep_args<T>(j, t);
      w2.lo = first;
      w2.hi = last;
      w2.clip = j->clip.as<double>();
      w2.center = slot64_t;
      w2.nv = nv_mom;
      w2.partials = j->partials2.as<double>();
      const int nrec2 = launch_sweep<T>(D, nt_mom, w2, s);
      FinishArgs g = base_finish(j, t, counter);
      g.flags = kFinishFit | kFinishAdaptEss;
      g.summary_count = N;
      g.center = slot64_t;
      g.prev_slot = slot64_t;
      g.next_slot64 = j->slots64.as<double>() + (size_t)(t + 1) * j->stride;
      g.next_slot_t = j->slotsT.as<T>() + (size_t)(t + 1) * j->stride;
      reduce(j->partials2, nrec2, nv_mom, j->reduced2, g);
    } else {  // DM-PMC
      a.k0 = 0;
      a.k1 = j->K - 1;
      a.gen_slot = 0;
      a.assign = j->assign.as<int32_t>() + (size_t)t * j->maxN;
      a.bank_ids = j->bank_ids_d.as<int32_t>();
      a.logmass = j->logmassT.as<T>() + (size_t)t * j->K;
      a.logtotal = T(0);
      tick(j, t, 0, s);
      if (const int mrc = mix(a, j->K)) return mrc;
      tick(j, t, 1, s);
      SweepArgs<T> w = sweep_args<T>(j, t);
      w.lo = first;
      w.hi = last;
      w.nv = 2;
      const int nrec = launch_sweep<T>(D, 0, w, s);
      double* summ = j->summary.as<double>() + (size_t)t * 5;
      FinishArgs f = base_finish(j, t, counter);
      f.flags = kFinishSummary;
      f.summary_count = N;
      f.summary_out = summ;
      reduce(j->partials, nrec, 2, j->reduced, f);
      // resampling: global (src/dm_pmc_execution.jl:267-295) from the round's CDF, or
      // local (:297-374), each proposal from its own group
      const T* lw = j->logweights.as<T>() + first;
      int32_t* anc = j->ancestors.as<int32_t>() + (size_t)t * j->K;
      if (j->prefilled_ancestors[t]) {
        // replay seam: this round's ancestors are given (e.g. another run's trajectory)
        hipLaunchKernelGGL((k_copy<int32_t>), dim3(cdiv(j->K, 64)), dim3(64), 0, s,
                           (const int32_t*)(j->pre_ancestors.as<int32_t>() + (size_t)t * j->K), (int64_t)j->K, anc);
      } else {
        if (j->resampling != IS_LOCAL)
          launch_cdf<T>(lw, N, summ, j->tile_total.as<double>(), j->cdf.as<double>(), j->cdf_status.as<int>() + t, s);
        double* u = j->uniforms.as<double>() + (size_t)t * j->K;
        if (j->prefilled_uniforms[t]) {
          u = j->pre_uniforms.as<double>() + (size_t)t * j->K;
        } else {
          hipLaunchKernelGGL(k_resampling_uniforms, dim3(cdiv(j->K, 64)), dim3(64), 0, s, j->key,
                             j->run_dev.as<unsigned long long>(), t, (int64_t)j->K, u);
        }
        if (j->resampling == IS_LOCAL)
          hipLaunchKernelGGL((k_local_select<T>), dim3(j->K), dim3(kLocalLanes), 0, s, lw,
                             (const int32_t*)(j->plan_group.as<int32_t>() + (size_t)t * (j->K + 1)), (const double*)u,
                             anc, j->cdf_status.as<int>() + t);
        else
          hipLaunchKernelGGL((k_select_ancestors<T>), dim3(cdiv(j->K, 64)), dim3(64), 0, s, j->cdf.as<double>(),
                             lw, N, u, (int64_t)j->K, j->resampling == IS_SYSTEMATIC ? 1 : 0, anc);
      }
      hipLaunchKernelGGL((k_gather_locations<T>), dim3(cdiv((int64_t)j->K * d, 64)), dim3(64), 0, s,
                         j->X.as<T>(), j->total, first, d, anc, j->K, j->slots64.as<double>(), j->slotsT.as<T>(),
                         j->stride);
    }
    HIP_TRY(hipGetLastError());
  }
  return IS_OK;
}

int phase_from_bits(uint64_t w1, uint64_t w2) {
  if (w2 != 0) return IS_PHASE_SAMPLING;
  const uint16_t bits = (uint16_t)(w1 & 0xffff);
  if (bits & kDrawReasons) return IS_PHASE_SAMPLING;
  if (bits & (kTargetNaN | kTargetPosInf)) return IS_PHASE_TARGET;
  if (bits & kLogweightInvalid) return IS_PHASE_WEIGHT;
  return IS_PHASE_DENOMINATOR;
}

void decode(uint64_t packed, int64_t& index, int32_t& block, uint16_t& bits) {
  const uint32_t inv_index = (uint32_t)(packed >> 32);
  const uint16_t inv_block = (uint16_t)((packed >> 16) & 0xffff);
  index = (int64_t)(uint32_t)(0xffffffffu - inv_index + 1u);
  block = (int32_t)(uint16_t)(0xffffu - inv_block + 1u);
  bits = (uint16_t)(packed & 0xffff);
}

template <class T>
int run_epilogue(is_job* j);

// Slot-table entries of the committed state: the bank's K slots (DM-PMC, static MIS) or slot 0.
size_t committed_slots(const is_job* j) {
  const bool bank = j->method == IS_DMPMC || j->method == IS_MIS;
  return (size_t)(bank ? j->K : 1) * j->stride;
}

// The committed state as the run's first slots, FP64 (h64) and in T (hT, the inverse
// diagonals that are not normal in T marked for the extended-range density).
template <class T>
void pack_committed(is_job* j, double* h64, T* hT) {
  const int d = j->d;
  const size_t n = committed_slots(j);
  if (j->method == IS_DMPMC || j->method == IS_MIS) {
    std::copy(j->bank_slots_init.begin(), j->bank_slots_init.begin() + n, h64);
    for (int k = 0; k < j->K; ++k)
      for (int c = 0; c < d; ++c) h64[(size_t)k * j->stride + c] = j->committed_locs[(size_t)k * d + c];
  } else {
    pack_slot(j->committed, d, 0.0, h64);
  }
  for (size_t i = 0; i < n; ++i) hT[i] = (T)h64[i];
  for (size_t k = 0; k < n / (size_t)j->stride; ++k)
    for (int r = 0; r < d; ++r) {
      const size_t at = k * j->stride + d + r;
      hT[at] = slot_inv_value<T>(h64[at]);
    }
}

template <class T>
int job_run_typed(is_job* j) {
  hipStream_t s = j->stream;
  const int T_ = j->T;
  // run word and zeroed status from the pinned image (one copy, no host wait)
  j->ctl_up.as<uint64_t>()[0] = (uint64_t)j->run_index << 32;
  HIP_TRY(hipMemcpyAsync(j->ctl.p, j->ctl_up.p, j->ctl_up_bytes, hipMemcpyHostToDevice, s));
  // committed state into slot 0 (AMIS / NPMC) or the bank slots (DM-PMC, static MIS);
  // plain IS and static MIS never change it, so it is uploaded once
  const int d = j->d;
  const bool fixed_state = j->method == IS_PLAIN || j->method == IS_MIS;
  if (!(fixed_state && j->slots_resident)) {
    double* h64 = j->slot_up.as<double>();
    const size_t n = committed_slots(j);
    T* hT = reinterpret_cast<T*>(h64 + n);
    pack_committed<T>(j, h64, hT);
    HIP_TRY(hipMemcpyAsync(j->slots64.p, h64, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(j->slotsT.p, hT, n * sizeof(T), hipMemcpyHostToDevice, s));
    j->slots_resident = fixed_state;
  }
  int rc = IS_OK;
  // per-round event timing records events between launches, so a timed run is issued
  // eagerly; otherwise the whole launch sequence is replayed from one captured graph
  if (j->use_graph && !j->timing && j->tkind != IS_TARGET_CALLBACK) {
    if (j->graph_dirty || !j->graph_exec) {
      CaptureGuard cg(capture_lock());
      if (j->graph_exec) (void)hipGraphExecDestroy(j->graph_exec);
      if (j->graph) (void)hipGraphDestroy(j->graph);
      j->graph_exec = nullptr;
      j->graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      rc = run_rounds<T>(j);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(IS_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
      j->graph = g;
      HIP_TRY(hipGraphInstantiate(&j->graph_exec, g, nullptr, nullptr, 0));
      j->graph_dirty = false;
    }
    HIP_TRY(hipGraphLaunch(j->graph_exec, s));
  } else {
    rc = run_rounds<T>(j);
    if (rc) return rc;
  }
  // status and round summaries (one transfer), plus the state a successful run commits
  HIP_TRY(hipMemcpyAsync(j->ctl_down.as<char>() + j->ctl_down_off, j->ctl.as<char>() + j->ctl_down_off,
                         j->ctl.bytes - j->ctl_down_off, hipMemcpyDeviceToHost, s));
  if (j->method == IS_AMIS || j->method == IS_NPMC)
    HIP_TRY(hipMemcpyAsync(j->slot_down.p, j->slots64.as<double>() + (size_t)T_ * j->stride,
                           (size_t)j->stride * sizeof(double), hipMemcpyDeviceToHost, s));
  else if (j->method == IS_DMPMC)
    HIP_TRY(hipMemcpyAsync(j->slot_down.p, j->slots64.p, (size_t)j->K * j->stride * sizeof(double),
                           hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return run_epilogue<T>(j);
}

// A run's outcome from its status image in j->ctl_down (and, for AMIS / NPMC / DM-PMC, the
// fitted state in j->slot_down): the failure snapshot, the extended-range replay of a run
// whose densities left the floating-point range, and the transactional commit.
template <class T>
int run_epilogue(is_job* j) {
  const int T_ = j->T;
  const int d = j->d;
  const char* down = j->ctl_down.as<const char>();
  const uint64_t* rec = reinterpret_cast<const uint64_t*>(down + ((const char*)j->fail_rec.p - j->ctl.as<char>()));
  const int* fit = reinterpret_cast<const int*>(down + ((const char*)j->fit_status.p - j->ctl.as<char>()));
  const int* cdfs = reinterpret_cast<const int*>(down + ((const char*)j->cdf_status.p - j->ctl.as<char>()));
  j->run_index += 1;
  if (j->timing) {
    j->round_ms.assign(T_, 0.0);
    for (int t = 0; t < T_; ++t) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, j->events[2 * t], j->events[2 * t + 1]) == hipSuccess) j->round_ms[t] = ms;
    }
  }
  is_failure f{};
  for (int t = 0; t < T_; ++t) {
    const uint64_t* w = rec + 3 * t;
    if (w[0] != 0) {
      f.round = t + 1;
      f.count = w[0];
      f.words[0] = w[0];
      f.words[1] = w[1];
      f.words[2] = w[2];
      decode(w[1], f.first_logical_index, f.first_block, f.reason_bits);
      if (w[2]) decode(w[2], f.draw_first_logical_index, f.draw_first_block, f.draw_reason_bits);
      f.phase = phase_from_bits(w[1], w[2]);
      break;
    }
    if (fit[t] != 0) {
      f.round = t + 1;
      f.phase = fit[t] == 1 ? IS_PHASE_MOMENT : IS_PHASE_FACTORIZATION;
      f.reason_bits = kMomentFitInvalid;
      break;
    }
    if (cdfs[t] != 0) {
      f.round = t + 1;
      f.phase = IS_PHASE_RESAMPLING;
      break;
    }
  }
  j->failure = f;
  if (f.round != 0 && !j->range_mode && d <= kMaxDim &&
      (f.phase == IS_PHASE_DENOMINATOR || f.phase == IS_PHASE_WEIGHT || f.phase == IS_PHASE_TARGET)) {
    // a density outside the ordinary floating-point range gives NaN in the regular
    // kernels: replay the run (same counters, same committed state) on the
    // extended-range kernels; the replay's outcome is the run's
    j->range_mode = true;
    j->graph_dirty = true;
    j->run_index -= 1;
    const int rc2 = job_run_typed<T>(j);
    if (rc2 != IS_OK) {
      j->range_mode = false;  // a failure of its own: later runs start on the regular kernels
      j->graph_dirty = true;
    }
    return rc2;
  }
  if (f.round != 0) {
    j->have_results = false;
    j->last_status = IS_EFAILED;
    return fail(IS_EFAILED, "sampler failed in round %d (phase %d, first index %lld)", f.round, f.phase,
                (long long)f.first_logical_index);
  }
  // transactional commit (src/adaptive_gaussian_execution.jl:1101-1111)
  if (j->method == IS_AMIS || j->method == IS_NPMC) {
    unpack_slot(j->slot_down.as<const double>(), d, j->committed);
  } else if (j->method == IS_DMPMC) {
    const double* slots = j->slot_down.as<const double>();
    for (int k = 0; k < j->K; ++k)
      for (int c = 0; c < d; ++c) j->committed_locs[(size_t)k * d + c] = slots[(size_t)k * j->stride + c];
  }
  j->have_results = true;
  j->last_status = IS_OK;
  return IS_OK;
}

// Values per moment record of d coordinates (weights, means, packed second moments); the
// weights alone past kMaxDim, where moments are refused.
inline int moment_values(int d) { return d <= kMaxDim ? 2 + d + d * (d + 1) / 2 : 2; }

template <class T>
int alloc_job_buffers(is_job* j) {
  const size_t cap = (size_t)j->total;
  const int d = j->d;
  int rc;
  if ((rc = j->X.alloc(cap * d * sizeof(T)))) return rc;
  if (j->has_tf && (rc = j->XL.alloc(cap * j->dl * sizeof(T)))) return rc;
  if ((rc = j->logtargets.alloc(cap * sizeof(T)))) return rc;
  if ((rc = j->lognums.alloc(cap * sizeof(T)))) return rc;
  if ((rc = j->logweights.alloc(cap * sizeof(T)))) return rc;
  if ((rc = j->round_ids.alloc(cap * sizeof(int32_t)))) return rc;
  if ((rc = j->proposal_ids.alloc(cap * sizeof(int32_t)))) return rc;
  HIP_TRY(locked_memset(j->round_ids.p, 0, cap * sizeof(int32_t)));
  if ((rc = j->slots64.alloc((size_t)j->nslots * j->stride * sizeof(double)))) return rc;
  if ((rc = j->slotsT.alloc((size_t)j->nslots * j->stride * sizeof(T)))) return rc;
  HIP_TRY(locked_memset(j->slots64.p, 0, j->slots64.bytes));
  HIP_TRY(locked_memset(j->slotsT.p, 0, j->slotsT.bytes));
  // partial records (column-major): one per sweep block over at most `cap` samples
  // (a fused AMIS round adds one record per 64 new samples)
  j->pcap = cdiv((int64_t)cap, kSweepPerBlock) + cdiv(std::max<int64_t>(1, j->maxN), 64) + 1;
  const int nv_mom = 2 + d + d * (d + 1) / 2;
  const int nv_res = std::max(nv_mom, moment_values(j->dl));  // results' moments over the logical samples
  if ((rc = j->partials.alloc((size_t)j->pcap * (1 + nv_mom) * sizeof(double)))) return rc;
  if ((rc = j->partials2.alloc((size_t)j->pcap * (1 + nv_res) * sizeof(double)))) return rc;
  if ((rc = j->reduced.alloc((size_t)(1 + nv_mom) * sizeof(double)))) return rc;
  if ((rc = j->reduced2.alloc((size_t)(1 + nv_res) * sizeof(double)))) return rc;
  if ((rc = j->summary.alloc((size_t)5 * std::max(1, j->T) * sizeof(double)))) return rc;
  // the clip value, then the radix select's state (zero histogram and ticket)
  if ((rc = j->clip.alloc(64 + sizeof(SelectState)))) return rc;
  HIP_TRY(locked_memset(j->clip.p, 0, 64 + sizeof(SelectState)));
  {
    // control block layout (bytes): run word 64 | counters 64 | failure records 24 T
    // | fit status 4 T | resampling status 4 T (padded to 8) | ESS 8 T | log Z 8 T | adaptation ESS 8 T
    const size_t Tn = (size_t)j->T;
    const size_t o_fail = 128, o_fit = o_fail + 24 * Tn, o_cdf = o_fit + 4 * Tn;
    const size_t o_ess = (o_cdf + 4 * Tn + 7) & ~(size_t)7, o_logz = o_ess + 8 * Tn, o_aess = o_logz + 8 * Tn;
    const size_t total = o_aess + 8 * Tn;
    if ((rc = j->ctl.alloc(total))) return rc;
    HIP_TRY(locked_memset(j->ctl.p, 0, total));
    char* base = j->ctl.as<char>();
    auto view = [&](DevView& v, size_t off, size_t n) {
      v.p = base + off;
      v.bytes = n;
    };
    view(j->run_dev, 0, 64);
    view(j->counter, 64, 64);
    view(j->fail_rec, o_fail, 24 * Tn);
    view(j->fit_status, o_fit, 4 * Tn);
    view(j->cdf_status, o_cdf, 4 * Tn);
    view(j->round_ess, o_ess, 8 * Tn);
    view(j->round_logz, o_logz, 8 * Tn);
    view(j->adapt_ess, o_aess, 8 * Tn);
    j->ctl_up_bytes = o_ess;
    j->ctl_down_off = o_fail;
    if ((rc = j->ctl_up.alloc(total)) || (rc = j->ctl_down.alloc(total))) return rc;
    const bool bank = j->method == IS_DMPMC || j->method == IS_MIS;
    const size_t nup = (size_t)(bank ? j->K : 1) * j->stride;
    if ((rc = j->slot_up.alloc(nup * (sizeof(double) + sizeof(T))))) return rc;
    const size_t ndown = j->method == IS_DMPMC ? (size_t)j->K * j->stride : (size_t)j->stride;
    if ((rc = j->slot_down.alloc(ndown * sizeof(double)))) return rc;
    j->slots_resident = false;
  }
  std::vector<double> zeros(std::max(j->nslots, 1), 0.0);
  if ((rc = upload_T<T>(j->zerosT, zeros))) return rc;
  if ((j->tkind == IS_TARGET_CALLBACK || (j->has_layout && (j->tkind == IS_TARGET_STANDARD || j->tkind == IS_TARGET_GAUSSIAN))) &&
      (rc = j->tpre.alloc((size_t)j->maxN * sizeof(T))))
    return rc;
  if (mix_on_mfma(j)) {
    const int kt = cdiv(d, 16);
    const size_t ncomp = (size_t)std::max(j->nslots, 1);
    if ((rc = j->mix_frag.alloc(ncomp * mix_pairs(kt) * 64 * sizeof(float))) ||
        (rc = j->mix_mu.alloc(ncomp * 16 * kt * sizeof(float))) || (rc = j->mix_lognorm.alloc(ncomp * sizeof(float))) ||
        (rc = j->mix_out.alloc((size_t)j->maxN * sizeof(float))))
      return rc;
  }
  if (j->tkind == IS_TARGET_LOGISTIC || j->tkind == IS_TARGET_LINREG) {
    if ((rc = j->tpre.alloc((size_t)j->maxN * sizeof(T)))) return rc;
    int64_t part = 1;
    for (int64_t N : j->schedule) part = std::max(part, target_splits(j->tn, N) * N);
    if ((rc = j->tpart.alloc((size_t)part * sizeof(double)))) return rc;
  }
  if (j->method == IS_DMPMC) {
    if ((rc = j->cdf.alloc((size_t)j->maxN * sizeof(double)))) return rc;
    // tile totals, then the tile carries and the grand total
    if ((rc = j->tile_total.alloc((size_t)(2 * cdf_tiles(j->maxN) + 2) * sizeof(double))))
      return rc;
    if ((rc = j->uniforms.alloc((size_t)j->K * j->T * sizeof(double)))) return rc;
    if ((rc = j->ancestors.alloc((size_t)j->K * j->T * sizeof(int32_t)))) return rc;
  }
  return IS_OK;
}

}  // namespace

// ---------------------------------------------------------------- centred BF16 logistic data
// The data of the centred BF16 logistic form (k_logistic_bf16), from the FP64 rows in
// j->tdata: each row split as x = xh + xl (BF16 each, xl the rounding of x - xh) in the
// MFMA A-operand order, x_i . x0 in FP64 rounded to FP32, the responses, the centre.
int upload_logistic_bf16(is_job* j, int p) {
  const int64_t n = j->tn;
  const int ks2 = p <= 32 ? 1 : 2;
  const int64_t ntiles = cdiv(n, 16);
  const double* X = j->tdata.data();
  const double* y = X + (size_t)n * p;
  // rows with y = 1 are stored negated (s = 1 - 2 y), so each row's term is
  // -max(s eta, 0) - log1p(e^-|eta|) without the response; other responses keep
  // the FP32 GEMM form
  for (int64_t i = 0; i < n; ++i)
    if (!(y[i] == 0.0 || y[i] == 1.0)) return IS_OK;
  std::vector<uint16_t> frag((size_t)ntiles * ks2 * 2 * 64 * 8, 0);
  for (int64_t t = 0; t < ntiles; ++t)
    for (int ks = 0; ks < ks2; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int64_t row = 16 * t + (l & 15);
        for (int e = 0; e < 8; ++e) {
          const int c = 32 * ks + 8 * (l >> 4) + e;
          if (row >= n || c >= p) continue;
          const double x = y[row] == 1.0 ? -X[(size_t)row * p + c] : X[(size_t)row * p + c];
          const uint16_t h = bf16_bits((float)x);
          const uint16_t lo = bf16_bits((float)(x - bf16_value(h)));
          frag[((((size_t)t * ks2 + ks) * 2 + 0) * 64 + l) * 8 + e] = h;
          frag[((((size_t)t * ks2 + ks) * 2 + 1) * 64 + l) * 8 + e] = lo;
        }
      }
  std::vector<float> base((size_t)ntiles * 16, 0.0f);
  for (int64_t i = 0; i < n; ++i) {
    double v = 0.0;
    for (int c = 0; c < p; ++c) v += X[(size_t)i * p + c] * j->t_x0[c];
    base[i] = (float)(y[i] == 1.0 ? -v : v);
  }
  int rc;
  if ((rc = j->t_xfrag.alloc(frag.size() * sizeof(uint16_t))) || (rc = j->t_base.alloc(base.size() * sizeof(float))) ||
      (rc = j->t_x0d.alloc((size_t)p * sizeof(double))))
    return rc;
  HIP_TRY(locked_memcpy(j->t_xfrag.p, frag.data(), frag.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIP_TRY(locked_memcpy(j->t_base.p, base.data(), base.size() * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(locked_memcpy(j->t_x0d.p, j->t_x0.data(), (size_t)p * sizeof(double), hipMemcpyHostToDevice));
  j->t_bf16 = true;
  j->t_ks2 = ks2;
  return IS_OK;
}

// ---------------------------------------------------------------- allocation-free scratch
// Device memory of is_functional_moments_device_noalloc: the column-major partial
// records (4 columns of cdiv(n, kFuncPerBlock) doubles) followed by the two reduced
// records (8 doubles).  Each GPU holds its own copy until the process ends.  In host
// code `scratch` names a host-side stand-in, never the device memory: the per-device
// address comes from hipGetSymbolAddress at is_ctx_create and is the only one used.
constexpr int64_t kScratchDoubles = 131072;  // 1 MiB
constexpr int64_t kNoAllocMaxSamples = (kScratchDoubles - 8) / 4 * kFuncPerBlock;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// With `device` current: looks up this device's copy of `scratch` once.
int bind_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return IS_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr)
    return fail(IS_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<double*>(ptr);
  return IS_OK;
}

double* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

// ===================================================================== C ABI
extern "C" {

const char* is_last_error(void) { return g_last_error.c_str(); }

int is_ctx_create(int device, is_ctx** out) {
  if (!out) return fail(IS_EINVAL, "null output pointer");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(IS_EDEVICE, "no HIP device is available");
  if (device < 0 || device >= count) return fail(IS_EINVAL, "device %d out of range", device);
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(IS_EUNSUPPORTED, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
  if (device >= kMaxDevices) return fail(IS_EUNSUPPORTED, "device %d: at most %d devices", device, kMaxDevices);
  if (int rc = bind_scratch(device)) return rc;
  is_ctx* c = new is_ctx();
  c->device = device;
  c->cus = std::max(1, prop.multiProcessorCount);
  *out = c;
  return IS_OK;
}

int is_ctx_destroy(is_ctx* ctx) {
  delete ctx;
  return IS_OK;
}

// _prepare_flat_transform_layout (src/proposal_composition.jl:179-240): complete, disjoint
// selectors; scalar transforms on Int selectors, simplices on ranges of K - 1 >= 1
// coordinates (src/proposal_composition.jl:140-160); the logical dimension dl.
int prepare_layout(is_job* j, const is_job_desc* desc) {
  const int d = desc->dim, nb = desc->transform_block_count;
  if (nb < 0 || !desc->transform_blocks) return fail(IS_EINVAL, "a transform layout needs its blocks");
  if (desc->transform_kinds) return fail(IS_EINVAL, "give per-coordinate transforms or a layout, not both");
  std::vector<char> selected(d, 0);
  int dl = 0;
  for (int b = 0; b < nb; ++b) {
    const is_transform_block& B = desc->transform_blocks[b];
    if (B.count < 1 || B.first < 0 || B.first > d - B.count)
      return fail(IS_EINVAL, "transform block %d: selector is out of bounds for dimension %d", b + 1, d);
    if (B.scalar != 0 && B.scalar != 1) return fail(IS_EINVAL, "transform block %d: scalar must be 0 or 1", b + 1);
    if (B.scalar && B.count != 1) return fail(IS_EINVAL, "transform block %d: an Int selector holds one coordinate", b + 1);
    if (B.kind < IS_TRANSFORM_IDENTITY || B.kind > IS_TRANSFORM_SIMPLEX)
      return fail(IS_EINVAL, "transform block %d: unknown transform kind %d", b + 1, B.kind);
    if (B.kind != IS_TRANSFORM_IDENTITY && B.kind != IS_TRANSFORM_SIMPLEX && !B.scalar)
      return fail(IS_EINVAL, "transform block %d: a scalar transform requires one unconstrained scalar", b + 1);
    if (B.kind == IS_TRANSFORM_SIMPLEX && B.scalar)
      return fail(IS_EINVAL, "transform block %d: SimplexTransform(%d) requires %d coordinates", b + 1, B.count + 1,
                  B.count);
    if (B.kind == IS_TRANSFORM_INTERVAL &&
        (B.lower != B.lower || B.upper != B.upper || !(B.lower < B.upper) || (std::isinf(B.lower) && std::isinf(B.upper))))
      return fail(IS_EINVAL, "transform block %d: interval needs lower < upper with a finite side", b + 1);
    for (int c = B.first; c < B.first + B.count; ++c) {
      if (selected[c]) return fail(IS_EINVAL, "flat transform selectors must be disjoint");
      selected[c] = 1;
    }
    dl += B.count + (B.kind == IS_TRANSFORM_SIMPLEX ? 1 : 0);
  }
  for (int c = 0; c < d; ++c)
    if (!selected[c]) return fail(IS_EINVAL, "flat transform selectors must collectively cover the base dimension");
  j->layout.assign(desc->transform_blocks, desc->transform_blocks + nb);
  j->dl = dl;
  j->has_tf = true;
  j->has_layout = true;
  return IS_OK;
}

extern "C++" {
// The device blocks of a layout, with the simplex constants of src/transforms.jl:341-345
// computed in the job's floating-point type T.
template <class T>
std::vector<LayoutBlock> layout_blocks(const std::vector<is_transform_block>& layout) {
  std::vector<LayoutBlock> out(layout.size());
  int row = 0;
  for (size_t b = 0; b < layout.size(); ++b) {
    const is_transform_block& B = layout[b];
    LayoutBlock& L = out[b];
    L.first = B.first;
    L.count = B.count;
    L.kind = B.kind;
    L.out = row;
    L.lower = B.lower;
    L.upper = B.upper;
    if (B.kind == IS_TRANSFORM_SIMPLEX) {
      const T K = (T)(B.count + 1);
      const T inv_root = T(1) / std::sqrt(K);
      L.inv_root = (double)inv_root;
      L.shared = (double)((T(1) + inv_root) / (T)B.count);
      L.half_log_k = (double)(T(0.5) * std::log(K));
    }
    row += B.count + (B.kind == IS_TRANSFORM_SIMPLEX ? 1 : 0);
  }
  return out;
}
}  // extern "C++"

// Logical image of one coordinate vector under the layout, in FP64 (the host's anchor
// for the FP32 regression offset); false when a block fails.
bool layout_forward_host(const std::vector<is_transform_block>& layout, const double* z, double* x) {
  int row = 0;
  for (const is_transform_block& B : layout) {
    const double* zb = z + B.first;
    double* xb = x + row;
    if (B.kind == IS_TRANSFORM_SIMPLEX) {
      const int K = B.count + 1;
      double sum = 0.0;
      for (int i = 0; i < B.count; ++i) sum += zb[i];
      const double inv_root = 1.0 / std::sqrt((double)K), shared = (1.0 + inv_root) / (double)B.count;
      double mx = inv_root * sum;
      xb[B.count] = mx;
      for (int i = 0; i < B.count; ++i) {
        xb[i] = zb[i] - shared * sum;
        mx = std::max(mx, xb[i]);
      }
      double es = 0.0;
      for (int i = 0; i < K; ++i) es += (xb[i] = std::exp(xb[i] - mx));
      for (int i = 0; i < K; ++i) xb[i] /= es;
      row += K;
    } else {
      for (int i = 0; i < B.count; ++i) {
        const double v = zb[i];
        double o = v;
        if (B.kind == IS_TRANSFORM_POSITIVE) {
          o = std::exp(v);
        } else if (B.kind == IS_TRANSFORM_SOFTPLUS) {
          o = v > 0.0 ? v + std::log1p(std::exp(-v)) : std::log1p(std::exp(v));
        } else if (B.kind == IS_TRANSFORM_INTERVAL) {
          if (std::isfinite(B.lower) && std::isfinite(B.upper))
            o = B.lower + (B.upper - B.lower) / (1.0 + std::exp(-v));
          else
            o = std::isfinite(B.lower) ? B.lower + std::exp(v) : B.upper - std::exp(v);
        }
        xb[i] = o;
      }
      row += B.count;
    }
  }
  for (int c = 0; c < row; ++c)
    if (!std::isfinite(x[c])) return false;
  return true;
}

int is_job_create(is_ctx* ctx, const is_job_desc* desc, is_job** out) {
  if (!ctx || !desc || !out) return fail(IS_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(ctx->device));
  const int d = desc->dim;
  if (d < 1 || d > kMaxDim) return fail(IS_EUNSUPPORTED, "dimension %d outside [1, %d]", d, kMaxDim);
  if (desc->scalar && d != 1) return fail(IS_EINVAL, "scalar proposals have dimension 1");
  if (desc->method < IS_PLAIN || desc->method > IS_MIS) return fail(IS_EINVAL, "unknown method");
  if (desc->family != IS_GAUSSIAN && desc->family != IS_STUDENT_T) return fail(IS_EINVAL, "unknown family");
  if (desc->precision != IS_FP64 && desc->precision != IS_FP32) return fail(IS_EINVAL, "unknown precision");
  if (desc->target_kind == IS_TARGET_CALLBACK && !desc->target_batch)
    return fail(IS_EINVAL, "a callback target needs target_batch");
  if (desc->target_kind < IS_TARGET_STANDARD || desc->target_kind > IS_TARGET_CALLBACK)
    return fail(IS_EINVAL, "unknown target kind %d", desc->target_kind);
  const int T = (desc->method == IS_PLAIN || desc->method == IS_MIS) ? 1 : desc->rounds;
  if (T < 1 || !desc->round_size) return fail(IS_EINVAL, "rounds must be positive");
  auto j = new is_job();
  std::unique_ptr<is_job> guard(j);
  j->ctx = ctx;
  j->method = desc->method;
  j->family = desc->family;
  j->kind = desc->scale_kind;
  j->scalar = desc->scalar;
  j->d = d;
  j->D = dim_bucket(d);
  j->T = T;
  j->dof = desc->dof;
  j->resampling = desc->resampling;
  if (desc->method == IS_DMPMC && (desc->resampling < IS_MULTINOMIAL || desc->resampling > IS_LOCAL))
    return fail(IS_EINVAL, "unknown resampling %d", desc->resampling);
  j->tkind = desc->target_kind;
  j->tbatch = desc->target_batch;
  j->tuser = desc->target_user;
  if (desc->transform_kinds) {
    if (!desc->transform_bounds) return fail(IS_EINVAL, "transforms need bounds (2 per coordinate)");
    const int dd = desc->dim;
    for (int c = 0; c < dd; ++c) {
      const int k = desc->transform_kinds[c];
      const double lo = desc->transform_bounds[2 * c], hi = desc->transform_bounds[2 * c + 1];
      if (k < IS_TRANSFORM_IDENTITY || k > IS_TRANSFORM_INTERVAL) return fail(IS_EINVAL, "unknown transform kind %d", k);
      if (k == IS_TRANSFORM_INTERVAL && (lo != lo || hi != hi || !(lo < hi) || (std::isinf(lo) && std::isinf(hi))))
        return fail(IS_EINVAL, "interval transform %d needs lower < upper with a finite side", c);
    }
    int rc;
    if ((rc = j->tf_kinds_d.alloc((size_t)dd * sizeof(int32_t))) || (rc = j->tf_bounds_d.alloc((size_t)2 * dd * sizeof(double))))
      return rc;
    HIP_TRY(locked_memcpy(j->tf_kinds_d.p, desc->transform_kinds, (size_t)dd * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(locked_memcpy(j->tf_bounds_d.p, desc->transform_bounds, (size_t)2 * dd * sizeof(double), hipMemcpyHostToDevice));
    j->has_tf = true;
  }
  j->dl = d;
  if (desc->transform_block_count != 0) {
    if (const int rc = prepare_layout(j, desc)) return rc;
  }
  j->precision = desc->precision;
  j->key = RngKey{desc->seed, desc->stream};
  j->stride = slot_stride(d);
  if (j->family == IS_STUDENT_T) {
    if (!(std::isfinite(j->dof) && j->dof > 0.0)) return fail(IS_EINVAL, "degrees of freedom must be finite and positive");
    j->logconstant = student_t_logconstant(j->dof, d);
    if ((j->method == IS_AMIS || j->method == IS_NPMC) && !(j->dof > 2.0))
      return fail(IS_EINVAL, "covariance adaptation requires Student-t degrees of freedom > 2");
  }
  j->schedule.assign(desc->round_size, desc->round_size + T);
  j->offsets.assign(T + 1, 0);
  for (int t = 0; t < T; ++t) {
    if (j->schedule[t] <= 0) return fail(IS_EINVAL, "round sizes must be positive");
    j->offsets[t + 1] = j->offsets[t] + j->schedule[t];
    j->maxN = std::max(j->maxN, j->schedule[t]);
  }
  j->total = j->offsets[T];
  if (j->total >= ((int64_t)1 << 31)) return fail(IS_EUNSUPPORTED, "sample budget exceeds 2^31 - 1");
  // proposals
  if (j->method == IS_DMPMC || j->method == IS_MIS) {
    const int Kc = desc->bank_size;
    if (Kc < 1 || !desc->bank_locations || !desc->bank_scales || !desc->bank_masses)
      return fail(IS_EINVAL, "%s needs a proposal bank", j->method == IS_MIS ? "static MIS" : "DM-PMC");
    std::vector<int> active;
    for (int k = 0; k < Kc; ++k) {
      const double m = desc->bank_masses[k];
      if (!(std::isfinite(m) && m >= 0.0)) return fail(IS_EINVAL, "masses must be finite and nonnegative");
      if (m != 0.0) active.push_back(k);
    }
    if (active.empty()) return fail(IS_EINVAL, "bank has no positive mass");
    std::stable_sort(active.begin(), active.end(),
                     [&](int a, int b) { return desc->bank_masses[a] < desc->bank_masses[b]; });
    const int K = (int)active.size();
    j->K = K;
    j->nslots = K;
    j->bank_slots_init.assign((size_t)K * j->stride, 0.0);
    j->committed_locs.assign((size_t)K * d, 0.0);
    std::vector<double> masses(K);
    const int scale_len = desc->scale_kind == IS_SPHERICAL ? 1 : desc->scale_kind == IS_DIAGONAL ? d : d * d;
    for (int s = 0; s < K; ++s) {
      const int k = active[s];
      HostProp p;
      std::string err = make_prop(j->family, j->dof, j->kind, d, desc->bank_locations + (size_t)k * d,
                                  desc->bank_scales + (size_t)k * scale_len, p);
      if (!err.empty()) return fail(IS_EINVAL, "%s", err.c_str());
      pack_slot(p, d, 0.0, j->bank_slots_init.data() + (size_t)s * j->stride);
      for (int c = 0; c < d; ++c) j->committed_locs[(size_t)s * d + c] = p.mu[c];
      j->bank_ids.push_back(k + 1);
      masses[s] = desc->bank_masses[k];
    }
    if (j->method == IS_MIS) {
      int rc = setup_mis(j, desc, masses);
      if (rc) return rc;
    } else {
      std::vector<u128> I;
      if (!exact_integer_masses(masses, I))
        return fail(IS_EUNSUPPORTED, "mass ratios too wide for the exact allocation plan");
      j->plan_counts.assign((size_t)K * T, 0);
      j->plan_logcoef.assign((size_t)K * T, 0.0);
      std::vector<int32_t> assign((size_t)T * j->maxN, 0);
      for (int t = 0; t < T; ++t) {
        std::vector<int64_t> cnt;
        if (!round_counts(I, j->schedule[t], t + 1, cnt)) return fail(IS_EINVAL, "invalid DM-PMC allocation remainder");
        int64_t pos = 0;
        for (int s = 0; s < K; ++s) {
          if (cnt[s] <= 0)
            return fail(IS_EINVAL, "round %d of size %lld cannot allocate at least one sample to every positive-mass proposal",
                        t + 1, (long long)j->schedule[t]);
          j->plan_counts[(size_t)t * K + s] = cnt[s];
          j->plan_logcoef[(size_t)t * K + s] = std::log((double)cnt[s] / (double)j->schedule[t]);
          for (int64_t q = 0; q < cnt[s]; ++q) assign[(size_t)t * j->maxN + pos++] = s;
        }
      }
      int rc;
      if ((rc = j->assign.alloc(assign.size() * sizeof(int32_t)))) return rc;
      HIP_TRY(locked_memcpy(j->assign.p, assign.data(), assign.size() * sizeof(int32_t), hipMemcpyHostToDevice));
      std::vector<int32_t> groups((size_t)T * (K + 1), 0);
      for (int t = 0; t < T; ++t)
        for (int k = 0; k < K; ++k)
          groups[(size_t)t * (K + 1) + k + 1] =
              groups[(size_t)t * (K + 1) + k] + (int32_t)j->plan_counts[(size_t)t * K + k];
      if ((rc = j->plan_group.alloc(groups.size() * sizeof(int32_t)))) return rc;
      HIP_TRY(locked_memcpy(j->plan_group.p, groups.data(), groups.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    int rc;
    if ((rc = j->bank_ids_d.alloc(j->bank_ids.size() * sizeof(int32_t)))) return rc;
    HIP_TRY(locked_memcpy(j->bank_ids_d.p, j->bank_ids.data(), j->bank_ids.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  } else {
    if (!desc->location || !desc->scale) return fail(IS_EINVAL, "proposal location and scale are required");
    std::string err = make_prop(j->family, j->dof, j->kind, d, desc->location, desc->scale, j->initial);
    if (!err.empty()) return fail(IS_EINVAL, "%s", err.c_str());
    j->committed = j->initial;
    j->nslots = j->method == IS_PLAIN ? 1 : T + 1;
  }
  // target (over the logical dimension dl: d, or the layout's)
  const int dl = j->dl;
  if (j->tkind == IS_TARGET_GAUSSIAN) {
    if (!desc->target_params || desc->target_params_len < (int64_t)dl + dl * dl + 1)
      return fail(IS_EINVAL, "Gaussian target needs mean, factor and lognormalizer");
    HostProp tp;
    tp.mu.assign(desc->target_params, desc->target_params + dl);
    tp.L.assign(desc->target_params + dl, desc->target_params + dl + dl * dl);
    tp.lognorm = desc->target_params[dl + dl * dl];
    for (int c = 0; c < dl; ++c)
      if (!(tp.L[c + (size_t)c * dl] != 0.0)) return fail(IS_EINVAL, "target factor diagonal must be nonzero");
    j->tslot.assign(slot_stride(dl), 0.0);
    pack_slot(tp, dl, 0.0, j->tslot.data());
  } else if (j->tkind == IS_TARGET_LOGISTIC || j->tkind == IS_TARGET_LINREG) {
    // [n, p, prior_var, X, y] (logistic) or [n, p, sigma, tau, X, y] (linear regression)
    const int head = j->tkind == IS_TARGET_LOGISTIC ? 3 : 4;
    const double* tp = desc->target_params;
    if (!tp || desc->target_params_len < head) return fail(IS_EINVAL, "regression target parameters are missing");
    const double nd = tp[0], pd = tp[1];
    if (!(nd >= 1.0) || nd != std::floor(nd) || pd != (double)dl)
      return fail(IS_EINVAL, "regression target needs n >= 1 rows and p == dim coefficients");
    if (dl > kMaxDim) return fail(IS_EUNSUPPORTED, "regression targets take at most %d coefficients", kMaxDim);
    const int64_t n = (int64_t)nd;
    if (desc->target_params_len < head + n * dl + n) return fail(IS_EINVAL, "regression data is shorter than n x p + n");
    j->tn = n;
    const double pi2 = 2.0 * 3.141592653589793;
    if (j->tkind == IS_TARGET_LOGISTIC) {
      const double pv = tp[2];
      if (!(pv > 0.0) || !std::isfinite(pv)) return fail(IS_EINVAL, "prior variance must be positive");
      j->t_inv_two_prior = 0.5 / pv;
      j->t_const = -0.5 * (double)dl * std::log(pi2 * pv);
    } else {
      const double sigma = tp[2], tau = tp[3];
      if (!(sigma > 0.0) || !(tau > 0.0) || !std::isfinite(sigma) || !std::isfinite(tau))
        return fail(IS_EINVAL, "sigma and tau must be positive");
      j->t_inv_two_var = 0.5 / (sigma * sigma);
      j->t_inv_two_prior = 0.5 / (tau * tau);
      j->t_const = -(double)n * std::log(sigma) - 0.5 * (double)n * std::log(pi2) - (double)dl * std::log(tau) -
                   0.5 * (double)dl * std::log(pi2);
    }
    j->tdata.assign(tp + head, tp + head + n * dl + n);
    if (j->precision == IS_FP32) {
      // FP32 log-targets of a large data set sit near log pi(x0) ~ -n; storing them
      // minus that FP64 constant keeps their FP32 rounding at the scale of the
      // differences between samples, which is all the weights depend on
      std::vector<double> x0(d, 0.0);
      if (j->method == IS_DMPMC) {
        const int K = (int)(j->committed_locs.size() / d);
        for (int k = 0; k < K; ++k)
          for (int c = 0; c < d; ++c) x0[c] += j->committed_locs[(size_t)k * d + c] / K;
      } else {
        x0 = j->initial.mu;
      }
      if (j->has_layout) {
        // the offset's anchor is the logical image of that location
        std::vector<double> xl(dl, 0.0);
        if (!layout_forward_host(j->layout, x0.data(), xl.data())) std::fill(xl.begin(), xl.end(), 0.0);
        x0.swap(xl);
      }
      const double c0 = regression_logtarget(j, x0.data());
      j->t_shift = std::isfinite(c0) ? c0 : 0.0;
      j->t_x0 = x0;
    }
  }
  {
    CaptureGuard cg(capture_lock());
    HIP_TRY(hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking));
  }
  int rc = j->precision == IS_FP32 ? alloc_job_buffers<float>(j) : alloc_job_buffers<double>(j);
  if (rc) return rc;
  // logmass tables: AMIS log N_t per slot; DM-PMC per-round log coefficients
  std::vector<double> logmass;
  if (j->method == IS_AMIS) {
    logmass.assign(j->nslots, 0.0);
    for (int t = 0; t < T; ++t) logmass[t] = std::log((double)j->schedule[t]);
  } else if (j->method == IS_DMPMC) {
    logmass = j->plan_logcoef;
  } else if (j->method == IS_MIS) {
    logmass = j->mis_logmass;
    if (!j->mis_gcoef.empty()) {
      rc = j->precision == IS_FP32 ? upload_T<float>(j->mis_gcoefT, j->mis_gcoef)
                                   : upload_T<double>(j->mis_gcoefT, j->mis_gcoef);
      if (rc) return rc;
    }
  } else {
    logmass.assign(std::max(j->nslots, 1), 0.0);
  }
  rc = j->precision == IS_FP32 ? upload_T<float>(j->logmassT, logmass) : upload_T<double>(j->logmassT, logmass);
  if (rc) return rc;
  if (j->has_layout) {
    const std::vector<LayoutBlock> lb =
        j->precision == IS_FP32 ? layout_blocks<float>(j->layout) : layout_blocks<double>(j->layout);
    if ((rc = j->layout_d.alloc(lb.size() * sizeof(LayoutBlock)))) return rc;
    HIP_TRY(locked_memcpy(j->layout_d.p, lb.data(), lb.size() * sizeof(LayoutBlock), hipMemcpyHostToDevice));
  }
  if (!j->tslot.empty()) {
    rc = j->precision == IS_FP32 ? upload_T<float>(j->tslotT, j->tslot) : upload_T<double>(j->tslotT, j->tslot);
    if (rc) return rc;
  }
  if (!j->tdata.empty()) {
    const size_t nx = (size_t)j->tn * j->dl;
    std::vector<double> xpart(j->tdata.begin(), j->tdata.begin() + nx), ypart(j->tdata.begin() + nx, j->tdata.end());
    rc = j->precision == IS_FP32 ? upload_T<float>(j->Xd, xpart) : upload_T<double>(j->Xd, xpart);
    if (!rc) rc = j->precision == IS_FP32 ? upload_T<float>(j->yd, ypart) : upload_T<double>(j->yd, ypart);
    if (rc) return rc;
    if (j->precision == IS_FP32 && j->tkind == IS_TARGET_LOGISTIC && dl <= 64 && (int)j->t_x0.size() == dl) {
      if ((rc = upload_logistic_bf16(j, dl))) return rc;
    }
    j->tdata.clear();
    j->tdata.shrink_to_fit();
  }
  j->prefilled.assign(T, 0);
  j->prefilled_radial.assign(T, 0);
  j->prefilled_uniforms.assign(T, 0);
  j->prefilled_ancestors.assign(T, 0);
  *out = guard.release();
  return IS_OK;
}

// Device bytes a job of this descriptor allocates (alloc_job_buffers plus the
// target data and the DM-PMC plan); the serving loop admits jobs against it.
int64_t is_job_device_bytes(const is_job_desc* desc) {
  if (!desc || !desc->round_size || desc->dim < 1 || desc->dim > kMaxDim) return -1;
  const int d = desc->dim;
  const int T = (desc->method == IS_PLAIN || desc->method == IS_MIS) ? 1 : desc->rounds;
  if (T < 1) return -1;
  int64_t total = 0, maxN = 0;
  for (int t = 0; t < T; ++t) {
    if (desc->round_size[t] <= 0) return -1;
    total += desc->round_size[t];
    maxN = std::max<int64_t>(maxN, desc->round_size[t]);
  }
  const int64_t ts = desc->precision == IS_FP32 ? 4 : 8;
  const int stride = slot_stride(d);
  const bool banked = desc->method == IS_DMPMC || desc->method == IS_MIS;
  const int K = banked ? std::max(1, desc->bank_size) : 1;
  const int nslots = banked ? K : desc->method == IS_PLAIN ? 1 : T + 1;
  const int nv_mom = 2 + d + d * (d + 1) / 2;
  const int64_t pcap = cdiv(total, kSweepPerBlock) + cdiv(maxN, 64) + 1;
  int64_t bytes = total * (d * ts + 3 * ts + 2 * 4);         // samples, log arrays, ids
  bytes += (int64_t)nslots * stride * (8 + ts);               // slot tables
  bytes += 2 * pcap * (1 + nv_mom) * 8 + 2 * (1 + nv_mom) * 8; // partial records, reduced records
  bytes += (int64_t)T * (5 * 8 + 3 * 8 + 2 * 4 + 3 * 8) + 4 * 64;
  if (desc->target_kind == IS_TARGET_CALLBACK) bytes += maxN * ts;
  if (desc->transform_kinds) bytes += total * d * ts + (int64_t)d * 20;
  int dl = d;
  if (desc->transform_block_count > 0 && desc->transform_blocks) {
    dl = 0;
    for (int b = 0; b < desc->transform_block_count; ++b)
      dl += std::max(0, desc->transform_blocks[b].count) + (desc->transform_blocks[b].kind == IS_TRANSFORM_SIMPLEX ? 1 : 0);
    bytes += total * dl * ts + (int64_t)desc->transform_block_count * (int64_t)sizeof(LayoutBlock);
    if (desc->target_kind == IS_TARGET_STANDARD || desc->target_kind == IS_TARGET_GAUSSIAN) bytes += maxN * ts;
  }
  if (desc->target_kind == IS_TARGET_GAUSSIAN) bytes += (int64_t)slot_stride(dl) * ts;
  if (desc->target_kind == IS_TARGET_LOGISTIC || desc->target_kind == IS_TARGET_LINREG) {
    const int64_t n = desc->target_params && desc->target_params_len > 0 ? (int64_t)desc->target_params[0] : 0;
    bytes += n * (dl + 1) * ts + maxN * ts;
    bytes += std::max<int64_t>(1, target_splits(n, maxN) * maxN) * 8;
    if (desc->target_kind == IS_TARGET_LOGISTIC && desc->precision == IS_FP32 && dl <= 64)
      bytes += cdiv(n, 16) * (dl <= 32 ? 1 : 2) * 2 * 64 * 16 + cdiv(n, 16) * 16 * 8 + (int64_t)dl * 8;
  }
  if (desc->method == IS_DMPMC) {
    bytes += (int64_t)T * maxN * 4 + maxN * 8 + (int64_t)K * T * 12;
    bytes += (2 * cdf_tiles(maxN) + 2) * 8;
  }
  if (desc->method == IS_MIS) bytes += maxN * 4 + (int64_t)K * (8 + 4 * 4 + ts) + 64;  // assignment, design tables
  return bytes;
}

int is_job_destroy(is_job* job) {
  if (!job) return IS_OK;
  if (job->stream) (void)hipStreamSynchronize(job->stream);
  CaptureGuard cg(capture_lock());
  if (job->graph_exec) (void)hipGraphExecDestroy(job->graph_exec);
  if (job->graph) (void)hipGraphDestroy(job->graph);
  for (auto& e : job->events) (void)hipEventDestroy(e);
  if (job->stream) (void)hipStreamDestroy(job->stream);
  delete job;
  return IS_OK;
}

int is_job_run(is_job* job) {
  if (!job) return fail(IS_EINVAL, "null job");
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  HIP_TRY(hipSetDevice(job->ctx->device));
  return job->precision == IS_FP32 ? job_run_typed<float>(job) : job_run_typed<double>(job);
}

int is_job_reset(is_job* job) {
  if (!job) return fail(IS_EINVAL, "null job");
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  job->committed = job->initial;
  if (job->method == IS_DMPMC)
    for (int k = 0; k < job->K; ++k)
      for (int c = 0; c < job->d; ++c)
        job->committed_locs[(size_t)k * job->d + c] = job->bank_slots_init[(size_t)k * job->stride + c];
  return IS_OK;
}

static int upload_prefill(DevBuf& buf, size_t total_doubles, size_t offset, const double* src, size_t n) {
  if (!buf.p) {
    int rc = buf.alloc(total_doubles * sizeof(double));
    if (rc) return rc;
    HIP_TRY(locked_memset(buf.p, 0, buf.bytes));
  }
  if (offset + n > total_doubles) return fail(IS_EINVAL, "prefilled buffer too large");
  HIP_TRY(locked_memcpy(buf.as<double>() + offset, src, n * sizeof(double), hipMemcpyHostToDevice));
  return IS_OK;
}

int is_job_set_normals(is_job* job, int32_t round, const double* normals, int64_t count) {
  if (!job || !normals || round < 0 || round >= job->T) return fail(IS_EINVAL, "bad arguments");
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  const int rs = radial_normal_stride(job->family, job->dof);
  if (job->method == IS_PLAIN) {
    const size_t need = (size_t)(job->d + rs) * job->maxN;
    if ((size_t)count != need) return fail(IS_EINVAL, "plain IS normals need %zu values", need);
    int rc = upload_prefill(job->pre_normals, need, 0, normals, need);
    if (rc) return rc;
  } else {
    const size_t per = (size_t)job->d * job->maxN;
    if (count < (int64_t)job->d * job->schedule[round] || (size_t)count > per)
      return fail(IS_EINVAL, "round normals need d * round_size values");
    int rc = upload_prefill(job->pre_normals, per * job->T, per * round, normals, count);
    if (rc) return rc;
  }
  job->prefilled[round] = 1;
  job->graph_dirty = true;
  return IS_OK;
}

int is_job_set_radial(is_job* job, int32_t round, const double* normals, int64_t ncount, const double* uniforms,
                      int64_t ucount) {
  if (!job || round < 0 || round >= job->T) return fail(IS_EINVAL, "bad arguments");
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  const int rs = radial_normal_stride(job->family, job->dof);
  const int us = radial_uniform_stride(job->family, job->dof);
  if (job->method == IS_PLAIN) {
    // plain IS interleaves radial normals into the normal buffer; only uniforms are separate
    if (us) {
      const size_t need = (size_t)us * job->maxN;
      if ((size_t)ucount != need || !uniforms) return fail(IS_EINVAL, "plain IS uniforms need %zu values", need);
      int rc = upload_prefill(job->pre_ru, need, 0, uniforms, need);
      if (rc) return rc;
    }
    job->prefilled_radial[round] = 1;
  job->graph_dirty = true;
    return IS_OK;
  }
  if (rs) {
    const size_t per = (size_t)rs * job->maxN;
    if (!normals || ncount > (int64_t)per) return fail(IS_EINVAL, "radial normals too long");
    int rc = upload_prefill(job->pre_rn, per * job->T, per * round, normals, ncount);
    if (rc) return rc;
  }
  if (us) {
    const size_t per = (size_t)us * job->maxN;
    if (!uniforms || ucount > (int64_t)per) return fail(IS_EINVAL, "radial uniforms too long");
    int rc = upload_prefill(job->pre_ru, per * job->T, per * round, uniforms, ucount);
    if (rc) return rc;
  }
  job->prefilled_radial[round] = 1;
  job->graph_dirty = true;
  return IS_OK;
}

int is_job_set_uniforms(is_job* job, int32_t round, const double* uniforms, int64_t count) {
  if (!job || !uniforms || round < 0 || round >= job->T || (job->method != IS_DMPMC && job->method != IS_MIS))
    return fail(IS_EINVAL, "bad arguments");
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  if (job->method == IS_MIS) {
    // static MIS: one assignment uniform per sample
    const int64_t N = job->schedule[0];
    if (count != N) return fail(IS_EINVAL, "static MIS assignment uniforms: one per sample (%lld)", (long long)N);
    int rc = upload_prefill(job->pre_uniforms, (size_t)N, 0, uniforms, (size_t)N);
    if (rc) return rc;
    job->prefilled_uniforms[0] = 1;
    job->graph_dirty = true;
    return IS_OK;
  }
  if (count < 1 || count > job->K) return fail(IS_EINVAL, "uniform count must be in [1, K]");
  int rc = upload_prefill(job->pre_uniforms, (size_t)job->K * job->T, (size_t)job->K * round, uniforms, count);
  if (rc) return rc;
  job->prefilled_uniforms[round] = 1;
  job->graph_dirty = true;
  return IS_OK;
}

int is_job_set_ancestors(is_job* job, int32_t round, const int64_t* ancestors, int64_t count) {
  if (!job || !ancestors || round < 0 || round >= job->T || job->method != IS_DMPMC)
    return fail(IS_EINVAL, "bad arguments");
  if (count != job->K) return fail(IS_EINVAL, "one ancestor per active proposal (%d)", job->K);
  const int64_t N = job->schedule[round];
  std::vector<int32_t> h((size_t)count);
  for (int64_t k = 0; k < count; ++k) {
    if (ancestors[k] < 1 || ancestors[k] > N) return fail(IS_EINVAL, "ancestor %lld out of [1, %lld]",
                                                          (long long)ancestors[k], (long long)N);
    h[(size_t)k] = (int32_t)ancestors[k];
  }
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  if (!job->pre_ancestors.p) {
    int rc = job->pre_ancestors.alloc((size_t)job->K * job->T * sizeof(int32_t));
    if (rc) return rc;
  }
  HIP_TRY(locked_memcpy(job->pre_ancestors.as<int32_t>() + (size_t)round * job->K, h.data(), h.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
  job->prefilled_ancestors[round] = 1;
  job->graph_dirty = true;
  return IS_OK;
}

int is_job_clear_prefilled(is_job* job) {
  if (!job) return fail(IS_EINVAL, "null job");
  std::fill(job->prefilled.begin(), job->prefilled.end(), 0);
  std::fill(job->prefilled_radial.begin(), job->prefilled_radial.end(), 0);
  std::fill(job->prefilled_uniforms.begin(), job->prefilled_uniforms.end(), 0);
  std::fill(job->prefilled_ancestors.begin(), job->prefilled_ancestors.end(), 0);
  job->graph_dirty = true;
  return IS_OK;
}

static int export_purpose(is_job* job, int32_t round, int64_t n, int count, uint64_t purpose, bool uniform,
                          std::vector<double>& host) {
  host.assign((size_t)n * count, 0.0);
  if (n * count == 0) return IS_OK;
  DevBuf tmp;
  int rc = tmp.alloc(host.size() * sizeof(double));
  if (rc) return rc;
  const uint64_t word = round_word(job, round);
  if (uniform)
    hipLaunchKernelGGL(k_export_uniforms, dim3(cdiv(n, 256)), dim3(256), 0, job->stream, job->key, word, n, count,
                       purpose, tmp.as<double>());
  else
    hipLaunchKernelGGL(k_export_normals, dim3(cdiv(n, 256)), dim3(256), 0, job->stream, job->key, word, n, count,
                       purpose, tmp.as<double>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(host.data(), tmp.p, host.size() * sizeof(double), hipMemcpyDeviceToHost, job->stream));
  HIP_TRY(hipStreamSynchronize(job->stream));
  return IS_OK;
}

int is_job_export_normals(is_job* job, int32_t round, double* out, int64_t count) {
  if (!job || !out || round < 0 || round >= job->T) return fail(IS_EINVAL, "bad arguments");
  const int d = job->d;
  const int rs = radial_normal_stride(job->family, job->dof);
  const int64_t N = job->schedule[round];
  std::vector<double> z, r;
  int rc = export_purpose(job, round, N, d, kPurposeNormals, false, z);
  if (rc) return rc;
  if (job->method == IS_PLAIN) {
    const int stride = d + rs;
    if (count != (int64_t)stride * N) return fail(IS_EINVAL, "plain IS export needs stride * n values");
    if ((rc = export_purpose(job, round, N, rs, kPurposeRadialNormals, false, r))) return rc;
    for (int64_t i = 0; i < N; ++i) {
      for (int c = 0; c < d; ++c) out[i * stride + c] = z[(size_t)i * d + c];
      for (int q = 0; q < rs; ++q) out[i * stride + d + q] = r[(size_t)i * rs + q];
    }
    return IS_OK;
  }
  if (count != (int64_t)d * N) return fail(IS_EINVAL, "round export needs d * round_size values");
  std::memcpy(out, z.data(), z.size() * sizeof(double));
  return IS_OK;
}

int is_job_export_radial(is_job* job, int32_t round, double* normals, int64_t ncount, double* uniforms,
                         int64_t ucount) {
  if (!job || round < 0 || round >= job->T) return fail(IS_EINVAL, "bad arguments");
  const int rs = radial_normal_stride(job->family, job->dof);
  const int us = radial_uniform_stride(job->family, job->dof);
  const int64_t N = job->schedule[round];
  std::vector<double> r, u;
  int rc;
  if (normals && rs) {
    if (ncount != (int64_t)rs * N) return fail(IS_EINVAL, "radial normal export needs rs * n values");
    if ((rc = export_purpose(job, round, N, rs, kPurposeRadialNormals, false, r))) return rc;
    std::memcpy(normals, r.data(), r.size() * sizeof(double));
  }
  if (uniforms && us) {
    if (ucount != (int64_t)us * N) return fail(IS_EINVAL, "radial uniform export needs us * n values");
    if ((rc = export_purpose(job, round, N, us, kPurposeRadialUniforms, true, u))) return rc;
    std::memcpy(uniforms, u.data(), u.size() * sizeof(double));
  }
  return IS_OK;
}

int is_job_export_uniforms(is_job* job, int32_t round, double* out, int64_t count) {
  if (!job || !out || round < 0 || round >= job->T || (job->method != IS_DMPMC && job->method != IS_MIS))
    return fail(IS_EINVAL, "bad arguments");
  if (job->method == IS_MIS) {
    const int64_t N = job->schedule[0];
    if (count != N) return fail(IS_EINVAL, "static MIS assignment uniforms: one per sample");
    std::vector<double> u;
    int rc = export_purpose(job, 0, N, 1, kPurposeAssignment, true, u);
    if (rc) return rc;
    std::memcpy(out, u.data(), u.size() * sizeof(double));
    return IS_OK;
  }
  if (count != job->K) return fail(IS_EINVAL, "resampling uniforms: K values");
  std::vector<double> u;
  int rc = export_purpose(job, round, job->K, 1, kPurposeResampling, true, u);
  if (rc) return rc;
  std::memcpy(out, u.data(), u.size() * sizeof(double));
  return IS_OK;
}

int is_job_set_timing(is_job* job, int32_t enable) {
  if (!job) return fail(IS_EINVAL, "null job");
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  job->timing = enable ? 1 : 0;
  if (job->timing && job->events.empty()) {
    job->events.resize((size_t)2 * job->T);
    for (auto& e : job->events) HIP_TRY(hipEventCreate(&e));
  }
  job->graph_dirty = true;
  return IS_OK;
}

int is_job_set_graph(is_job* job, int32_t enable) {
  if (!job) return fail(IS_EINVAL, "null job");
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  job->use_graph = enable ? 1 : 0;
  job->graph_dirty = true;
  return IS_OK;
}

int is_job_round_kernel_ms(is_job* job, double* out) {
  if (!job || !out) return fail(IS_EINVAL, "null argument");
  if (!job->timing || job->round_ms.size() != (size_t)job->T) return fail(IS_ESTATE, "timing was not enabled for the last run");
  std::memcpy(out, job->round_ms.data(), job->round_ms.size() * sizeof(double));
  return IS_OK;
}

int32_t is_job_logical_dim(const is_job* job) { return job ? job->dl : 0; }

int64_t is_job_total(const is_job* job) { return job ? job->total : 0; }

int is_job_failure(const is_job* job, is_failure* out) {
  if (!job || !out) return fail(IS_EINVAL, "bad arguments");
  *out = job->failure;
  return IS_OK;
}

#define REQUIRE_RESULTS(job)                                                   \
  do {                                                                         \
    if (!(job)) return fail(IS_EINVAL, "null job");                            \
    if (!(job)->have_results) return fail(IS_ESTATE, "no successful run yet"); \
  } while (0)

int is_job_get_samples(is_job* job, double* out) {
  REQUIRE_RESULTS(job);
  const int64_t n = job->total;
  const int dl = job->dl;
  DevBuf tmp;
  int rc = tmp.alloc((size_t)n * dl * sizeof(double));
  if (rc) return rc;
  if (job->precision == IS_FP32)
    hipLaunchKernelGGL((k_soa_to_aos<float>), dim3(cdiv(n * dl, 256)), dim3(256), 0, job->stream,
                       (job->has_tf ? job->XL : job->X).as<float>(), n, n, dl, tmp.as<double>());
  else
    hipLaunchKernelGGL((k_soa_to_aos<double>), dim3(cdiv(n * dl, 256)), dim3(256), 0, job->stream,
                       (job->has_tf ? job->XL : job->X).as<double>(), n, n, dl, tmp.as<double>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, tmp.p, tmp.bytes, hipMemcpyDeviceToHost, job->stream));
  HIP_TRY(hipStreamSynchronize(job->stream));
  return IS_OK;
}

static int copy_as_double(is_job* job, const DevBuf& src, int64_t n, double* out) {
  if (job->precision == IS_FP64) {
    HIP_TRY(hipMemcpyAsync(out, src.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, job->stream));
  } else {
    DevBuf tmp;
    int rc = tmp.alloc((size_t)n * sizeof(double));
    if (rc) return rc;
    hipLaunchKernelGGL((k_to_double<float>), dim3(cdiv(n, 256)), dim3(256), 0, job->stream, src.as<float>(), n,
                       tmp.as<double>());
    HIP_TRY(hipMemcpyAsync(out, tmp.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, job->stream));
    HIP_TRY(hipStreamSynchronize(job->stream));
    return IS_OK;
  }
  HIP_TRY(hipStreamSynchronize(job->stream));
  return IS_OK;
}

int is_job_get_logweights(is_job* job, double* out) {
  REQUIRE_RESULTS(job);
  const int rc = copy_as_double(job, job->logweights, job->total, out);
  if (rc == IS_OK && job->t_shift != 0.0)
    for (int64_t i = 0; i < job->total; ++i) out[i] += job->t_shift;
  return rc;
}

int is_job_logweight_offset(const is_job* job, double* out) {
  if (!job || !out) return fail(IS_EINVAL, "null argument");
  *out = job->t_shift;
  return IS_OK;
}

static int copy_ids(is_job* job, const DevBuf& src, int64_t* out) {
  std::vector<int32_t> h((size_t)job->total);
  HIP_TRY(locked_memcpy(h.data(), src.p, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < h.size(); ++i) out[i] = h[i];
  return IS_OK;
}

int is_job_get_round_ids(is_job* job, int64_t* out) {
  REQUIRE_RESULTS(job);
  return copy_ids(job, job->round_ids, out);
}

int is_job_get_proposal_ids(is_job* job, int64_t* out) {
  REQUIRE_RESULTS(job);
  if (job->method != IS_DMPMC && job->method != IS_MIS)
    return fail(IS_ESTATE, "proposal ids exist for DM-PMC and static MIS only");
  return copy_ids(job, job->proposal_ids, out);
}

int is_job_get_round_summary(is_job* job, double* ess, double* lognormalizers, double* adaptation_ess) {
  REQUIRE_RESULTS(job);
  // the run's own read-back holds them (the round kernels are their only writers)
  const char* down = job->ctl_down.as<const char>();
  auto host_copy = [&](const DevView& v) {
    return reinterpret_cast<const double*>(down + ((const char*)v.p - job->ctl.as<char>()));
  };
  const size_t n = (size_t)job->T;
  if (ess) std::memcpy(ess, host_copy(job->round_ess), n * sizeof(double));
  if (lognormalizers) {
    std::memcpy(lognormalizers, host_copy(job->round_logz), n * sizeof(double));
    for (size_t t = 0; t < n; ++t) lognormalizers[t] += job->t_shift;
  }
  if (adaptation_ess) std::memcpy(adaptation_ess, host_copy(job->adapt_ess), n * sizeof(double));
  return IS_OK;
}

int is_job_get_proposal(is_job* job, double* location, double* factor, double* lognormalizer) {
  if (!job) return fail(IS_EINVAL, "null job");
  if (job->method == IS_PLAIN || job->method == IS_DMPMC || job->method == IS_MIS)
    return fail(IS_ESTATE, "no adaptive proposal");
  const int d = job->d;
  if (location) std::memcpy(location, job->committed.mu.data(), d * sizeof(double));
  if (factor) {
    if (job->scalar)
      factor[0] = job->committed.L[0];
    else
      std::memcpy(factor, job->committed.L.data(), (size_t)d * d * sizeof(double));
  }
  if (lognormalizer) *lognormalizer = job->committed.lognorm;
  return IS_OK;
}

int is_job_get_history(is_job* job, double* means, double* factors, double* lognormalizers) {
  REQUIRE_RESULTS(job);
  if (job->method != IS_AMIS && job->method != IS_NPMC) return fail(IS_ESTATE, "no proposal history");
  const int d = job->d;
  std::vector<double> slots((size_t)job->nslots * job->stride);
  HIP_TRY(locked_memcpy(slots.data(), job->slots64.p, slots.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int s = 0; s < job->nslots; ++s) {
    HostProp p;
    unpack_slot(slots.data() + (size_t)s * job->stride, d, p);
    if (means) std::memcpy(means + (size_t)s * d, p.mu.data(), d * sizeof(double));
    if (factors) std::memcpy(factors + (size_t)s * d * d, p.L.data(), (size_t)d * d * sizeof(double));
    if (lognormalizers) lognormalizers[s] = p.lognorm;
  }
  return IS_OK;
}

int is_job_get_locations(is_job* job, double* out) {
  if (!job || !out) return fail(IS_EINVAL, "bad arguments");
  if (job->method != IS_DMPMC) return fail(IS_ESTATE, "locations exist for DM-PMC only");
  std::memcpy(out, job->committed_locs.data(), job->committed_locs.size() * sizeof(double));
  return IS_OK;
}

int is_job_get_ancestors(is_job* job, int64_t* out) {
  REQUIRE_RESULTS(job);
  if (job->method != IS_DMPMC) return fail(IS_ESTATE, "ancestors exist for DM-PMC only");
  std::vector<int32_t> h((size_t)job->K * job->T);
  HIP_TRY(locked_memcpy(h.data(), job->ancestors.p, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < h.size(); ++i) out[i] = h[i];
  return IS_OK;
}

int is_job_device_results(is_job* job, const void** samples, int64_t* coord_stride, const void** logweights,
                          int32_t* is_float) {
  REQUIRE_RESULTS(job);
  if (samples) *samples = job->has_tf ? job->XL.p : job->X.p;  // logical samples under a transform
  if (coord_stride) *coord_stride = job->total;
  if (logweights) *logweights = job->logweights.p;
  if (is_float) *is_float = job->precision == IS_FP32;
  return IS_OK;
}

extern "C++" {
template <class T>
static int job_lognormalizer_typed(is_job* job, double* out3) {
  const int64_t n = job->total;
  SweepArgs<T> w = sweep_args<T>(job, 0);
  w.lo = 0;
  w.hi = n;
  w.nv = 2;
  w.fail = nullptr;
  w.partials = job->partials2.as<double>();
  if (job->D == 0) w.d = 1;  // results of more than 64 logical coordinates: the weights alone
  const int blocks = launch_sweep<T>(job->D > 0 ? job->D : 1, 0, w, job->stream);
  FinishArgs f{};
  f.d = w.d;
  f.flags = kFinishSummary;
  f.summary_count = n;
  f.summary_out = job->summary.as<double>();
  f.counter = job->counter.as<unsigned int>() + 1;
  hipLaunchKernelGGL((k_reduce_finish<0, false>), dim3(2), dim3(kBlock), 0, job->stream,
                     job->partials2.as<double>(), (int64_t)blocks, job->pcap, 2, job->reduced2.as<double>(), f);
  launch_finish(w.d, job->reduced2.as<double>(), f, job->stream);
  HIP_TRY(hipGetLastError());
  double s[5];
  HIP_TRY(hipMemcpyAsync(s, job->summary.p, sizeof(s), hipMemcpyDeviceToHost, job->stream));
  HIP_TRY(hipStreamSynchronize(job->stream));
  out3[0] = s[4] + job->t_shift;
  out3[1] = s[0] + job->t_shift;
  out3[2] = s[1];
  return IS_OK;
}
}  // extern "C++"

int is_job_lognormalizer(is_job* job, double* out3) {
  REQUIRE_RESULTS(job);
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  return job->precision == IS_FP32 ? job_lognormalizer_typed<float>(job, out3)
                                   : job_lognormalizer_typed<double>(job, out3);
}

extern "C++" {
template <class T>
static int job_moments_typed(is_job* job, double* mean, double* var, double* cov) {
  const int64_t n = job->total;
  const int d = job->dl;  // the logical samples' dimension
  if (d > kMaxDim) return fail(IS_EUNSUPPORTED, "moments of more than %d logical coordinates", kMaxDim);
  const int nv = 2 + d + d * (d + 1) / 2;
  // centre at the FP64 proposal mean of the first slot to keep the single pass well conditioned
  DevBuf center;
  int rc = center.alloc((size_t)d * sizeof(double));
  if (rc) return rc;
  std::vector<double> c0(d, 0.0);
  if (job->method == IS_AMIS || job->method == IS_NPMC || job->method == IS_PLAIN)
    c0 = job->committed.mu;
  else
    for (int c = 0; c < d; ++c) c0[c] = job->committed_locs[c];
  if (job->has_tf) {
    // logical samples: centre at the first sample's logical coordinates
    hipLaunchKernelGGL((k_soa_to_aos<T>), dim3(1), dim3(256), 0, job->stream, job->XL.as<T>(), job->total, (int64_t)1,
                       d, center.as<double>());
    HIP_TRY(hipMemcpyAsync(c0.data(), center.p, d * sizeof(double), hipMemcpyDeviceToHost, job->stream));
  } else {
    HIP_TRY(locked_memcpy(center.p, c0.data(), d * sizeof(double), hipMemcpyHostToDevice));
  }
  SweepArgs<T> w = sweep_args<T>(job, 0);
  w.lo = 0;
  w.hi = n;
  w.nv = nv;
  w.fail = nullptr;
  w.center = center.as<double>();
  if (job->has_tf) w.X = job->XL.as<T>();
  w.d = d;
  w.partials = job->partials2.as<double>();
  const int blocks = launch_sweep<T>(dim_bucket(d), sweep_nt(d, true), w, job->stream);
  FinishArgs f{};
  f.d = d;
  f.flags = 0;
  f.counter = job->counter.as<unsigned int>() + 2;
  hipLaunchKernelGGL((k_reduce_finish<0, false>), dim3(nv), dim3(kBlock), 0, job->stream,
                     job->partials2.as<double>(), (int64_t)blocks, job->pcap, nv, job->reduced2.as<double>(), f);
  HIP_TRY(hipGetLastError());
  std::vector<double> red(1 + nv);
  HIP_TRY(hipMemcpyAsync(red.data(), job->reduced2.p, red.size() * sizeof(double), hipMemcpyDeviceToHost, job->stream));
  HIP_TRY(hipStreamSynchronize(job->stream));
  const double s0 = red[1];
  if (!(s0 > 0.0)) return fail(IS_EFAILED, "all weights are zero");
  std::vector<double> delta(d);
  for (int c = 0; c < d; ++c) {
    delta[c] = red[3 + c] / s0;
    if (mean) mean[c] = c0[c] + delta[c];
  }
  for (int r = 0; r < d; ++r)
    for (int c = 0; c <= r; ++c) {
      const double v = red[3 + d + r * (r + 1) / 2 + c] / s0 - delta[r] * delta[c];
      if (cov) {
        cov[r + (size_t)c * d] = v;
        cov[c + (size_t)r * d] = v;
      }
      if (var && r == c) var[r] = v;
    }
  return IS_OK;
}
}  // extern "C++"

int is_job_moments(is_job* job, double* mean, double* var, double* cov) {
  REQUIRE_RESULTS(job);
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  return job->precision == IS_FP32 ? job_moments_typed<float>(job, mean, var, cov)
                                   : job_moments_typed<double>(job, mean, var, cov);
}

extern "C++" {
// resample (src/resampling.jl:245-283) of the job's device log-weights: the CDF in
// the engine's fixed order, the ancestors of the host uniforms (count, or one for
// systematic); `ancestors` (host, 1-based) and `gathered` (a results-only job with
// `count` samples: the ancestors' samples and zero log-weights, the UnweightedSamples
// of src/resampling.jl:263-283) are each optional.
template <class T>
static int job_resample_typed(is_job* job, const double* uniforms, int64_t count, int32_t method, int64_t* ancestors,
                              is_job* gathered) {
  const int64_t n = job->total;
  double lz[3];
  int rc = job_lognormalizer_typed<T>(job, lz);
  if (rc) return rc;
  if (lz[1] == -__builtin_huge_val()) return fail(IS_EFAILED, "all weights are zero");
  DevBuf cdf, tiles, u, anc, st;
  const int64_t ucount = method == IS_SYSTEMATIC ? 1 : count;
  if ((rc = cdf.alloc((size_t)n * sizeof(double))) || (rc = tiles.alloc((size_t)(2 * cdf_tiles(n) + 2) * sizeof(double))) ||
      (rc = u.alloc((size_t)ucount * sizeof(double))) || (rc = anc.alloc((size_t)count * sizeof(int32_t))) ||
      (rc = st.alloc(sizeof(int))))
    return rc;
  HIP_TRY(locked_memcpy(u.p, uniforms, (size_t)ucount * sizeof(double), hipMemcpyHostToDevice));
  launch_cdf<T>(job->logweights.as<T>(), n, job->summary.as<double>(), tiles.as<double>(), cdf.as<double>(),
                st.as<int>(), job->stream);
  hipLaunchKernelGGL((k_select_ancestors<T>), dim3(cdiv(count, 256)), dim3(256), 0, job->stream, cdf.as<double>(),
                     job->logweights.as<T>(), n, u.as<double>(), count, method == IS_SYSTEMATIC ? 1 : 0,
                     anc.as<int32_t>());
  if (gathered) {
    const T* src = (job->has_tf ? job->XL : job->X).as<T>();
    hipLaunchKernelGGL((k_gather_samples<T>), dim3(cdiv(count * job->dl, 256)), dim3(256), 0, job->stream, src, n,
                       job->dl, anc.as<int32_t>(), count, gathered->X.as<T>());
    hipLaunchKernelGGL((k_fill<T>), dim3(cdiv(count, 256)), dim3(256), 0, job->stream, gathered->logweights.as<T>(),
                       count, T(0));
  }
  HIP_TRY(hipGetLastError());
  std::vector<int32_t> h(ancestors ? (size_t)count : 0);
  if (ancestors)
    HIP_TRY(hipMemcpyAsync(h.data(), anc.p, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, job->stream));
  HIP_TRY(hipStreamSynchronize(job->stream));
  for (size_t i = 0; i < h.size(); ++i) ancestors[i] = h[i];
  return IS_OK;
}
}  // extern "C++"

int is_job_resample(is_job* job, const double* uniforms, int64_t count, int32_t method, int64_t* ancestors) {
  REQUIRE_RESULTS(job);
  if (!uniforms || !ancestors || count < 1) return fail(IS_EINVAL, "resampling count must be positive");
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  return job->precision == IS_FP32 ? job_resample_typed<float>(job, uniforms, count, method, ancestors, nullptr)
                                   : job_resample_typed<double>(job, uniforms, count, method, ancestors, nullptr);
}

int is_cdf_device(is_ctx* ctx, const double* lw, int64_t n, double* cdf, void* stream) {
  if (!ctx || !lw || !cdf || n < 1) return fail(IS_EINVAL, "bad arguments");
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipSetDevice(ctx->device));
  // maximum via the summary sweep
  const int64_t pcap = cdiv(n, kSweepPerBlock) + 1;
  DevBuf partials, reduced, summary, counter, tiles, st;
  int rc;
  if ((rc = partials.alloc((size_t)pcap * 3 * sizeof(double))) || (rc = reduced.alloc(4 * sizeof(double))) ||
      (rc = summary.alloc(5 * sizeof(double))) || (rc = counter.alloc(64)) ||
      (rc = tiles.alloc((size_t)(2 * cdf_tiles(n) + 2) * sizeof(double))) || (rc = st.alloc(sizeof(int))))
    return rc;
  HIP_TRY(hipMemsetAsync(counter.p, 0, 64, s));
  SweepArgs<double> w{};
  w.d = 1;
  w.lo = 0;
  w.hi = n;
  w.logweights = const_cast<double*>(lw);
  w.partials = partials.as<double>();
  w.pcap = pcap;
  w.nv = 2;
  const int blocks = launch_sweep<double>(dim_bucket(1), 0, w, s);
  FinishArgs f{};
  f.d = 1;
  f.flags = kFinishSummary;
  f.summary_count = n;
  f.summary_out = summary.as<double>();
  f.counter = counter.as<unsigned int>();
  hipLaunchKernelGGL((k_reduce_finish<0, false>), dim3(2), dim3(kBlock), 0, s, partials.as<double>(),
                     (int64_t)blocks, pcap, 2, reduced.as<double>(), f);
  launch_finish(1, reduced.as<double>(), f, s);
  launch_cdf<double>(lw, n, summary.as<double>(), tiles.as<double>(), cdf, st.as<int>(), s);
  HIP_TRY(hipGetLastError());
  int status = 0;
  HIP_TRY(hipMemcpyAsync(&status, st.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (status) return fail(IS_EFAILED, "all weights are zero");
  return IS_OK;
}

int is_select_device(is_ctx* ctx, const double* cdf, const double* lw, int64_t n, const double* uniforms, int64_t count,
                     int32_t method, int32_t* ancestors, void* stream) {
  if (!ctx || !cdf || !lw || !uniforms || !ancestors || n < 1 || count < 1) return fail(IS_EINVAL, "bad arguments");
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipSetDevice(ctx->device));
  hipLaunchKernelGGL((k_select_ancestors<double>), dim3(cdiv(count, 256)), dim3(256), 0, s, cdf, lw, n, uniforms,
                     count, method == IS_SYSTEMATIC ? 1 : 0, ancestors);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return IS_OK;
}

// Statistics.mean(f, r) / var(f, r) on device arrays: pass 0 records and their
// merge give the mean; pass 1 (centre read from the pass-0 merge on the device)
// gives the variance.  One host wait, for both results.  `partials` holds
// 4 * cdiv(n, kFuncPerBlock) doubles, `red` 8; both are written before they are
// read, so neither needs clearing.
extern "C++" {
template <class T, class V>
static int functional_moments_in(const void* lw, const void* val, int64_t n, double* mean, double* var,
                                 hipStream_t s, double* partials, double* red) {
  const int64_t blocks = cdiv(n, kFuncPerBlock);
  const int64_t pcap = blocks;
  double* r0 = red;
  double* r1 = r0 + 4;
  FinishArgs f{};
  hipLaunchKernelGGL((k_functional<T, V>), dim3((unsigned)blocks), dim3(kBlock), 0, s, (const T*)lw, (const V*)val, n,
                     (const double*)nullptr, partials, pcap);
  hipLaunchKernelGGL((k_reduce_finish<0, false>), dim3(3), dim3(kBlock), 0, s, (const double*)partials, blocks, pcap,
                     3, r0, f);
  if (var) {
    hipLaunchKernelGGL((k_functional<T, V>), dim3((unsigned)blocks), dim3(kBlock), 0, s, (const T*)lw, (const V*)val,
                       n, (const double*)r0, partials, pcap);
    hipLaunchKernelGGL((k_reduce_finish<0, false>), dim3(3), dim3(kBlock), 0, s, (const double*)partials, blocks,
                       pcap, 3, r1, f);
  }
  HIP_TRY(hipGetLastError());
  double h[8];
  HIP_TRY(hipMemcpyAsync(h, r0, sizeof(h), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (!(h[1] > 0.0)) return fail(IS_EFAILED, "all weights are zero");
  if (mean) *mean = h[3] / h[1];
  if (var) *var = h[7] / h[5];
  return IS_OK;
}

static int functional_dispatch_in(const void* lw, int lw_fp64, const void* val, int val_fp64, int64_t n,
                                  double* mean, double* var, hipStream_t s, double* partials, double* red) {
  if (lw_fp64)
    return val_fp64 ? functional_moments_in<double, double>(lw, val, n, mean, var, s, partials, red)
                    : functional_moments_in<double, float>(lw, val, n, mean, var, s, partials, red);
  return val_fp64 ? functional_moments_in<float, double>(lw, val, n, mean, var, s, partials, red)
                  : functional_moments_in<float, float>(lw, val, n, mean, var, s, partials, red);
}

static int functional_dispatch(const void* lw, int lw_fp64, const void* val, int val_fp64, int64_t n, double* mean,
                               double* var, hipStream_t s) {
  DevBuf partials, red;
  int rc;
  if ((rc = partials.alloc((size_t)cdiv(n, kFuncPerBlock) * 4 * sizeof(double))) ||
      (rc = red.alloc(8 * sizeof(double))))
    return rc;
  return functional_dispatch_in(lw, lw_fp64, val, val_fp64, n, mean, var, s, partials.as<double>(),
                                red.as<double>());
}
}  // extern "C++"

int is_functional_moments_device(is_ctx* ctx, const void* logweights_device, int32_t logweights_fp64,
                                 const void* values_device, int32_t values_fp64, int64_t n, double* mean, double* var,
                                 void* stream) {
  if (!ctx || !logweights_device || !values_device || n < 1) return fail(IS_EINVAL, "bad arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  return functional_dispatch(logweights_device, logweights_fp64, values_device, values_fp64, n, mean, var,
                             (hipStream_t)stream);
}

int is_functional_moments_device_noalloc(is_ctx* ctx, const void* logweights_device, int32_t logweights_fp64,
                                         const void* values_device, int32_t values_fp64, int64_t n, double* mean,
                                         double* var, void* stream) {
  if (!ctx || !logweights_device || !values_device || n < 1) return fail(IS_EINVAL, "bad arguments");
  if (n > kNoAllocMaxSamples)
    return fail(IS_EINVAL, "n = %lld exceeds the allocation-free limit of %lld samples", (long long)n,
                (long long)kNoAllocMaxSamples);
  double* ptr = scratch_ptr(ctx->device);
  if (!ptr) return fail(IS_ESTATE, "device %d has no scratch binding (create the context first)", ctx->device);
  HIP_TRY(hipSetDevice(ctx->device));
  const int64_t records = 4 * cdiv(n, kFuncPerBlock);
  return functional_dispatch_in(logweights_device, logweights_fp64, values_device, values_fp64, n, mean, var,
                                (hipStream_t)stream, ptr, ptr + records);
}

int is_job_functional_moments(is_job* job, const void* values, int32_t values_fp64, int32_t values_on_device,
                              double* mean, double* var) {
  REQUIRE_RESULTS(job);
  if (!values) return fail(IS_EINVAL, "null values");
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  HIP_TRY(hipSetDevice(job->ctx->device));
  const void* dv = values;
  DevBuf up;
  if (!values_on_device) {
    const size_t bytes = (size_t)job->total * (values_fp64 ? sizeof(double) : sizeof(float));
    int rc = up.alloc(bytes);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(up.p, values, bytes, hipMemcpyHostToDevice, job->stream));
    dv = up.p;
  }
  // the device log-weights carry the offset t_shift, which normalised weights cancel
  return functional_dispatch(job->logweights.p, job->precision == IS_FP64, dv, values_fp64, job->total, mean, var,
                             job->stream);
}

int is_cdf(is_ctx* ctx, const double* logweights, int64_t n, double* cdf) {
  if (!ctx || !logweights || !cdf || n < 1) return fail(IS_EINVAL, "bad arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  DevBuf lw, out;
  int rc;
  if ((rc = lw.alloc((size_t)n * sizeof(double))) || (rc = out.alloc((size_t)n * sizeof(double)))) return rc;
  HIP_TRY(locked_memcpy(lw.p, logweights, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  rc = is_cdf_device(ctx, lw.as<double>(), n, out.as<double>(), nullptr);
  if (rc) return rc;
  HIP_TRY(locked_memcpy(cdf, out.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  return IS_OK;
}

int is_select(is_ctx* ctx, const double* cdf, const double* logweights, int64_t n, const double* uniforms,
              int64_t count, int32_t method, int64_t* ancestors) {
  if (!ctx || !cdf || !logweights || !uniforms || !ancestors || n < 1 || count < 1)
    return fail(IS_EINVAL, "bad arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  const int64_t ucount = method == IS_SYSTEMATIC ? 1 : count;
  DevBuf c, lw, u, a;
  int rc;
  if ((rc = c.alloc((size_t)n * sizeof(double))) || (rc = lw.alloc((size_t)n * sizeof(double))) ||
      (rc = u.alloc((size_t)ucount * sizeof(double))) || (rc = a.alloc((size_t)count * sizeof(int32_t))))
    return rc;
  HIP_TRY(locked_memcpy(c.p, cdf, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(locked_memcpy(lw.p, logweights, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(locked_memcpy(u.p, uniforms, (size_t)ucount * sizeof(double), hipMemcpyHostToDevice));
  rc = is_select_device(ctx, c.as<double>(), lw.as<double>(), n, u.as<double>(), count, method, a.as<int32_t>(),
                        nullptr);
  if (rc) return rc;
  std::vector<int32_t> h((size_t)count);
  HIP_TRY(locked_memcpy(h.data(), a.p, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < count; ++i) ancestors[i] = h[i];
  return IS_OK;
}

// ---------------------------------------------------------------- results with their own storage
// A WeightedSamples of a run owns its storage, which later runs of the sampler do not
// change (src/methods/importance_sampling.jl:904-906).  Here a result is a
// results-only job: the run's logical samples and log-weights copied on the device,
// the centre its moment pass uses, its own reduction buffers and stream.
}  // extern "C"

struct is_result {
  is_job* j = nullptr;
  bool unweighted = false;  // UnweightedSamples (src/resampling.jl:24-48): no weights
};

namespace {
// A results-only job of n samples shaped like `src`; allocates the sample and
// log-weight arrays (not filled) and the reduction buffers.
template <class T>
int result_shell(const is_job* src, int64_t n, std::unique_ptr<is_job>& out) {
  std::unique_ptr<is_job> r(new is_job());
  r->ctx = src->ctx;
  r->method = IS_PLAIN;
  r->family = src->family;
  r->kind = src->kind;
  r->scalar = src->scalar;
  r->d = src->dl;  // the logical samples
  r->dl = src->dl;
  r->D = src->dl <= kMaxDim ? dim_bucket(src->dl) : 0;
  r->T = 1;
  r->dof = src->dof;
  r->precision = src->precision;
  r->stride = src->stride;
  r->total = n;
  r->maxN = n;
  r->schedule = {n};
  r->use_graph = 0;
  HIP_TRY(hipSetDevice(src->ctx->device));
  HIP_TRY(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
  const int d = r->d;
  int rc;
  if ((rc = r->X.alloc((size_t)n * d * sizeof(T))) || (rc = r->logweights.alloc((size_t)n * sizeof(T)))) {
    is_job_destroy(r.release());
    return rc;
  }
  r->pcap = cdiv(n, kSweepPerBlock) + 1;
  const int nv_mom = moment_values(d);
  if ((rc = r->partials2.alloc((size_t)r->pcap * (1 + nv_mom) * sizeof(double))) ||
      (rc = r->reduced2.alloc((size_t)(1 + nv_mom) * sizeof(double))) || (rc = r->summary.alloc(5 * sizeof(double))) ||
      (rc = r->ctl.alloc(128))) {
    is_job_destroy(r.release());
    return rc;
  }
  if (locked_memset(r->ctl.p, 0, 128) != hipSuccess) {
    is_job_destroy(r.release());
    return fail(IS_EDEVICE, "hipMemset failed");
  }
  r->counter.p = r->ctl.as<char>() + 64;
  r->counter.bytes = 64;
  r->have_results = true;
  out = std::move(r);
  return IS_OK;
}

// The centre of the job's moment pass (as is_job_moments chooses it), FP64 on the host.
template <class T>
int moment_center(is_job* job, std::vector<double>& c0) {
  const int d = job->dl;
  c0.assign(d, 0.0);
  if (job->has_tf) {
    DevBuf center;
    int rc = center.alloc((size_t)d * sizeof(double));
    if (rc) return rc;
    hipLaunchKernelGGL((k_soa_to_aos<T>), dim3(1), dim3(256), 0, job->stream, job->XL.as<T>(), job->total, (int64_t)1,
                       d, center.as<double>());
    HIP_TRY(hipMemcpyAsync(c0.data(), center.p, d * sizeof(double), hipMemcpyDeviceToHost, job->stream));
    HIP_TRY(hipStreamSynchronize(job->stream));
  } else if (job->method == IS_AMIS || job->method == IS_NPMC || job->method == IS_PLAIN) {
    c0 = job->committed.mu;
  } else {
    for (int c = 0; c < d; ++c) c0[c] = job->committed_locs[c];
  }
  return IS_OK;
}

template <class T>
int snapshot_typed(is_job* job, std::unique_ptr<is_job>& out) {
  std::unique_ptr<is_job> r;
  int rc = result_shell<T>(job, job->total, r);
  if (rc) return rc;
  if ((rc = moment_center<T>(job, r->committed.mu))) {
    is_job_destroy(r.release());
    return rc;
  }
  r->t_shift = job->t_shift;
  const size_t n = (size_t)job->total;
  const DevBuf& xs = job->has_tf ? job->XL : job->X;
  if (hipMemcpyAsync(r->X.p, xs.p, n * job->dl * sizeof(T), hipMemcpyDeviceToDevice, job->stream) != hipSuccess ||
      hipMemcpyAsync(r->logweights.p, job->logweights.p, n * sizeof(T), hipMemcpyDeviceToDevice, job->stream) !=
          hipSuccess ||
      hipStreamSynchronize(job->stream) != hipSuccess) {
    is_job_destroy(r.release());
    return fail(IS_EDEVICE, "result copy failed");
  }
  out = std::move(r);
  return IS_OK;
}

template <class T>
int result_resample_typed(is_result* res, const double* uniforms, int64_t count, int32_t method, int64_t* ancestors,
                          is_result** unweighted) {
  is_job* j = res->j;
  std::unique_ptr<is_job> g;
  int rc;
  if (unweighted) {
    if ((rc = result_shell<T>(j, count, g))) return rc;
    g->committed.mu = j->committed.mu;
  }
  rc = job_resample_typed<T>(j, uniforms, count, method, ancestors, g.get());
  if (rc) {
    if (g) is_job_destroy(g.release());
    return rc;
  }
  if (unweighted) {
    is_result* u = new is_result();
    u->j = g.release();
    u->unweighted = true;
    *unweighted = u;
  }
  return IS_OK;
}
}  // namespace

extern "C" {

#define REQUIRE_RESULT(res)                                   \
  do {                                                        \
    if (!(res) || !(res)->j) return fail(IS_EINVAL, "null result"); \
  } while (0)

int is_job_result(is_job* job, is_result** out) {
  REQUIRE_RESULTS(job);
  if (!out) return fail(IS_EINVAL, "null output pointer");
  BusyGuard g(job);
  if (!g.ok) return fail(IS_EBUSY, "sampler is already running");
  HIP_TRY(hipSetDevice(job->ctx->device));
  std::unique_ptr<is_job> r;
  const int rc = job->precision == IS_FP32 ? snapshot_typed<float>(job, r) : snapshot_typed<double>(job, r);
  if (rc) return rc;
  is_result* res = new is_result();
  res->j = r.release();
  *out = res;
  return IS_OK;
}

int is_result_destroy(is_result* result) {
  if (!result) return IS_OK;
  is_job_destroy(result->j);
  delete result;
  return IS_OK;
}

int64_t is_result_count(const is_result* result) { return (result && result->j) ? result->j->total : 0; }

int is_result_is_weighted(const is_result* result) { return (result && result->j && !result->unweighted) ? 1 : 0; }

int is_result_get_samples(is_result* result, double* out) {
  REQUIRE_RESULT(result);
  if (!out) return fail(IS_EINVAL, "null output");
  return is_job_get_samples(result->j, out);
}

int is_result_get_logweights(is_result* result, double* out) {
  REQUIRE_RESULT(result);
  if (result->unweighted) return fail(IS_ESTATE, "unweighted samples carry no log-weights");
  if (!out) return fail(IS_EINVAL, "null output");
  return is_job_get_logweights(result->j, out);
}

int is_result_device_arrays(is_result* result, const void** samples, int64_t* coord_stride, const void** logweights,
                            int32_t* is_float) {
  REQUIRE_RESULT(result);
  if (logweights) *logweights = result->unweighted ? nullptr : result->j->logweights.p;
  return is_job_device_results(result->j, samples, coord_stride, nullptr, is_float);
}

int is_result_logweight_offset(const is_result* result, double* out) {
  REQUIRE_RESULT(result);
  return is_job_logweight_offset(result->j, out);
}

int is_result_lognormalizer(is_result* result, double* out3) {
  REQUIRE_RESULT(result);
  if (result->unweighted) return fail(IS_ESTATE, "unweighted samples carry no log-weights");
  if (!out3) return fail(IS_EINVAL, "null output");
  return is_job_lognormalizer(result->j, out3);
}

int is_result_moments(is_result* result, double* mean, double* var, double* cov) {
  REQUIRE_RESULT(result);
  return is_job_moments(result->j, mean, var, cov);
}

int is_result_functional_moments(is_result* result, const void* values, int32_t values_fp64,
                                 int32_t values_on_device, double* mean, double* var) {
  REQUIRE_RESULT(result);
  return is_job_functional_moments(result->j, values, values_fp64, values_on_device, mean, var);
}

int is_result_resample(is_result* result, const double* uniforms, int64_t count, int32_t method,
                       int64_t* ancestors, is_result** unweighted) {
  REQUIRE_RESULT(result);
  if (result->unweighted) return fail(IS_ESTATE, "unweighted samples carry no weights to resample");
  if (!uniforms || count < 1) return fail(IS_EINVAL, "resampling count must be positive");
  if (count > INT32_MAX) return fail(IS_EINVAL, "resampling count exceeds 2^31 - 1");
  BusyGuard g(result->j);
  if (!g.ok) return fail(IS_EBUSY, "result is in use");
  HIP_TRY(hipSetDevice(result->j->ctx->device));
  return result->j->precision == IS_FP32
             ? result_resample_typed<float>(result, uniforms, count, method, ancestors, unweighted)
             : result_resample_typed<double>(result, uniforms, count, method, ancestors, unweighted);
}

}  // extern "C"

namespace {
// logdensityof(proposal, x) for n points (d x n column-major FP64), evaluated in T by
// the density kernels' own slot code (the 64-coordinate instantiation, any d <= 64).
template <class T>
__global__ void k_proposal_logdensity(const T* __restrict__ slot, int d, Family<T> fam, const double* __restrict__ pts,
                                      int64_t n, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T x[kMaxDim], y[kMaxDim];
#pragma unroll
  for (int c = 0; c < kMaxDim; ++c) x[c] = c < d ? (T)pts[i * d + c] : T(0);
  out[i] = (double)slot_logdensity<T, kMaxDim, true>(slot, d, fam, x, y);
}

template <class T>
int proposal_logdensity_typed(is_ctx* ctx, const HostProp& p, int family, double dof, int d, const double* points,
                              int64_t n, double* out) {
  std::vector<double> s64((size_t)slot_stride(d));
  pack_slot(p, d, 0.0, s64.data());
  std::vector<T> sT(s64.begin(), s64.end());
  for (int r = 0; r < d; ++r) sT[(size_t)d + r] = slot_inv_value<T>(s64[(size_t)d + r]);
  DevBuf slot, pts, res;
  int rc;
  if ((rc = slot.alloc(sT.size() * sizeof(T))) || (rc = pts.alloc((size_t)n * d * sizeof(double))) ||
      (rc = res.alloc((size_t)n * sizeof(double))))
    return rc;
  HIP_TRY(locked_memcpy(slot.p, sT.data(), sT.size() * sizeof(T), hipMemcpyHostToDevice));
  HIP_TRY(locked_memcpy(pts.p, points, (size_t)n * d * sizeof(double), hipMemcpyHostToDevice));
  Family<T> fam;
  fam.student = family == IS_STUDENT_T;
  fam.dof = (T)dof;
  fam.half_dof_plus_d = (T)((dof + (double)d) / 2.0);
  hipLaunchKernelGGL((k_proposal_logdensity<T>), dim3((unsigned)cdiv(n, 256)), dim3(256), 0, 0, slot.as<T>(), d, fam,
                     pts.as<double>(), n, res.as<double>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(locked_memcpy(out, res.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  return IS_OK;
}
}  // namespace

extern "C" {

int is_proposal_logdensity(is_ctx* ctx, int32_t family, double dof, int32_t scale_kind, int32_t dim,
                           const double* location, const double* scale, const double* points, int64_t n,
                           int32_t precision, double* out) {
  if (!ctx || !location || !scale || !points || !out || n < 1 || dim < 1 || dim > kMaxDim)
    return fail(IS_EINVAL, "bad arguments");
  HostProp p;
  const std::string err = make_prop(family, dof, scale_kind, dim, location, scale, p);
  if (!err.empty()) return fail(IS_EINVAL, "%s", err.c_str());
  HIP_TRY(hipSetDevice(ctx->device));
  return precision == IS_FP32 ? proposal_logdensity_typed<float>(ctx, p, family, dof, dim, points, n, out)
                              : proposal_logdensity_typed<double>(ctx, p, family, dof, dim, points, n, out);
}

int is_rng_raw(is_ctx* ctx, uint64_t key0, uint64_t key1, uint64_t c1, uint64_t c2, uint64_t c3, int64_t n,
               uint64_t* out) {
  if (!ctx || !out || n < 1) return fail(IS_EINVAL, "bad arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  DevBuf tmp;
  int rc = tmp.alloc((size_t)n * 4 * sizeof(uint64_t));
  if (rc) return rc;
  hipLaunchKernelGGL(k_rng_raw, dim3(cdiv(n, 256)), dim3(256), 0, 0, key0, key1, c1, c2, c3, n,
                     tmp.as<unsigned long long>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(locked_memcpy(out, tmp.p, tmp.bytes, hipMemcpyDeviceToHost));
  return IS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- job groups
// A group advances `count` same-shape jobs (plain IS or AMIS with a built-in standard or
// Gaussian target, d <= 15, no transforms) together: each round is one k_round_group
// launch over every member (blockIdx.y = member) and one k_reduce_finish_group, the
// whole run one captured graph between a staging upload and one status download.  The
// members' launch arguments are the stand-alone runs' own (round_plan), so every member's
// results equal its stand-alone run's.
namespace {

// Per member: its control block and slot tables.
struct GroupMember {
  uint64_t* ctl;
  double* slot0_64;
  void* slot0_t;
  const double* slotT_64;
};

// Run start for every member (block = member): the run word and the zeroed status words
// [1, up_words) of the control block, and the committed slot images staged by the host
// (n64 FP64 values, then the T copy of tbytes bytes).
__global__ __launch_bounds__(256) void k_group_begin(const GroupMember* __restrict__ mem, const uint64_t* __restrict__ stage,
                                                     int members, int up_words, int n64, int tbytes) {
  const int m = blockIdx.x;
  const GroupMember g = mem[m];
  const uint64_t* img64 = stage + members + (size_t)m * n64;
  const int twords = (tbytes + 7) / 8;
  const uint64_t* imgT = stage + members + (size_t)members * n64 + (size_t)m * twords;
  for (int q = threadIdx.x; q < up_words; q += blockDim.x) g.ctl[q] = q == 0 ? stage[m] : 0ull;
  uint64_t* s64 = reinterpret_cast<uint64_t*>(g.slot0_64);
  for (int q = threadIdx.x; q < n64; q += blockDim.x) s64[q] = img64[q];
  unsigned char* sT = reinterpret_cast<unsigned char*>(g.slot0_t);
  const unsigned char* iT = reinterpret_cast<const unsigned char*>(imgT);
  for (int q = threadIdx.x; q < tbytes; q += blockDim.x) sT[q] = iT[q];
}

// Run end for every member: its status words [down_word, total_words) and, for AMIS, the
// fitted final slot (n64 FP64 values) into its row of the gather buffer.
__global__ __launch_bounds__(256) void k_group_end(const GroupMember* __restrict__ mem, uint64_t* __restrict__ gather,
                                                   int down_word, int total_words, int n64, int row_words) {
  const int m = blockIdx.x;
  const GroupMember g = mem[m];
  uint64_t* out = gather + (size_t)m * row_words;
  const int nd = total_words - down_word;
  for (int q = threadIdx.x; q < nd; q += blockDim.x) out[q] = g.ctl[down_word + q];
  if (g.slotT_64) {
    const uint64_t* src = reinterpret_cast<const uint64_t*>(g.slotT_64);
    for (int q = threadIdx.x; q < n64; q += blockDim.x) out[nd + q] = src[q];
  }
}

struct GroupRound {
  int nb_new = 0, nrec_old = 0, nt = 0, nv = 2;
  int64_t nrec = 0;
  size_t lds = 0;
};

}  // namespace

struct is_group {
  is_ctx* ctx = nullptr;
  std::vector<is_job*> jobs;
  int method = IS_PLAIN, precision = IS_FP64, d = 0, D = 0, T = 1, student = 0;
  hipStream_t stream = nullptr;
  std::vector<GroupRound> rounds;
  DevBuf round_args, reduce_args, members, stage, gather;
  HostPinned stage_up, gather_down;
  size_t round_bytes = 0, sweep_bytes = 0;
  int up_words = 0, down_word = 0, total_words = 0, n64 = 0, tbytes = 0, row_words = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::atomic<bool> busy{false};
  bool ran = false;  // a run has completed (status images hold its values)
  ~is_group() {
    if (stream) (void)hipStreamSynchronize(stream);
    CaptureGuard cg(capture_lock());
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

template <class T>
int group_prepare(is_group* g) {
  const int M = (int)g->jobs.size();
  is_job* j0 = g->jobs[0];
  g->round_bytes = sizeof(RoundArgs<T>);
  g->sweep_bytes = sizeof(SweepArgs<T>);
  g->rounds.assign(g->T, GroupRound{});
  std::vector<char> rargs((size_t)g->T * M * (g->round_bytes + g->sweep_bytes));
  std::vector<GroupReduce> red((size_t)g->T * M);
  for (int t = 0; t < g->T; ++t) {
    for (int m = 0; m < M; ++m) {
      is_job* j = g->jobs[m];
      RoundPlan<T> p = round_plan<T>(j, t, j->counter.as<unsigned int>());
      GroupRound& r = g->rounds[t];
      p.a.slot_chunk = round_slot_chunk(g->D, p.nt, p.a.stride, (int)sizeof(T), p.a.tkind == kTGaussian, p.ncomp, &r.lds);
      r.nb_new = p.nb_new;
      r.nrec_old = p.nrec_old;
      r.nt = p.nt;
      r.nv = p.nv;
      r.nrec = (int64_t)p.nrec_old + p.nb_new;
      char* base = rargs.data() + (size_t)t * M * (g->round_bytes + g->sweep_bytes);
      std::memcpy(base + (size_t)m * g->round_bytes, &p.a, g->round_bytes);
      std::memcpy(base + (size_t)M * g->round_bytes + (size_t)m * g->sweep_bytes, &p.wo, g->sweep_bytes);
      GroupReduce& gr = red[(size_t)t * M + m];
      gr.partials = j->partials.as<double>();
      gr.reduced = j->reduced.as<double>();
      gr.f = p.f;
    }
  }
  int rc;
  if ((rc = g->round_args.alloc(rargs.size())) || (rc = g->reduce_args.alloc(red.size() * sizeof(GroupReduce))))
    return rc;
  HIP_TRY(locked_memcpy(g->round_args.p, rargs.data(), rargs.size(), hipMemcpyHostToDevice));
  HIP_TRY(locked_memcpy(g->reduce_args.p, red.data(), red.size() * sizeof(GroupReduce), hipMemcpyHostToDevice));
  // member table, staging and gather layouts
  std::vector<GroupMember> mem(M);
  for (int m = 0; m < M; ++m) {
    is_job* j = g->jobs[m];
    mem[m].ctl = j->ctl.as<uint64_t>();
    mem[m].slot0_64 = j->slots64.as<double>();
    mem[m].slot0_t = j->slotsT.p;
    mem[m].slotT_64 = j->method == IS_AMIS ? j->slots64.as<double>() + (size_t)j->T * j->stride : nullptr;
  }
  if ((rc = g->members.alloc(mem.size() * sizeof(GroupMember)))) return rc;
  HIP_TRY(locked_memcpy(g->members.p, mem.data(), mem.size() * sizeof(GroupMember), hipMemcpyHostToDevice));
  g->up_words = (int)(j0->ctl_up_bytes / 8);
  g->down_word = (int)(j0->ctl_down_off / 8);
  g->total_words = (int)(j0->ctl.bytes / 8);
  g->n64 = j0->stride;
  g->tbytes = j0->stride * (int)sizeof(T);
  g->row_words = (g->total_words - g->down_word) + (j0->method == IS_AMIS ? g->n64 : 0);
  const size_t stage_words = (size_t)M + (size_t)M * g->n64 + (size_t)M * ((g->tbytes + 7) / 8);
  if ((rc = g->stage.alloc(stage_words * 8)) || (rc = g->stage_up.alloc(stage_words * 8)) ||
      (rc = g->gather.alloc((size_t)M * g->row_words * 8)) || (rc = g->gather_down.alloc((size_t)M * g->row_words * 8)))
    return rc;
  std::memset(g->stage_up.p, 0, stage_words * 8);
  return IS_OK;
}

template <class T>
int group_issue(is_group* g) {
  const int M = (int)g->jobs.size();
  hipStream_t s = g->stream;
  hipLaunchKernelGGL(k_group_begin, dim3(M), dim3(256), 0, s, g->members.as<GroupMember>(), g->stage.as<uint64_t>(), M,
                     g->up_words, g->n64, g->tbytes);
  const size_t per_round = (size_t)M * (g->round_bytes + g->sweep_bytes);
  for (int t = 0; t < g->T; ++t) {
    const GroupRound& r = g->rounds[t];
    const char* base = g->round_args.as<char>() + (size_t)t * per_round;
    const RoundArgs<T>* ga = reinterpret_cast<const RoundArgs<T>*>(base);
    const SweepArgs<T>* gw = reinterpret_cast<const SweepArgs<T>*>(base + (size_t)M * g->round_bytes);
    bool done = false;
    with_dim(g->D, [&](auto k) {
      done = DimKernels<decltype(k)::value>::round_group(r.nt, g->student != 0, ga, gw, M, r.nb_new, r.nrec_old, r.lds,
                                                          s);
    });
    if (!done) return fail(IS_EUNSUPPORTED, "no fused round for dimension %d", g->d);
    const GroupReduce* gr = g->reduce_args.as<GroupReduce>() + (size_t)t * M;
    const int64_t pcap = g->jobs[0]->pcap;
    switch (g->d) {
#define IS_GROUP_REDUCE(v)                                                                                     \
  case v:                                                                                                      \
    hipLaunchKernelGGL((k_reduce_finish_group<v, true>), dim3(r.nv, M), dim3(kBlock), 0, s, gr, r.nrec, pcap, r.nv); \
    break;
      IS_EXACT_DIMS(IS_GROUP_REDUCE)
#undef IS_GROUP_REDUCE
      default:
        return fail(IS_EUNSUPPORTED, "no grouped reduction for dimension %d", g->d);
    }
  }
  hipLaunchKernelGGL(k_group_end, dim3(M), dim3(256), 0, s, g->members.as<GroupMember>(), g->gather.as<uint64_t>(),
                     g->down_word, g->total_words, g->n64, g->row_words);
  HIP_TRY(hipGetLastError());
  return IS_OK;
}

template <class T>
int group_run_typed(is_group* g) {
  const int M = (int)g->jobs.size();
  hipStream_t s = g->stream;
  // stage the members' run words and committed slot images
  uint64_t* st = g->stage_up.as<uint64_t>();
  const int twords = (g->tbytes + 7) / 8;
  for (int m = 0; m < M; ++m) {
    is_job* j = g->jobs[m];
    st[m] = (uint64_t)j->run_index << 32;
    double* h64 = reinterpret_cast<double*>(st + M + (size_t)m * g->n64);
    T* hT = reinterpret_cast<T*>(st + M + (size_t)M * g->n64 + (size_t)m * twords);
    pack_committed<T>(j, h64, hT);
  }
  HIP_TRY(hipMemcpyAsync(g->stage.p, g->stage_up.p, g->stage_up.bytes, hipMemcpyHostToDevice, s));
  if (!g->graph_exec) {
    CaptureGuard cg(capture_lock());
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = group_issue<T>(g);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &gr);
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return fail(IS_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    g->graph = gr;
    HIP_TRY(hipGraphInstantiate(&g->graph_exec, gr, nullptr, nullptr, 0));
  }
  HIP_TRY(hipGraphLaunch(g->graph_exec, s));
  HIP_TRY(hipMemcpyAsync(g->gather_down.p, g->gather.p, g->gather_down.bytes, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  // each member's outcome from its status image, as a stand-alone run's
  g->ran = true;
  int result = IS_OK;
  const size_t down_bytes = (size_t)(g->total_words - g->down_word) * 8;
  for (int m = 0; m < M; ++m) {
    is_job* j = g->jobs[m];
    const char* row = g->gather_down.as<const char>() + (size_t)m * g->row_words * 8;
    std::memcpy(j->ctl_down.as<char>() + j->ctl_down_off, row, down_bytes);
    if (j->method == IS_AMIS) std::memcpy(j->slot_down.p, row + down_bytes, (size_t)g->n64 * 8);
    j->slots_resident = false;
    const int rc = run_epilogue<T>(j);
    if (rc == IS_EFAILED)
      result = IS_EFAILED;
    else if (rc != IS_OK)
      return rc;
  }
  return result;
}

}  // namespace

extern "C" {

int is_group_create(is_ctx* ctx, const is_job_desc* descs, int32_t count, is_group** out) {
  if (!ctx || !descs || !out || count < 1) return fail(IS_EINVAL, "bad arguments");
  const is_job_desc& a = descs[0];
  for (int32_t m = 0; m < count; ++m) {
    const is_job_desc& b = descs[m];
    if (b.method != IS_PLAIN && b.method != IS_AMIS)
      return fail(IS_EUNSUPPORTED, "group member %d: groups run plain IS and AMIS", m);
    if (b.target_kind != IS_TARGET_STANDARD && b.target_kind != IS_TARGET_GAUSSIAN)
      return fail(IS_EUNSUPPORTED, "group member %d: groups take the built-in standard and Gaussian targets", m);
    if (b.transform_kinds || b.transform_block_count)
      return fail(IS_EUNSUPPORTED, "group member %d: groups run without transforms", m);
    if (b.dim < 1 || b.dim > 15) return fail(IS_EUNSUPPORTED, "group member %d: groups take d <= 15", m);
    const int Tb = b.method == IS_PLAIN ? 1 : b.rounds;
    const int Ta = a.method == IS_PLAIN ? 1 : a.rounds;
    if (b.method != a.method || b.dim != a.dim || b.family != a.family || b.precision != a.precision || Tb != Ta ||
        !b.round_size)
      return fail(IS_EINVAL, "group member %d: members share method, dimension, family, precision and rounds", m);
    for (int t = 0; t < Tb; ++t)
      if (b.round_size[t] != a.round_size[t]) return fail(IS_EINVAL, "group member %d: round sizes differ", m);
  }
  std::unique_ptr<is_group> g(new is_group());
  g->ctx = ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  for (int32_t m = 0; m < count; ++m) {
    is_job* j = nullptr;
    const int rc = is_job_create(ctx, &descs[m], &j);
    if (rc) {
      for (is_job* q : g->jobs) is_job_destroy(q);
      return rc;
    }
    g->jobs.push_back(j);
  }
  is_job* j0 = g->jobs[0];
  g->method = j0->method;
  g->precision = j0->precision;
  g->d = j0->d;
  g->D = j0->D;
  g->T = j0->T;
  g->student = j0->family == IS_STUDENT_T;
  {
    CaptureGuard cg(capture_lock());
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
      for (is_job* q : g->jobs) is_job_destroy(q);
      return fail(IS_EDEVICE, "hipStreamCreate failed");
    }
  }
  const int rc = g->precision == IS_FP32 ? group_prepare<float>(g.get()) : group_prepare<double>(g.get());
  if (rc) {
    for (is_job* q : g->jobs) is_job_destroy(q);
    return rc;
  }
  *out = g.release();
  return IS_OK;
}

int is_group_run(is_group* group) {
  if (!group) return fail(IS_EINVAL, "null group");
  bool expected = false;
  if (!group->busy.compare_exchange_strong(expected, true)) return fail(IS_EBUSY, "group is already running");
  std::vector<std::unique_ptr<BusyGuard>> guards;
  for (is_job* j : group->jobs) {
    guards.emplace_back(new BusyGuard(j));
    if (!guards.back()->ok) {
      guards.clear();
      group->busy = false;
      return fail(IS_EBUSY, "a member is running on its own");
    }
  }
  (void)hipSetDevice(group->ctx->device);
  const int rc = group->precision == IS_FP32 ? group_run_typed<float>(group) : group_run_typed<double>(group);
  guards.clear();
  group->busy = false;
  return rc;
}

int32_t is_group_size(const is_group* group) { return group ? (int32_t)group->jobs.size() : 0; }

int is_group_reset(is_group* group) {
  if (!group) return fail(IS_EINVAL, "null group");
  if (group->busy) return fail(IS_EBUSY, "group is already running");
  for (is_job* j : group->jobs) {
    const int rc = is_job_reset(j);
    if (rc) return rc;
  }
  return IS_OK;
}

int is_group_round_summary(is_group* group, int32_t round, double* ess, double* lognormalizers) {
  if (!group || round < 0 || round >= group->T) return fail(IS_EINVAL, "bad arguments");
  if (!group->ran) return fail(IS_ESTATE, "the group has not run yet");
  for (size_t m = 0; m < group->jobs.size(); ++m) {
    const is_job* j = group->jobs[m];
    // the member's status image of the group's last run (copied in by group_run_typed)
    const char* down = j->ctl_down.as<const char>();
    auto at = [&](const DevView& v) {
      return reinterpret_cast<const double*>(down + ((const char*)v.p - j->ctl.as<char>())) + round;
    };
    if (ess) ess[m] = *at(j->round_ess);
    if (lognormalizers) lognormalizers[m] = *at(j->round_logz) + j->t_shift;
  }
  return IS_OK;
}

int is_group_job(is_group* group, int32_t index, is_job** out) {
  if (!group || !out || index < 0 || index >= (int32_t)group->jobs.size()) return fail(IS_EINVAL, "bad arguments");
  *out = group->jobs[index];
  return IS_OK;
}

int is_group_destroy(is_group* group) {
  if (!group) return IS_OK;
  if (group->stream) (void)hipStreamSynchronize(group->stream);
  for (is_job* j : group->jobs) is_job_destroy(j);
  group->jobs.clear();
  delete group;
  return IS_OK;
}

}  // extern "C"
