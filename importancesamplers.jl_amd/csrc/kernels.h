// gfx950 kernels of the importance-sampling engine.
//
// Data layout in HBM (per job):
//   samples      SoA, coordinate-major: X[c * cap + i]  (T = double or float)
//   logtargets, lognumerators, logweights: length cap (T)
//   round_ids, proposal_ids: int32, length cap
//   slots        packed proposal slots (see device.h), FP64 master copy and a T copy
//   partials     one record per block: [m, s0, s2, s1(d), S(d(d+1)/2)] (FP64)
//   fail         3 x u64 failure record per round
#pragma once

#include "device.h"

namespace issampler {

constexpr int kBlock = 256;
constexpr int kTile = 64;  // samples per LDS moment sub-tile
constexpr int kMaxDim = 64;

enum TargetKind { kTStandard = 0, kTGaussian = 1, kTLogistic = 2, kTLinreg = 3 };
enum RoundMode { kDrawFused = 0, kDrawOnly = 1, kPreTarget = 2 };

template <class T>
struct RoundArgs {
  int d;
  int round;          // 0-based round index (provenance)
  const unsigned long long* run_hi;  // device word holding (run index << 32); RNG counter = run_hi | round
  int64_t first;      // global index of the first new sample
  int64_t n_new;
  int64_t n_old;      // retained samples [0, n_old) updated by the append path
  int old_blocks;     // blocks [0, old_blocks) run the append path
  // proposal slots
  const T* slots;     // slot table (T copy)
  int stride;         // slot stride in elements
  int k0, k1;         // mixture components [k0, k1] for new samples
  int slot_chunk;     // components staged in LDS at a time
  int append_slot;    // component appended to retained samples
  const T* logmass;   // per-slot log mixture mass (indexed by slot)
  int gen_slot;       // generating slot when assign == nullptr
  const int32_t* assign;    // per-new-sample generating slot (DM-PMC, static MIS)
  const int32_t* bank_ids;  // per-slot proposal id (DM-PMC, static MIS)
  // per-sample mixture terms (static MIS): 0 the slots [k0, k1] for every sample;
  // 1 the generating slot alone (log coefficient 0); 2 the members of the
  // generating slot's group, mis_gmem[mis_goff[g] .. mis_goff[g + 1]) with log
  // coefficients mis_gcoef, g = mis_slot_group[generating slot]
  int mis_terms;
  const int32_t* mis_slot_group;
  const int32_t* mis_goff;
  const int32_t* mis_gmem;
  const T* mis_gcoef;
  T logtotal;
  Family<T> fam;
  double dof;
  // randomness (prefilled buffers when non-null; otherwise Philox)
  RngKey key;
  const double* pre_normals;
  int nstride;
  const double* pre_rn;
  int rn_stride;
  const double* pre_ru;
  int ru_stride;
  int rs, us;         // radial normal / uniform counts per sample
  int record_nonfinite_after_exhaust;
  // per-coordinate transforms (null: none) and the logical-sample store
  const int32_t* tf_kinds;
  const double* tf_bounds;
  T* XL;
  int tf_carry;       // the pre-target pass reads the forward log Jacobian from lognums
  // target
  int tkind;
  const T* tslot;           // Gaussian target slot
  const T* pre_logtarget;   // batch-target values for new samples (kPreTarget)
  const T* pre_mix;         // mixture log densities of the new samples (k_mix_mfma), or null
  // storage
  T* X;
  int64_t cap;
  T* logtargets;
  T* lognums;
  T* logweights;
  int32_t* round_ids;
  int32_t* proposal_ids;
  unsigned long long* fail;
  // partials
  double* partials;
  int nv;             // values per record (excluding the max)
  int moments;        // accumulate s1 / S about `center`
  const double* center;
};

// packed lower pair index p -> (r, c), r >= c
__device__ inline void pair_index(int p, int& r, int& c) {
  r = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
  while ((r + 1) * (r + 2) / 2 <= p) ++r;
  while (r * (r + 1) / 2 > p) --r;
  c = p - r * (r + 1) / 2;
}

template <class T, int D, bool RANGE = false, class P>
__device__ __forceinline__ T inline_target(int tkind, int d, P tslot, const T (&x)[D], T (&y)[D]) {
  if (tkind == kTStandard) {
    T s = T(0);
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (c < d) s += x[c] * x[c];
    return -s / T(2);
  }
  // Gaussian target: lognorm - r^2/2 with r = L^{-1}(x - m)
  Family<T> g;
  g.student = 0;
  g.dof = T(0);
  g.half_dof_plus_d = T(0);
  return slot_logdensity<T, D, RANGE>(tslot, d, g, x, y);
}

#include "mix.h"
#include "target.h"
#include "mix_mfma.h"
#include "layout.h"

// ---------------------------------------------------------------- partial reduction + finish
struct FinishArgs {
  int d;
  int flags;               // kFinish* bits
  int round;               // 0-based round index
  int64_t summary_count;   // n for log Z = M + log s0 - log n
  // fit
  const double* center;    // d
  const double* prev_slot; // FP64 slot of the previous proposal (ridge source)
  double* next_slot64;     // FP64 slot to write
  void* next_slot_t;       // T copy (float or double) to write
  int t_is_float;
  int student;
  double dof;
  double logconstant;
  double* round_ess;
  double* round_logz;
  double* adapt_ess;
  double* summary_out;     // [M, s0, s2, ess, logz]
  int* fit_status;         // per round: 0 ok, 1 all-zero weights, 2 not positive definite, 3 invalid
  unsigned int* counter;
};
constexpr int kFinishSummary = 1;
constexpr int kFinishFit = 2;
constexpr int kFinishAdaptEss = 4;


// grid = nv blocks; block k reduces value k over all partial records.  Partial
// records are stored column-major: column 0 holds the block maxima, column
// 1 + v value v, each column `pcap` long, so every read here is coalesced.
// Each thread keeps its own running maximum and rescaled sum over batches of
// kReduceUnroll records whose loads are all issued before any is used, so up
// to kBlock * kReduceUnroll records cost one memory round trip; the threads
// are then merged in a fixed order (deterministic).
// With DR > 0 the block that finishes last (a ticket on f.counter, which it
// returns to zero) also runs the summary and fit of k_finish_small<DR> in its
// first wave, so a round needs no separate finish launch.
constexpr int kReduceUnroll = 8;
__device__ inline void finish_summary(double M, double s0, double s2, const FinishArgs& f);
template <int DR>
struct FitPrefetch;
template <int DR, bool EXACT>
__device__ __forceinline__ void fit_prefetch(const FinishArgs& f, FitPrefetch<DR>& pf);
template <int DR, bool EXACT>
__device__ __forceinline__ void fit_small(const double* __restrict__ red, const FinishArgs& f,
                                          const FitPrefetch<DR>& pf, bool summary);

template <int DR, bool EXACT>
__device__ __forceinline__ void reduce_finish_block(const double* __restrict__ partials, int64_t nrec, int64_t pcap,
                                                    int nv, double* __restrict__ reduced, const FinishArgs& f) {
  __shared__ double wm[kBlock / 64], ws[kBlock / 64];
  const int tid = threadIdx.x;
  const int k = blockIdx.x;
  const bool squared = (k == 1);  // value 1 is the sum of squared weights
  const double ninf = -__builtin_huge_val();
  const double* maxima = partials;
  const double* column = partials + (int64_t)(1 + k) * pcap;
  // the fit's inputs from earlier launches, fetched now in case this block is last
  FitPrefetch<(DR > 0 ? DR : 1)> pf;
  if constexpr (DR > 0) {
    if (tid < 64) fit_prefetch<DR, EXACT>(f, pf);
  }
  double m = ninf, s = 0.0;
  for (int64_t base = 0; base < nrec; base += (int64_t)kBlock * kReduceUnroll) {
    double mv[kReduceUnroll], cv[kReduceUnroll];
#pragma unroll
    for (int u = 0; u < kReduceUnroll; ++u) {
      const int64_t r = base + (int64_t)u * kBlock + tid;
      mv[u] = r < nrec ? maxima[r] : ninf;
      cv[u] = r < nrec ? column[r] : 0.0;
    }
    double mb = m;
#pragma unroll
    for (int u = 0; u < kReduceUnroll; ++u) mb = fmax(mb, mv[u]);
    if (mb == ninf) continue;
    double sc = (m == ninf) ? 0.0 : exp(m - mb);
    s *= squared ? sc * sc : sc;
#pragma unroll
    for (int u = 0; u < kReduceUnroll; ++u) {
      if (mv[u] == ninf) continue;
      double e = exp(mv[u] - mb);
      if (squared) e *= e;
      s += cv[u] * e;
    }
    m = mb;
  }
  // merge the threads: block maximum, then rescaled sums in thread order
  double M = wave_max(m);
  if ((tid & 63) == 0) wm[tid >> 6] = M;
  __syncthreads();
  M = wm[0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) M = fmax(M, wm[w]);
  double v = 0.0;
  if (m != ninf) {
    const double sc = exp(m - M);
    v = s * (squared ? sc * sc : sc);
  }
  v = wave_sum(v);
  if ((tid & 63) == 0) ws[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) tot += ws[w];
    reduced[1 + k] = tot;
    if (k == 0) reduced[0] = M;
  }
  if constexpr (DR > 0) {
    __shared__ int last;
    if (tid == 0) {
      // release: this block's values before the ticket; acquire: every earlier block's
      const unsigned int t = __hip_atomic_fetch_add(f.counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      last = t == gridDim.x - 1;
      if (last) __hip_atomic_store(f.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (last) {
      __atomic_thread_fence(__ATOMIC_ACQUIRE);
      // wave 0 fits; the summary runs beside it on wave 1
      if (tid < 64)
        fit_small<DR, EXACT>(reduced, f, pf, false);
      else if (tid == 64 && (f.flags & (kFinishSummary | kFinishAdaptEss)))
        finish_summary(reduced[0], reduced[1], reduced[2], f);
    }
  } else {
    (void)f;
  }
}

template <int DR, bool EXACT>
static __global__ __launch_bounds__(kBlock) void k_reduce_finish(const double* __restrict__ partials,
                                                          int64_t nrec, int64_t pcap, int nv,
                                                          double* __restrict__ reduced, FinishArgs f) {
  reduce_finish_block<DR, EXACT>(partials, nrec, pcap, nv, reduced, f);
}

// One job's reduction in a group launch (grid nv x jobs).
struct GroupReduce {
  const double* partials;
  double* reduced;
  FinishArgs f;
};
template <int DR, bool EXACT>
static __global__ __launch_bounds__(kBlock) void k_reduce_finish_group(const GroupReduce* __restrict__ g, int64_t nrec,
                                                                int64_t pcap, int nv) {
  const GroupReduce r = g[blockIdx.y];
  reduce_finish_block<DR, EXACT>(r.partials, nrec, pcap, nv, r.reduced, r.f);
}

// Weighted functional partial records for Statistics.mean(f, r) and var(f, r)
// (src/statistics.jl:17-31,50-64,170-199) over device arrays: block b covers
// samples [b kFuncPerBlock, ...) and writes the column-major record
// [M, sum e, sum e^2, sum e g], e = exp(lw - M), with g = f (pass 0: zero-weight
// samples enter as 0 f, so a non-finite f there gives NaN, as dot(weights,
// values) does) or, when `center_rec` is given (pass 1), the squared deviation
// about the centre center_rec[3] / center_rec[1], zero-weight samples skipped and
// a non-finite deviation recomputed as (sqrt(e) f - sqrt(e) c)^2.  A block whose
// weights are all zero stores the finite maximum -DBL_MAX so the merge keeps its
// sums (and any NaN).  k_reduce_finish<0> merges the records (nv = 3).
constexpr int kFuncPerThread = 8;
constexpr int64_t kFuncPerBlock = (int64_t)kBlock * kFuncPerThread;
template <class T, class V>
static __global__ __launch_bounds__(kBlock) void k_functional(const T* __restrict__ lw, const V* __restrict__ val,
                                                              int64_t n, const double* __restrict__ center_rec,
                                                              double* __restrict__ partials, int64_t pcap) {
  __shared__ double wm[kBlock / 64], w0[kBlock / 64], w2[kBlock / 64], w1[kBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double ninf = -__builtin_huge_val();
  const bool dev_pass = center_rec != nullptr;
  const double c = dev_pass ? center_rec[3] / center_rec[1] : 0.0;
  const int64_t base = (int64_t)blockIdx.x * kFuncPerBlock;
  double l[kFuncPerThread], v[kFuncPerThread];
#pragma unroll
  for (int u = 0; u < kFuncPerThread; ++u) {
    const int64_t i = base + (int64_t)u * kBlock + tid;
    l[u] = i < n ? (double)lw[i] : ninf;
    v[u] = i < n ? (double)val[i] : 0.0;
  }
  double m = ninf;
#pragma unroll
  for (int u = 0; u < kFuncPerThread; ++u) m = fmax(m, l[u]);
  double s0 = 0.0, s2 = 0.0, s1 = 0.0;
#pragma unroll
  for (int u = 0; u < kFuncPerThread; ++u) {
    const double e = (l[u] == ninf) ? 0.0 : exp(l[u] - m);
    s0 += e;
    s2 += e * e;
    if (!dev_pass) {
      s1 += e * v[u];
    } else if (e != 0.0) {
      const double dv = v[u] - c;
      double t;
      if (isfinite(dv)) {
        t = dv * (e * dv);
      } else {
        const double r = sqrt(e);
        const double q = r * v[u] - r * c;
        t = q * q;
      }
      s1 += t;
    }
  }
  // block merge: the block maximum, each thread rescaled to it, sums in thread order
  double M = wave_max(m);
  if (lane == 0) wm[wave] = M;
  __syncthreads();
  M = wm[0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) M = fmax(M, wm[w]);
  const double sc = (m == ninf) ? 0.0 : exp(m - M);
  s0 = wave_sum(s0 * sc);
  s2 = wave_sum(s2 * sc * sc);
  s1 = wave_sum(s1 * sc);
  if (lane == 0) {
    w0[wave] = s0;
    w2[wave] = s2;
    w1[wave] = s1;
  }
  __syncthreads();
  if (tid == 0) {
    double t0 = 0.0, t2 = 0.0, t1 = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
      t0 += w0[w];
      t2 += w2[w];
      t1 += w1[w];
    }
    double* out = partials + blockIdx.x;
    out[0] = (M == ninf) ? -1.7976931348623157e308 : M;
    out[1 * pcap] = t0;
    out[2 * pcap] = t2;
    out[3 * pcap] = t1;
  }
}

// Summary (src/execution.jl:176-186): ESS = s0^2 / s2, log Z = M + log s0 - log n.
__device__ inline void finish_summary(double M, double s0, double s2, const FinishArgs& f) {
  double ess, logz;
  if (s0 == 0.0 || M == -__builtin_huge_val()) {
    ess = 0.0;
    logz = -__builtin_huge_val();
  } else {
    ess = (s0 * s0) / s2;
    logz = M + log(s0) - log((double)f.summary_count);
  }
  if (f.flags & kFinishSummary) {
    if (f.round_ess) f.round_ess[f.round] = ess;
    if (f.round_logz) f.round_logz[f.round] = logz;
    if (f.summary_out) {
      f.summary_out[0] = M;
      f.summary_out[1] = s0;
      f.summary_out[2] = s2;
      f.summary_out[3] = ess;
      f.summary_out[4] = logz;
    }
  }
  if ((f.flags & kFinishAdaptEss) && f.adapt_ess) f.adapt_ess[f.round] = ess;
}

__device__ inline double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

constexpr int kFinishStride = kMaxDim + 1;  // LDS row stride of the covariance

// One wavefront: summary (src/execution.jl:176-186) and FP64 moment fit
// (src/adaptive_gaussian_execution.jl:303-425): mean = c + s1/s0,
// cov = S/s0 - (s1/s0)(s1/s0)^T + ridge I, Cholesky, Student-t rescale,
// log-det.  Lane r owns row r (d <= 64).  All global inputs are fetched in one
// batch into LDS; the Cholesky is the left-looking column order of
// _gaussian_potrf! (src/adaptive_gaussian_execution.jl:427-434) with the rows
// of each column computed in parallel, so a fit is d dependent steps of
// (dot product, sqrt, divide) and no workgroup barrier.
static __global__ __launch_bounds__(64) void k_finish(const double* __restrict__ red, FinishArgs f) {
  __shared__ double A[kMaxDim * kFinishStride];
  __shared__ double meanv[kMaxDim];
  const int lane = threadIdx.x;
  const int d = f.d;
  const double M = red[0], s0 = red[1], s2 = red[2];
  if (lane == 0 && (f.flags & (kFinishSummary | kFinishAdaptEss))) finish_summary(M, s0, s2, f);
  if (!(f.flags & kFinishFit)) return;
  if (!(s0 > 0.0) || !isfinite(s0) || M == -__builtin_huge_val()) {
    if (lane == 0) f.fit_status[f.round] = 1;
    return;
  }
  const bool row = lane < d;
  const int npk = d * (d + 1) / 2;
  // ---- one batch of loads: S pairs and the previous factor (ridge), s1 and the centre
  double tr = 0.0;
  for (int p = lane; p < npk; p += 64) {
    int r, c;
    pair_index(p, r, c);
    const double l = f.prev_slot[2 * d + p];
    A[r * kFinishStride + c] = red[3 + d + p];
    tr += l * l;
  }
  const double s1 = row ? red[3 + lane] : 0.0;
  const double cr = row ? f.center[lane] : 0.0;
  const double mean = s1 / s0;
  meanv[lane] = mean;
  const double mult = f.student ? f.dof / (f.dof - 2.0) : 1.0;
  // scale-aware ridge (src/adaptive_gaussian_execution.jl:283-301): sqrt(eps) tr(L L^T) mult / d
  const double reg = 1.4901161193847656e-08;  // sqrt(eps(Float64))
  double ridge = reg * (wave_sum(tr) * mult) / (double)d;
  if (!isfinite(ridge)) {
    const double sc = sqrt(reg) * sqrt(mult / (double)d);
    double part = 0.0;
    for (int p = lane; p < npk; p += 64) {
      const double v = f.prev_slot[2 * d + p] * sc;
      part += v * v;
    }
    ridge = wave_sum(part);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // ---- centred covariance, row `lane`
  if (row) {
    double* ar = A + lane * kFinishStride;
    for (int c = 0; c <= lane; ++c) {
      double v = ar[c] / s0 - mean * meanv[c];
      if (c == lane) v += ridge;
      ar[c] = v;
    }
  }
  // ---- Cholesky, column by column
  int status = 0;
  for (int c = 0; c < d; ++c) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double v = 0.0;
    if (lane >= c && row) {
      const double* ar = A + lane * kFinishStride;
      const double* ac = A + c * kFinishStride;
      v = ar[c];
      int k = 0;
      for (; k + 4 <= c; k += 4) {
        const double a0 = ar[k], a1 = ar[k + 1], a2 = ar[k + 2], a3 = ar[k + 3];
        const double b0 = ac[k], b1 = ac[k + 1], b2 = ac[k + 2], b3 = ac[k + 3];
        v -= a0 * b0;
        v -= a1 * b1;
        v -= a2 * b2;
        v -= a3 * b3;
      }
      for (; k < c; ++k) v -= ar[k] * ac[k];
    }
    const double piv = readlane_f64(v, c);
    if (!(piv > 0.0) || !isfinite(piv)) status = 2;
    const double dg = sqrt(piv);
    if (row && lane > c) A[lane * kFinishStride + c] = v / dg;
    if (lane == c) A[c * kFinishStride + c] = dg;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // ---- outputs: lane r writes mean r, 1 / L_rr and row r of L (Student-t rescaled)
  const double scale = f.student ? sqrt(1.0 / mult) : 1.0;
  bool ok = true;
  double lg = 0.0, diag = 1.0;
  double* out = f.next_slot64;
  const bool tf = f.t_is_float != 0;
  float* of = reinterpret_cast<float*>(f.next_slot_t);
  double* od = reinterpret_cast<double*>(f.next_slot_t);
  if (row) {
    const double* ar = A + lane * kFinishStride;
    const double mu = cr + mean;
    diag = ar[lane] * scale;
    ok = isfinite(mu) && diag > 0.0;
    lg = log(fabs(diag));
    out[lane] = mu;
    out[d + lane] = slot_inv_entry<double>(diag);
    if (tf) {
      of[lane] = (float)mu;
      of[d + lane] = slot_inv_entry<float>(diag);
    } else {
      od[lane] = mu;
      od[d + lane] = slot_inv_entry<double>(diag);
    }
    const int base = 2 * d + lane * (lane + 1) / 2;
    for (int c = 0; c <= lane; ++c) {
      const double v = ar[c] * scale;
      ok &= isfinite(v);
      out[base + c] = v;
      if (tf)
        of[base + c] = (float)v;
      else
        od[base + c] = v;
    }
  }
  const double logabsdet = wave_sum(lg);
  const bool all_ok = __all(ok);
  if (lane == 0) {
    const double lognorm = f.student ? f.logconstant - logabsdet
                                     : -0.5 * (double)d * log(2.0 * 3.141592653589793) - logabsdet;
    int st = status;
    if (st == 0 && (!all_ok || !isfinite(lognorm))) st = 3;
    out[slot_lognorm_off(d)] = lognorm;
    if (tf)
      of[slot_lognorm_off(d)] = (float)lognorm;
    else
      od[slot_lognorm_off(d)] = lognorm;
    f.fit_status[f.round] = st;
  }
}

// The same fit for d <= DR with the covariance held in registers: lane r keeps
// row r and the Cholesky takes row c of the factor from lane c by readlane, so
// no LDS round trip or wave barrier sits between the d column steps.  Each
// column step scales by 1 / sqrt(pivot) from the hardware reciprocal square
// root and two Newton steps (relative error a few ulp) instead of a square root
// and a division; the covariance scales by 1 / s0 once.  The previous factor's
// row and the centre (written by earlier launches) are fetched by fit_prefetch
// before the reduced record is available.
template <int DR>
struct FitPrefetch {
  double l[DR];
  double cr;
};

template <int DR, bool EXACT>
__device__ __forceinline__ void fit_prefetch(const FinishArgs& f, FitPrefetch<DR>& pf) {
  const int lane = threadIdx.x & 63;
  const int d = EXACT ? DR : f.d;
  const bool row = lane < d && (f.flags & kFinishFit) != 0;
  const int rl = 2 * d + lane * (lane + 1) / 2;
#pragma unroll
  for (int c = 0; c < DR; ++c) pf.l[c] = (row && c <= lane) ? f.prev_slot[rl + c] : 0.0;
  pf.cr = row ? f.center[lane] : 0.0;
}

__device__ __forceinline__ double rsqrt_nr(double v) {
  double r = __builtin_amdgcn_rsq(v);
  const double h = 0.5 * v;
  r = r * fma(-h * r, r, 1.5);
  r = r * fma(-h * r, r, 1.5);
  return r;
}

template <int DR, bool EXACT>
__device__ __forceinline__ void fit_small(const double* __restrict__ red, const FinishArgs& f,
                                          const FitPrefetch<DR>& pf, bool summary) {
  const int lane = threadIdx.x & 63;
  const int d = EXACT ? DR : f.d;
  const bool row = lane < d;
  const bool fit = (f.flags & kFinishFit) != 0;
  // ---- one batch of loads: summary, row `lane` of S, s1
  const double M = red[0], s0 = red[1], s2 = red[2];
  double a[DR];
  const int rs = 3 + d + lane * (lane + 1) / 2;
#pragma unroll
  for (int c = 0; c < DR; ++c) a[c] = (fit && row && c <= lane) ? red[rs + c] : 0.0;
  const double s1 = (fit && row) ? red[3 + lane] : 0.0;
  if (summary && lane == 0 && (f.flags & (kFinishSummary | kFinishAdaptEss))) finish_summary(M, s0, s2, f);
  if (!fit) return;
  if (!(s0 > 0.0) || !isfinite(s0) || M == -__builtin_huge_val()) {
    if (lane == 0) f.fit_status[f.round] = 1;
    return;
  }
  const double is0 = 1.0 / s0;
  const double mean = s1 * is0;
  const double mult = f.student ? f.dof / (f.dof - 2.0) : 1.0;
  // scale-aware ridge (src/adaptive_gaussian_execution.jl:283-301): sqrt(eps) tr(L L^T) mult / d
  const double reg = 1.4901161193847656e-08;  // sqrt(eps(Float64))
  double tr = 0.0;
#pragma unroll
  for (int c = 0; c < DR; ++c) tr += pf.l[c] * pf.l[c];
  double ridge = reg * (wave_sum(tr) * mult) / (double)d;
  if (!isfinite(ridge)) {
    const double sc = sqrt(reg) * sqrt(mult / (double)d);
    double part = 0.0;
#pragma unroll
    for (int c = 0; c < DR; ++c) {
      const double v = pf.l[c] * sc;
      part += v * v;
    }
    ridge = wave_sum(part);
  }
  // ---- centred covariance, row `lane`
#pragma unroll
  for (int c = 0; c < DR; ++c) {
    if (c < d) {
      const double mc = readlane_f64(mean, c);
      double v = a[c] * is0 - mean * mc;
      if (c == lane) v += ridge;
      if (row && c <= lane) a[c] = v;
    }
  }
  // ---- Cholesky, column by column (left-looking order of _gaussian_potrf!,
  // src/adaptive_gaussian_execution.jl:427-434); lane c's a[k] (k < c) already
  // holds L[c][k]
  int status = 0;
  double rdiag = 0.0;  // 1 / L[lane][lane]
#pragma unroll
  for (int c = 0; c < DR; ++c) {
    if (c < d) {
      double v = a[c];
#pragma unroll
      for (int k = 0; k < c; ++k) v -= a[k] * readlane_f64(a[k], c);
      const double piv = readlane_f64(v, c);
      if (!(piv > 0.0) || !isfinite(piv)) status = 2;
      const double r = rsqrt_nr(piv);
      if (row && lane > c) a[c] = v * r;
      if (lane == c) {
        a[c] = piv * r;
        rdiag = r;
      }
    }
  }
  // ---- outputs: lane r writes mean r, 1 / L_rr and row r of L (Student-t rescaled)
  const double scale = f.student ? sqrt(1.0 / mult) : 1.0;
  bool ok = true;
  double lg = 0.0;
  double* out = f.next_slot64;
  const bool tf = f.t_is_float != 0;
  float* of = reinterpret_cast<float*>(f.next_slot_t);
  double* od = reinterpret_cast<double*>(f.next_slot_t);
  if (row) {
    double dl = 0.0;
#pragma unroll
    for (int c = 0; c < DR; ++c)
      if (c == lane) dl = a[c];
    const double mu = pf.cr + mean;
    const double diag = dl * scale;
    const double idiag = f.student ? rdiag / scale : rdiag;
    ok = isfinite(mu) && diag > 0.0;
    lg = log(fabs(diag));
    out[lane] = mu;
    out[d + lane] = slot_inv_value<double>(idiag);
    if (tf) {
      of[lane] = (float)mu;
      of[d + lane] = slot_inv_value<float>(idiag);
    } else {
      od[lane] = mu;
      od[d + lane] = slot_inv_value<double>(idiag);
    }
    const int base = 2 * d + lane * (lane + 1) / 2;
#pragma unroll
    for (int c = 0; c < DR; ++c) {
      if (c <= lane) {
        const double v = a[c] * scale;
        ok &= isfinite(v);
        out[base + c] = v;
        if (tf)
          of[base + c] = (float)v;
        else
          od[base + c] = v;
      }
    }
  }
  const double logabsdet = wave_sum(lg);
  const bool all_ok = __all(ok);
  if (lane == 0) {
    const double lognorm = f.student ? f.logconstant - logabsdet
                                     : -0.5 * (double)d * log(2.0 * 3.141592653589793) - logabsdet;
    int st = status;
    if (st == 0 && (!all_ok || !isfinite(lognorm))) st = 3;
    out[slot_lognorm_off(d)] = lognorm;
    if (tf)
      of[slot_lognorm_off(d)] = (float)lognorm;
    else
      od[slot_lognorm_off(d)] = lognorm;
    f.fit_status[f.round] = st;
  }
}

// EXACT: d == DR, compiled without per-column guards (d <= 16)
template <int DR, bool EXACT>
static __global__ __launch_bounds__(64) void k_finish_small(const double* __restrict__ red, FinishArgs f) {
  FitPrefetch<DR> pf;
  fit_prefetch<DR, EXACT>(f, pf);
  fit_small<DR, EXACT>(red, f, pf, true);
}

// ---------------------------------------------------------------- static MIS assignment
// Generating slot of every sample (src/methods/static_mis.jl:172-218): the
// sample's uniform u (counter RNG, or the prefilled buffer), mapped to
// (i + u) / n under stratified assignment or kept as is, then the first slot
// whose cumulative mass is >= the mapped value (binary search over the K
// active slots; the last cumulative mass is exactly 1).
static __global__ __launch_bounds__(256) void k_mis_assign(RngKey key, const unsigned long long* __restrict__ run_hi,
                                                          int64_t n, int stratified,
                                                          const double* __restrict__ cdf, int K,
                                                          const double* __restrict__ pre_uniforms,
                                                          int32_t* __restrict__ assign) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double u = pre_uniforms ? pre_uniforms[i] : rng_uniform(key, (uint64_t)i, *run_hi, kPurposeAssignment, 0);
  const double v = stratified ? ((double)i + u) / (double)n : u;
  int lo = 0, hi = K - 1;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (v <= cdf[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  assign[i] = lo;
}

// ---------------------------------------------------------------- order statistic (NPMC clip)
// Exact k-th smallest (1-based) of lw[first, first+n) by an 8-pass radix select
// on order-preserving 64-bit keys, each pass one launch over the whole chip.
__device__ inline uint64_t order_key(double v) {
  const uint64_t b = __double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double key_value(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

// State of the multi-workgroup radix select, in device memory: the key prefix fixed by
// the passes so far, the rank left within it, the global histogram of the current
// pass and the workgroups' ticket.  The histogram and ticket are zero between passes
// (the last workgroup of every pass clears them; the job's setup zeroes them once).
struct SelectState {
  unsigned long long prefix;
  long long k;
  unsigned int ticket, pad_;
  unsigned int hist[256];
};
constexpr int kSelectBlock = 256;

// One pass (digit `pass`, bits 56 - 8 pass .. + 7) over lw[first, first + n): every
// workgroup histograms the digit of the keys that carry the prefix in LDS and adds its
// counts to the global histogram; the last workgroup to finish (a ticket) picks the
// bin holding rank k (kth at pass 0), moves the prefix and the rank on, clears the
// histogram and the ticket, and after the last pass writes the k-th smallest value.
// The counts are exact, so the result does not depend on the launch shape.
template <class T>
__global__ __launch_bounds__(kSelectBlock) void k_select_pass(const T* __restrict__ lw, int64_t first, int64_t n,
                                                              int pass, int64_t kth, SelectState* st,
                                                              double* __restrict__ out) {
  __shared__ unsigned int h[256];
  __shared__ int last;
  const int tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const int shift = 56 - 8 * pass;
  const uint64_t prefix =
      pass == 0 ? 0ull : __hip_atomic_load(&st->prefix, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t hmask = pass == 0 ? 0ull : (~0ull << (shift + 8));
  for (int64_t i = (int64_t)blockIdx.x * kSelectBlock + tid; i < n; i += (int64_t)gridDim.x * kSelectBlock) {
    const uint64_t key = order_key((double)lw[first + i]);
    if ((key & hmask) == prefix) atomicAdd(&h[(key >> shift) & 0xff], 1u);
  }
  __syncthreads();
  if (h[tid]) atomicAdd(&st->hist[tid], h[tid]);
  __syncthreads();
  if (tid == 0) {
    // release: this workgroup's counts before its ticket; acquire: every earlier one's
    const unsigned int t = __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = t == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __atomic_thread_fence(__ATOMIC_ACQUIRE);
  h[tid] = __hip_atomic_load(&st->hist[tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (tid == 0) {
    int64_t k = pass == 0 ? kth : (int64_t)__hip_atomic_load(&st->k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int b = 0;
    for (; b < 256; ++b) {
      if (k <= (int64_t)h[b]) break;
      k -= h[b];
    }
    if (b > 255) b = 255;
    const uint64_t np = prefix | ((uint64_t)b << shift);
    __hip_atomic_store(&st->k, (long long)k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&st->prefix, (unsigned long long)np, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&st->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (pass == 7) *out = key_value(np);
  }
  __hip_atomic_store(&st->hist[tid], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- deterministic blocked CDF
// Tile = kCdfThreads x kCdfPer elements.  Thread j sums its kCdfPer weights
// sequentially; thread prefixes are a sequential scan of the thread totals;
// tile carries are a sequential scan of the tile totals; value = carry +
// (thread prefix + local sum); finally divide by the last value.  The CPU
// oracle replicates this order exactly.
constexpr int kCdfThreads = 256;
constexpr int kCdfPer = 16;

template <class T>
__device__ inline void cdf_thread_sums(const T* lw, int64_t n, double mx, int64_t tile, double* local,
                                       double& total) {
  const int64_t base = tile * (kCdfThreads * kCdfPer) + (int64_t)threadIdx.x * kCdfPer;
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < kCdfPer; ++e) {
    const int64_t i = base + e;
    const double w = i < n ? det_exp((double)lw[i] - mx) : 0.0;
    s = s + w;
    local[e] = s;
  }
  total = s;
}

// Wavefront scans in the fixed association the oracle restates (bit-exact CDF):
// lane l of a wave ends with v_0 + ... + v_l from six shuffle-up steps
// (Hillis-Steele: at step `off`, v_l <- v_{l-off} + v_l for l >= off).
__device__ __forceinline__ double wave_incl_scan(double v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double up = __shfl_up(v, off, 64);
    if (lane >= off) v = up + v;
  }
  return v;
}

// Exclusive prefix of one value per thread over a kCdfThreads-thread block and the
// block total: within each wave the inclusive scan shifted by one lane (lane 0: 0);
// across the waves W_0 + W_1 + ... in wave order, added to the lane's prefix.
// `wtot` is kCdfThreads / 64 doubles of LDS.
__device__ __forceinline__ void block_excl_scan(double tau, double* wtot, double& pre, double& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double incl = wave_incl_scan(tau, lane);
  double ex = __shfl_up(incl, 1, 64);
  if (lane == 0) ex = 0.0;
  if (lane == 63) wtot[wave] = incl;
  __syncthreads();
  double run = 0.0, mine = 0.0;
#pragma unroll
  for (int w = 0; w < kCdfThreads / 64; ++w) {
    if (w == wave) mine = run;
    run = run + wtot[w];
  }
  pre = mine + ex;
  total = run;
}

// Tile totals: tile b covers elements [b kCdfThreads kCdfPer, ...); each thread sums
// its kCdfPer elements in order, the block scan adds the thread totals.  The thread
// holding element n - 1 also writes its prefix plus its total, the part of the last
// element's running value below the tile carry, to tile_total[ntiles].
template <class T>
__global__ __launch_bounds__(kCdfThreads) void k_cdf_tiles(const T* __restrict__ lw, int64_t n,
                                                           const double* __restrict__ maxp,
                                                           double* __restrict__ tile_total) {
  __shared__ double wtot[kCdfThreads / 64];
  double local[kCdfPer], tot, pre, total;
  cdf_thread_sums(lw, n, *maxp, blockIdx.x, local, tot);
  block_excl_scan(tot, wtot, pre, total);
  if (threadIdx.x == 0) tile_total[blockIdx.x] = total;
  const int64_t base = (int64_t)blockIdx.x * (kCdfThreads * kCdfPer) + (int64_t)threadIdx.x * kCdfPer;
  if (base <= n - 1 && n - 1 < base + kCdfPer) tile_total[gridDim.x] = pre + tot;
}

// Tile carries (exclusive prefix of the tile totals) by the same scheme in one block:
// thread j sums the totals of tiles [j C, (j + 1) C), C = ceil(ntiles / kCdfThreads),
// in order; carry[b] = (block prefix of thread j) + (its running sum before b).
// carry[ntiles] is the divisor, the last element's running value (its tile's carry
// plus tile_total[ntiles]), so the last CDF entry is exactly 1 (src/resampling.jl:222);
// status = 1 when the grand total is not finite and positive (every weight zero).
static __global__ __launch_bounds__(kCdfThreads) void k_cdf_carry(const double* __restrict__ tile_total, int64_t ntiles,
                                                           double* __restrict__ carry, int* __restrict__ status) {
  __shared__ double wtot[kCdfThreads / 64];
  const int64_t C = (ntiles + kCdfThreads - 1) / kCdfThreads;
  const int64_t b0 = (int64_t)threadIdx.x * C;
  const int64_t b1 = b0 + C < ntiles ? b0 + C : ntiles;
  double s = 0.0;
  for (int64_t b = b0; b < b1; ++b) s = s + tile_total[b];
  double pre, total;
  block_excl_scan(s, wtot, pre, total);
  double run = 0.0;
  for (int64_t b = b0; b < b1; ++b) {
    const double c = pre + run;
    carry[b] = c;
    if (b == ntiles - 1) carry[ntiles] = c + tile_total[ntiles];
    run = run + tile_total[b];
  }
  if (threadIdx.x == 0 && status) *status = (isfinite(total) && total > 0.0) ? 0 : 1;
}

// cdf[i] = (carry[b] + (thread prefix + running sum)) / carry[ntiles], the tile's sums
// ordered as in k_cdf_tiles.
template <class T>
__global__ __launch_bounds__(kCdfThreads) void k_cdf_write(const T* __restrict__ lw, int64_t n,
                                                           const double* __restrict__ maxp,
                                                           const double* __restrict__ carry, int64_t ntiles,
                                                           double* __restrict__ cdf) {
  __shared__ double wtot[kCdfThreads / 64];
  double local[kCdfPer], tot, p, tile_sum;
  cdf_thread_sums(lw, n, *maxp, blockIdx.x, local, tot);
  block_excl_scan(tot, wtot, p, tile_sum);
  const double c = carry[blockIdx.x], last = carry[ntiles];
  const int64_t base = (int64_t)blockIdx.x * (kCdfThreads * kCdfPer) + (int64_t)threadIdx.x * kCdfPer;
#pragma unroll
  for (int e = 0; e < kCdfPer; ++e) {
    const int64_t i = base + e;
    if (i < n) cdf[i] = (c + (p + local[e])) / last;
  }
}

// Launches the three CDF kernels on `s`; `work` holds 2 ntiles + 2 doubles (tile
// totals, the last element's in-tile part, the carries and the divisor).
inline int64_t cdf_tiles(int64_t n) { return (n + (int64_t)kCdfThreads * kCdfPer - 1) / ((int64_t)kCdfThreads * kCdfPer); }
template <class T>
inline void launch_cdf(const T* lw, int64_t n, const double* maxp, double* work, double* cdf, int* status,
                       hipStream_t s) {
  const int64_t ntiles = cdf_tiles(n);
  double* carry = work + ntiles + 1;
  hipLaunchKernelGGL((k_cdf_tiles<T>), dim3((unsigned)ntiles), dim3(kCdfThreads), 0, s, lw, n, maxp, work);
  hipLaunchKernelGGL(k_cdf_carry, dim3(1), dim3(kCdfThreads), 0, s, (const double*)work, ntiles, carry, status);
  hipLaunchKernelGGL((k_cdf_write<T>), dim3((unsigned)ntiles), dim3(kCdfThreads), 0, s, lw, n, maxp,
                     (const double*)carry, ntiles, cdf);
}

// UnweightedSamples of a resample (src/resampling.jl:120-164): sample k of the output
// is input sample ancestors[k] (1-based), coordinate-major with stride `count`.
template <class T>
__global__ void k_gather_samples(const T* __restrict__ X, int64_t cap, int d, const int32_t* __restrict__ ancestors,
                                 int64_t count, T* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= count * d) return;
  const int64_t k = idx % count;
  const int c = (int)(idx / count);
  out[idx] = X[(int64_t)c * cap + (ancestors[k] - 1)];
}

template <class T>
__global__ void k_copy(const T* __restrict__ src, int64_t n, T* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

template <class T>
__global__ void k_fill(T* __restrict__ out, int64_t n, T value) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = value;
}

// ---------------------------------------------------------------- ancestor selection
// src/resampling.jl:92-118 (binary search + zero-mass plateau fix-up); the
// systematic method uses points (j + u)/count with one uniform u.
template <class T>
__global__ void k_select_ancestors(const double* __restrict__ cdf, const T* __restrict__ lw, int64_t n,
                                   const double* __restrict__ uniforms, int64_t count, int systematic,
                                   int32_t* __restrict__ ancestors) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const double u = systematic ? ((double)k + uniforms[0]) / (double)count : uniforms[k];
  int64_t first = 1, last = n;
  while (first < last) {
    const int64_t middle = first + ((last - first) >> 1);
    if (u < cdf[middle - 1])
      last = middle;
    else
      first = middle + 1;
  }
  const T ninf = -T(__builtin_huge_val());
  if (lw[first - 1] == ninf) {
    while (first > 1 && lw[first - 1] == ninf) --first;
    while (first < n && lw[first - 1] == ninf) ++first;
  }
  ancestors[k] = (int32_t)first;
}

// LocalResampling (src/dm_pmc_execution.jl:80-265): block k (kLocalLanes threads)
// picks proposal k's ancestor from its own group of the round's samples,
// [group[k], group[k + 1]).  The order is that of the reference's cooperative device
// kernel (:173-246): the group maximum; lane l sums exp(lw - max) over its contiguous
// chunk of ceil(count / kLocalLanes) samples; lane 0 adds the chunk totals in lane
// order, draws u_k * total and finds the first chunk whose running total exceeds it
// (else the last non-empty chunk, with its own total as remainder); that chunk's lane
// returns the first sample whose running sum exceeds the remainder (else its last).
// The exponentials are det_exp in FP64.  A group without positive finite weight gives
// ancestor 0 and sets *status (AllZeroWeightsError in the resampling phase).
constexpr int kLocalLanes = 256;
template <class T>
__global__ __launch_bounds__(kLocalLanes) void k_local_select(const T* __restrict__ lw,
                                                              const int32_t* __restrict__ group,
                                                              const double* __restrict__ uniforms,
                                                              int32_t* __restrict__ ancestors,
                                                              int* __restrict__ status) {
  __shared__ double red[kLocalLanes];
  __shared__ double chunk[kLocalLanes];
  __shared__ int chosen_s;
  __shared__ double thr_s;
  const int k = blockIdx.x, l = threadIdx.x;
  const int64_t g0 = group[k], g1 = group[k + 1];
  const int64_t count = g1 - g0;
  const double ninf = -__builtin_huge_val();
  double m = ninf;
  for (int64_t i = g0 + l; i < g1; i += kLocalLanes) m = fmax(m, (double)lw[i]);
  red[l] = m;
  __syncthreads();
  for (int off = kLocalLanes / 2; off > 0; off >>= 1) {
    if (l < off) red[l] = fmax(red[l], red[l + off]);
    __syncthreads();
  }
  const double M = red[0];
  const int64_t cs = (count + kLocalLanes - 1) / kLocalLanes;
  const int64_t c0 = g0 + (int64_t)l * cs;
  const int64_t c1 = c0 + cs < g1 ? c0 + cs : g1;
  double tot = 0.0;
  if (M != ninf)
    for (int64_t i = c0; i < c1; ++i) tot = tot + det_exp((double)lw[i] - M);
  chunk[l] = tot;
  __syncthreads();
  if (l == 0) {
    double total = 0.0;
    for (int j = 0; j < kLocalLanes; ++j) total = total + chunk[j];
    int chosen = -1;
    double rem = 0.0;
    if (isfinite(total) && total > 0.0) {
      const double draw = uniforms[k] * total;
      chosen = (int)(count < kLocalLanes ? count : kLocalLanes) - 1;
      rem = chunk[chosen];
      double prefix = 0.0;
      for (int j = 0; j < kLocalLanes; ++j) {
        const double next = prefix + chunk[j];
        if (draw < next) {
          chosen = j;
          rem = draw - prefix;
          break;
        }
        prefix = next;
      }
    }
    chosen_s = chosen;
    thr_s = rem;
    if (chosen < 0) {
      ancestors[k] = 0;
      *status = 1;
    }
  }
  __syncthreads();
  if (l == chosen_s) {
    double cum = 0.0;
    int64_t anc = c1 - 1;
    for (int64_t i = c0; i < c1; ++i) {
      cum = cum + det_exp((double)lw[i] - M);
      if (thr_s < cum) {
        anc = i;
        break;
      }
    }
    ancestors[k] = (int32_t)(anc + 1);
  }
}

// DM-PMC: new locations are the ancestors' samples (src/resampling.jl:120-164);
// writes both slot tables.  An ancestor of 0 (a failed local group) gathers nothing.
template <class T>
__global__ void k_gather_locations(const T* __restrict__ X, int64_t cap, int64_t first, int d,
                                   const int32_t* __restrict__ ancestors, int K, double* __restrict__ slots64,
                                   T* __restrict__ slots_t, int stride) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= K * d) return;
  const int k = idx / d, c = idx % d;
  if (ancestors[k] < 1) return;
  const int64_t a = first + ancestors[k] - 1;
  const T v = X[(int64_t)c * cap + a];
  slots64[(int64_t)k * stride + c] = (double)v;
  slots_t[(int64_t)k * stride + c] = v;
}

// ---------------------------------------------------------------- RNG export (test seam)
static __global__ void k_export_normals(RngKey key, uint64_t round, int64_t n, int count, uint64_t purpose,
                                 double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < count; ++q) out[i * count + q] = rng_normal(key, (uint64_t)i, round, purpose, q);
}
static __global__ void k_export_uniforms(RngKey key, uint64_t round, int64_t n, int count, uint64_t purpose,
                                  double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < count; ++q) out[i * count + q] = rng_uniform(key, (uint64_t)i, round, purpose, q);
}
static __global__ void k_rng_raw(uint64_t k0, uint64_t k1, uint64_t c1, uint64_t c2, uint64_t c3, int64_t n,
                          unsigned long long* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const U64x4 w = philox4x64((uint64_t)i, c1, c2, c3, k0, k1);
  for (int q = 0; q < 4; ++q) out[i * 4 + q] = w.v[q];
}

// Resampling uniforms of a round: counter (k, round, purpose 3, 0), word 0.
static __global__ void k_resampling_uniforms(RngKey key, const unsigned long long* run_hi, int round, int64_t count,
                                      double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  out[k] = rng_uniform(key, (uint64_t)k, *run_hi | (uint64_t)round, kPurposeResampling, 0);
}

// ---------------------------------------------------------------- layout helpers
template <class T>
__global__ void k_soa_to_aos(const T* __restrict__ X, int64_t cap, int64_t n, int d, double* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * d) return;
  const int64_t i = idx / d;
  const int c = (int)(idx % d);
  out[idx] = (double)X[(int64_t)c * cap + i];
}
template <class T>
__global__ void k_to_double(const T* __restrict__ src, int64_t n, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)src[i];
}

}  // namespace issampler
