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
  uint64_t rng_word;  // RNG counter word: (run << 32) | round
  int64_t first;      // global index of the first new sample
  int64_t n_new;
  int64_t n_old;      // retained samples [0, n_old) updated by the append path
  int old_blocks;     // blocks [0, old_blocks) run the append path
  // proposal slots
  const T* slots;     // slot table (T copy)
  int stride;         // slot stride in elements
  int k0, k1;         // mixture components [k0, k1] for new samples
  int slot_chunk;     // components staged in LDS per chunk
  int append_slot;    // component appended to retained samples
  const T* logmass;   // per-slot log mixture mass (indexed by slot)
  int gen_slot;       // generating slot when assign == nullptr
  const int32_t* assign;    // per-new-sample generating slot (DM-PMC)
  const int32_t* bank_ids;  // per-slot proposal id (DM-PMC)
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
  // target
  int tkind;
  const T* tslot;           // Gaussian target slot
  const T* pre_logtarget;   // batch-target values for new samples (kPreTarget)
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

// ---------------------------------------------------------------- block partials
// Reduces (lw, y = x - center) of the block's samples into one partial record
// [M, s0 = sum e, s2 = sum e^2, s1 = sum e y, S = sum e y y^T] with e = exp(lw - M)
// and M the block maximum.  lw = -inf marks an empty lane.  Deterministic:
// fixed sub-tile order and fixed per-value accumulation order.
template <int D>
__device__ inline void block_partials(double lw, const double (&y)[D], int d, int nv, int moments,
                                      double* __restrict__ rec, double* lds) {
  __shared__ double red[kBlock / 64];
  __shared__ double bmax_s;
  const int tid = threadIdx.x;
  double m = wave_max(lw);
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    double mm = red[0];
    for (int w = 1; w < kBlock / 64; ++w) mm = fmax(mm, red[w]);
    bmax_s = mm;
  }
  __syncthreads();
  const double M = bmax_s;
  const double e = (M == -__builtin_huge_val() || lw == -__builtin_huge_val()) ? 0.0 : exp(lw - M);
  constexpr int kMaxPer = (2 + D + D * (D + 1) / 2 + kBlock - 1) / kBlock;
  double acc[kMaxPer];
  int pr[kMaxPer], pc[kMaxPer];
#pragma unroll
  for (int q = 0; q < kMaxPer; ++q) {
    acc[q] = 0.0;
    const int v = tid + q * kBlock;
    pr[q] = pc[q] = 0;
    if (v >= 2 + d && v < nv) pair_index(v - 2 - d, pr[q], pc[q]);
  }
  const int rowlen = moments ? d + 1 : 1;
  for (int sub = 0; sub < kBlock / kTile; ++sub) {
    __syncthreads();
    if ((tid / kTile) == sub) {
      double* row = lds + (tid % kTile) * rowlen;
      row[0] = e;
      if (moments) {
#pragma unroll
        for (int c = 0; c < D; ++c)
          if (c < d) row[1 + c] = y[c];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMaxPer; ++q) {
      const int v = tid + q * kBlock;
      if (v < nv) {
        double s = acc[q];
        if (v == 0) {
          for (int i = 0; i < kTile; ++i) s += lds[i * rowlen];
        } else if (v == 1) {
          for (int i = 0; i < kTile; ++i) {
            const double ei = lds[i * rowlen];
            s += ei * ei;
          }
        } else if (v < 2 + d) {
          const int r = v - 2;
          for (int i = 0; i < kTile; ++i) s += lds[i * rowlen] * lds[i * rowlen + 1 + r];
        } else {
          const int r = pr[q], c = pc[q];
          for (int i = 0; i < kTile; ++i) {
            const double* row = lds + i * rowlen;
            s += (row[0] * row[1 + r]) * row[1 + c];
          }
        }
        acc[q] = s;
      }
    }
  }
  if (tid == 0) rec[0] = M;
#pragma unroll
  for (int q = 0; q < kMaxPer; ++q) {
    const int v = tid + q * kBlock;
    if (v < nv) rec[1 + v] = acc[q];
  }
}

// ---------------------------------------------------------------- normals for one sample
template <class T, int D>
__device__ inline bool sample_normals(const RoundArgs<T>& a, int64_t j, T (&z)[D], int64_t logical) {
  double zz[D];
  const int d = a.d;
#pragma unroll
  for (int c = 0; c < D; ++c) zz[c] = 0.0;
  if (a.pre_normals) {
    const double* src = a.pre_normals + j * a.nstride;
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (c < d) zz[c] = src[c];
  } else {
    rng_normals<D>(a.key, (uint64_t)j, a.rng_word, kPurposeNormals, d, zz);
  }
  bool ok = true;
  if (a.fam.student) {
    double rn[kGammaAttempts], ru[kGammaAttempts + 1];
    for (int q = 0; q < a.rs; ++q)
      rn[q] = a.pre_rn ? a.pre_rn[j * a.rn_stride + q]
                       : rng_normal(a.key, (uint64_t)j, a.rng_word, kPurposeRadialNormals, q);
    for (int q = 0; q < a.us; ++q)
      ru[q] = a.pre_ru ? a.pre_ru[j * a.ru_stride + q]
                       : rng_uniform(a.key, (uint64_t)j, a.rng_word, kPurposeRadialUniforms, q);
    double mult;
    if (!student_t_multiplier(a.dof, rn, ru, a.us > 0, &mult)) {
      record_failure(a.fail, logical, 0, kDrawExhausted);
      mult = __builtin_nan("");
      ok = false;
    }
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (c < d) zz[c] *= mult;
  }
#pragma unroll
  for (int c = 0; c < D; ++c) z[c] = T(zz[c]);
  return ok;
}

template <class T, int D>
__device__ inline T inline_target(int tkind, int d, const T* tslot_lds, const T (&x)[D], T (&y)[D]) {
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
  return slot_logdensity<T, D>(tslot_lds, d, g, x, y);
}

// ---------------------------------------------------------------- the round kernel
// Blocks [0, old_blocks): AMIS append over retained samples
//   (src/amis_execution.jl:37-65 with the weight formation of :107-132).
// Blocks [old_blocks, ...): new samples: draw, target, mixture log-sum-exp and
//   weight (src/mis_execution.jl:601-771; fused plain IS src/native_execution.jl:741-812).
// LDS: [moment tile: kTile*(d+1) doubles][target slot][component chunk].
template <class T, int D, int MODE>
__global__ __launch_bounds__(kBlock) void k_round(RoundArgs<T> a) {
  extern __shared__ double lds_raw[];
  const int d = a.d;
  const int tid = threadIdx.x;
  double* tile_lds = lds_raw;
  T* slot_lds = reinterpret_cast<T*>(lds_raw + kTile * (d + 1));
  double lw_acc = -__builtin_huge_val();
  double yc[D];
#pragma unroll
  for (int c = 0; c < D; ++c) yc[c] = 0.0;

  if ((int)blockIdx.x < a.old_blocks) {
    // ------------------------------------------------ append path
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const T* s = a.slots + (int64_t)a.append_slot * a.stride;
    for (int q = tid; q < a.stride; q += kBlock) slot_lds[q] = s[q];
    __syncthreads();
    if (i < a.n_old) {
      T x[D], y[D];
#pragma unroll
      for (int c = 0; c < D; ++c) x[c] = (c < d) ? a.X[(int64_t)c * a.cap + i] : T(0);
      const T q = slot_logdensity<T, D>(slot_lds, d, a.fam, x, y);
      uint16_t reason = (q != q) ? kProposalInvalid : 0;
      T num = a.lognums[i];
      if (!reason) {
        const T value = logaddexp<T>(num, a.logmass[a.append_slot] + q);
        reason = proposal_reason(value);
        if (!reason) {
          num = value;
          a.lognums[i] = value;
        }
      }
      if (reason) record_failure(a.fail, i + 1, 0, reason);
      if (a.round_ids[i] != 0) {
        const T lw = a.logtargets[i] - (num - a.logtotal);
        const uint16_t wr = logweight_reason(lw);
        if (wr) {
          record_failure(a.fail, i + 1, 0, wr);
        } else {
          a.logweights[i] = lw;
          lw_acc = (double)lw;
        }
      }
      if (a.moments) {
#pragma unroll
        for (int c = 0; c < D; ++c)
          if (c < d) yc[c] = (double)x[c] - a.center[c];
      }
    }
  } else {
    // ------------------------------------------------ new-sample path
    const int64_t j = (int64_t)(blockIdx.x - a.old_blocks) * kBlock + tid;
    const bool active = j < a.n_new;
    const int64_t gi = a.first + j;
    T x[D], y[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      x[c] = T(0);
      y[c] = T(0);
    }
    bool ok = active;
    int gen = a.gen_slot;
    T tv = T(0);
    const bool staged_target = (a.tkind == kTGaussian && MODE == kDrawFused);
    T* tslot_lds = slot_lds;
    T* comp_lds = slot_lds + (staged_target ? a.stride : 0);
    if (staged_target) {
      for (int q = tid; q < a.stride; q += kBlock) tslot_lds[q] = a.tslot[q];
    }
    __syncthreads();
    if (active) {
      if (a.assign) gen = a.assign[j];
      if (MODE != kDrawOnly) {
        a.logweights[gi] = -T(__builtin_huge_val());
        a.logtargets[gi] = -T(__builtin_huge_val());
        a.lognums[gi] = T(0);
        a.round_ids[gi] = 0;
        if (a.proposal_ids) a.proposal_ids[gi] = 0;
      }
    }
    if (MODE == kPreTarget) {
      if (active) {
#pragma unroll
        for (int c = 0; c < D; ++c) x[c] = (c < d) ? a.X[(int64_t)c * a.cap + gi] : T(0);
        tv = a.pre_logtarget[j];
        const uint16_t tr = target_reason(tv);
        if (tr) {
          record_failure(a.fail, j + 1, 0, tr);
          ok = false;
        }
      }
    } else if (active) {
      T z[D];
      const bool drawn = sample_normals<T, D>(a, j, z, j + 1);
      if (!drawn && !a.record_nonfinite_after_exhaust) {
        ok = false;
      } else {
        const T* gs = a.slots + (int64_t)gen * a.stride;
        const bool valid = slot_draw<T, D>(gs, d, z, x);
#pragma unroll
        for (int c = 0; c < D; ++c)
          if (c < d) a.X[(int64_t)c * a.cap + gi] = x[c];
        if (!valid) {
          record_failure(a.fail, j + 1, 0, kGeneratedNonfinite);
          ok = false;
        } else if (MODE == kDrawFused) {
          tv = inline_target<T, D>(a.tkind, d, tslot_lds, x, y);
          const uint16_t tr = target_reason(tv);
          if (tr) {
            record_failure(a.fail, j + 1, 0, tr);
            ok = false;
          }
        }
      }
    }
    if (MODE != kDrawOnly) {
      // mixture denominator over components [k0, k1], staged through LDS in chunks
      LseAcc<T> acc;
      acc.init();
      T generating = T(0);
      const int ncomp = a.k1 - a.k0 + 1;
      for (int base = 0; base < ncomp; base += a.slot_chunk) {
        const int cnt = min(a.slot_chunk, ncomp - base);
        __syncthreads();
        const T* src = a.slots + (int64_t)(a.k0 + base) * a.stride;
        for (int q = tid; q < cnt * a.stride; q += kBlock) comp_lds[q] = src[q];
        __syncthreads();
        if (ok) {
          for (int kk = 0; kk < cnt; ++kk) {
            const int k = a.k0 + base + kk;
            const T q = slot_logdensity<T, D>(comp_lds + kk * a.stride, d, a.fam, x, y);
            if (k == gen) generating = q;
            acc.add(a.logmass[k] + q);
          }
        }
      }
      if (ok) {
        const T value = acc.value();
        if (generating != generating || generating == -T(__builtin_huge_val()) || value != value ||
            value == -T(__builtin_huge_val())) {
          record_failure(a.fail, j + 1, 0, kProposalInvalid);
        } else {
          const T lw = tv - (value - a.logtotal);
          const uint16_t wr = logweight_reason(lw);
          if (wr) {
            record_failure(a.fail, j + 1, 0, wr);
          } else {
            a.logtargets[gi] = tv;
            a.lognums[gi] = value;
            a.logweights[gi] = lw;
            a.round_ids[gi] = a.round + 1;
            if (a.proposal_ids) a.proposal_ids[gi] = a.bank_ids ? a.bank_ids[gen] : 0;
            lw_acc = (double)lw;
            if (a.moments) {
#pragma unroll
              for (int c = 0; c < D; ++c)
                if (c < d) yc[c] = (double)x[c] - a.center[c];
            }
          }
        }
      }
    }
  }
  if (MODE != kDrawOnly && a.partials) {
    block_partials<D>(lw_acc, yc, d, a.nv, a.moments, a.partials + (int64_t)blockIdx.x * (1 + a.nv),
                      tile_lds);
  }
}

// ---------------------------------------------------------------- moment partials over a range
// NPMC clipped fit and the results API: lw_i = min(lw_i, clip) over [first, first + n).
template <class T, int D>
__global__ __launch_bounds__(kBlock) void k_moment_partials(const T* __restrict__ X, int64_t cap,
                                                            int64_t x_stride_sample, const T* __restrict__ lw,
                                                            int64_t first, int64_t n, int d,
                                                            const double* __restrict__ clip,
                                                            const double* __restrict__ center, int moments,
                                                            int nv, double* __restrict__ partials) {
  extern __shared__ double lds_raw[];
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double l = -__builtin_huge_val();
  double y[D];
#pragma unroll
  for (int c = 0; c < D; ++c) y[c] = 0.0;
  if (j < n) {
    const int64_t i = first + j;
    l = (double)lw[i];
    if (clip) l = fmin(l, *clip);
    if (moments) {
#pragma unroll
      for (int c = 0; c < D; ++c)
        if (c < d) {
          const double xv = x_stride_sample == 1 ? (double)X[(int64_t)c * cap + i]
                                                 : (double)X[i * x_stride_sample + c];
          y[c] = xv - center[c];
        }
    }
  }
  block_partials<D>(l, y, d, nv, moments, partials + (int64_t)blockIdx.x * (1 + nv), lds_raw);
}

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

__device__ void finish_block(const double* red, const FinishArgs& f, double* lds);

// grid = nv blocks; block k reduces value k over all records, then the last
// block to finish runs the summary / fit on the reduced values.
__global__ __launch_bounds__(kBlock) void k_reduce_finish(const double* __restrict__ partials,
                                                          int64_t nrec, int nv,
                                                          double* __restrict__ reduced, FinishArgs f) {
  __shared__ double red[kBlock / 64];
  __shared__ double Ms;
  __shared__ bool last_s;
  extern __shared__ double lds_raw[];
  const int tid = threadIdx.x;
  const int k = blockIdx.x;
  const int rs = 1 + nv;
  double m = -__builtin_huge_val();
  for (int64_t r = tid; r < nrec; r += kBlock) m = fmax(m, partials[r * rs]);
  m = wave_max(m);
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    double mm = red[0];
    for (int w = 1; w < kBlock / 64; ++w) mm = fmax(mm, red[w]);
    Ms = mm;
  }
  __syncthreads();
  const double M = Ms;
  double s = 0.0;
  if (M != -__builtin_huge_val()) {
    for (int64_t r = tid; r < nrec; r += kBlock) {
      const double mr = partials[r * rs];
      if (mr == -__builtin_huge_val()) continue;
      double sc = exp(mr - M);
      if (k == 1) sc *= sc;
      s += partials[r * rs + 1 + k] * sc;
    }
  }
  s = wave_sum(s);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) {
    double tot = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) tot += red[w];
    reduced[1 + k] = tot;
    if (k == 0) reduced[0] = M;
    __threadfence();
    const unsigned prev = atomicAdd(f.counter, 1u);
    last_s = (prev == (unsigned)(gridDim.x - 1));
  }
  __syncthreads();
  if (!last_s) return;
  __threadfence();
  if (tid == 0) *f.counter = 0u;
  finish_block(reduced, f, lds_raw);
}

// Summary (src/execution.jl:176-186) and FP64 moment fit
// (src/adaptive_gaussian_execution.jl:303-425): mean = c + s1/s0,
// cov = S/s0 - (s1/s0)(s1/s0)^T + ridge I, Cholesky, Student-t rescale, log-det.
__device__ void finish_block(const double* red, const FinishArgs& f, double* lds) {
  const int tid = threadIdx.x;
  const int d = f.d;
  volatile const double* rv = red;
  const double M = rv[0], s0 = rv[1], s2 = rv[2];
  if (tid == 0 && (f.flags & (kFinishSummary | kFinishAdaptEss))) {
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
  if (!(f.flags & kFinishFit)) return;
  __shared__ int status_s;
  if (tid == 0) status_s = 0;
  __syncthreads();
  if (!(s0 > 0.0) || !isfinite(s0) || M == -__builtin_huge_val()) {
    if (tid == 0) f.fit_status[f.round] = 1;
    return;
  }
  double* A = lds;           // d*d, row-major full
  double* mean = lds + d * d;
  // mean and centred covariance
  for (int c = tid; c < d; c += kBlock) mean[c] = rv[3 + c] / s0;
  __syncthreads();
  for (int p = tid; p < d * (d + 1) / 2; p += kBlock) {
    int r, c;
    pair_index(p, r, c);
    const double v = rv[3 + d + p] / s0 - mean[r] * mean[c];
    A[r * d + c] = v;
    A[c * d + r] = v;
  }
  __syncthreads();
  // scale-aware ridge from the previous proposal's factor (src/adaptive_gaussian_execution.jl:283-301)
  if (tid == 0) {
    const double reg = 1.4901161193847656e-08;  // sqrt(eps(Float64))
    const double mult = f.student ? f.dof / (f.dof - 2.0) : 1.0;
    const double* L = f.prev_slot + 2 * d;
    double trace = 0.0;
    for (int r = 0; r < d; ++r)
      for (int c = 0; c <= r; ++c) trace += L[r * (r + 1) / 2 + c] * L[r * (r + 1) / 2 + c];
    double ridge = reg * (trace * mult) / (double)d;
    if (!isfinite(ridge)) {
      const double sc = sqrt(reg) * sqrt(mult / (double)d);
      ridge = 0.0;
      for (int r = 0; r < d; ++r)
        for (int c = 0; c <= r; ++c) {
          const double v = L[r * (r + 1) / 2 + c] * sc;
          ridge += v * v;
        }
    }
    for (int c = 0; c < d; ++c) A[c * d + c] += ridge;
  }
  __syncthreads();
  // left-looking Cholesky, lower factor stored in A's lower triangle
  for (int c = 0; c < d; ++c) {
    if (tid == 0) {
      double v = A[c * d + c];
      for (int k = 0; k < c; ++k) v -= A[c * d + k] * A[c * d + k];
      if (!(v > 0.0) || !isfinite(v)) status_s = 2;
      A[c * d + c] = sqrt(v);
    }
    __syncthreads();
    const double diag = A[c * d + c];
    for (int r = c + 1 + tid; r < d; r += kBlock) {
      double v = A[r * d + c];
      for (int k = 0; k < c; ++k) v -= A[r * d + k] * A[c * d + k];
      A[r * d + c] = v / diag;
    }
    __syncthreads();
  }
  if (tid == 0) {
    int status = status_s;
    const double mult = f.student ? f.dof / (f.dof - 2.0) : 1.0;
    const double scale = f.student ? sqrt(1.0 / mult) : 1.0;
    double logabsdet = 0.0;
    bool ok = status == 0;
    double* out = f.next_slot64;
    for (int r = 0; r < d; ++r) {
      const double mu = f.center[r] + mean[r];
      out[r] = mu;
      ok &= isfinite(mu);
    }
    for (int r = 0; r < d; ++r) {
      for (int c = 0; c <= r; ++c) {
        const double v = A[r * d + c] * scale;
        out[2 * d + r * (r + 1) / 2 + c] = v;
        ok &= isfinite(v);
      }
      const double diag = out[2 * d + r * (r + 1) / 2 + r];
      ok &= diag > 0.0;
      out[d + r] = 1.0 / diag;
      logabsdet += log(fabs(diag));
    }
    const double lognorm = f.student ? f.logconstant - logabsdet
                                     : -0.5 * (double)d * log(2.0 * 3.141592653589793) - logabsdet;
    ok &= isfinite(lognorm);
    out[slot_lognorm_off(d)] = lognorm;
    if (status == 0 && !ok) status = 3;
    f.fit_status[f.round] = status;
  }
  __syncthreads();
  // T copy of the new slot (logmass is owned by the host schedule)
  const int nfill = slot_lognorm_off(d) + 1;
  for (int q = tid; q < nfill; q += kBlock) {
    if (f.t_is_float)
      reinterpret_cast<float*>(f.next_slot_t)[q] = (float)f.next_slot64[q];
    else
      reinterpret_cast<double*>(f.next_slot_t)[q] = f.next_slot64[q];
  }
}

// ---------------------------------------------------------------- order statistic (NPMC clip)
// Exact k-th smallest (1-based) of lw[first, first+n) by an 8-pass radix select
// on order-preserving 64-bit keys, one workgroup.
__device__ inline uint64_t order_key(double v) {
  const uint64_t b = __double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double key_value(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

template <class T>
__global__ __launch_bounds__(1024) void k_select_kth(const T* __restrict__ lw, int64_t first, int64_t n,
                                                     int64_t kth, double* __restrict__ out) {
  __shared__ unsigned int hist[256];
  __shared__ uint64_t prefix_s;
  __shared__ int64_t k_s;
  const int tid = threadIdx.x;
  if (tid == 0) {
    prefix_s = 0;
    k_s = kth;
  }
  for (int pass = 0; pass < 8; ++pass) {
    const int shift = 56 - 8 * pass;
    for (int b = tid; b < 256; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    const uint64_t prefix = prefix_s;
    const uint64_t hmask = pass == 0 ? 0ull : (~0ull << (shift + 8));
    for (int64_t i = tid; i < n; i += blockDim.x) {
      const uint64_t key = order_key((double)lw[first + i]);
      if ((key & hmask) == prefix) atomicAdd(&hist[(key >> shift) & 0xff], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      int64_t k = k_s;
      int b = 0;
      for (; b < 256; ++b) {
        if (k <= (int64_t)hist[b]) break;
        k -= hist[b];
      }
      if (b > 255) b = 255;
      k_s = k;
      prefix_s = prefix | ((uint64_t)b << shift);
    }
    __syncthreads();
  }
  if (tid == 0) *out = key_value(prefix_s);
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

template <class T>
__global__ __launch_bounds__(kCdfThreads) void k_cdf_tiles(const T* __restrict__ lw, int64_t n,
                                                           const double* __restrict__ maxp,
                                                           double* __restrict__ tile_total) {
  __shared__ double tau[kCdfThreads];
  double local[kCdfPer], tot;
  cdf_thread_sums(lw, n, *maxp, blockIdx.x, local, tot);
  tau[threadIdx.x] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    for (int j = 0; j < kCdfThreads; ++j) run = run + tau[j];
    tile_total[blockIdx.x] = run;
  }
}

template <class T>
__global__ __launch_bounds__(kCdfThreads) void k_cdf_write(const T* __restrict__ lw, int64_t n,
                                                           const double* __restrict__ maxp,
                                                           const double* __restrict__ tile_total,
                                                           int64_t ntiles, double* __restrict__ cdf,
                                                           int* __restrict__ status) {
  __shared__ double tau[kCdfThreads];
  __shared__ double pre[kCdfThreads];
  __shared__ double carry_s, total_s;
  double local[kCdfPer], tot;
  cdf_thread_sums(lw, n, *maxp, blockIdx.x, local, tot);
  tau[threadIdx.x] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    for (int j = 0; j < kCdfThreads; ++j) {
      pre[j] = run;
      run = run + tau[j];
    }
    double carry = 0.0;
    for (int64_t t = 0; t < (int64_t)blockIdx.x; ++t) carry = carry + tile_total[t];
    double all = 0.0;
    for (int64_t t = 0; t < ntiles - 1; ++t) all = all + tile_total[t];
    carry_s = carry;
    total_s = all + tile_total[ntiles - 1];
    if (blockIdx.x == 0 && status) *status = (isfinite(total_s) && total_s > 0.0) ? 0 : 1;
  }
  __syncthreads();
  const double carry = carry_s, total = total_s, p = pre[threadIdx.x];
  const int64_t base = (int64_t)blockIdx.x * (kCdfThreads * kCdfPer) + (int64_t)threadIdx.x * kCdfPer;
#pragma unroll
  for (int e = 0; e < kCdfPer; ++e) {
    const int64_t i = base + e;
    if (i < n) cdf[i] = (carry + (p + local[e])) / total;
  }
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

// DM-PMC: new locations are the ancestors' samples (src/resampling.jl:120-164);
// writes both slot tables.
template <class T>
__global__ void k_gather_locations(const T* __restrict__ X, int64_t cap, int64_t first, int d,
                                   const int32_t* __restrict__ ancestors, int K, double* __restrict__ slots64,
                                   T* __restrict__ slots_t, int stride) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= K * d) return;
  const int k = idx / d, c = idx % d;
  const int64_t a = first + ancestors[k] - 1;
  const T v = X[(int64_t)c * cap + a];
  slots64[(int64_t)k * stride + c] = (double)v;
  slots_t[(int64_t)k * stride + c] = v;
}

// ---------------------------------------------------------------- RNG export (test seam)
__global__ void k_export_normals(RngKey key, uint64_t round, int64_t n, int count, uint64_t purpose,
                                 double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < count; ++q) out[i * count + q] = rng_normal(key, (uint64_t)i, round, purpose, q);
}
__global__ void k_export_uniforms(RngKey key, uint64_t round, int64_t n, int count, uint64_t purpose,
                                  double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < count; ++q) out[i * count + q] = rng_uniform(key, (uint64_t)i, round, purpose, q);
}
__global__ void k_rng_raw(uint64_t k0, uint64_t k1, uint64_t c1, uint64_t c2, uint64_t c3, int64_t n,
                          unsigned long long* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const U64x4 w = philox4x64((uint64_t)i, c1, c2, c3, k0, k1);
  for (int q = 0; q < 4; ++q) out[i * 4 + q] = w.v[q];
}

// Resampling uniforms of a round: counter (k, round, purpose 3, 0), word 0.
__global__ void k_resampling_uniforms(RngKey key, uint64_t round, int64_t count, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  out[k] = rng_uniform(key, (uint64_t)k, round, kPurposeResampling, 0);
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
