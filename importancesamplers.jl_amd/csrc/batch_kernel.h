// Device code of the batched AMIS engine (see batch.h); included only by the
// kernel translation units batch_k.hip.
//
// One workgroup of W waves runs one job for all of its rounds.  Per round t:
//   pass A  each wave takes 64 of the round's N new samples at a time (one per
//           lane): counter-RNG normals and the Student-t radial multiplier,
//           the affine draw from slot t, the regression target on MFMA
//           (X_data beta for the wave's 64 samples, streamed over the data rows
//           in 16-row tiles), the t+1 term deterministic-mixture log-sum-exp
//           over the slots staged in LDS, the weight, and the sample's
//           contribution to the summary and the moments;
//   pass B  the waves stream the t N retained samples: append of component t
//           to the AMIS numerator (src/amis_execution.jl:37-65), weight
//           re-formation (src/amis_execution.jl:107-132), summary and moments;
//   finish  the waves' records are merged in wave order; wave 0 computes the
//           round's ESS and log Z (src/execution.jl:176-186) and the FP64 fit
//           (src/adaptive_gaussian_execution.jl:303-434): mean, covariance,
//           scale-aware ridge, Cholesky (one row per lane, in registers),
//           Student-t rescale and log-determinant, written as slot t + 1.
//
// Every loop over coordinates runs to the job's padded dimension P = d rounded
// up to 4, a compile-time constant (one instantiation per P, chosen per job
// inside the persistent kernel).  Padded coordinates are exact zeros: their
// slot rows are mu = 0, L row = 0, 1/L_rr = 1, their data columns are zero,
// so they add nothing to any sum, and the hot loops carry no per-coordinate
// guards.  Moments accumulate about the centre mu_t on FP64 MFMA (16x16x4)
// from the staged products sqrt(e)(x - mu_t).
#pragma once

#include "batch.h"
#include "kernels.h"

namespace issampler {

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 16-byte vectors: 4 floats or 2 doubles.
template <class T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float __attribute__((ext_vector_type(4))) type;
  static constexpr int E = 4;
};
template <>
struct Vec16<double> {
  typedef double __attribute__((ext_vector_type(2))) type;
  static constexpr int E = 2;
};

// Student-t radial multiplier of one sample (the 8 buffered gamma attempts of
// _buffered_unit_gamma, src/native_execution.jl:147-180) with each attempt's values
// generated when the acceptance loop reaches it: attempts 4 b .. 4 b + 3 take Philox
// block b of the radial normals (Box-Muller pairs (0, 1) and (2, 3)) and of the radial
// uniforms; a pair's logarithm and sine/cosine are computed for the lanes still
// searching.  The values and the result of the buffered form, without the 17 values in
// registers and without the second pair for the samples accepted at attempt 0 or 1.
__device__ inline void batch_radial_pair(const U64x4& w, int p, double& n0, double& n1) {
  const double r = sqrt(-2.0 * log(u53_open(w.v[2 * p])));
  double sn, cs;
  sincospi(2.0 * u53_open(w.v[2 * p + 1]), &sn, &cs);
  n0 = r * cs;
  n1 = r * sn;
}

__device__ inline bool batch_multiplier(const RngKey& key, uint64_t sample, uint64_t round, int us, double dof,
                                        double* out) {
  if (us == 0 || dof == 1.0) {
    const U64x4 w = philox4x64(sample, round, kPurposeRadialNormals, 0, key.k0, key.k1);
    double n0, n1;
    batch_radial_pair(w, 0, n0, n1);
    const double den = fabs(n0);
    if (!(isfinite(den) && den > 0.0)) return false;
    *out = 1.0 / den;
    return true;
  }
  const GammaSetup g = gamma_setup(dof);
#pragma unroll 1
  for (int blk = 0; blk < kGammaAttempts / 4; ++blk) {
    const U64x4 wn = philox4x64(sample, round, kPurposeRadialNormals, (uint64_t)blk, key.k0, key.k1);
    const U64x4 wu = philox4x64(sample, round, kPurposeRadialUniforms, (uint64_t)blk, key.k0, key.k1);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      double rn[2];
      batch_radial_pair(wn, p, rn[0], rn[1]);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double root = 1.0 + g.coefficient * rn[h];
        if (!(root > 0.0)) continue;
        const double u = u53(wu.v[2 * p + h]);
        // the squeeze accepts most attempts: the logarithms only when a lane needs them
        const double n2 = rn[h] * rn[h];
        bool acc = u < 1.0 - 0.0331 * (n2 * n2);
        if (__any(!acc)) {
          if (!acc) acc = gamma_accepts(g, rn[h], root, u);
        }
        if (!acc) continue;
        return gamma_multiplier(
            g, dof, root,
            [&] {
              const int b = kGammaAttempts / 4;  // the extra uniform: block 2, word 0
              return u53(philox4x64(sample, round, kPurposeRadialUniforms, (uint64_t)b, key.k0, key.k1).v[0]);
            },
            out);
      }
    }
  }
  return false;
}

// Whitened density of a batch slot (layout in batch.h, padded dimension P)
// at x; y receives L^{-1}(x - mu).  Same operation order as slot_logdensity.
// The rows are software-pipelined: row r + 1 is loaded from LDS while row r
// is reduced, and a scheduling barrier per row keeps the compiler from
// hoisting every row's loads (which would hold the whole factor in registers).
template <class T, int P>
__device__ inline T bslot_logdensity(const T* s, const Family<T>& fam, const T (&x)[P], T (&y)[P]) {
  typedef typename Vec16<T>::type V;
  constexpr int E = Vec16<T>::E;
  constexpr int NV = P / E;
  const T* inv = s + P;
  const T* L = s + 2 * P;
  V buf[2][NV];
  T mu_n = s[0], inv_n = inv[0];
  T r2 = T(0);
#pragma unroll
  for (int r = 0; r < P; ++r) {
    const T mu_r = mu_n, inv_r = inv_n;
    if (r + 1 < P) {
      const V* p = reinterpret_cast<const V*>(L + bslot_rowoff(r + 1));
#pragma unroll
      for (int q = 0; q < (r + E) / E; ++q) buf[(r + 1) & 1][q] = p[q];
      mu_n = s[r + 1];
      inv_n = inv[r + 1];
    }
    T v = x[r] - mu_r;
#pragma unroll
    for (int q = 0; q < (r + E - 1) / E; ++q)
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (q * E + e < r) v -= buf[r & 1][q][e] * y[q * E + e];
    v *= inv_r;
    y[r] = v;
    r2 += v * v;
    __builtin_amdgcn_sched_barrier(0);
  }
  return radial_logdensity<T, P>(fam, s[bslot_lognorm(P)], r2, y, P);
}

// acc[0..2]: moment tiles (0,0), (1,0), (1,1) of sum e (x - c)(x - c)^T;
// acc[3..4]: first-order sums sum e (x - c) for coordinates 0..15 and 16..31
// (every column of these tiles holds the same sums).
constexpr int kBatchTiles = 5;
struct BatchAcc {
  double s0, s2;
  double4_t acc[kBatchTiles];
};

// Adds one sample per lane (log weight lw, coordinates x) to the wave's
// summary monoid (src/execution.jl:159-186) and moments about the centre.
template <class T, int P>
__device__ inline void batch_accumulate(T lw, const T (&x)[P], const T* centre, T* stage, int lane, T& mw,
                                        BatchAcc& A) {
  constexpr int NT = (P + 15) / 16;
  const T ninf = -T(__builtin_huge_val());
  if (__any(lw > mw)) {
    const T wm = wave_max(lw);
    const double f = (mw == ninf) ? 0.0 : exp((double)mw - (double)wm);
    A.s0 *= f;
    A.s2 *= f * f;
#pragma unroll
    for (int p = 0; p < kBatchTiles; ++p) A.acc[p] *= f;
    mw = wm;
  }
  const T se = (lw == ninf) ? T(0) : exp(T(0.5) * (lw - mw));
  const double sd = (double)se;
  const double e = sd * sd;
  A.s0 += e;
  A.s2 += e * e;
  T* row = stage + lane * kBatchStage;
  // when P is not a multiple of 16, sqrt(e) itself is staged as coordinate P, so the
  // first-order sums come out as row P of the second-moment tiles (1 or 3 MFMAs per
  // 4 samples instead of 2 or 5); the tiles stay Gram matrices of the staged vectors,
  // products exact in FP64, so they stay positive semidefinite
  constexpr bool embed = P % 16 != 0;
#pragma unroll
  for (int c = 0; c < 16 * NT; ++c) {
    if constexpr (sizeof(T) == 4) {
      // FP32 jobs stage in FP32 arithmetic (the stage holds floats either way)
      T v = T(0);
      if (c < P) v = (se != T(0)) ? se * (x[c] - centre[c]) : T(0);
      if (embed && c == P) v = se;
      row[c] = v;
    } else {
      double v = 0.0;
      if (c < P) v = (sd != 0.0) ? sd * ((double)x[c] - (double)centre[c]) : 0.0;
      if (embed && c == P) v = sd;
      row[c] = (T)v;
    }
  }
  if (!embed) row[kBatchMaxDim] = se;
  wave_sync();
  // v_mfma_f64_16x16x4f64 over 4-sample steps: A = the staged column block of
  // coordinates 16 ti + l % 16, B = the same block (second moments) or the
  // samples' sqrt(e) (first moments)
#pragma unroll
  for (int st = 0; st < 16; ++st) {
    const T* rs = stage + (4 * st + (lane >> 4)) * kBatchStage;
    const double a0 = (double)rs[lane & 15];
    A.acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, A.acc[0], 0, 0, 0);
    if (!embed) A.acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, (double)rs[kBatchMaxDim], A.acc[3], 0, 0, 0);
    if (NT > 1) {
      const double a1 = (double)rs[16 + (lane & 15)];
      A.acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a0, A.acc[1], 0, 0, 0);
      A.acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a1, A.acc[2], 0, 0, 0);
      if (!embed) A.acc[4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, (double)rs[kBatchMaxDim], A.acc[4], 0, 0, 0);
    }
    // four steps' operands in flight at a time, not all sixteen
    if ((st & 3) == 3) __builtin_amdgcn_sched_barrier(0);
  }
  wave_sync();
}

// The wave's 64 staged samples as MFMA B operands: breg[q][k] = coordinate 4 k + l / 16
// of sample 16 q + l % 16 (lane l).
template <class T, int P>
__device__ inline void batch_operands(const T* stage, int lane, T (&breg)[4][P / 4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int k = 0; k < P / 4; ++k) breg[q][k] = stage[(16 * q + (lane & 15)) * kBatchStage + 4 * k + (lane >> 4)];
}

// Regression log-likelihood of the wave's 64 samples (lane = sample), from their B
// operands: eta = X_data beta on MFMA over 16-row tiles, epilogue per element, FP32
// partial sums folded into FP64 every 8 tiles.
template <class T, int P>
__device__ inline double batch_target(const BatchJob<T>& J, const T (&breg)[4][P / 4], int lane) {
  typedef MfmaTile<T> M;
  typedef T V4 __attribute__((ext_vector_type(4)));
  constexpr int KS = P / 4;
  const T* xp = J.xpack + lane;
  const V4* yp = reinterpret_cast<const V4*>(J.ypack) + lane;
  const bool logistic = J.tkind == kTLogistic;
  double ll[4] = {0.0, 0.0, 0.0, 0.0};
  T part[4] = {T(0), T(0), T(0), T(0)};
  T an[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) an[k] = xp[k * 64];
  V4 yn = yp[0];
  for (int64_t rt = 0; rt < J.row_tiles; ++rt) {
    T av[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) av[k] = an[k];
    const V4 yv = yn;
    if (rt + 1 < J.row_tiles) {
      const T* xn = xp + (rt + 1) * (KS * 64);
#pragma unroll
      for (int k = 0; k < KS; ++k) an[k] = xn[k * 64];
      yn = yp[(rt + 1) * 64];
    }
    typename M::acc_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = M::zero();
#pragma unroll
    for (int k = 0; k < KS; ++k)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = M::mma(av[k], breg[q][k], acc[q]);
    if (!logistic) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const T res = yv[r] - acc[q][r];
          part[q] += res * res;
        }
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (16 * rt + M::row(lane, r) < J.n_rows) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const T eta = acc[q][r];
            part[q] += yv[r] * eta - log1pexp(eta);
          }
        }
      }
    }
    if ((rt & 7) == 7) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ll[q] += (double)part[q];
        part[q] = T(0);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ll[q] += (double)part[q];
    ll[q] += __shfl_xor(ll[q], 16, 64);
    ll[q] += __shfl_xor(ll[q], 32, 64);
  }
  const int qq = lane >> 4;
  const double mine = qq == 0 ? ll[0] : qq == 1 ? ll[1] : qq == 2 ? ll[2] : ll[3];
  return logistic ? mine : -J.inv_two_var * mine;
}

// FP32 linear regression, P >= 8: the log-likelihood of the wave's 64 staged samples
// about the centre x0, sum_i (r0_i - x_i . delta)^2 with r0 = y - X x0 (FP64, packed) and
// delta = beta - x0, x_i . delta as three BF16 products on v_mfma_f32_16x16x32_bf16
// (xh dh + xh dl + xl dh, each operand split into a BF16 value and the BF16 rounding of
// its remainder), FP32 residual sums folded into FP64 every 8 row tiles.
template <int P>
__device__ inline double batch_target_bf16(const BatchJob<float>& J, const float* stage, int lane) {
  typedef float V4 __attribute__((ext_vector_type(4)));
  // B operands: lane l, sample 16 q + l % 16, coordinates 8 (l / 16) + e
  bf16x8_t bh[4], bl[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float* row = stage + (16 * q + (lane & 15)) * kBatchStage;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int c = 8 * (lane >> 4) + e;
      float dv = 0.0f;
      if (c < P) dv = (float)((double)row[c] - J.x0[c]);
      const __bf16 h = (__bf16)dv;
      bh[q][e] = h;
      bl[q][e] = (__bf16)(dv - (float)h);
    }
  }
  const u32x4_t* xp = reinterpret_cast<const u32x4_t*>(J.xbf) + lane;
  // the residuals at the centre, one per row: this lane's output rows 4 (l / 16) .. + 3
  const V4* yp = reinterpret_cast<const V4*>(J.ypack) + (lane >> 4);
  double ll[4] = {0.0, 0.0, 0.0, 0.0};
  float part[4] = {0.f, 0.f, 0.f, 0.f};
  u32x4_t nh = xp[0], nl = xp[64];
  V4 yn = yp[0];
  for (int64_t rt = 0; rt < J.row_tiles; ++rt) {
    const bf16x8_t ah = __builtin_bit_cast(bf16x8_t, nh), al = __builtin_bit_cast(bf16x8_t, nl);
    const V4 r0 = yn;
    if (rt + 1 < J.row_tiles) {
      nh = xp[(rt + 1) * 128];
      nl = xp[(rt + 1) * 128 + 64];
      yn = yp[(rt + 1) * 4];
    }
    float4_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[q], (float4_t){0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[q], acc[q], 0, 0, 0);
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[q], acc[q], 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float res = r0[r] - acc[q][r];
        part[q] = fmaf(res, res, part[q]);
      }
    if ((rt & 7) == 7) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ll[q] += (double)part[q];
        part[q] = 0.f;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ll[q] += (double)part[q];
    ll[q] += __shfl_xor(ll[q], 16, 64);
    ll[q] += __shfl_xor(ll[q], 32, 64);
  }
  const int qq = lane >> 4;
  const double mine = qq == 0 ? ll[0] : qq == 1 ? ll[1] : qq == 2 ? ll[2] : ll[3];
  return -J.inv_two_var * mine;
}

// FP32 jobs: r^2 = |M (x - mu)|^2 of the wave's 64 samples under one slot on
// v_mfma_f32_16x16x4f32 (A = the slot's factor fragments, B = the samples' centred
// coordinates, four 16-sample tiles q), the squares summed over each lane's tile rows
// and then over the four lanes of a sample by a two-step exchange that leaves lane l
// with the sum of its own sample (16 q + l % 16 with q = l / 16).
template <int P>
__device__ inline float batch_mfma_r2(const float* frag, const float* mu, const float (&bx)[4][P / 4], int lane) {
  constexpr int KS = P / 4, KT = (P + 15) / 16;
  float4_t acc[4][KT];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int i = 0; i < KT; ++i) acc[q][i] = (float4_t){0.f, 0.f, 0.f, 0.f};
  float bm[KS];
#pragma unroll
  for (int kk = 0; kk < KS; ++kk) bm[kk] = mu[4 * kk + (lane >> 4)];
  int pidx = 0;
#pragma unroll
  for (int i = 0; i < KT; ++i)
#pragma unroll
    for (int kk = 0; kk < (4 * (i + 1) < KS ? 4 * (i + 1) : KS); ++kk) {
      const float a = frag[pidx * 64 + lane];
      ++pidx;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        acc[q][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bx[q][kk] - bm[kk], acc[q][i], 0, 0, 0);
    }
  float r2q[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    float v = 0.0f;
#pragma unroll
    for (int i = 0; i < KT; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) v = fmaf(acc[q][i][r], acc[q][i][r], v);
    r2q[q] = v;
  }
  // lanes l and l ^ 32 swap halves (keep tiles q with q / 2 = l / 32), then l and l ^ 16
  const bool h5 = (lane & 32) != 0, h4 = (lane & 16) != 0;
  float k0 = h5 ? r2q[2] : r2q[0], k1 = h5 ? r2q[3] : r2q[1];
  k0 += __shfl_xor(h5 ? r2q[0] : r2q[2], 32, 64);
  k1 += __shfl_xor(h5 ? r2q[1] : r2q[3], 32, 64);
  const float keep = h4 ? k1 : k0;
  return keep + __shfl_xor(h4 ? k0 : k1, 16, 64);
}

// Log density of a slot from r^2 (the finite-range values of radial_logdensity); a
// non-finite r^2 gives NaN (the sample is then failed as an invalid proposal).  For
// r^2 below the smallest normal, log1p of the ratio is as far below one unit in the
// last place of the log normaliser as the density kernels' rescaled sum.
__device__ inline float batch_radial_r2(const Family<float>& fam, float lognorm, float r2) {
  if (!fam.student) return isfinite(r2) ? lognorm - 0.5f * r2 : __builtin_nanf("");
  const float ratio = r2 / fam.dof;
  float radial;
  if (isfinite(ratio))
    radial = log1p(ratio);
  else if (isfinite(r2))
    radial = log(r2) - log(fam.dof);
  else
    return __builtin_nanf("");
  return lognorm - fam.half_dof_plus_d * radial;
}

// Wave 0 (FP32 jobs): the factor fragments of one slot from its FP64 factor, row r in
// lane r's Lrow (entries past the diagonal unread) and 1 / L_rr in lane r's inv.  Lane c
// solves L m = e_c by forward substitution in FP64, m_r = (delta_rc - sum_{k<r} L_rk m_k)
// / L_rr (the order of the density kernels' solve), reading L_rk from lane r with
// v_readlane (no memory operations, so nothing can hoist a row's reads ahead of the
// previous rows' arithmetic); the columns are transposed through LDS (Ms, P x P floats)
// into the fragments.
template <int P>
__device__ inline void batch_frags(const double (&Lrow)[P], double inv, float* Ms, float* frag, int lane) {
  double m[P];
#pragma unroll
  for (int r = 0; r < P; ++r) {
    double v = (r == lane) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < r; ++k) v -= readlane_f64(Lrow[k], r) * m[k];
    m[r] = v * readlane_f64(inv, r);
  }
  if (lane < P) {
#pragma unroll
    for (int r = 0; r < P; ++r) Ms[r * P + lane] = (float)m[r];
  }
  wave_sync();
  int pidx = 0;
#pragma unroll
  for (int i = 0; i < (P + 15) / 16; ++i) {
    const int row = 16 * i + (lane & 15);
#pragma unroll
    for (int kk = 0; kk < (4 * (i + 1) < P / 4 ? 4 * (i + 1) : P / 4); ++kk) {
      frag[pidx * 64 + lane] = row < P ? Ms[row * P + 4 * kk + (lane >> 4)] : 0.0f;
      ++pidx;
    }
  }
  wave_sync();
}

// All waves, round end: the W wave records rec(w) = [mw, s0, s2, s1(32), tiles 3 x 256]
// (doubles in wave w's stage) merged entry by entry into rec(0), each entry of the
// first-order sums and moment tiles scaled by exp(mw - M) and summed in wave order;
// thread tid owns entries tid, tid + 64 W, ...  The summary words stay for batch_finish.
template <int W>
__device__ inline void batch_merge(double* rec0, int rec_stride, int tid) {
  const double ninf = -__builtin_huge_val();
  double M = rec0[0];
#pragma unroll
  for (int w = 1; w < W; ++w) M = fmax(M, rec0[w * rec_stride]);
  double f[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const double mw = rec0[w * rec_stride];
    f[w] = (mw == ninf) ? 0.0 : exp(mw - M);
  }
  for (int e = 3 + tid; e < 35 + 3 * 256; e += 64 * W) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) v += f[w] * rec0[w * rec_stride + e];
    rec0[e] = v;
  }
}

// Wave 0, after batch_merge: the round summary and the fit of slot t + 1.
// FP32 jobs also get slot t + 1's factor fragments (frags: the LDS fragment area;
// scratch: P x P floats of LDS, free at this point).
template <class T, int W, int P>
__device__ inline void batch_finish(const BatchJob<T>& J, int t, double* rec0, int rec_stride, T* slots_lds,
                                    float* frags, float* scratch, int lane) {
  const int d = J.d;
  const int SS = J.slot_elems;
  const double ninf = -__builtin_huge_val();
  double M = rec0[0];
#pragma unroll
  for (int w = 1; w < W; ++w) M = fmax(M, rec0[w * rec_stride]);
  double s0 = 0.0, s2 = 0.0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const double mw = rec0[w * rec_stride];
    const double f = (mw == ninf) ? 0.0 : exp(mw - M);
    s0 += f * rec0[w * rec_stride + 1];
    s2 += f * f * rec0[w * rec_stride + 2];
  }
  // summary (src/execution.jl:176-186)
  const int64_t count = (int64_t)(t + 1) * J.N;
  double ess, logz;
  if (s0 == 0.0 || M == ninf) {
    ess = 0.0;
    logz = ninf;
  } else {
    ess = (s0 * s0) / s2;
    logz = M + log(s0) - log((double)count);
  }
  if (lane == 0) {
    J.round_ess[t] = ess;
    J.round_logz[t] = logz;
  }
  if (!(s0 > 0.0) || !isfinite(s0) || M == ninf) {
    if (lane == 0) J.fit_status[t] = 1;
    return;
  }
  // ---- FP64 fit about the centre c = mu_t (the staged T values)
  const T* slot_t = slots_lds + t * SS;
  const bool row = lane < d;
  const double* prev = J.slots64 + (int64_t)t * SS;
  const double mean = row ? rec0[3 + lane] / s0 : 0.0;
  const double cr = row ? (double)slot_t[lane] : 0.0;
  const double mult = J.student ? J.dof / (J.dof - 2.0) : 1.0;
  // scale-aware ridge (src/adaptive_gaussian_execution.jl:283-301): sqrt(eps) tr(L L^T) mult / d
  double tr = 0.0;
  if (row) {
    const double* lr = prev + 2 * P + bslot_rowoff(lane);
    for (int c = 0; c <= lane; ++c) tr += lr[c] * lr[c];
  }
  const double reg = 1.4901161193847656e-08;  // sqrt(eps(Float64))
  double ridge = reg * (wave_sum(tr) * mult) / (double)d;
  if (!isfinite(ridge)) {
    const double sc = sqrt(reg) * sqrt(mult / (double)d);
    double part = 0.0;
    if (row) {
      const double* lr = prev + 2 * P + bslot_rowoff(lane);
      for (int c = 0; c <= lane; ++c) {
        const double v = lr[c] * sc;
        part += v * v;
      }
    }
    ridge = wave_sum(part);
  }
  // centred covariance, row `lane`, in registers
  double a[P];
#pragma unroll
  for (int c = 0; c < P; ++c) {
    a[c] = 0.0;
    if (c < d && row && c <= lane) {
      const int r = lane;
      const int ti = r >> 4, tj = c >> 4, ra = r & 15, cb = c & 15;
      const int p = ti * (ti + 1) / 2 + tj;
      const int ln = cb + 16 * (ra & 3), reg_i = ra >> 2;
      const double S = rec0[35 + p * 256 + ln * 4 + reg_i];
      double v = S / s0 - mean * (rec0[3 + c] / s0);
      if (c == lane) v += ridge;
      a[c] = v;
    }
  }
  // Cholesky, left-looking by columns (the order of _gaussian_potrf!,
  // src/adaptive_gaussian_execution.jl:427-434); row c of L sits in lane c
  int status = 0;
#pragma unroll
  for (int c = 0; c < P; ++c) {
    if (c < d) {
      double v = a[c];
#pragma unroll
      for (int k = 0; k < c; ++k) v -= a[k] * readlane_f64(a[k], c);
      if (!(lane >= c && row)) v = 0.0;
      const double piv = readlane_f64(v, c);
      if (!(piv > 0.0) || !isfinite(piv)) status = 2;
      const double dg = sqrt(piv);
      if (row && lane > c) a[c] = v / dg;
      if (lane == c) a[c] = dg;
    }
  }
  // outputs: lane r < d writes mu_r, 1 / L_rr and row r of L (Student-t
  // rescaled); lanes d <= r < P write the zero padding (mu 0, row 0, 1/L_rr 1)
  const double scale = J.student ? sqrt(1.0 / mult) : 1.0;
  double* o64 = J.slots64 + (int64_t)(t + 1) * SS;
  T* oT = J.slotsT + (int64_t)(t + 1) * SS;
  T* oL = (t + 1 < J.rounds) ? slots_lds + (t + 1) * SS : nullptr;
  bool ok = true;
  double lg = 0.0;
  double Lrow[P], inv_out = 1.0;  // the new factor's row `lane` and 1 / L_rr (FP32 fragments)
#pragma unroll
  for (int c = 0; c < P; ++c) Lrow[c] = 0.0;
  if (lane < P) {
    double mu = 0.0, invd = 1.0;
    if (row) {
      double diag = 0.0;
#pragma unroll
      for (int c = 0; c < P; ++c)
        if (c == lane) diag = a[c] * scale;
      mu = cr + mean;
      ok = isfinite(mu) && diag > 0.0;
      lg = log(fabs(diag));
      invd = 1.0 / diag;
    }
    o64[lane] = mu;
    oT[lane] = (T)mu;
    // a reciprocal that is not normal in T: NaN, so the job fails with an invalid proposal
    o64[P + lane] = slot_inv_value<double>(invd);
    oT[P + lane] = slot_inv_value<T>(invd);
    if (oL) {
      oL[lane] = (T)mu;
      oL[P + lane] = slot_inv_value<T>(invd);
      inv_out = slot_inv_value<double>(invd);
    }
    const int base = 2 * P + bslot_rowoff(lane);
#pragma unroll
    for (int c = 0; c < P; ++c) {
      if (c <= lane) {
        const double v = row ? a[c] * scale : 0.0;
        ok &= isfinite(v);
        o64[base + c] = v;
        oT[base + c] = (T)v;
        if (oL) oL[base + c] = (T)v;
        Lrow[c] = v;
      }
    }
  }
  const double logabsdet = wave_sum(lg);
  const bool all_ok = __all(ok);
  if (lane == 0) {
    const double lognorm =
        J.student ? J.logconstant - logabsdet : -0.5 * (double)d * log(2.0 * 3.141592653589793) - logabsdet;
    int st = status;
    if (st == 0 && (!all_ok || !isfinite(lognorm))) st = 3;
    const int off = bslot_lognorm(P);
    o64[off] = lognorm;
    oT[off] = (T)lognorm;
    if (oL) oL[off] = (T)lognorm;
    J.fit_status[t] = st;
  }
  if (oL && frags) batch_frags<P>(Lrow, inv_out, scratch, frags + (t + 1) * bfrag_elems(P), lane);
}

// One job, all rounds, by the whole workgroup; P = the job's padded dimension.
template <class T, int W, int P>
__device__ inline void batch_job(const BatchJob<T>& J, uint64_t run_hi, T* lds_t, int slots_off) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const T ninf = -T(__builtin_huge_val());
  const int SS = J.slot_elems;
  T* stage = lds_t + wave * 64 * kBatchStage;
  T* slots = lds_t + slots_off;
  const RngKey key{J.key0, J.key1};
  Family<T> fam;
  fam.student = J.student;
  fam.dof = J.fdof;
  fam.half_dof_plus_d = J.fhalf;
  // FP32: the slots' factor fragments after the slots; wave 1's stage is the LDS
  // scratch of their transposes (free whenever wave 0 computes them)
  constexpr bool kMfma = sizeof(T) == 4;
  constexpr int FE = bfrag_elems(P);
  float* frags = kMfma ? reinterpret_cast<float*>(slots + J.rounds * SS) : nullptr;
  float* scratch = reinterpret_cast<float*>(lds_t + 64 * kBatchStage);
  // slot 0 (set from the committed proposal by k_batch_begin)
  for (int q = tid; q < SS; q += 64 * W) slots[q] = J.slotsT[q];
  if (kMfma && wave == 0) {
    const double* s64 = J.slots64;
    double Lrow[P], inv = 1.0;
#pragma unroll
    for (int c = 0; c < P; ++c) Lrow[c] = 0.0;
    if (lane < P) {
#pragma unroll
      for (int c = 0; c < P; ++c)
        if (c <= lane) Lrow[c] = s64[2 * P + bslot_rowoff(lane) + c];
      inv = s64[P + lane];
    }
    batch_frags<P>(Lrow, inv, scratch, frags, lane);
  }
  __syncthreads();
  for (int t = 0; t < J.rounds; ++t) {
    const uint64_t rng_word = run_hi | (uint64_t)t;
    const T* slot_t = slots + t * SS;
    const int64_t first = (int64_t)t * J.N;
    const T logtotal = (T)log((double)((int64_t)(t + 1) * J.N));
    unsigned long long* fail_t = J.fail + 3 * t;
    T mw = ninf;
    BatchAcc A;
    A.s0 = 0.0;
    A.s2 = 0.0;
#pragma unroll
    for (int p = 0; p < kBatchTiles; ++p) A.acc[p] = (double4_t){0.0, 0.0, 0.0, 0.0};

    // ---- pass A: the round's new samples, 64 per wave at a time
    for (int64_t j0 = 64 * (int64_t)wave; j0 < J.N; j0 += 64 * W) {
      const int64_t j = j0 + lane;
      const bool active = j < J.N;
      T x[P];
      bool ok = active;
      {
        double mult = 1.0;
        if (J.student) {
          if (!batch_multiplier(key, (uint64_t)j, rng_word, J.us, J.dof, &mult)) {
            if (active) record_failure(fail_t, j + 1, 0, kDrawExhausted);
            mult = __builtin_nan("");
            ok = false;
          }
        }
        // x = mu + L z, accumulated column by column as each Philox block
        // yields four normals: row r still receives mu_r + L_r0 z_0 + L_r1 z_1
        // + ... in that order (the order of slot_draw).  One block at a time.
        const T* Ls = slot_t + 2 * P;
#pragma unroll
        for (int r = 0; r < P; ++r) x[r] = slot_t[r];
#pragma unroll
        for (int blk = 0; blk < P / 4; ++blk) {
          const U64x4 w = philox4x64((uint64_t)j, rng_word, kPurposeNormals, (uint64_t)blk, key.k0, key.k1);
          T zq[4];
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            if constexpr (sizeof(T) == 4) {
              // FP32 jobs: the Box-Muller pair in FP32 from the same 53-bit uniforms
              // (relative differences from the FP64 pair of order 1e-7)
              const float r = sqrtf(-2.0f * logf((float)u53_open(w.v[2 * p])));
              float sn, cs;
              sincospif(2.0f * (float)u53_open(w.v[2 * p + 1]), &sn, &cs);
              const float m = (float)mult;
              zq[2 * p] = r * cs * m;
              zq[2 * p + 1] = r * sn * m;
            } else {
              const double r = sqrt(-2.0 * log(u53_open(w.v[2 * p])));
              double sn, cs;
              sincospi(2.0 * u53_open(w.v[2 * p + 1]), &sn, &cs);
              zq[2 * p] = T(r * cs * mult);
              zq[2 * p + 1] = T(r * sn * mult);
            }
          }
          // rows r >= 4 blk: the row's four entries of this column block in one
          // 16-byte-aligned LDS read (entries past the diagonal are not added),
          // a few rows' reads in flight at a time
          typedef typename Vec16<T>::type V;
          constexpr int E = Vec16<T>::E;
          constexpr int RG = 16 / E;  // rows per scheduling group
#pragma unroll
          for (int r = 4 * blk; r < P; ++r) {
            V ch[4 / E];
#pragma unroll
            for (int h = 0; h < 4 / E; ++h)
              ch[h] = *reinterpret_cast<const V*>(Ls + bslot_rowoff(r) + 4 * blk + E * h);
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (4 * blk + e <= r) x[r] += ch[e / E][e % E] * zq[e];
            if ((r - 4 * blk) % RG == RG - 1) __builtin_amdgcn_sched_barrier(0);
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        bool valid = true;
#pragma unroll
        for (int r = 0; r < P; ++r) valid &= isfinite(x[r]);
        if (!active) {
#pragma unroll
          for (int r = 0; r < P; ++r) x[r] = T(0);
        }
        T* rowp = stage + lane * kBatchStage;
#pragma unroll
        for (int c = 0; c < P; ++c) rowp[c] = x[c];
        if (active) {
          T* xb = J.X + bx_base(first + j, P);
#pragma unroll
          for (int c = 0; c < P; ++c) xb[c * 64] = x[c];
          if (!valid) record_failure(fail_t, j + 1, 0, kGeneratedNonfinite);
        }
        ok = ok && valid;
      }
      // target on MFMA for the wave's 64 samples (coordinates staged in LDS)
      wave_sync();
      T bx[4][P / 4];
      double ll;
      if constexpr (kMfma && P >= 8) {
        if (J.tbf16) {
          ll = batch_target_bf16<P>(J, stage, lane);
          batch_operands<T, P>(stage, lane, bx);
        } else {
          batch_operands<T, P>(stage, lane, bx);
          ll = batch_target<T, P>(J, bx, lane);
        }
      } else {
        batch_operands<T, P>(stage, lane, bx);
        ll = batch_target<T, P>(J, bx, lane);
      }
      wave_sync();
      double b2 = 0.0;
#pragma unroll
      for (int c = 0; c < P; ++c) b2 += (double)x[c] * (double)x[c];
      const T tv = (T)(ll - b2 * J.inv_two_prior + J.tconst);
      if (ok) {
        const uint16_t tr = target_reason(tv);
        if (tr) {
          record_failure(fail_t, j + 1, 0, tr);
          ok = false;
        }
      }
      // deterministic mixture over slots 0..t (src/mis_execution.jl:30-95); FP32:
      // the whitened coordinates of the whole wave on MFMA, per slot
      T m = ninf, sum = T(0), sumc = T(0);  // FP32: compensated sum
      bool genbad = false;
      if constexpr (kMfma) {
        for (int k = 0; k <= t; ++k) {
          const T* sk = slots + k * SS;
          const T r2 = batch_mfma_r2<P>(frags + k * FE, sk, bx, lane);
          const T q = batch_radial_r2(fam, sk[bslot_lognorm(P)], r2);
          if (k == t && (q != q || q == ninf)) genbad = true;
          lse_push(J.logN + q, m, sum, sumc);
        }
        sum += sumc;
      } else if (ok) {
        T y[P];
        for (int k = 0; k <= t; ++k) {
          const T q = bslot_logdensity<T, P>(slots + k * SS, fam, x, y);
          if (k == t && (q != q || q == ninf)) genbad = true;
          const T v = J.logN + q;
          if (v > m) {
            sum = sum * LseAcc<T>::fast_exp(m - v) + T(1);
            m = v;
          } else {
            sum += LseAcc<T>::fast_exp(v - m);
          }
        }
      }
      T ltv = ninf, lnv = T(0), lwv = ninf;
      if (ok) {
        const T value = m + log(sum);
        if (genbad || value != value || value == ninf) {
          record_failure(fail_t, j + 1, 0, kProposalInvalid);
        } else {
          const T w = tv - (value - logtotal);
          const uint16_t wr = logweight_reason(w);
          if (wr) {
            record_failure(fail_t, j + 1, 0, wr);
          } else {
            ltv = tv;
            lnv = value;
            lwv = w;
          }
        }
      }
      if (active) {
        J.lt[first + j] = ltv;
        J.ln[first + j] = lnv;
        J.lw[first + j] = lwv;
      }
      batch_accumulate<T, P>(active ? lwv : ninf, x, slot_t, stage, lane, mw, A);
    }

    // ---- pass B: append component t to the t N retained samples
    for (int64_t base = 64 * (int64_t)wave; base < first; base += 64 * W) {
      const int64_t i = base + lane;
      const bool v = i < first;
      T cx[P];
      {
        const T* xb = J.X + bx_base(v ? i : 0, P);
#pragma unroll
        for (int c = 0; c < P; ++c) cx[c] = v ? xb[c * 64] : T(0);
      }
      const T cln = v ? J.ln[i] : T(0);
      const T clt = v ? J.lt[i] : ninf;
      T qm = T(0);
      if constexpr (kMfma) {
        // B operands straight from the sample pool: base is a multiple of 64, so the
        // wave's samples are one 64-sample block (lanes past `first` read samples of
        // round t, whose columns are ignored)
        const T* xblk = J.X + bx_base(base, P);
        T bb[4][P / 4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int k = 0; k < P / 4; ++k) bb[q][k] = xblk[64 * (4 * k + (lane >> 4)) + 16 * q + (lane & 15)];
        qm = batch_radial_r2(fam, slot_t[bslot_lognorm(P)], batch_mfma_r2<P>(frags + t * FE, slot_t, bb, lane));
      }
      T lwv = ninf;
      if (v) {
        T q;
        if constexpr (kMfma) {
          q = qm;
        } else {
          T y[P];
          q = bslot_logdensity<T, P>(slot_t, fam, cx, y);
        }
        uint16_t reason = (q != q) ? kProposalInvalid : 0;
        T num = cln;
        if (!reason) {
          const T value = logaddexp<T>(num, J.logN + q);
          reason = proposal_reason(value);
          if (!reason) {
            num = value;
            J.ln[i] = value;
          }
        }
        if (reason) record_failure(fail_t, i + 1, 0, reason);
        const T w = clt - (num - logtotal);
        const uint16_t wr = logweight_reason(w);
        if (wr) {
          record_failure(fail_t, i + 1, 0, wr);
        } else {
          J.lw[i] = w;
          lwv = w;
        }
      }
      batch_accumulate<T, P>(lwv, cx, slot_t, stage, lane, mw, A);
    }

    // ---- round end: wave records, merge and fit on wave 0
    const double s0 = wave_sum(A.s0);
    const double s2 = wave_sum(A.s2);
    __syncthreads();  // every wave is done with its stage and the slots of this round
    const int rec_stride = 64 * kBatchStage * (int)sizeof(T) / 8;
    double* rec = reinterpret_cast<double*>(lds_t) + wave * rec_stride;
    if (lane == 0) {
      rec[0] = (double)mw;
      rec[1] = s0;
      rec[2] = s2;
    }
    // tile entry (row i, column j) sits in lane j + 16 (i & 3), register i >> 2
    if constexpr (P % 16 != 0) {
      // first-order sums: row P % 16 of tile (0, 0) (P < 16) or of tiles (1, 0), (1, 1)
      constexpr int ra = P & 15;
      if ((lane >> 4) == (ra & 3)) {
        const int cb = lane & 15;
        if (P < 16) {
          rec[3 + cb] = cb < P ? A.acc[0][ra >> 2] : 0.0;
          rec[3 + 16 + cb] = 0.0;
        } else {
          rec[3 + cb] = A.acc[1][ra >> 2];
          rec[3 + 16 + cb] = 16 + cb < P ? A.acc[2][ra >> 2] : 0.0;
        }
      }
    } else if ((lane & 15) == 0) {
      // first-order sums: column 0 of tiles 3 and 4, entry (row i, column 0) in lane 16 (i & 3)
      const int g = lane >> 4;
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int q = 0; q < 4; ++q) rec[3 + 16 * ti + 4 * q + g] = A.acc[3 + ti][q];
    }
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int i = 0; i < 4; ++i) rec[35 + p * 256 + lane * 4 + i] = A.acc[p][i];
    __syncthreads();
    batch_merge<W>(reinterpret_cast<double*>(lds_t), rec_stride, tid);
    __syncthreads();
    if (wave == 0)
      batch_finish<T, W, P>(J, t, reinterpret_cast<double*>(lds_t), rec_stride, slots, frags, scratch, lane);
    __syncthreads();
  }
}

// Padded dimensions with their own job runner (every multiple of 4 up to 32).
#define IS_BATCH_DIMS(X) X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32)

// Persistent kernel: every workgroup pulls jobs from the queue until it is empty.
template <class T, int W>
__global__ __launch_bounds__(64 * W) void k_batch_amis(BatchArgs<T> a) {
  extern __shared__ double lds_raw[];
  __shared__ int item_s;
  T* lds_t = reinterpret_cast<T*>(lds_raw);
  const BatchJob<T>* __restrict__ jobs = a.jobs;
  const uint64_t run_hi = *a.run_hi;
  for (;;) {
    if (threadIdx.x == 0) item_s = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const int item = item_s;
    __syncthreads();
    if (item >= a.njobs) break;
    const int jid = __builtin_amdgcn_readfirstlane(a.order[item]);
    // the descriptor is uniform: copy it word by word through the constant
    // address space so it lands in scalar registers
    BatchJob<T> J;
    {
      const cptr<uint32_t> src = as_const(reinterpret_cast<const uint32_t*>(jobs + jid));
      uint32_t* dst = reinterpret_cast<uint32_t*>(&J);
#pragma unroll
      for (int q = 0; q < (int)(sizeof(BatchJob<T>) / 4); ++q) dst[q] = src[q];
    }
    switch (J.dp) {
#define IS_BATCH_CASE(v)                               \
  case v:                                              \
    batch_job<T, W, v>(J, run_hi, lds_t, a.slots_off); \
    break;
      IS_BATCH_DIMS(IS_BATCH_CASE)
#undef IS_BATCH_CASE
      default:
        break;
    }
  }
}

// Before the persistent launch: clear each job's failure words and fit
// status, put the committed proposal into slot 0, reset the queue head.
template <class T>
__global__ void k_batch_begin(const BatchJob<T>* jobs, int njobs, unsigned int* queue) {
  const int j = blockIdx.x;
  if (j == 0 && threadIdx.x == 0) *queue = 0u;
  if (j >= njobs) return;
  const BatchJob<T>& J = jobs[j];
  for (int q = threadIdx.x; q < 3 * J.rounds; q += blockDim.x) J.fail[q] = 0ull;
  for (int q = threadIdx.x; q < J.rounds; q += blockDim.x) J.fit_status[q] = 0;
  for (int q = threadIdx.x; q < J.slot_elems; q += blockDim.x) {
    J.slots64[q] = J.committed64[q];
    J.slotsT[q] = J.committedT[q];
  }
}

// After the launch: per-job status (0 ok, else the first failing round) and,
// when `commit`, the transactional commit of the last fit
// (src/adaptive_gaussian_execution.jl:1101-1111); advances the run word.
template <class T>
__global__ void k_batch_end(const BatchJob<T>* jobs, int njobs, int commit, int* status,
                            unsigned long long* run_word) {
  __shared__ int st_s;
  const int j = blockIdx.x;
  if (j >= njobs) return;
  const BatchJob<T>& J = jobs[j];
  if (threadIdx.x == 0) {
    int st = 0;
    for (int t = 0; t < J.rounds && st == 0; ++t)
      if (J.fail[3 * t] != 0ull || J.fit_status[t] != 0) st = t + 1;
    st_s = st;
    status[j] = st;
  }
  __syncthreads();
  if (commit && st_s == 0) {
    const int64_t off = (int64_t)J.rounds * J.slot_elems;
    for (int q = threadIdx.x; q < J.slot_elems; q += blockDim.x) {
      J.committed64[q] = J.slots64[off + q];
      J.committedT[q] = J.slotsT[off + q];
    }
  }
  if (commit && j == 0 && threadIdx.x == 0) *run_word += (1ull << 32);
}

}  // namespace issampler
