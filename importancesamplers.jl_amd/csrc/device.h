// Device-side building blocks for the gfx950 importance-sampling engine:
// counter-based RNG, LogExpFunctions-compatible log-sum-exp helpers, packed
// proposal slots and their log densities, the Student-t radial multiplier,
// built-in target log densities and the packed failure record.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace issampler {

// ---------------------------------------------------------------- constants
// Failure reason bits (src/execution.jl:758-767).
constexpr uint16_t kTargetNaN = 0x0100;
constexpr uint16_t kTargetPosInf = 0x0200;
constexpr uint16_t kProposalInvalid = 0x0400;
constexpr uint16_t kGeneratedNonfinite = 0x0800;
constexpr uint16_t kLogweightInvalid = 0x1000;
constexpr uint16_t kDrawExhausted = 0x0010;
constexpr uint16_t kMomentFitInvalid = 0x2000;  // src/adaptive_gaussian_execution.jl:9
constexpr uint16_t kDrawReasons = 0x000f | kGeneratedNonfinite | kDrawExhausted;
constexpr int kGammaAttempts = 8;  // src/execution.jl:557

// RNG stream purposes (third counter word).
constexpr uint64_t kPurposeNormals = 0;
constexpr uint64_t kPurposeRadialNormals = 1;
constexpr uint64_t kPurposeRadialUniforms = 2;
constexpr uint64_t kPurposeResampling = 3;
constexpr uint64_t kPurposeAssignment = 4;  // static MIS slot assignment, one uniform per sample

// ---------------------------------------------------------------- Philox4x64-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11).
struct U64x4 {
  uint64_t v[4];
};

__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

__host__ __device__ inline U64x4 philox4x64(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                                            uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t hi0 = mulhi64(M0, c0), lo0 = M0 * c0;
    const uint64_t hi1 = mulhi64(M1, c2), lo1 = M1 * c2;
    const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  U64x4 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

// Open-interval uniform (0, 1) from the top 53 bits (Box-Muller input).
__host__ __device__ inline double u53_open(uint64_t w) {
  return ((double)(w >> 11) + 0.5) * 0x1.0p-53;
}
// Half-open uniform [0, 1) from the top 53 bits.
__host__ __device__ inline double u53(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }

struct RngKey {
  uint64_t k0, k1;
};

// Normal number `index` of (sample, round, purpose): Box-Muller on word pairs
// (0,1) and (2,3) of the Philox block index/4.
__device__ inline double rng_normal(const RngKey& key, uint64_t sample, uint64_t round,
                                    uint64_t purpose, int index) {
  const U64x4 w = philox4x64(sample, round, purpose, (uint64_t)(index >> 2), key.k0, key.k1);
  const int lane = index & 3;
  const uint64_t a = (lane < 2) ? w.v[0] : w.v[2];
  const uint64_t b = (lane < 2) ? w.v[1] : w.v[3];
  const double r = sqrt(-2.0 * log(u53_open(a)));
  double s, c;
  sincospi(2.0 * u53_open(b), &s, &c);
  return (lane & 1) ? r * s : r * c;
}

// Fills out[0..count) with the normals of (sample, round, purpose), four per Philox call.
template <int MAXN>
__device__ inline void rng_normals(const RngKey& key, uint64_t sample, uint64_t round,
                                   uint64_t purpose, int count, double* out) {
#pragma unroll
  for (int blk = 0; blk < (MAXN + 3) / 4; ++blk) {
    if (blk * 4 < count) {
      const U64x4 w = philox4x64(sample, round, purpose, (uint64_t)blk, key.k0, key.k1);
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const double r = sqrt(-2.0 * log(u53_open(w.v[2 * p])));
        double s, c;
        sincospi(2.0 * u53_open(w.v[2 * p + 1]), &s, &c);
        const int i0 = blk * 4 + 2 * p;
        if (i0 < MAXN) out[i0] = r * c;
        if (i0 + 1 < MAXN) out[i0 + 1] = r * s;
      }
    }
  }
}

__device__ inline double rng_uniform(const RngKey& key, uint64_t sample, uint64_t round,
                                     uint64_t purpose, int index) {
  const U64x4 w = philox4x64(sample, round, purpose, (uint64_t)(index >> 2), key.k0, key.k1);
  return u53(w.v[index & 3]);
}

// ---------------------------------------------------------------- log-sum-exp helpers
// log1p(t) for t in [0, 1] as 2 atanh(s), s = t / (2 + t) <= 1/3:
// 2 s (1 + s^2/3 + s^4/5 + ...), NC terms after the leading one.  The first
// omitted term is below 1.3e-10 of the sum with NC = 8 and at FP64 rounding
// with NC = 16.  Branch-free: s from one reciprocal and Newton / exact-residual
// corrections, so that its rounding is unbiased.
template <int NC>
__device__ inline double log1p_unit(double t) {
  const double w = 2.0 + t;
  double r = __builtin_amdgcn_rcp(w);
  r = fma(r, fma(-w, r, 1.0), r);
  r = fma(r, fma(-w, r, 1.0), r);
  double sv = t * r;
  sv = fma(fma(-sv, w, t), r, sv);
  const double s2 = sv * sv;
  double q = 1.0 / (2 * NC + 1);
#pragma unroll
  for (int k = NC - 1; k >= 1; --k) q = fma(q, s2, 1.0 / (2 * k + 1));
  const double s2v = 2.0 * sv;
  return fma(s2v * s2, q, s2v);
}
__device__ inline float log1p_unit_f32(float t) {
  const float w = 2.0f + t;
  const float w_lo = t - (w - 2.0f);  // 2 + t = w + w_lo exactly
  const float r = __builtin_amdgcn_rcpf(w);
  const float s0 = t * r;
  const float sv = fmaf(fmaf(-s0, w, t) - s0 * w_lo, r, s0);
  const float s2 = sv * sv;
  float q = 1.0f / 17.0f;
#pragma unroll
  for (int k = 7; k >= 1; --k) q = fmaf(q, s2, 1.0f / (2 * k + 1));
  const float s2v = 2.0f * sv;
  return fmaf(s2v * s2, q, s2v);
}

// LogExpFunctions 1.0.1 log1pexp with the Maechler thresholds for the
// working precision (53 or 24 mantissa bits).
__device__ inline double log1pexp(double x) {
  const double x0 = -36.7368005696771;    // -53 log 2
  const double x1 = 18.021826694558577;   // 52 log 2 / 2
  const double x2 = 33.23111882352963;    // -x0 - log(-x0) (1 + 1/x0)
  if (x < x0) return exp(x);
  if (x < x1) return log1p(exp(x));
  if (x < x2) return x + exp(-x);
  return x;
}
__device__ inline float log1pexp(float x) {
  const float x0 = -16.635532f;   // -24 log 2
  const float x1 = 7.9711638f;    // 23 log 2 / 2
  const float x2 = 13.8989f;
  if (x < x0) return __expf(x);
  if (x < x1) return log1pf(__expf(x));
  if (x < x2) return x + __expf(-x);
  return x;
}

// log(1 + e^x) for x <= 0 (or NaN), as logaddexp needs it.  FP32: branch-free
// through the unbiased 8-term series (for x below -24 log 2 it returns e^x to
// within e^2x, as the threshold branch would).
__device__ inline double log1pexp_nonpos(double x) { return log1pexp(x); }
__device__ inline float log1pexp_nonpos(float x) { return log1p_unit_f32(__expf(x)); }

template <class T>
__device__ inline T logaddexp(T a, T b) {
  T diff, mx;
  if (a < b) {
    diff = a - b;
    mx = b;
  } else {
    diff = (a == b) ? T(0) : b - a;
    mx = (b == b) ? a : b;
  }
  return mx + log1pexp_nonpos(diff);
}

// Online (running max, scaled sum) log-sum-exp accumulator.
template <class T>
struct LseAcc {
  T m, s;
  __device__ inline void init() {
    m = -__builtin_huge_val();
    s = T(0);
  }
  __device__ inline void add(T a) {
    if (a > m) {
      s = s * fast_exp(m - a) + T(1);
      m = a;
    } else {
      s += fast_exp(a - m);
    }
  }
  __device__ static inline T fast_exp(T v);
  __device__ inline T value() const { return m + log(s); }
};
template <>
__device__ inline double LseAcc<double>::fast_exp(double v) {
  return exp(v);
}
template <>
__device__ inline float LseAcc<float>::fast_exp(float v) {
  return __expf(v);
}

// One term v of a running log-sum-exp (maximum m, sum s of exp(v - m)).  FP32: the
// sum carries a compensation c, Fast2Sum against s >= 1 >= exp(v - m) and, when a new
// maximum rescales the sum, the exact product error and TwoSum, so K terms cost about
// one rounding instead of K; the caller adds c to s before the logarithm.  FP64: the
// plain recurrence (c stays 0).
template <class T>
__device__ __forceinline__ void lse_push(T v, T& m, T& s, T& c) {
  if (v > m) {
    const T f = LseAcc<T>::fast_exp(m - v);
    if constexpr (sizeof(T) == 4) {
      const T a = s * f;
      const T ae = fmaf(s, f, -a);
      const T t = a + T(1);
      const T bv = t - a;
      c = c * f + (ae + ((a - (t - bv)) + (T(1) - bv)));
      s = t;
    } else {
      s = s * f + T(1);
    }
    m = v;
  } else {
    const T e = LseAcc<T>::fast_exp(v - m);
    if constexpr (sizeof(T) == 4) {
      const T t = s + e;
      c += (s - t) + e;
      s = t;
    } else {
      s += e;
    }
  }
}

// Deterministic exponential for the resampling CDF: Cody-Waite reduction with
// a two-part ln 2, degree-13 Taylor polynomial by fused multiply-add Horner,
// scaling by 2^k.  The CPU oracle evaluates the same operation sequence.
__host__ __device__ inline double det_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_huge_val();
  if (x < -745.2) return 0.0;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  const double k = rint(x * inv_ln2);
  double r = fma(-k, ln2_hi, x);
  r = fma(-k, ln2_lo, r);
  double p = 1.0 / 6227020800.0;
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  const int ki = (int)k;
  if (ki < -1000) return ldexp(ldexp(p, ki + 600), -600);
  return ldexp(p, ki);
}

// ---------------------------------------------------------------- failure record
// src/execution.jl:669-699: count, atomic max of (inverted index, inverted
// block, reason), and the same maximum restricted to draw failures.
__host__ __device__ inline uint64_t pack_failure(int64_t logical_index, int block, uint16_t reason) {
  const uint32_t inv_index = (uint32_t)(0xffffffffu - (uint32_t)logical_index + 1u);
  const uint16_t inv_block = (uint16_t)(0xffffu - (uint16_t)block + 1u);
  return ((uint64_t)inv_index << 32) | ((uint64_t)inv_block << 16) | (uint64_t)reason;
}

__device__ inline void record_failure(unsigned long long* rec, int64_t logical_index, int block,
                                      uint16_t reason) {
  const unsigned long long packed = pack_failure(logical_index, block, reason);
  atomicAdd(rec, 1ull);
  atomicMax(rec + 1, packed);
  if (reason & kDrawReasons) atomicMax(rec + 2, packed);
}

// Pointer into the constant address space: loads through it whose address is
// uniform across the wave are issued as scalar loads.
template <class T>
using cptr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ inline cptr<T> as_const(const T* p) {
  return (cptr<T>)p;
}

// Dimension instantiations: every d <= kExactDims has its own kernels (D == d,
// loops fully unrolled with no guards); larger d run in a bucket D >= d with
// the runtime d guarding each coordinate.
constexpr int kExactDims = 16;
template <int D>
__host__ __device__ constexpr int exact_dim(int d) {
  return D <= kExactDims ? D : d;
}

// ---------------------------------------------------------------- scalar sample transforms
// Forward map z -> x with log|dx/dz| (src/transforms.jl:160-235, helpers :295-333);
// false when z, x or the log Jacobian is not finite or x leaves the support.
// kind: 0 identity, 1 positive, 2 softplus, 3 interval (lo, hi; an infinite side is open).
template <class T>
__device__ __forceinline__ T tf_softplus(T z) {
  return z > T(0) ? z + log1p(exp(-z)) : log1p(exp(z));
}
template <class T>
__device__ __forceinline__ T tf_logistic(T z) {
  if (z >= T(0)) return T(1) / (T(1) + exp(-z));
  const T e = exp(z);
  return e / (T(1) + e);
}
template <class T>
__device__ __forceinline__ T tf_log_positive_difference(T upper, T lower) {
  if (lower < T(0) && upper > T(0)) {
    const T a = -lower, b = upper, larger = fmax(a, b), smaller = fmin(a, b);
    return log(larger) + log1p(smaller / larger);
  }
  return log(upper - lower);
}
// Transform failure reasons (src/transforms.jl:106-109), part of kDrawReasons.
constexpr uint16_t kTransformNonfiniteInput = 0x0001;
constexpr uint16_t kTransformNonfiniteOutput = 0x0002;
constexpr uint16_t kTransformNonfiniteLogjac = 0x0004;
constexpr uint16_t kTransformOutsideSupport = 0x0008;

// Returns the reason bits (0 on success) in the order of the reference's checks
// (_native_transform_with_logjac, src/transforms.jl:160-235).  Interval bounds are
// FP64 and, as the reference promotes to the bounds' type, the interval value,
// its log Jacobian and the support check are computed in FP64 before x and lj are
// rounded to T.
template <class T>
__device__ __forceinline__ uint16_t transform_forward(int kind, double lo, double hi, T z, T& x, T& lj) {
  x = z;
  lj = kind == 0 ? T(0) : z;
  if (!isfinite(z)) return kTransformNonfiniteInput;
  if (kind == 0) return 0;
  if (kind == 1) {
    x = exp(z);
    if (!isfinite(x)) return kTransformNonfiniteOutput;
    return x > T(0) ? 0 : kTransformOutsideSupport;
  }
  if (kind == 2) {
    x = tf_softplus(z);
    lj = z - x;
    if (!isfinite(x)) return kTransformNonfiniteOutput;
    if (!isfinite(lj)) return kTransformNonfiniteLogjac;
    return x > T(0) ? 0 : kTransformOutsideSupport;
  }
  const bool hl = isfinite(lo), hh = isfinite(hi);
  if (hl != hh) {
    // one bound: the positive transform, then lower + distance or upper - distance
    const T dist = exp(z);
    x = dist;
    if (!isfinite(dist)) return kTransformNonfiniteOutput;
    if (!(dist > T(0))) return kTransformOutsideSupport;
    const double xd = hl ? lo + (double)dist : hi - (double)dist;
    x = (T)xd;
    if (!isfinite(xd)) return kTransformNonfiniteOutput;
    return (hl ? xd > lo : xd < hi) ? 0 : kTransformOutsideSupport;
  }
  if (!hl) return 0;  // no bound (refused at preparation)
  const T p = tf_logistic(z);
  double xd;
  if (lo < 0.0 && hi > 0.0)
    xd = (double)(T(1) - p) * lo + (double)p * hi;
  else
    xd = p <= T(0.5) ? lo + (double)p * (hi - lo) : hi - (double)(T(1) - p) * (hi - lo);
  const double ljd = (tf_log_positive_difference(hi, lo) + (double)(-tf_softplus(-z))) + (double)(-tf_softplus(z));
  x = (T)xd;
  lj = (T)ljd;
  if (!isfinite(xd)) return kTransformNonfiniteOutput;
  if (!isfinite(ljd)) return kTransformNonfiniteLogjac;
  return (lo < xd && xd < hi) ? 0 : kTransformOutsideSupport;
}

// ---------------------------------------------------------------- proposal slots
// A slot packs one proposal for the density kernels:
//   [ mu(d) | inv_diag(d) | L lower triangle incl. diagonal, row-major packed | lognorm | logmass ]
// inv_diag holds 1/L_rr so the forward substitution multiplies instead of divides.
__host__ __device__ inline int slot_stride(int d) {
  const int s = 2 * d + d * (d + 1) / 2 + 2;
  return (s + 1) & ~1;
}
__host__ __device__ inline int slot_lognorm_off(int d) { return 2 * d + d * (d + 1) / 2; }
__host__ __device__ inline int slot_logmass_off(int d) { return 2 * d + d * (d + 1) / 2 + 1; }

template <class T>
struct Family {
  int student;  // 0 Gaussian, 1 Student-t
  T dof;
  T half_dof_plus_d;  // (dof + d) / 2
};

// _radial_logdensity (src/proposals.jl:24-55) for the finite-range paths.
template <class T, int D>
__device__ __forceinline__ T radial_logdensity(const Family<T>& fam, T lognorm, T r2, const T (&y)[D], int d) {
  if (!fam.student) {
    T half = T(0.5) * r2;
    if (!isfinite(r2)) {
      half = T(0);
#pragma unroll
      for (int c = 0; c < D; ++c)
        if (c < d) half += (T(0.5) * y[c]) * y[c];
    }
    return lognorm - half;
  }
  const T ratio = r2 / fam.dof;
  T radial;
  const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
  if (r2 < tiny) {
    const T s = sqrt(fam.dof);
    T scaled = T(0);
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (c < d) {
        const T v = y[c] / s;
        scaled += v * v;
      }
    radial = log1p(scaled);
  } else if (isfinite(ratio)) {
    radial = log1p(ratio);
  } else if (isfinite(r2)) {
    radial = log(r2) - log(fam.dof);
  } else {
    T s = T(0);
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (c < d) s = fmax(s, fabs(y[c]));
    if (!isfinite(s)) return T(__builtin_nan(""));
    T scaled = T(0);
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (c < d) {
        const T v = y[c] / s;
        scaled += v * v;
      }
    radial = log1pexp(T(2) * log(s) + log(scaled) - log(fam.dof));
  }
  return lognorm - fam.half_dof_plus_d * radial;
}

// ---------------------------------------------------------------- extended-range densities
// Values as (mantissa, binary exponent) pairs, so that a standardized coordinate
// whose plain value over- or underflows keeps its magnitude
// (src/student_t_range.jl:55-74: _range_subtract, _range_multiply, _range_divide).
template <class T>
struct RangeVal {
  T m;
  int e;
};
__device__ inline double rv_frexp(double v, int* e) { return frexp(v, e); }
__device__ inline float rv_frexp(float v, int* e) { return frexpf(v, e); }
__device__ inline double rv_ldexp(double v, int e) { return ldexp(v, e); }
__device__ inline float rv_ldexp(float v, int e) { return ldexpf(v, e); }
template <class T>
__device__ inline RangeVal<T> range_of(T v) {
  RangeVal<T> r;
  r.m = rv_frexp(v, &r.e);
  return r;
}
template <class T>
__device__ inline RangeVal<T> range_subtract(RangeVal<T> a, RangeVal<T> b) {
  if (a.m == T(0)) return RangeVal<T>{-b.m, b.e};
  if (b.m == T(0)) return a;
  const int ex = a.e > b.e ? a.e : b.e;
  int sh;
  const T m = rv_frexp(rv_ldexp(a.m, a.e - ex) - rv_ldexp(b.m, b.e - ex), &sh);
  return RangeVal<T>{m, m == T(0) ? 0 : ex + sh};
}
template <class T>
__device__ inline RangeVal<T> range_multiply(RangeVal<T> v, T factor) {
  int ef, sh;
  const T mf = rv_frexp(factor, &ef);
  const T p = rv_frexp(v.m * mf, &sh);
  return RangeVal<T>{p, p == T(0) ? 0 : v.e + ef + sh};
}
template <class T>
__device__ inline RangeVal<T> range_divide(RangeVal<T> v, T divisor) {
  int ed, sh;
  const T md = rv_frexp(divisor, &ed);
  const T q = rv_frexp(v.m / md, &sh);
  return RangeVal<T>{q, q == T(0) ? 0 : v.e - ed + sh};
}

// The slot density where the plain forward substitution does not give it: the
// standardized coordinates by division as the reference forms them, rows whose
// plain value is not finite (from finite inputs) recomputed in extended range
// (_repair_factor_squared_radius! / _factor_standardized_coordinate,
// src/student_t_range.jl:76-116); Gaussian: lognorm - r^2/2, or - sum (y/2) y when
// r^2 overflows (src/proposals.jl:24-34); Student-t: the log radius in extended range
// (_factor_range_logradius!, src/student_t_range.jl:137-153) through
// log1pexp(log r^2 - log nu) (src/proposals.jl:36-55).  It runs only for samples (or
// slots) at the edges of the floating-point range: its loops stay rolled and its arrays
// live in private memory (volatile), so it adds few registers to the hot loops that
// inline it.
template <class T, int D>
__device__ __forceinline__ T slot_logdensity_range(const T* s, int d, const Family<T>& fam, const T (&x)[D]) {
  const T* mu = s;
  auto L = [&](int r, int c) { return s[2 * d + r * (r + 1) / 2 + c]; };
  volatile T y[D];  // private memory: no registers taken from the hot loops
  T r2 = T(0);
#pragma nounroll
  for (int r = 0; r < d; ++r) {
    T v = x[r] - mu[r];
#pragma nounroll
    for (int c = 0; c < r; ++c) v -= L(r, c) * y[c];
    v /= L(r, r);
    y[r] = v;
    r2 += v * v;
  }
  if (!isfinite(r2)) {
    r2 = T(0);
#pragma nounroll
    for (int r = 0; r < d; ++r) {
      T v = x[r] - mu[r];
#pragma nounroll
      for (int c = 0; c < r; ++c) v -= L(r, c) * y[c];
      v /= L(r, r);
      if (!isfinite(v) && isfinite(x[r]) && isfinite(mu[r])) {
        RangeVal<T> val = range_subtract(range_of(x[r]), range_of(mu[r]));
        bool done = false;
#pragma nounroll
        for (int c = 0; c < r && !done; ++c) {
          if (!isfinite(y[c])) {
            v = y[c];  // overflow of a Gaussian tail stays an overflow
            done = true;
          } else {
            val = range_subtract(val, range_multiply(range_of(y[c]), L(r, c)));
          }
        }
        if (!done) {
          const RangeVal<T> q = range_divide(val, L(r, r));
          v = rv_ldexp(q.m, q.e);
        }
      }
      y[r] = v;
      r2 += v * v;
    }
  }
  const T lognorm = s[slot_lognorm_off(d)];
  if (!fam.student) {
    T half = T(0.5) * r2;
    if (!isfinite(r2)) {
      half = T(0);
#pragma nounroll
      for (int r = 0; r < d; ++r) half += (T(0.5) * y[r]) * y[r];
    }
    return lognorm - half;
  }
  volatile T mant[D];
  volatile int expo[D];
  T total = -T(__builtin_huge_val());
  const T ln2 = T(0.6931471805599453);
#pragma nounroll
  for (int r = 0; r < d; ++r) {
    RangeVal<T> val = range_subtract(range_of(x[r]), range_of(mu[r]));
#pragma nounroll
    for (int c = 0; c < r; ++c) val = range_subtract(val, range_multiply(RangeVal<T>{mant[c], expo[c]}, L(r, c)));
    const RangeVal<T> q = range_divide(val, L(r, r));
    mant[r] = q.m;
    expo[r] = q.e;
    const T logabs = log(fabs(q.m)) + T(q.e) * ln2;
    total = logaddexp(total, T(2) * logabs);
  }
  return lognorm - fam.half_dof_plus_d * log1pexp(total - log(fam.dof));
}

// The inverse-diagonal entry of a packed slot: 1/L_rr when that is a normal number
// in T, else NaN, which sends every density of the slot to slot_logdensity_range
// (the multiply-by-reciprocal form would lose the value there).
template <class T>
__host__ __device__ inline T slot_inv_value(double inv) {
  const T v = (T)inv;
  const T a = v < T(0) ? -v : v;
  const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
  const T big = sizeof(T) == 8 ? T(1.7976931348623157e308) : T(3.40282347e38f);
  return (a >= tiny && a <= big) ? v : T(__builtin_nan(""));
}
template <class T>
__host__ __device__ inline T slot_inv_entry(double diag) {
  return slot_inv_value<T>(1.0 / diag);
}

// Whitened squared radius r^2 = |L^{-1}(x - mu)|^2 by forward substitution
// (src/native_execution.jl:346-384), then the radial log density.
// P is `const T*` for per-lane slots, or a constant-address-space pointer
// (cptr<T>) for a slot shared by the whole wave, whose values then arrive by
// scalar loads and feed the FMAs as SGPR operands.  Samples whose r^2 is not
// finite (a Gaussian), or not finite, subnormal or with r^2/nu not finite (a
// Student-t), need slot_logdensity_range: the RANGE instantiations compute it; the
// others return NaN, which fails the run, and the host replays such a run on the
// RANGE kernels (their private arrays would cost the hot kernels occupancy).
template <class T, int D, bool RANGE = false, class P>
__device__ __forceinline__ T slot_logdensity(P s, int d, const Family<T>& fam, const T (&x)[D], T (&y)[D]) {
  T r2 = T(0);
#pragma unroll
  for (int r = 0; r < D; ++r) {
    if (r < d) {
      T v = x[r] - s[r];
      const int row = 2 * d + r * (r + 1) / 2;
#pragma unroll
      for (int c = 0; c < r; ++c) v -= s[row + c] * y[c];
      v *= s[d + r];
      y[r] = v;
      r2 += v * v;
    }
  }
  const T lognorm = s[slot_lognorm_off(d)];
  if (!fam.student) {
    if (__builtin_expect(isfinite(r2), 1)) return lognorm - T(0.5) * r2;
  } else {
    const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
    const T ratio = r2 / fam.dof;
    if (__builtin_expect(r2 >= tiny && isfinite(ratio), 1)) return lognorm - fam.half_dof_plus_d * log1p(ratio);
  }
  if constexpr (RANGE) {
    return slot_logdensity_range<T, D>((const T*)s, d, fam, x);
  } else {
    return T(__builtin_nan(""));
  }
}

// slot_logdensity for a slot staged in LDS, at an exact dimension (D == d): the
// same operations in the same order, with the slot's values read into registers in
// a few batches (centre and inverse diagonal, then blocks of L's rows) issued ahead
// of their use, so the substitution waits on a few LDS round trips rather than one
// per pair of values.
__host__ __device__ constexpr int lds_row_batch_end(int r, int D, int batch) {
  int r1 = r + 1;
  while (r1 < D && (r1 + 1) * (r1 + 2) / 2 - r * (r + 1) / 2 <= batch) ++r1;
  return r1;
}

template <class T, int D, int R>
__device__ __forceinline__ void lds_rows(const T* __restrict__ s, const T (&x)[D], const T (&mu)[D],
                                         const T (&inv)[D], T (&y)[D], T& r2) {
  constexpr int R1 = lds_row_batch_end(R, D, sizeof(T) == 4 ? 48 : 24);
  constexpr int B0 = R * (R + 1) / 2, B1 = R1 * (R1 + 1) / 2;
  T lv[B1 - B0];
#pragma unroll
  for (int i = 0; i < B1 - B0; ++i) lv[i] = s[2 * D + B0 + i];
#pragma unroll
  for (int rr = R; rr < R1; ++rr) {
    T v = x[rr] - mu[rr];
    const int row = rr * (rr + 1) / 2 - B0;
#pragma unroll
    for (int c = 0; c < rr; ++c) v -= lv[row + c] * y[c];
    v *= inv[rr];
    y[rr] = v;
    r2 += v * v;
  }
  if constexpr (R1 < D) lds_rows<T, D, R1>(s, x, mu, inv, y, r2);
}

template <class T, int D, bool RANGE = false>
__device__ __forceinline__ T slot_logdensity_lds(const T* __restrict__ s, const Family<T>& fam, const T (&x)[D],
                                                 T (&y)[D]) {
  static_assert(D <= kExactDims, "exact dimensions only");
  T mu[D], inv[D];
#pragma unroll
  for (int r = 0; r < D; ++r) {
    mu[r] = s[r];
    inv[r] = s[D + r];
  }
  T r2 = T(0);
  lds_rows<T, D, 0>(s, x, mu, inv, y, r2);
  const T lognorm = s[slot_lognorm_off(D)];
  if (!fam.student) {
    if (__builtin_expect(isfinite(r2), 1)) return lognorm - T(0.5) * r2;
  } else {
    const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
    const T ratio = r2 / fam.dof;
    if (__builtin_expect(r2 >= tiny && isfinite(ratio), 1)) return lognorm - fam.half_dof_plus_d * log1p(ratio);
  }
  if constexpr (RANGE) {
    return slot_logdensity_range<T, D>(s, D, fam, x);
  } else {
    return T(__builtin_nan(""));
  }
}

// Packed affine draw x = mu + L z (src/adaptive_gaussian_execution.jl:49-62).
template <class T, int D, class P>
__device__ __forceinline__ bool slot_draw(P s, int d, const T (&z)[D], T (&x)[D]) {
  bool valid = true;
#pragma unroll
  for (int r = 0; r < D; ++r) {
    if (r < d) {
      T v = s[r];
      const int row = 2 * d + r * (r + 1) / 2;
#pragma unroll
      for (int c = 0; c <= r; ++c) v += s[row + c] * z[c];
      x[r] = v;
      valid &= isfinite(v);
    }
  }
  return valid;
}

// ---------------------------------------------------------------- Student-t radial multiplier
// _buffered_unit_gamma / _student_t_radial_multiplier (src/native_execution.jl:147-180)
// with the gamma candidate of src/proposals.jl:738-751, in three steps shared by the
// buffered form below and the batched kernel's form (which generates its attempts
// block by block): the constants of the shape dof / 2, the acceptance test of one
// attempt whose root 1 + c normal is positive, and the multiplier of the accepted one.
struct GammaSetup {
  double shape, offset, coefficient;
};
__device__ inline GammaSetup gamma_setup(double dof) {
  GammaSetup g;
  g.shape = dof / 2.0;
  const double adjusted = g.shape < 1.0 ? g.shape + 1.0 : g.shape;
  g.offset = adjusted - 1.0 / 3.0;
  g.coefficient = (1.0 / sqrt(g.offset)) / 3.0;
  return g;
}
__device__ inline bool gamma_accepts(const GammaSetup& g, double normal, double root, double uniform) {
  const double cube = root * root * root;
  const double n2 = normal * normal;
  return uniform < 1.0 - 0.0331 * (n2 * n2) || log(uniform) < 0.5 * n2 + g.offset * (1.0 - cube + log(cube));
}
// boost(): the extra uniform, read only for a shape below 1
template <class Boost>
__device__ inline bool gamma_multiplier(const GammaSetup& g, double dof, double root, Boost boost, double* out) {
  double gamma = g.offset * (root * root * root);
  if (g.shape < 1.0) gamma *= pow(boost(), 1.0 / g.shape);
  if (!(isfinite(gamma) && gamma > 0.0)) return false;
  *out = sqrt(dof / (2.0 * gamma));
  return true;
}

// Over buffered attempts: rn[kGammaAttempts], ru[kGammaAttempts + 1].
__device__ inline bool student_t_multiplier(double dof, const double* rn, const double* ru,
                                            bool have_uniforms, double* out) {
  if (!have_uniforms || dof == 1.0) {
    const double den = fabs(rn[0]);
    if (!(isfinite(den) && den > 0.0)) return false;
    *out = 1.0 / den;
    return true;
  }
  const GammaSetup g = gamma_setup(dof);
  for (int a = 0; a < kGammaAttempts; ++a) {
    const double root = 1.0 + g.coefficient * rn[a];
    if (!(root > 0.0)) continue;
    if (!gamma_accepts(g, rn[a], root, ru[a])) continue;
    return gamma_multiplier(g, dof, root, [&] { return ru[kGammaAttempts]; }, out);
  }
  return false;
}

__host__ __device__ inline int radial_normal_stride(int student, double dof) {
  return student ? (dof == 1.0 ? 1 : kGammaAttempts) : 0;
}
__host__ __device__ inline int radial_uniform_stride(int student, double dof) {
  return student ? (dof == 1.0 ? 0 : kGammaAttempts + (dof < 2.0 ? 1 : 0)) : 0;
}

// ---------------------------------------------------------------- reasons
template <class T>
__device__ inline uint16_t target_reason(T v) {
  if (v != v) return kTargetNaN;
  if (v == T(__builtin_huge_val())) return kTargetPosInf;
  return 0;
}
template <class T>
__device__ inline uint16_t proposal_reason(T v) {
  return (v != v || v == -T(__builtin_huge_val())) ? kProposalInvalid : 0;
}
template <class T>
__device__ inline uint16_t logweight_reason(T v) {
  return (v != v || v == T(__builtin_huge_val())) ? kLogweightInvalid : 0;
}

// ---------------------------------------------------------------- wavefront reductions
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace issampler
