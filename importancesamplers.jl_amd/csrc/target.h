// Batched built-in targets: Bayesian logistic and linear regression over an
// n x p data matrix, evaluated for one round's new samples (the coefficient
// vectors beta_s).  Included from kernels.h inside namespace issampler.
//
// eta = X beta is a dense contraction and runs on MFMA (16x16x4, FP32 in FP32
// mode, FP64 in FP64 mode).  A block owns 64 samples (four 16-column tiles)
// and a slice of the data rows.  The samples' coefficients are loaded once
// into registers as the B operands.  The data rows stream through two LDS
// buffers in 64-row chunks; wave w takes a chunk's 16-row tile w.  The
// epilogue of one chunk (each eta to its log-likelihood term, accumulated per
// sample in FP64) runs beside the next chunk's MFMAs; partial sums per row
// slice are combined in a fixed order.
#pragma once

typedef float float4_t __attribute__((ext_vector_type(4)));

template <class T>
struct MfmaTile;
template <>
struct MfmaTile<float> {
  typedef float4_t acc_t;
  __device__ static inline acc_t zero() { return (acc_t){0.f, 0.f, 0.f, 0.f}; }
  __device__ static inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // output register r of lane l holds row 4 (l / 16) + r, column l % 16
  __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct MfmaTile<double> {
  typedef double4_t acc_t;
  __device__ static inline acc_t zero() { return (acc_t){0.0, 0.0, 0.0, 0.0}; }
  __device__ static inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output register r of lane l holds row 4 r + l / 16, column l % 16
  __device__ static inline int row(int lane, int r) { return 4 * r + (lane >> 4); }
};

template <class T>
struct TargetArgs {
  int kind;              // kTLogistic or kTLinreg
  int p;                 // coefficients per sample (= d)
  int64_t n;             // data rows
  const T* Xd;           // n x p, row-major
  const T* yd;           // n
  double inv_two_var;    // linear regression: 1 / (2 sigma^2)
  const T* beta;         // sample coordinates, SoA: beta[c * cap + first + s]
  int64_t cap, first, n_new;
  int64_t rows_per_split;  // multiple of 64
  double* part;          // [split][n_new] partial log-likelihoods
};

__host__ __device__ constexpr int target_row_stride(int ks) { return 4 * ks + 1; }

// One row's log-likelihood term for linear predictor eta and response y,
// branch-free.  Logistic: y eta - log(1 + e^eta) = y eta - max(eta, 0) -
// log1p(e^-|eta|); linear regression: -(y - eta)^2 / (2 sigma^2), the
// constants added once per sample.
//   FP64: e^-|eta| from exp and the 16-term series.
//   FP32: e^-|eta| = 2^(-|eta| log2 e) from the hardware base-2 exponential,
//   log2 e split in two floats (the product's rounding carried as a
//   first-order correction), and the 8-term series.  Every step rounds without
//   bias: summed over 1e5 rows, a per-row bias of a tenth of an FP32 ulp would
//   move log pi by 5e-4, and the hardware logarithm and a float log1p carry
//   biases of that size; this form's is about 4e-11 per row.
template <class T, int KIND>
__device__ inline T target_term(T y, T eta, double inv_two_var);
template <>
__device__ inline double target_term<double, kTLogistic>(double y, double eta, double) {
  return fma(y, eta, -fmax(eta, 0.0)) - log1p_unit<16>(exp(-fabs(eta)));
}
template <>
__device__ inline float target_term<float, kTLogistic>(float y, float eta, double) {
  const float a = fabsf(eta);
  const float L2E_HI = 1.44269502162933349609375f, L2E_LO = 1.925962989e-08f;
  const float e_hi = -a * L2E_HI;
  const float e_lo = fmaf(-a, L2E_HI, -e_hi) - a * L2E_LO;
  const float t0 = __builtin_amdgcn_exp2f(e_hi);
  const float t = fmaf(t0 * 0.6931471805599453f, e_lo, t0);
  // y eta - max(eta, 0) is exact for y in {0, 1}
  return fmaf(y, eta, -fmaxf(eta, 0.0f)) - log1p_unit_f32(t);
}
template <>
__device__ inline double target_term<double, kTLinreg>(double y, double eta, double inv_two_var) {
  const double res = y - eta;
  return -(res * res) * inv_two_var;
}
template <>
__device__ inline float target_term<float, kTLinreg>(float y, float eta, double inv_two_var) {
  const float res = y - eta;
  return -(res * res) * (float)inv_two_var;
}

// LDS of the target kernel: two chunk buffers, each 65 rows of PS values (row
// 64 takes the staging writes that fall outside the chunk) and 64 responses.
__host__ __device__ constexpr int target_buf_values(int ks) { return 65 * target_row_stride(ks) + 64; }
__host__ __device__ constexpr size_t target_lds_bytes(int ks, int tsize) {
  return (size_t)2 * target_buf_values(ks) * tsize > 256 * 8 ? (size_t)2 * target_buf_values(ks) * tsize
                                                                : (size_t)256 * 8;
}

// One block's work with KS = ceil(p / 4) K steps of the 16x16x4 MFMA.
// A chunk of 64 rows is the contiguous span X[r0 p, (r0 + 64) p): thread tid
// stages span values tid + 256 u (u < KS), value e going to LDS row e / p (for
// a bucket, a multiply-high by ceil(2^32 / p), exact for e p < 2^32; an exact
// instantiation divides by its constant p), column e mod p; values past the
// span go to the spare row 64.  The padding columns [p, 4 KS) stay zero.
// Software pipeline, one barrier per chunk: iteration c reads chunk c's A
// operands (buffer c mod 2), stages chunk c + 1 (loaded from global memory
// during iteration c - 1) into the other buffer, issues the loads of chunk
// c + 2, runs chunk c's MFMAs and the epilogue of chunk c - 1.  A wave does
// not issue vector work while its own MFMAs run; the matrix core and the
// vector ALU of a SIMD overlap across the waves of different blocks.  Only the
// last chunk can be short; its rows past the end are masked.
template <class T, int KS, bool EXACT, int KIND>
__device__ inline void target_tile(const TargetArgs<T>& a, int p) {
  typedef MfmaTile<T> M;
  typedef typename M::acc_t acc_t;
  constexpr int PS = target_row_stride(KS);
  constexpr int BV = target_buf_values(KS);
  extern __shared__ double lds_raw[];
  T* const lds = reinterpret_cast<T*>(lds_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double inv_two_var = a.inv_two_var;
  const int64_t s0 = (int64_t)blockIdx.x * 64;
  const int64_t r_lo = (int64_t)blockIdx.y * a.rows_per_split;
  const int64_t r_hi = min(a.n, r_lo + a.rows_per_split);
  const int64_t nchunks = r_hi > r_lo ? (r_hi - r_lo + 63) >> 6 : 0;
  const int span = 64 * p;
  const uint32_t pmagic = EXACT ? 0u : 0xffffffffu / (uint32_t)p + 1u;
  for (int q = tid; q < 2 * BV; q += 256) lds[q] = T(0);
  // staging destinations, the same for every chunk
  int soff[KS];
#pragma unroll
  for (int u = 0; u < KS; ++u) {
    const int e = tid + 256 * u;
    const int row = EXACT ? e / p : (int)__umulhi((uint32_t)e, pmagic);
    soff[u] = e < span ? row * PS + (e - row * p) : 64 * PS;
  }

  // B operands: lane l holds beta[s0 + 16 q + l % 16][4 k + l / 16]
  T breg[4][KS];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t s = s0 + 16 * q + (lane & 15);
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int c = 4 * k + (lane >> 4);
      breg[q][k] = (s < a.n_new && c < p) ? a.beta[(int64_t)c * a.cap + a.first + s] : T(0);
    }
  }
  double ll[4] = {0.0, 0.0, 0.0, 0.0};

  T pre[KS], prey;
  // loads of chunk c (clamped into the slice, so a fetch past the end is harmless)
  auto fetch = [&](int64_t c) {
    const int64_t r0 = r_lo + 64 * min(c, nchunks - 1);
    const int rows = (int)min((int64_t)64, r_hi - r0);
    const int last = rows * p - 1;
    const T* src = a.Xd + r0 * p;
#pragma unroll
    for (int u = 0; u < KS; ++u) pre[u] = src[min(tid + 256 * u, last)];
    prey = a.yd[r0 + min(lane, rows - 1)];
  };
  auto stage = [&](T* buf) {
#pragma unroll
    for (int u = 0; u < KS; ++u) buf[soff[u]] = pre[u];
    buf[65 * PS + lane] = prey;
  };
  const int arow = (16 * wave + (lane & 15)) * PS + (lane >> 4);
  // log-likelihood terms of a finished chunk, accumulated per sample in FP64
  // (a lane's four rows of a column summed in the working precision)
  auto epilogue = [&](const acc_t (&acc)[4], const T (&yv)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      T part = target_term<T, KIND>(yv[0], acc[q][0], inv_two_var);
#pragma unroll
      for (int r = 1; r < 4; ++r) part += target_term<T, KIND>(yv[r], acc[q][r], inv_two_var);
      ll[q] += (double)part;
    }
  };
  auto epilogue_masked = [&](const acc_t (&acc)[4], const T (&yv)[4], int valid_rows) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      T part = T(0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool valid = 16 * wave + M::row(lane, r) < valid_rows;
        const T term = target_term<T, KIND>(yv[r], acc[q][r], inv_two_var);
        part += valid ? term : T(0);
      }
      ll[q] += (double)part;
    }
  };
  // chunk c: its A operands and responses from LDS, the staging of chunk c + 1,
  // the loads of chunk c + 2 and its MFMAs
  auto chunk = [&](int64_t c, acc_t (&acc)[4], T (&yv)[4]) {
    const T* cur = lds + (c & 1) * BV;
    T av[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) av[k] = cur[arow + 4 * k];  // A lane l: X[16 w + l % 16][4 k + l / 16]
#pragma unroll
    for (int r = 0; r < 4; ++r) yv[r] = cur[65 * PS + 16 * wave + M::row(lane, r)];
    stage(lds + ((c + 1) & 1) * BV);
    fetch(c + 2);
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = M::zero();
#pragma unroll
    for (int k = 0; k < KS; ++k)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = M::mma(av[k], breg[q][k], acc[q]);
  };

  if (nchunks > 0) {
    acc_t accP[4], accC[4];
    T yP[4], yC[4];
    fetch(0);
    __syncthreads();  // zero fill done
    stage(lds);
    fetch(1);
    __syncthreads();
    chunk(0, accP, yP);
    __syncthreads();
    for (int64_t c = 1; c < nchunks; ++c) {
      chunk(c, accC, yC);
      epilogue(accP, yP);  // chunk c - 1, full
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; ++q) accP[q] = accC[q];
#pragma unroll
      for (int r = 0; r < 4; ++r) yP[r] = yC[r];
    }
    epilogue_masked(accP, yP, (int)(r_hi - (r_lo + 64 * (nchunks - 1))));
  }
  // sum the four row groups of each column (lanes l, l^16, l^32, l^48), then the waves
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ll[q] += __shfl_xor(ll[q], 16, 64);
    ll[q] += __shfl_xor(ll[q], 32, 64);
  }
  __syncthreads();
  double* red = lds_raw;  // 4 waves x 64 samples
  if (lane < 16) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[wave * 64 + 16 * q + lane] = ll[q];
  }
  __syncthreads();
  if (tid < 64 && s0 + tid < a.n_new) {
    const double v = ((red[tid] + red[64 + tid]) + red[128 + tid]) + red[192 + tid];
    a.part[(int64_t)blockIdx.y * a.n_new + s0 + tid] = v;
  }
}

template <class T, int KS, bool EXACT>
__device__ inline void target_kind(const TargetArgs<T>& a, int p) {
  if (a.kind == kTLogistic)
    target_tile<T, KS, EXACT, kTLogistic>(a, p);
  else
    target_tile<T, KS, EXACT, kTLinreg>(a, p);
}

// grid = (ceil(n_new / 64), splits), block = 256 (4 waves).  D is the dimension
// instantiation; a bucket (D > 16) serves the p above the next smaller
// instantiation, at most four K-step counts, each with its own unrolled body.
// FP32 asks for 3 waves per SIMD (a few registers spill outside the chunk
// loop); FP64, whose B operands take twice the registers, runs 1.
template <class T, int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1))) void k_target_gemm(
    TargetArgs<T> a) {
  constexpr int KS = (D + 3) / 4;
  const int p = exact_dim<D>(a.p);
  if constexpr (D <= kExactDims) {
    target_kind<T, KS, true>(a, p);
  } else {
    switch ((p + 3) >> 2) {
      case KS - 3: target_kind<T, KS - 3, false>(a, p); break;
      case KS - 2: target_kind<T, KS - 2, false>(a, p); break;
      case KS - 1: target_kind<T, KS - 1, false>(a, p); break;
      default: target_kind<T, KS, false>(a, p); break;
    }
  }
}

// ---------------------------------------------------------------- centred BF16 logistic target
// FP32 jobs with a logistic target and p <= 64: eta = X beta as
//   eta_i = x_i . x0 + x_i . (beta - x0),
// with x_i . x0 (the centre x0: the job's initial location) computed once per job in
// FP64, and the centred product on the BF16 matrix cores as three products of split
// operands, X = Xh + Xl and delta = beta - x0 = dh + dl (each part BF16):
//   X delta ~ Xh dh + Xh dl + Xl dh, FP32 accumulation,
// whose dropped terms are about 2^-16 |x_i| |delta|: small because delta is a
// draw's distance from the centre, not the coefficient itself.  Each row's term
// y eta - max(eta, 0) - log1p(e^-|eta|) uses the one eta, so an error in eta enters
// the sum weighted by the gradient y - sigmoid(eta), small near the posterior mode.
// Responses are 0 or 1 (other data keep the FP32 GEMM form), and rows with y = 1
// are stored negated (rows and centre term), so with eta_s = (1 - 2 y) eta the term is
// -max(eta_s, 0) - log1p(e^-|eta_s|), exactly, with no response read per row.
// The log1p terms of a lane's rows are summed as one logarithm of the product of
// u = 1 + t over eight rows, t = e^-|eta|, with each u's rounding t - (u - 1) added
// exactly (a t below half an FP32 ulp of 1 is kept that way), and the logarithm by an
// unbiased atanh series after range reduction to [sqrt(1/2), sqrt(2)).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

struct LogisticBf16Args {
  int p;                    // coefficients, <= 32 KS2
  int64_t n;                // data rows
  const u32x4_t* xfrag;     // [row tile][K step][Xh, Xl][lane]: 8 BF16 of row l % 16, columns 8 (l / 16) + e
  const float* base;        // ntiles * 16: x_i . x0 (zero past n)
  const double* x0;         // p
  const float* beta;        // samples, SoA: beta[c * cap + first + s]
  int64_t cap, first, n_new;
  int64_t ntiles;           // row tiles of 16 rows
  int64_t tiles_per_split;
  double* part;             // [split][n_new]
};

// log(v) for v in [1, 2^24): v = m 2^e with m in [sqrt(1/2), sqrt(2)), then
// 2 atanh(s), s = (m - 1) / (m + 1) in [-0.172, 0.172], through the s^12 term (the
// first omitted term is below 2e-11); s by one reciprocal and an exact-residual step,
// e ln 2 with ln 2 split in two floats.
__device__ __forceinline__ float log_unbiased_f32(float v) {
  int e;
  float m = frexpf(v, &e);  // m in [0.5, 1)
  if (m < 0.70710678f) {
    m *= 2.0f;
    e -= 1;
  }
  const float num = m - 1.0f, den = m + 1.0f;
  const float r = __builtin_amdgcn_rcpf(den);
  const float s0 = num * r;
  const float sv = fmaf(fmaf(-s0, den, num), r, s0);
  const float s2 = sv * sv;
  float q = 1.0f / 13.0f;
  q = fmaf(q, s2, 1.0f / 11.0f);
  q = fmaf(q, s2, 1.0f / 9.0f);
  q = fmaf(q, s2, 1.0f / 7.0f);
  q = fmaf(q, s2, 1.0f / 5.0f);
  q = fmaf(q, s2, 1.0f / 3.0f);
  const float s2v = 2.0f * sv;
  const float lm = fmaf(s2v * s2, q, s2v);
  const float fe = (float)e;
  return fmaf(fe, 0.693145751953125f, fmaf(fe, 1.428606765330187e-06f, lm));
}

template <int KS2>
__global__ __launch_bounds__(256) void k_logistic_bf16(LogisticBf16Args a) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t s0 = (int64_t)blockIdx.x * 64;
  const int64_t t_lo = (int64_t)blockIdx.y * a.tiles_per_split;
  const int64_t t_hi = min(a.ntiles, t_lo + a.tiles_per_split);
  // B operands: lane l holds delta[8 (l / 16) + e][sample s0 + 16 q + l % 16], split
  bf16x8_t bh[4][KS2], bl[4][KS2];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t s = s0 + 16 * q + (lane & 15);
#pragma unroll
    for (int ks = 0; ks < KS2; ++ks) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int c = 32 * ks + 8 * (lane >> 4) + e;
        float dv = 0.0f;
        if (c < a.p && s < a.n_new) dv = (float)((double)a.beta[(int64_t)c * a.cap + a.first + s] - a.x0[c]);
        const __bf16 h = (__bf16)dv;
        bh[q][ks][e] = h;
        bl[q][ks][e] = (__bf16)(dv - (float)h);
      }
    }
  }
  const float L2E_HI = 1.44269502162933349609375f, L2E_LO = 1.925962989e-08f;
  double ll[4] = {0.0, 0.0, 0.0, 0.0};
  // per sample column: the product of u, the roundings of u, the sum of y eta - max(eta, 0)
  float pm[4] = {1.0f, 1.0f, 1.0f, 1.0f}, perr[4] = {0.0f, 0.0f, 0.0f, 0.0f}, pmax[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  int pending = 0;  // tiles folded into pm since the last logarithm
  // this lane's output rows of a tile: 4 (l / 16) + r
  const int rbase = 4 * (lane >> 4);
  u32x4_t ah[KS2], al[KS2];
  float4_t bs;
  auto fetch = [&](int64_t t) {
    const int64_t tc = t < t_hi ? t : t_hi - 1;
#pragma unroll
    for (int ks = 0; ks < KS2; ++ks) {
      ah[ks] = a.xfrag[((tc * KS2 + ks) * 2 + 0) * 64 + lane];
      al[ks] = a.xfrag[((tc * KS2 + ks) * 2 + 1) * 64 + lane];
    }
    bs = *reinterpret_cast<const float4_t*>(a.base + tc * 16 + rbase);
  };
  auto flush = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ll[q] += (double)(pmax[q] - (log_unbiased_f32(pm[q]) + perr[q]));
      pm[q] = 1.0f;
      perr[q] = 0.0f;
      pmax[q] = 0.0f;
    }
    pending = 0;
  };
  int64_t t = t_lo + wave;
  if (t < t_hi) fetch(t);
  for (; t < t_hi; t += 4) {
    u32x4_t ch[KS2], cl[KS2];
    float4_t cb = bs;
#pragma unroll
    for (int ks = 0; ks < KS2; ++ks) {
      ch[ks] = ah[ks];
      cl[ks] = al[ks];
    }
    if (t + 4 < t_hi) fetch(t + 4);
    float4_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = (float4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS2; ++ks) {
      const bf16x8_t xh = __builtin_bit_cast(bf16x8_t, ch[ks]);
      const bf16x8_t xl = __builtin_bit_cast(bf16x8_t, cl[ks]);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, bh[q][ks], acc[q], 0, 0, 0);
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, bl[q][ks], acc[q], 0, 0, 0);
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, bh[q][ks], acc[q], 0, 0, 0);
      }
    }
    // rows past n are zero (x, x0 term and y), so eta = 0 there exactly: each adds
    // y eta - max(eta, 0) = 0, u = 2 exactly and no rounding, i.e. -log 2 to the sum,
    // which the last slice adds back
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float eta = cb[r] + acc[q][r];
        const float av = fabsf(eta);
        // e^-|eta| = 2^(-|eta| log2 e): the rounding of the product with log2 e's high
        // part and its low part carried as a first-order correction (without them the
        // mean shift over 1e5 rows measures 1.6e-4)
        const float e_hi = -av * L2E_HI;
        const float e_lo = fmaf(-av, L2E_HI, -e_hi) - av * L2E_LO;
        const float t0 = __builtin_amdgcn_exp2f(e_hi);
        const float tv = fmaf(t0 * 0.6931471805599453f, e_lo, t0);
        const float u = 1.0f + tv;
        // rows with y = 1 are stored negated, so y eta - max(eta, 0) = -max(eta_s, 0)
        pmax[q] -= fmaxf(eta, 0.0f);
        perr[q] += tv - (u - 1.0f);
        pm[q] *= u;
      }
    }
    if (++pending == 2) flush();
  }
  if (pending) flush();
  // sum the four row groups of each column (lanes l, l^16, l^32, l^48), then the waves
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ll[q] += __shfl_xor(ll[q], 16, 64);
    ll[q] += __shfl_xor(ll[q], 32, 64);
  }
  __shared__ double red[4 * 64];
  if (lane < 16) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[wave * 64 + 16 * q + lane] = ll[q];
  }
  __syncthreads();
  if (tid < 64 && s0 + tid < a.n_new) {
    double v = ((red[tid] + red[64 + tid]) + red[128 + tid]) + red[192 + tid];
    if (t_hi == a.ntiles) v += (double)(a.ntiles * 16 - a.n) * 0.6931471805599453;  // the padding rows' -log 2
    a.part[(int64_t)blockIdx.y * a.n_new + s0 + tid] = v;
  }
}

// Per new sample: the row-slice partials in slice order, plus the log prior
// and normalising constants, into the target value the mixture kernel reads.
template <class T>
__global__ __launch_bounds__(256) void k_target_combine(const double* __restrict__ part, int splits, int64_t n_new,
                                                        const T* __restrict__ beta, int64_t cap, int64_t first,
                                                        int p, double inv_two_prior_var, double constant,
                                                        T* __restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_new) return;
  double ll = 0.0;
  for (int r = 0; r < splits; ++r) ll += part[(int64_t)r * n_new + s];
  double b2 = 0.0;
  for (int c = 0; c < p; ++c) {
    const double b = (double)beta[(int64_t)c * cap + first + s];
    b2 += b * b;
  }
  out[s] = (T)(ll - b2 * inv_two_prior_var + constant);
}
