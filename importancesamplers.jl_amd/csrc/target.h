// Batched built-in targets: Bayesian logistic and linear regression over an
// n x p data matrix, evaluated for one round's new samples (the coefficient
// vectors beta_s).  Included from kernels.h inside namespace issampler.
//
// eta = X beta is a dense contraction and runs on MFMA (16x16x4, FP32 in FP32
// mode, FP64 in FP64 mode).  A block owns 64 samples (four 16-column tiles)
// and a slice of the data rows.  The samples' coefficients are loaded once
// into registers as the B operands.  The data rows stream through LDS in
// 64-row chunks (the next chunk is prefetched into registers while the
// current one is multiplied); wave w takes the chunk's 16-row tile w.  The
// epilogue maps each eta to its log-likelihood term and accumulates it per
// sample in FP64; partial sums per row slice are combined in a fixed order.
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

// grid = (ceil(n_new / 64), splits), block = 256 (4 waves); D = dimension
// instantiation, KS = ceil(D / 4) K steps of the 16x16x4 MFMA
template <class T, int D>
__global__ __launch_bounds__(256) void k_target_gemm(TargetArgs<T> a) {
  typedef MfmaTile<T> M;
  constexpr int KS = (D + 3) / 4;
  constexpr int PS = target_row_stride(KS);
  constexpr int PER = (64 * 4 * KS + 255) / 256;  // staged values per thread per chunk
  extern __shared__ double lds_raw[];
  T* xs = reinterpret_cast<T*>(lds_raw);  // 64 x PS
  T* ys = xs + 64 * PS;                   // 64
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int p = exact_dim<D>(a.p);
  const int64_t s0 = (int64_t)blockIdx.x * 64;
  const int64_t r_lo = (int64_t)blockIdx.y * a.rows_per_split;
  const int64_t r_hi = min(a.n, r_lo + a.rows_per_split);

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

  // chunk staging: value v = 64-row chunk element (row v / (4 KS), coefficient v % (4 KS))
  T pre[PER], prey = T(0);
  auto fetch = [&](int64_t r0) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int v = tid + 256 * u;
      const int row = v / (4 * KS), c = v - row * (4 * KS);
      const int64_t r = r0 + row;
      pre[u] = (v < 64 * 4 * KS && r < r_hi && c < p) ? a.Xd[r * p + c] : T(0);
    }
    prey = (tid < 64 && r0 + tid < r_hi) ? a.yd[r0 + tid] : T(0);
  };
  if (r_lo < r_hi) fetch(r_lo);
  for (int64_t r0 = r_lo; r0 < r_hi; r0 += 64) {
    __syncthreads();  // the previous chunk's LDS reads are done
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int v = tid + 256 * u;
      if (v < 64 * 4 * KS) {
        const int row = v / (4 * KS), c = v - row * (4 * KS);
        xs[row * PS + c] = pre[u];
      }
    }
    if (tid < 64) ys[tid] = prey;
    __syncthreads();
    if (r0 + 64 < r_hi) fetch(r0 + 64);
    typename M::acc_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = M::zero();
    const T* arow = xs + (16 * wave + (lane & 15)) * PS + (lane >> 4);
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const T av = arow[4 * k];  // A lane l: X[row 16 w + l % 16][coefficient 4 k + l / 16]
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = M::mma(av, breg[q][k], acc[q]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * wave + M::row(lane, r);
      if (r0 + i < r_hi) {
        const T y = ys[i];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const T eta = acc[q][r];
          double term;
          if (a.kind == kTLogistic) {
            term = (double)(y * eta) - (double)log1pexp(eta);
          } else {
            const double res = (double)y - (double)eta;
            term = -(res * res) * a.inv_two_var;
          }
          ll[q] += term;
        }
      }
    }
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

__host__ __device__ constexpr size_t target_lds_bytes(int ks, int tsize) {
  return (size_t)(64 * target_row_stride(ks) + 64) * tsize > 256 * 8 ? (size_t)(64 * target_row_stride(ks) + 64) * tsize
                                                                     : (size_t)256 * 8;
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
