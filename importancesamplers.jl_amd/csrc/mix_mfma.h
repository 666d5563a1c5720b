// The deterministic-mixture log density of a round's new samples on the FP32 matrix
// cores (FP32 jobs whose samples are drawn before the mixture pass, d >= 16, a bank
// of components; src/mis_execution.jl:30-95, src/dm_pmc_execution.jl).  Included
// from kernels.h inside namespace issampler.
//
// Each component's whitened coordinates y = L^{-1} (x - mu) are a dense product of
// the inverse factor M = L^{-1} (lower triangular, KP x KP with KP = 16 KT) and the
// 64 samples' centred coordinates:
//   k_mix_prep   one wave per component: M by forward substitution of L M = I in FP64
//                (each lane one column, L's rows read across lanes), stored in FP32 as v_mfma_f32_16x16x4f32
//                A-operand fragments of the lower-triangular 16-row tiles only;
//   k_mix_mfma   one block per 64 samples, wave w the 16 samples 16 w .. 16 w + 15:
//                per component, the B operands x - mu (four dimensions per K step,
//                one value per lane), the tiles' MFMAs, r^2 from the squares of the
//                tiles summed over the four lanes of a sample, the radial log density
//                and the running log-sum-exp of logmass + log q in component order.
// The component's fragments are staged in LDS, the next component's loaded while the
// current one's MFMAs run.  A sample whose generating component has a NaN or -inf
// density gets a NaN mixture value (the new-sample kernel fails it as invalid, as its
// own loop does).
#pragma once

__device__ inline double readlane_f64(double v, int lane);  // kernels.h

// Fragments of one component: lower-triangular (row tile i, K step kk) pairs, kk < 4 (i + 1),
// each 64 lanes x 1 float; pairs in (i, kk) order.
__host__ __device__ constexpr int mix_pairs(int kt) { return 4 * kt * (kt + 1) / 2; }

struct MixPrepArgs {
  int d, kp;                 // dimension, padded dimension 16 KT
  int k0, ncomp;             // components [k0, k0 + ncomp) of the slot table
  const double* slots64;     // FP64 slot table
  int stride;
  float* frag;               // ncomp x mix_pairs(KT) x 64
  float* mu;                 // ncomp x kp
  float* lognorm;            // ncomp
};

// One wave per component.  Lane c < d solves L m = e_c for column c of M:
// m_r = (delta_rc - sum_{k<r} L_rk m_k) / L_rr (m_r = 0 for r < c), in FP64, the column
// in registers and L_rk read from lane r (which holds row r of L) with v_readlane, so
// the solve has no memory operations; the columns are transposed through LDS.
template <int KT>
__global__ __launch_bounds__(64) void k_mix_prep(MixPrepArgs a) {
  constexpr int KP = 16 * KT;
  __shared__ float Ms[KP][KP + 1];  // Ms[r][c] = M[r][c]
  const int comp = blockIdx.x;
  const int lane = threadIdx.x;
  const double* s = a.slots64 + (int64_t)(a.k0 + comp) * a.stride;
  const int d = a.d;
  double Lrow[KP], inv = 1.0;
#pragma unroll
  for (int c = 0; c < KP; ++c) Lrow[c] = 0.0;
  if (lane < d) {
    const int row = 2 * d + lane * (lane + 1) / 2;
#pragma unroll
    for (int c = 0; c < KP; ++c)
      if (c <= lane) Lrow[c] = s[row + c];
    inv = s[d + lane];  // the slot holds 1 / L_rr
  }
  double m[KP];
#pragma unroll
  for (int r = 0; r < KP; ++r) {
    m[r] = 0.0;
    if (r < d) {
      double v = (r == lane) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < r; ++k) v -= readlane_f64(Lrow[k], r) * m[k];
      m[r] = v * readlane_f64(inv, r);
    }
  }
  if (lane < KP) {
#pragma unroll
    for (int r = 0; r < KP; ++r) Ms[r][lane] = (float)m[r];
  }
  __syncthreads();
  // fragments: pair (i, kk), lane l: M[16 i + l % 16][4 kk + l / 16]
  float* out = a.frag + (int64_t)comp * mix_pairs(KT) * 64;
  int pidx = 0;
#pragma unroll
  for (int i = 0; i < KT; ++i)
#pragma unroll
    for (int kk = 0; kk < 4 * (i + 1); ++kk) {
      out[pidx * 64 + lane] = Ms[16 * i + (lane & 15)][4 * kk + (lane >> 4)];
      ++pidx;
    }
  for (int q = lane; q < KP; q += 64) a.mu[(int64_t)comp * KP + q] = q < d ? (float)s[q] : 0.0f;
  if (lane == 0) a.lognorm[comp] = (float)s[slot_lognorm_off(d)];
}

struct MixMfmaArgs {
  int d, kp;
  int ncomp, k0;
  const float* frag;         // from k_mix_prep
  const float* mu;
  const float* lognorm;
  const float* logmass;      // indexed by slot
  Family<float> fam;
  const float* X;            // samples, SoA: X[c * cap + first + j]
  int64_t cap, first, n_new;
  const int32_t* assign;     // generating slot per sample (null: gen_slot)
  int gen_slot;
  float* out;                // n_new: log sum_k exp(logmass_k + log q_k(x_j)), NaN if the generator's density is invalid
};

template <int KT>
__global__ __launch_bounds__(256) void k_mix_mfma(MixMfmaArgs a) {
  constexpr int KP = 16 * KT;
  constexpr int NP = mix_pairs(KT);
  constexpr int FR = NP * 64;           // floats of one component's fragments
  constexpr int PER = (FR + 255) / 256;  // fragment floats staged per thread
  __shared__ float frag_s[2][FR];
  __shared__ float mu_s[2][KP + 2];  // the centre, then the log normaliser and the log mass
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t j = (int64_t)blockIdx.x * 64 + 16 * wave + (lane & 15);
  const bool active = j < a.n_new;
  const int64_t jc = active ? j : a.n_new - 1;
  // this lane's coordinates: dims 4 kk + l / 16 of its sample
  float xv[KP / 4];
#pragma unroll
  for (int kk = 0; kk < KP / 4; ++kk) {
    const int c = 4 * kk + (lane >> 4);
    xv[kk] = c < a.d ? a.X[(int64_t)c * a.cap + a.first + jc] : 0.0f;
  }
  const int gen = a.assign ? a.assign[jc] : a.gen_slot;
  float pre[PER], premu = 0.0f;
  auto fetch = [&](int k) {
    const float* src = a.frag + (int64_t)k * FR;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int q = tid + 256 * u;
      pre[u] = q < FR ? src[q] : 0.0f;
    }
    if (tid < KP)
      premu = a.mu[(int64_t)k * KP + tid];
    else if (tid == KP)
      premu = a.lognorm[k];
    else if (tid == KP + 1)
      premu = a.logmass[a.k0 + k];
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int q = tid + 256 * u;
      if (q < FR) frag_s[buf][q] = pre[u];
    }
    if (tid < KP + 2) mu_s[buf][tid] = premu;
  };
  const float ninf = -__builtin_huge_valf();
  float m = ninf, sum = 0.0f, sumc = 0.0f;  // compensated sum
  bool bad = false;
  fetch(0);
  stage(0);
  __syncthreads();
  for (int k = 0; k < a.ncomp; ++k) {
    const int buf = k & 1;
    if (k + 1 < a.ncomp) fetch(k + 1);
    float4_t acc[KT];
#pragma unroll
    for (int i = 0; i < KT; ++i) acc[i] = (float4_t){0.f, 0.f, 0.f, 0.f};
    int pidx = 0;
#pragma unroll
    for (int i = 0; i < KT; ++i)
#pragma unroll
      for (int kk = 0; kk < 4 * (i + 1); ++kk) {
        const float b = xv[kk] - mu_s[buf][4 * kk + (lane >> 4)];
        acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(frag_s[buf][pidx * 64 + lane], b, acc[i], 0, 0, 0);
        ++pidx;
      }
    // r^2 of sample l % 16: its rows 4 (l / 16) + r of every tile, then the lanes l ^ 16, l ^ 32
    float r2 = 0.0f;
#pragma unroll
    for (int i = 0; i < KT; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) r2 = fmaf(acc[i][r], acc[i][r], r2);
    r2 += __shfl_xor(r2, 16, 64);
    r2 += __shfl_xor(r2, 32, 64);
    const int slot = a.k0 + k;
    const float lnorm = mu_s[buf][KP];
    float q;
    if (!a.fam.student) {
      q = isfinite(r2) ? lnorm - 0.5f * r2 : __builtin_nanf("");
    } else {
      const float tiny = 1.17549435e-38f;
      const float ratio = r2 / a.fam.dof;
      q = (r2 >= tiny && isfinite(ratio)) ? lnorm - a.fam.half_dof_plus_d * log1pf(ratio) : __builtin_nanf("");
    }
    if (slot == gen && (q != q || q == ninf)) bad = true;
    const float v = mu_s[buf][KP + 1] + q;
    lse_push(v, m, sum, sumc);
    __syncthreads();  // every wave is done with buffer buf ^ 1 (component k - 1)
    if (k + 1 < a.ncomp) stage(buf ^ 1);
    __syncthreads();
  }
  if (active && lane < 16) a.out[j] = bad ? __builtin_nanf("") : m + logf(sum + sumc);
}
