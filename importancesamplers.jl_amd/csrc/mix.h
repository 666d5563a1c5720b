// New-sample kernel and the streaming sweep with FP64 MFMA moments.
// Included from kernels.h (uses RoundArgs, sample helpers and device.h).
#pragma once

// (included inside namespace issampler)

typedef double double4_t __attribute__((ext_vector_type(4)));

// Merge two running (max, scaled-sum) log-sum-exp states.
template <class T>
__device__ inline void lse_merge(T& m, T& s, T m2, T s2) {
  if (m2 > m) {
    s = s * LseAcc<T>::fast_exp(m - m2) + s2;
    m = m2;
  } else if (m > m2) {
    s = s + s2 * LseAcc<T>::fast_exp(m2 - m);
  } else {
    s = s + s2;
  }
}

// ---------------------------------------------------------------- new samples
// One round's new samples: draw from the generating slot, evaluate the target,
// the deterministic-mixture log-sum-exp over slots [k0, k1] and the weight
// (src/mis_execution.jl:601-771; plain IS src/native_execution.jl:741-812).
// A block of W waves owns 64 consecutive samples, one per lane:
//   A  the normals of the 64 samples, one Philox block (two Box-Muller pairs)
//      per thread, and their Student-t radial multipliers are spread over all
//      64 W threads and staged in LDS, while the target slot and the first chunk
//      of mixture slots with their log masses are copied into LDS (one memory
//      round trip, overlapped with the Philox work);
//   B  wave 0 forms x = mu + L z, stores X and shares x through LDS;
//   C  wave w evaluates the components kk with (kk + 1) mod W == w, so every
//      slot read is a wave-uniform LDS broadcast; wave 0 also evaluates the
//      target;
//   D  wave 0 merges the W partial log-sum-exp states in wave order and writes
//      the sample's outputs.
// LDS: normals 64 x (D + 1) doubles | multipliers 64 | x 64 x (D + 1) T |
//      W x 64 partial (m, s, flag) | 64 draw flags | target slot | slot chunk |
//      the chunk's log masses.
// Threads per block the register budget allows: x, y and z of D values each.
__host__ __device__ constexpr int new_block_limit(int D, int tsize, bool st) {
  return (!st && D * tsize <= 64) ? 1024 : (D * tsize <= 128 ? 512 : 256);
}

// Byte offset of the slot area (target slot, then the component chunk).
__host__ __device__ constexpr size_t new_lds_fixed(int D, int W, int tsize) {
  return (((size_t)(64 * (D + 1) + 64) * 8 + (size_t)64 * (D + 1) * tsize + (size_t)W * 64 * (2 * tsize + 4) +
           64 * 4) + 15) & ~(size_t)15;
}

// The body of one new-sample block `blk` (samples 64 blk .. 64 blk + 63).  On
// return, wave 0's active lanes hold their sample in out_x and its log-weight
// (-inf when it failed) in out_lw; other lanes leave both untouched.
template <class T, int D, int MODE, bool ST, bool RG = false>
__device__ __forceinline__ void new_block(const RoundArgs<T>& a, int64_t blk, T& out_lw, T (&out_x)[D]) {
  extern __shared__ double lds[];
  constexpr int ZS = D + 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int W = blockDim.x >> 6;
  const int d = exact_dim<D>(a.d);
  const int64_t j0 = blk * 64;
  const int64_t j = j0 + lane;
  const bool active = j < a.n_new;
  const int64_t gi = a.first + j;
  double* zs = lds;
  double* mults = zs + 64 * ZS;
  T* xs = reinterpret_cast<T*>(mults + 64);
  T* part_m = xs + 64 * ZS;
  T* part_s = part_m + W * 64;
  int* part_f = reinterpret_cast<int*>(part_s + W * 64);
  int* drawn_f = part_f + W * 64;
  T* tslot = reinterpret_cast<T*>(reinterpret_cast<char*>(lds) + new_lds_fixed(D, W, sizeof(T)));
  const bool staged_target = (MODE == kDrawFused && a.tkind == kTGaussian);
  T* comp = tslot + (staged_target ? a.stride : 0);
  const int ncomp = a.k1 - a.k0 + 1;
  const int chunk = min(ncomp, a.slot_chunk);
  // the chunk's log mixture masses follow its slots
  T* cmass = comp + chunk * a.stride;

  // ---- A: slot staging, normals and radial multipliers
  if (staged_target)
    for (int q = tid; q < a.stride; q += blockDim.x) tslot[q] = a.tslot[q];
  {
    const T* src = a.slots + (int64_t)a.k0 * a.stride;
    for (int q = tid; q < chunk * a.stride; q += blockDim.x) comp[q] = src[q];
    if (a.mis_terms == 0)
      for (int q = tid; q < chunk; q += blockDim.x) cmass[q] = a.logmass[a.k0 + q];
  }
  if (MODE != kPreTarget) {
    const uint64_t rng_word = *a.run_hi | (uint64_t)a.round;
    const int np = (d + 1) >> 1;
    // one job per Philox block: pairs 2b and 2b + 1 of a sample take words (0, 1) and
    // (2, 3) of its block b
    const int nb = (np + 1) >> 1;
    const int njobs = 64 * nb + (ST ? 64 : 0);
    for (int q = tid; q < njobs; q += blockDim.x) {
      if (q < 64 * nb) {
        const int s = q / nb, b = q - s * nb;
        const int64_t js = j0 + s;
        const bool second = 2 * b + 1 < np;
        double z[4] = {0.0, 0.0, 0.0, 0.0};
        if (js < a.n_new) {
          if (a.pre_normals) {
            const double* src = a.pre_normals + js * a.nstride;
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (4 * b + u < d) z[u] = src[4 * b + u];
          } else {
            const U64x4 w = philox4x64((uint64_t)js, rng_word, kPurposeNormals, (uint64_t)b, a.key.k0, a.key.k1);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              if (h == 0 || second) {
                const double r = sqrt(-2.0 * log(u53_open(w.v[2 * h])));
                double sn, cs;
                sincospi(2.0 * u53_open(w.v[2 * h + 1]), &sn, &cs);
                z[2 * h] = r * cs;
                z[2 * h + 1] = r * sn;
              }
            }
          }
        }
        zs[s * ZS + 4 * b] = z[0];
        zs[s * ZS + 4 * b + 1] = z[1];
        if (second) {
          zs[s * ZS + 4 * b + 2] = z[2];
          zs[s * ZS + 4 * b + 3] = z[3];
        }
      } else if (ST) {
        const int s = q - 64 * nb;
        const int64_t js = j0 + s;
        double mult = 1.0;
        int drawn = 1;
        if (js < a.n_new) {
          double rn[kGammaAttempts], ru[kGammaAttempts + 1];
          for (int t = 0; t < a.rs; ++t)
            rn[t] = a.pre_rn ? a.pre_rn[js * a.rn_stride + t]
                             : rng_normal(a.key, (uint64_t)js, rng_word, kPurposeRadialNormals, t);
          for (int t = 0; t < a.us; ++t)
            ru[t] = a.pre_ru ? a.pre_ru[js * a.ru_stride + t]
                             : rng_uniform(a.key, (uint64_t)js, rng_word, kPurposeRadialUniforms, t);
          if (!student_t_multiplier(a.dof, rn, ru, a.us > 0, &mult)) {
            record_failure(a.fail, js + 1, 0, kDrawExhausted);
            mult = __builtin_nan("");
            drawn = 0;
          }
        }
        mults[s] = mult;
        drawn_f[s] = drawn;
      }
    }
  }
  __syncthreads();

  // ---- B: affine draw and X store (wave 0), shared through LDS
  T x[D], y[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    x[c] = T(0);
    y[c] = T(0);
  }
  bool ok = active;
  int gen = a.gen_slot;
  if (active && a.assign) gen = a.assign[j];
  T tf_lj = T(0);  // pre-target pass under a transform: the draw pass's log Jacobian
  if (MODE == kPreTarget) {
    // the draw-only pass marks an invalid draw (already recorded) with a NaN numerator,
    // and under a transform leaves the forward log Jacobian there
    if (active) {
      const T mark = a.lognums[gi];
      if (a.tf_carry) tf_lj = mark;
      if (mark != mark) {
        ok = false;
      } else {
#pragma unroll
        for (int c = 0; c < D; ++c)
          if (c < d) x[c] = a.X[(int64_t)c * a.cap + gi];
      }
    }
  } else {
    int good = 0;
    if (wave == 0 && active) {
      double mult = 1.0;
      bool drawn = true;
      if (ST) {
        mult = mults[lane];
        drawn = drawn_f[lane] != 0;
      }
      if (drawn || a.record_nonfinite_after_exhaust) {
        T z[D];
#pragma unroll
        for (int c = 0; c < D; ++c) z[c] = (c < d) ? T(zs[lane * ZS + c] * mult) : T(0);
        // the generating slot is read from the staged chunk when it is there
        const int gl = gen - a.k0;
        const int64_t goff = (int64_t)gen * a.stride;
        bool valid;
        if (gl >= 0 && gl < chunk)
          valid = slot_draw<T, D>(comp + gl * a.stride, d, z, x);
        else if (a.assign)
          valid = slot_draw<T, D>(a.slots + goff, d, z, x);
        else
          valid = slot_draw<T, D>(as_const(a.slots + goff), d, z, x);
#pragma unroll
        for (int c = 0; c < D; ++c)
          if (c < d) {
            a.X[(int64_t)c * a.cap + gi] = x[c];
            xs[lane * ZS + c] = x[c];
          }
        if (!valid) record_failure(a.fail, j + 1, 0, kGeneratedNonfinite);
        good = valid ? 1 : 0;
      }
      drawn_f[lane] = good;
      if (MODE == kDrawOnly) {
        T mark = good ? T(0) : T(__builtin_nan(""));
        if (good && a.tf_kinds) {
          // the batched target sees the logical sample x = T(z): write it and carry the
          // forward log Jacobian to the pre-target pass in the numerator slot
          T lj = T(0);
          uint16_t treason = 0;
          int tblock = 0;
#pragma unroll
          for (int c = 0; c < D; ++c)
            if (c < d) {
              T xc, ljc;
              const uint16_t r =
                  transform_forward<T>(a.tf_kinds[c], a.tf_bounds[2 * c], a.tf_bounds[2 * c + 1], x[c], xc, ljc);
              if (r && !treason) {
                treason = r;
                tblock = c + 1;
              }
              lj += ljc;
              a.XL[(int64_t)c * a.cap + gi] = xc;
            }
          if (treason) {
            record_failure(a.fail, j + 1, tblock, treason);
            mark = T(__builtin_nan(""));
          } else {
            mark = lj;
          }
        }
        a.lognums[gi] = mark;
      }
    }
    if (MODE == kDrawOnly) return;
    __syncthreads();
    if (active) {
      ok = drawn_f[lane] != 0;
      if (wave != 0) {
#pragma unroll
        for (int c = 0; c < D; ++c)
          if (c < d) x[c] = xs[lane * ZS + c];
      }
    }
  }

  // ---- C: target (wave 0) and this wave's share of the mixture components
  T tv = T(0);
  if (wave == 0 && ok) {
    if (MODE == kPreTarget) {
      tv = a.pre_logtarget[j] + tf_lj;
    } else if (a.tf_kinds) {
      // the target at the logical sample x = T(z) plus the forward log Jacobian; the
      // first failing coordinate (one transform block each) fails the draw with its
      // transform reason (src/native_execution.jl:689-712,836-840)
      T xl[D];
      T lj = T(0);
      uint16_t treason = 0;
      int tblock = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        xl[c] = T(0);
        if (c < d) {
          T xc, ljc;
          const uint16_t r =
              transform_forward<T>(a.tf_kinds[c], a.tf_bounds[2 * c], a.tf_bounds[2 * c + 1], x[c], xc, ljc);
          if (r && !treason) {
            treason = r;
            tblock = c + 1;
          }
          xl[c] = xc;
          lj += ljc;
          a.XL[(int64_t)c * a.cap + gi] = xc;
        }
      }
      if (treason) {
        record_failure(a.fail, j + 1, tblock, treason);
        ok = false;
      } else {
        tv = inline_target<T, D, RG>(a.tkind, d, tslot, xl, y) + lj;
      }
    } else {
      tv = inline_target<T, D, RG>(a.tkind, d, tslot, x, y);
    }
    const uint16_t tr = target_reason(tv);
    if (tr) {
      record_failure(a.fail, j + 1, 0, tr);
      ok = false;
    }
  }
  T m = -T(__builtin_huge_val()), sum = T(0), sumc = T(0);
  int genbad = 0;
  if (a.mis_terms != 0) {
    // static MIS: each sample's own terms (the generating slot, or its group;
    // src/methods/static_mis.jl:383-468), evaluated by wave 0 from the slot table
    if (wave == 0 && ok) {
      int mb = gen, me = gen + 1;
      if (a.mis_terms == 2) {
        const int g = a.mis_slot_group[gen];
        mb = a.mis_goff[g];
        me = a.mis_goff[g + 1];
      }
      for (int mi = mb; mi < me; ++mi) {
        const int k = a.mis_terms == 2 ? a.mis_gmem[mi] : gen;
        const T q = slot_logdensity<T, D, RG>(a.slots + (int64_t)k * a.stride, d, a.fam, x, y);
        if (k == gen && (q != q || q == -T(__builtin_huge_val()))) genbad = 1;
        const T v = (a.mis_terms == 2 ? a.mis_gcoef[mi] : T(0)) + q;
        lse_push(v, m, sum, sumc);
      }
    }
  } else if (a.pre_mix) {
    // the mixture log density from the matrix-core pass (NaN: the generating
    // component's density is not valid)
    if (wave == 0 && ok) {
      const T v = a.pre_mix[j];
      if (v != v) {
        genbad = 1;
      } else {
        m = v;
        sum = T(1);
      }
    }
  } else {
    for (int base = 0; base < ncomp; base += chunk) {
      const int cnt = min(chunk, ncomp - base);
      if (base > 0) {
        __syncthreads();
        const T* src = a.slots + (int64_t)(a.k0 + base) * a.stride;
        for (int q = tid; q < cnt * a.stride; q += blockDim.x) comp[q] = src[q];
        for (int q = tid; q < cnt; q += blockDim.x) cmass[q] = a.logmass[a.k0 + base + q];
        __syncthreads();
      }
      if (ok) {
        // first kk >= base with (kk + 1) mod W == wave
        int kk = base + ((wave + W - 1 - base % W) % W + W) % W;
        for (; kk < base + cnt; kk += W) {
          const T* sl = comp + (kk - base) * a.stride;
          T q;
          if constexpr (D <= kExactDims)
            q = slot_logdensity_lds<T, D, RG>(sl, a.fam, x, y);
          else
            q = slot_logdensity<T, D, RG>(sl, d, a.fam, x, y);
          if (a.k0 + kk == gen && (q != q || q == -T(__builtin_huge_val()))) genbad = 1;
          const T v = cmass[kk - base] + q;
          lse_push(v, m, sum, sumc);
        }
      }
    }
  }
  sum += sumc;  // the compensation of the FP32 sum (zero in FP64)
  if (W > 1) {
    part_m[wave * 64 + lane] = m;
    part_s[wave * 64 + lane] = sum;
    part_f[wave * 64 + lane] = genbad;
    __syncthreads();
  }
  if (wave != 0 || !active) return;

  // ---- D: merge in wave order and write the sample
  for (int w = 1; w < W; ++w) {
    lse_merge(m, sum, part_m[w * 64 + lane], part_s[w * 64 + lane]);
    genbad |= part_f[w * 64 + lane];
  }
  T lt = -T(__builtin_huge_val()), ln = T(0), lwv = -T(__builtin_huge_val());
  int32_t rid = 0, pid = 0;
  if (ok) {
    const T value = m + log(sum);
    if (genbad || value != value || value == -T(__builtin_huge_val())) {
      record_failure(a.fail, j + 1, 0, kProposalInvalid);
    } else {
      const T lw = tv - (value - a.logtotal);
      const uint16_t wr = logweight_reason(lw);
      if (wr) {
        record_failure(a.fail, j + 1, 0, wr);
      } else {
        lt = tv;
        ln = value;
        lwv = lw;
        rid = a.round + 1;
        pid = a.bank_ids ? a.bank_ids[gen] : 0;
      }
    }
  }
  a.logtargets[gi] = lt;
  a.lognums[gi] = ln;
  a.logweights[gi] = lwv;
  a.round_ids[gi] = rid;
  if (a.proposal_ids) a.proposal_ids[gi] = pid;
  out_lw = lwv;
#pragma unroll
  for (int c = 0; c < D; ++c) out_x[c] = x[c];
}

template <class T, int D, int MODE, bool ST, bool RG = false>
__global__ __launch_bounds__(new_block_limit(D, sizeof(T), ST)) void k_new(RoundArgs<T> a) {
  T lw;
  T x[D];
  new_block<T, D, MODE, ST, RG>(a, blockIdx.x, lw, x);
}

// ---------------------------------------------------------------- streaming sweep
// Over samples [lo, hi): samples below append_hi receive the AMIS append of one
// new mixture component (src/amis_execution.jl:37-65) and their weights are
// re-formed (src/amis_execution.jl:107-132); every sample then contributes its
// weight e = exp(lw - m) (running wave maximum m) to the summary monoid
// (src/execution.jl:159-186) and, when NT > 0, to the weighted moments about
// `center`.  For the moments each lane stages w = sqrt(e) (x - center, 1) in
// LDS; sum_s w_s w_s^T then accumulates on FP64 MFMA (16x16x4: the A and B
// operands of a 4-sample step are the same register), NT tiles per side,
// lower tile pairs only.  The appended slot and the centre are uniform, so
// they arrive by scalar loads.  The next 64 samples' loads are issued before
// the current ones are processed.  One column-major partial record per block.
template <class T>
struct SweepArgs {
  int d;
  int64_t lo, hi, append_hi;
  int append;
  const T* slot;
  int stride;
  T logmass, logtotal;
  Family<T> fam;
  const T* X;
  int64_t cap;
  const T* logtargets;
  T* lognums;
  T* logweights;
  const int32_t* round_ids;
  unsigned long long* fail;
  const double* clip;
  const double* center;
  double* partials;
  int64_t pcap;
  int nv;
  int64_t per_block;
};

__host__ __device__ constexpr int sweep_row_stride(int nt) { return 16 * nt + 1; }

template <class T, int D>
struct SweepLoad {
  T x[D];
  T num, tgt, lw;
  int32_t rid;
};

// Loads sample i's inputs; a lane past the block's end (i >= bend) reads the block's
// last sample instead (finite values, so its zero weight contributes exact zeros),
// which keeps every load unconditional.
template <class T, int D, int NT>
__device__ __forceinline__ void sweep_load(const SweepArgs<T>& a, int64_t i, int64_t bend, SweepLoad<T, D>& L) {
  const int d = exact_dim<D>(a.d);
  const int64_t ic = i < bend ? i : bend - 1;
  const bool app = a.append && ic < a.append_hi;
#pragma unroll
  for (int c = 0; c < D; ++c) L.x[c] = T(0);
  if (NT > 0 || app) {
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (c < d) L.x[c] = a.X[(int64_t)c * a.cap + ic];
  }
  if (app) {
    L.num = a.lognums[ic];
    L.tgt = a.logtargets[ic];
    L.rid = a.round_ids[ic];
  } else {
    L.lw = a.logweights[ic];
  }
}

// e^v for v <= 0 (or -inf): FP64 through exp; FP32 as 2^(v log2 e) from the
// hardware base-2 exponential with log2 e split in two floats (the product's
// rounding carried as a first-order correction), within a few ulp.
__device__ __forceinline__ double exp_nonpos(double v) { return exp(v); }
__device__ __forceinline__ float exp_nonpos(float v) {
  const float L2E_HI = 1.44269502162933349609375f, L2E_LO = 1.925962989e-08f;
  const float e_hi = v * L2E_HI;
  const float e_lo = fmaf(v, L2E_HI, -e_hi) + v * L2E_LO;
  const float t0 = __builtin_amdgcn_exp2f(e_hi);
  return fmaf(t0 * 0.6931471805599453f, e_lo, t0);
}

// Per-wave running state of the summary monoid and the moment tiles.
// With one moment tile (NT = 1) the tile accumulates in two independent MFMA chains
// (even and odd 4-sample steps of each 64-sample step), summed when the wave's state
// is read out; with more tiles, their own chains are independent already.
template <class T, int NT>
struct SweepAcc {
  static constexpr int P = NT * (NT + 1) / 2;
  T mw;
  double s0, s2;
  double4_t acc[P > 0 ? P : 1];
  double4_t acc2[1];
};

template <class T, int NT>
__device__ __forceinline__ void sweep_acc_init(SweepAcc<T, NT>& st) {
  st.mw = -T(__builtin_huge_val());
  st.s0 = 0.0;
  st.s2 = 0.0;
#pragma unroll
  for (int p = 0; p < (SweepAcc<T, NT>::P > 0 ? SweepAcc<T, NT>::P : 1); ++p) st.acc[p] = (double4_t){0.0, 0.0, 0.0, 0.0};
  st.acc2[0] = (double4_t){0.0, 0.0, 0.0, 0.0};
}

// One 64-sample step of a wave: lane `lane` contributes log-weight lw (-inf: no
// contribution) of sample x.  `rows` is the wave's 64 x (16 NT + 1) LDS staging
// area, whose columns above d are zero.
template <class T, int D, int NT, class CP>
__device__ __forceinline__ void sweep_accumulate(SweepAcc<T, NT>& st, T lw, const T (&x)[D], int d, CP cen,
                                                 double* rows, int lane) {
  constexpr int RS = sweep_row_stride(NT);
  constexpr int P = SweepAcc<T, NT>::P;
  const T ninf = -T(__builtin_huge_val());
  // running wave maximum, recomputed only when some lane exceeds it
  if (__any(lw > st.mw)) {
    const T wm = wave_max(lw);
    const double f = (st.mw == ninf) ? 0.0 : exp((double)st.mw - (double)wm);
    st.s0 *= f;
    st.s2 *= f * f;
#pragma unroll
    for (int p = 0; p < (P > 0 ? P : 1); ++p) st.acc[p] *= f;
    if (P == 1) st.acc2[0] *= f;
    st.mw = wm;
  }
  const T se = (lw == ninf) ? T(0) : exp_nonpos(T(0.5) * (lw - st.mw));
  const double sd = (double)se;
  const double e = sd * sd;
  st.s0 += e;
  st.s2 += e * e;
  if (NT > 0) {
    double* row = rows + lane * RS;
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (c < d) row[c] = sd * ((double)x[c] - cen[c]);
    row[d] = sd;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int stp = 0; stp < 16; ++stp) {
      const double* r = rows + (4 * stp + (lane >> 4)) * RS + (lane & 15);
      double av[NT > 0 ? NT : 1];
#pragma unroll
      for (int ti = 0; ti < NT; ++ti) av[ti] = r[ti * 16];
      int p = 0;
#pragma unroll
      for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj <= ti; ++tj) {
          if (P == 1 && (stp & 1))
            st.acc2[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ti], av[tj], st.acc2[0], 0, 0, 0);
          else
            st.acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ti], av[tj], st.acc[p], 0, 0, 0);
          ++p;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Block reduction of the waves' states into one column-major partial record at
// `out` (column stride pcap): rescale the waves to the block maximum, sum in
// wave order.  Every thread of the block calls it (it holds barriers).
template <class T, int NT>
__device__ __forceinline__ void sweep_epilogue(SweepAcc<T, NT>& st, int d, int nv, double* out, int64_t pcap,
                                               double* lds) {
  constexpr int P = SweepAcc<T, NT>::P;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = blockDim.x >> 6;
  double s0 = wave_sum(st.s0);
  double s2 = wave_sum(st.s2);
  __syncthreads();
  double* wmax = lds;                    // nw
  double* ws0 = lds + 4;                 // nw
  double* ws2 = lds + 8;                 // nw
  double* tiles = lds + 16;              // nw * P * 256
  if (lane == 0) {
    wmax[wave] = (double)st.mw;
    ws0[wave] = s0;
    ws2[wave] = s2;
  }
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      tiles[((size_t)wave * P + p) * 256 + lane * 4 + r] = P == 1 ? st.acc[p][r] + st.acc2[0][r] : st.acc[p][r];
  __syncthreads();
  double M = wmax[0];
  for (int w = 1; w < nw; ++w) M = fmax(M, wmax[w]);
  double fw[4];
  for (int w = 0; w < 4; ++w)
    fw[w] = (w < nw && wmax[w] != -__builtin_huge_val()) ? exp(wmax[w] - M) : 0.0;
  // summed tiles go after the per-wave tiles
  double* sumt = tiles + (size_t)nw * P * 256;
  for (int q = tid; q < P * 256; q += blockDim.x) {
    double v = 0.0;
    for (int w = 0; w < nw; ++w) v += fw[w] * tiles[((size_t)w * P) * 256 + q];
    sumt[q] = v;
  }
  __syncthreads();
  if (tid == 0) out[0] = M;
  for (int v = tid; v < nv; v += blockDim.x) {
    double val;
    if (v == 1) {
      val = 0.0;
      for (int w = 0; w < nw; ++w) val += fw[w] * fw[w] * ws2[w];
    } else if (NT == 0) {
      val = 0.0;
      for (int w = 0; w < nw; ++w) val += fw[w] * ws0[w];
    } else {
      // moment entry C[r][c], r >= c, of the augmented matrix: s0 = C[d][d],
      // s1[r] = C[d][r], S pairs C[r][c]
      int r, c;
      if (v == 0) {
        r = c = d;
      } else if (v < 2 + d) {
        r = d;
        c = v - 2;
      } else {
        pair_index(v - 2 - d, r, c);
      }
      // v_mfma_f64_16x16x4f64 result layout: lane l, register i holds tile
      // entry (row 4 i + l / 16, column l % 16)
      const int ti = r >> 4, tj = c >> 4, ra = r & 15, cb = c & 15;
      const int p = ti * (ti + 1) / 2 + tj;
      const int ln = cb + 16 * (ra & 3), reg = ra >> 2;
      val = sumt[(size_t)p * 256 + ln * 4 + reg];
    }
    out[(int64_t)(1 + v) * pcap] = val;
  }
}

// One wave's state as the column-major partial record at `out` (the layout
// sweep_epilogue writes), when no other wave of the block holds samples: its
// values are the block's own, so each lane stores the tile entries it holds.
template <class T, int NT>
__device__ __forceinline__ void wave_record(const SweepAcc<T, NT>& st, int d, double* out, int64_t pcap, int lane) {
  static_assert(NT <= 1, "one moment tile per side");
  const double s0 = wave_sum(st.s0);
  const double s2 = wave_sum(st.s2);
  if (lane == 0) {
    out[0] = (double)st.mw;
    out[2 * pcap] = s2;
    if (NT == 0) out[pcap] = s0;
  }
  if (NT > 0) {
    // v_mfma_f64_16x16x4f64 result layout: register i of lane l holds entry
    // (row 4 i + l / 16, column l % 16); value 0 is C[d][d], 2 + c is C[d][c]
    // (c < d) and 2 + d + r (r + 1) / 2 + c is C[r][c] (c <= r < d)
    const int c = lane & 15;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * i + (lane >> 4);
      if (r <= d && c <= r) {
        const int v = r == d ? (c == d ? 0 : 2 + c) : 2 + d + r * (r + 1) / 2 + c;
        out[(int64_t)(1 + v) * pcap] = st.acc[0][i] + st.acc2[0][i];
      }
    }
  }
}

// The body of sweep block `blk` over [a.lo + blk * per_block, ...).
template <class T, int D, int NT, bool RG = false>
__device__ __forceinline__ void sweep_block(const SweepArgs<T>& a, int64_t blk) {
  extern __shared__ double lds[];
  constexpr int RS = sweep_row_stride(NT);
  const T ninf = -T(__builtin_huge_val());
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = blockDim.x >> 6;
  const int d = exact_dim<D>(a.d);
  double* rows = lds + wave * (64 * RS);
  if (NT > 0) {
    // columns above d stay zero for the whole kernel
    for (int q = lane; q < 64 * RS; q += 64) rows[q] = 0.0;
  }
  // the centre is uniform: scalar loads, read where it is used
  const cptr<double> cen = as_const(a.center);
  const cptr<T> slot = as_const(a.slot);
  const double clip = a.clip ? *a.clip : 0.0;
  SweepAcc<T, NT> st;
  sweep_acc_init(st);
  const int64_t bstart = a.lo + blk * a.per_block;
  const int64_t bend = min(a.hi, bstart + a.per_block);
  const int64_t step = (int64_t)nw * 64;
  // one 64-sample step of the wave on the loaded inputs L of samples base + lane
  auto step_on = [&](const SweepLoad<T, D>& L, int64_t base) {
    const int64_t i = base + lane;
    T lw = ninf;
    if (i < bend) {
      if (a.append && i < a.append_hi) {
        T yv[D];
        const T q = slot_logdensity<T, D, RG>(slot, d, a.fam, L.x, yv);
        uint16_t reason = (q != q) ? kProposalInvalid : 0;
        T num = L.num;
        if (!reason) {
          const T value = logaddexp<T>(num, a.logmass + q);
          reason = proposal_reason(value);
          if (!reason) {
            num = value;
            a.lognums[i] = value;
          }
        }
        if (reason) record_failure(a.fail, i + 1, 0, reason);
        if (L.rid != 0) {
          const T w = L.tgt - (num - a.logtotal);
          const uint16_t wr = logweight_reason(w);
          if (wr) {
            record_failure(a.fail, i + 1, 0, wr);
          } else {
            a.logweights[i] = w;
            lw = w;
          }
        }
      } else {
        lw = L.lw;
      }
      if (a.clip) lw = fmin(lw, (T)clip);
    }
    sweep_accumulate<T, D, NT>(st, lw, L.x, d, cen, rows, lane);
  };
  // two load buffers in turn: the next step's inputs are in flight while a step runs
  SweepLoad<T, D> la, lb;
  int64_t base = bstart + (int64_t)wave * 64;
  if (base < bend) sweep_load<T, D, NT>(a, base + lane, bend, la);
  while (base < bend) {
    if (base + step < bend) sweep_load<T, D, NT>(a, base + step + lane, bend, lb);
    step_on(la, base);
    base += step;
    if (base >= bend) break;
    if (base + step < bend) sweep_load<T, D, NT>(a, base + step + lane, bend, la);
    step_on(lb, base);
    base += step;
  }
  sweep_epilogue<T, NT>(st, d, a.nv, a.partials + blk, a.pcap, lds);
}

template <class T, int D, int NT, bool RG = false>
__global__ __launch_bounds__(256) void k_sweep(SweepArgs<T> a) {
  sweep_block<T, D, NT, RG>(a, blockIdx.x);
}

// One AMIS round in one launch (samples are disjoint between the two roles):
//   blocks [0, nb_new)   draw the round's new samples (new_block, 64 per block,
//                        4 waves); wave 0, which holds them, accumulates and
//                        writes the moment partial record of the 64 samples at
//                        nrec_old + block;
//   blocks [nb_new, ...) sweep the retained samples [w.lo, w.hi) (append of the
//                        newest component, weights, summary, moments), records
//                        [0, nrec_old).
// The new-sample blocks (few, and bound by their mixture latency) then run
// beside the sweep instead of before it.
template <class T, int D, int NT, bool ST>
__device__ __forceinline__ void round_block(const RoundArgs<T>& a, const SweepArgs<T>& w, int nb_new, int nrec_old) {
  extern __shared__ double lds[];
  if ((int)blockIdx.x < nb_new) {
    constexpr int RS = sweep_row_stride(NT);
    T lw = -T(__builtin_huge_val());
    T x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = T(0);
    new_block<T, D, kDrawFused, ST>(a, blockIdx.x, lw, x);
    // after the mixture's barrier only wave 0 touches LDS: it alone holds the
    // block's samples, accumulates them and writes the record
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wave != 0) return;
    const int d = exact_dim<D>(a.d);
    SweepAcc<T, NT> st;
    sweep_acc_init(st);
    if (NT > 0)
      for (int q = lane; q < 64 * RS; q += 64) lds[q] = 0.0;
    sweep_accumulate<T, D, NT>(st, lw, x, d, as_const(w.center), lds, lane);
    wave_record<T, NT>(st, d, w.partials + nrec_old + blockIdx.x, w.pcap, lane);
  } else {
    sweep_block<T, D, NT>(w, (int64_t)blockIdx.x - nb_new);
  }
}

// Four waves per SIMD (at most 128 VGPRs): the LDS of four blocks per CU allows no more,
// and the new-sample role's interleaved densities would otherwise take more registers.
template <class T, int D, int NT, bool ST>
__global__ __launch_bounds__(256) void k_round(RoundArgs<T> a,
                                                                                        SweepArgs<T> w, int nb_new,
                                                                                        int nrec_old) {
  round_block<T, D, NT, ST>(a, w, nb_new, nrec_old);
}

// The same round for a group of same-shape jobs: blockIdx.y selects the job, whose
// arguments are read (scalar loads, before any store) from the group's argument tables.
template <class T, int D, int NT, bool ST>
__global__ __launch_bounds__(256) void k_round_group(const RoundArgs<T>* __restrict__ ga,
                                                     const SweepArgs<T>* __restrict__ gw, int nb_new,
                                                     int nrec_old) {
  const RoundArgs<T> a = ga[blockIdx.y];
  const SweepArgs<T> w = gw[blockIdx.y];
  round_block<T, D, NT, ST>(a, w, nb_new, nrec_old);
}
