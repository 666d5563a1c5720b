// Named flat transform layouts: the sampling coordinates z of a draw, split into named
// blocks (a scalar transform on one coordinate, an identity range, or a K-simplex on K - 1
// coordinates), mapped to the logical sample the target sees
// (src/proposal_composition.jl:179-240, src/native_execution.jl:556-712), and the built-in
// standard / Gaussian targets over those logical rows.
#pragma once

constexpr int kTransformSimplex = 4;   // IS_TRANSFORM_SIMPLEX
constexpr int kMaxLogical = 2 * kMaxDim;  // every block holds at least one coordinate

// One layout block on the device.  The simplex constants are computed on the host in the
// job's floating-point type (src/transforms.jl:341-345): 1/sqrt(K), (1 + 1/sqrt(K)) / (K - 1)
// and log(K)/2, the first term of the log Jacobian.
struct LayoutBlock {
  int32_t first;   // first sampling coordinate
  int32_t count;   // sampling coordinates (a simplex: K - 1)
  int32_t kind;    // transform kind
  int32_t out;     // first logical row
  double lower, upper;
  double inv_root, shared, half_log_k;
};

// _compensated_add (src/transforms.jl:347-352)
template <class T>
__device__ __forceinline__ void compensated_add(T& total, T& correction, T value) {
#pragma clang fp contract(off)
  const T corrected = value - correction;
  const T updated = total + corrected;
  correction = (updated - total) - corrected;
  total = updated;
}

// _native_simplex_transform! (src/native_execution.jl:556-626): K - 1 coordinates z (stride
// cap) to K weights x (stride cap) with the forward log Jacobian in lj; the checks in the
// reference's order, returning the first failure's reason bit.  No contraction into FMAs,
// so each rounding is the reference's.
template <class T>
__device__ uint16_t simplex_forward(const LayoutBlock& B, const T* __restrict__ z, T* __restrict__ x, int64_t cap,
                                    T& lj) {
#pragma clang fp contract(off)
  const int km1 = B.count;
  const int K = km1 + 1;
  T sum = T(0), sum_c = T(0);
  lj = T(0);
  for (int i = 0; i < km1; ++i) {
    const T v = z[i * cap];
    if (!isfinite(v)) return kTransformNonfiniteInput;
    x[i * cap] = v;
    compensated_add(sum, sum_c, v);
  }
  if (!isfinite(sum)) return kTransformNonfiniteOutput;
  const T inv_root = (T)B.inv_root, shared = (T)B.shared;
  T maxl = inv_root * sum;
  x[km1 * cap] = maxl;
  for (int i = 0; i < km1; ++i) {
    const T logit = x[i * cap] - shared * sum;
    if (!isfinite(logit)) return kTransformNonfiniteOutput;
    x[i * cap] = logit;
    maxl = logit > maxl ? logit : maxl;
  }
  T es = T(0), es_c = T(0);
  for (int i = 0; i < K; ++i) {
    const T w = exp(x[i * cap] - maxl);
    x[i * cap] = w;
    compensated_add(es, es_c, w);
  }
  if (!isfinite(es)) return kTransformNonfiniteOutput;
  T acc = (T)B.half_log_k, acc_c = T(0);
  for (int i = 0; i < K; ++i) {
    const T w = x[i * cap] / es;
    if (!(w > T(0))) return kTransformOutsideSupport;
    x[i * cap] = w;
    compensated_add(acc, acc_c, (T)log(w));
  }
  lj = acc;
  if (!isfinite(acc)) return kTransformNonfiniteLogjac;
  return 0;
}

// _native_transform_flat_blocks! (src/native_execution.jl:640-712) over one round's draws:
// z (rows of X) to the logical rows of XL block by block, the blocks' log Jacobians summed as
// the reference's recursion does (block 1 + (block 2 + (... + 0))), into lognums, where the
// pre-target pass takes it.  The first failing block fails the draw (1-based block index and
// its reason bits); a draw that already failed (NaN mark) is left as it is.  One thread per
// sample, grid over n.
template <class T>
__global__ __launch_bounds__(256) void k_layout_forward(const LayoutBlock* __restrict__ blocks, int nb,
                                                        const T* __restrict__ X, T* __restrict__ XL, int64_t cap,
                                                        int64_t first, int64_t n, T* __restrict__ lognums,
                                                        unsigned long long* fail) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t gi = first + i;
  const T mark = lognums[gi];
  if (mark != mark) return;
  T acc = T(0);
  uint16_t reason = 0;
  int failed = 0;
  for (int b = nb - 1; b >= 0; --b) {
    const LayoutBlock B = blocks[b];
    const T* z = X + (int64_t)B.first * cap + gi;
    T* x = XL + (int64_t)B.out * cap + gi;
    T lj = T(0);
    uint16_t r = 0;
    if (B.kind == kTransformSimplex) {
      r = simplex_forward<T>(B, z, x, cap, lj);
    } else if (B.kind == 0) {
      for (int c = 0; c < B.count; ++c) {
        const T v = z[c * cap];
        x[c * cap] = v;
        if (!isfinite(v) && !r) r = kTransformNonfiniteInput;
      }
    } else {
      T xv;
      r = transform_forward<T>(B.kind, B.lower, B.upper, z[0], xv, lj);
      x[0] = xv;
    }
    if (r) {
      reason = r;
      failed = b + 1;
    }
    acc = lj + acc;
  }
  if (reason) {
    record_failure(fail, i + 1, failed, reason);
    lognums[gi] = T(__builtin_nan(""));
  } else {
    lognums[gi] = acc;
  }
}

// The built-in standard and Gaussian targets over the logical rows of XL (dl of them): the
// arithmetic of inline_target on a packed target slot of dimension dl, one thread per sample.
template <class T>
__global__ __launch_bounds__(256) void k_target_logical(int tkind, int dl, const T* __restrict__ tslot,
                                                        const T* __restrict__ XL, int64_t cap, int64_t first,
                                                        int64_t n, T* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T* x = XL + first + i;
  if (tkind == kTStandard) {
    T s = T(0);
    for (int c = 0; c < dl; ++c) {
      const T v = x[c * cap];
      s += v * v;
    }
    out[i] = -s / T(2);
    return;
  }
  T y[kMaxLogical];
  T r2 = T(0);
  for (int r = 0; r < dl; ++r) {
    T v = x[r * cap] - tslot[r];
    const int row = 2 * dl + r * (r + 1) / 2;
    for (int c = 0; c < r; ++c) v -= tslot[row + c] * y[c];
    v *= tslot[dl + r];
    y[r] = v;
    r2 += v * v;
  }
  out[i] = isfinite(r2) ? tslot[slot_lognorm_off(dl)] - T(0.5) * r2 : T(__builtin_nan(""));
}
