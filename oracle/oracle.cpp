// CPU FP64 restatement of the ImportanceSamplers.jl sampling hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the
// smoke check and the CPU-baseline leg of the benchmark call it as a checker.
// Nothing in the product links, loads or calls it.
//
// Every routine cites the reference file:line it restates (paths relative to
// the reference repository root).  The arithmetic follows the reference's
// association order wherever the reference fixes one; summations that the
// reference leaves to Julia's `sum`/BLAS are done sequentially here (their
// rounding order is not specified by the reference, so parity with them is a
// tolerance, not bits).
//
// Parity status of each function (what pins it) is recorded in the header
// comment of oracle/__init__.py.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

// Reason bits: src/execution.jl:758-767.
constexpr uint16_t kTargetNaN = 0x0100;
constexpr uint16_t kTargetPosInf = 0x0200;
constexpr uint16_t kProposalInvalid = 0x0400;
constexpr uint16_t kGeneratedNonfinite = 0x0800;
constexpr uint16_t kLogweightInvalid = 0x1000;
constexpr uint16_t kDrawExhausted = 0x0010;
constexpr uint16_t kTransformReasons = 0x000f;
constexpr uint16_t kDrawReasons = kTransformReasons | kGeneratedNonfinite | kDrawExhausted;
constexpr int kGammaAttempts = 8;  // src/execution.jl:557

// ---------------------------------------------------------------------------
// LogExpFunctions 1.0.1 (third-party, pinned in benchmark/comparison/Manifest.toml:964-969):
// log1pexp uses the Maechler (2012) thresholds x0 = -prec*log(2),
// x1 = (prec-1)*log(2)/2, x2 = -x0 - log(-x0)*(1 + 1/x0); logaddexp(a, b) =
// max(a, b) + log1pexp(-|a - b|) with diff = 0 when a == b (so +-Inf pairs give
// max + log 2) and NaN propagation through `max`.
inline double log1pexp(double x) {
  const double x0 = -53.0 * M_LN2;
  const double x1 = 52.0 * M_LN2 / 2.0;
  const double x2 = -x0 - std::log(-x0) * (1.0 + 1.0 / x0);
  if (x < x0) return std::exp(x);
  if (x < x1) return std::log1p(std::exp(x));
  if (x < x2) return x + std::exp(-x);
  return x;
}

inline double logaddexp(double a, double b) {
  double diff, mx;
  if (a < b) {
    diff = a - b;
    mx = b;
  } else {
    diff = (a == b) ? 0.0 : b - a;
    mx = !std::isnan(b) ? a : b;
  }
  return mx + log1pexp(diff);
}

// ---------------------------------------------------------------------------
// Failure record: src/execution.jl:669-699 (pack, atomic count/max).
inline uint64_t pack_failure(int64_t logical_index, int block, uint16_t reason) {
  uint32_t inv_index = (uint32_t)(0xffffffffu - (uint32_t)logical_index + 1u);
  uint16_t inv_block = (uint16_t)(0xffffu - (uint16_t)block + 1u);
  return ((uint64_t)inv_index << 32) | ((uint64_t)inv_block << 16) | (uint64_t)reason;
}

struct FailureRecord {
  uint64_t w[3] = {0, 0, 0};
  void record(int64_t logical_index, int block, uint16_t reason) {
    uint64_t packed = pack_failure(logical_index, block, reason);
    w[0] += 1;
    w[1] = std::max(w[1], packed);
    if (reason & kDrawReasons) w[2] = std::max(w[2], packed);
  }
  void merge(const FailureRecord& o) {
    w[0] += o.w[0];
    w[1] = std::max(w[1], o.w[1]);
    w[2] = std::max(w[2], o.w[2]);
  }
};

// Chunked worker split of src/execution.jl:9-21 (cld(n, nthreads) chunks).
template <class F>
void parallel_for(int64_t n, int nthreads, F&& body) {
  if (n <= 0) return;
  if (nthreads <= 1 || n < 2) {
    body(0, n, 0);
    return;
  }
  int64_t tasks = std::min<int64_t>(nthreads, n);
  int64_t chunk = (n + tasks - 1) / tasks;
  std::vector<std::thread> workers;
  int worker = 0;
  for (int64_t first = 0; first < n; first += chunk, ++worker) {
    int64_t last = std::min(n, first + chunk);
    workers.emplace_back([&body, first, last, worker] { body(first, last, worker); });
  }
  for (auto& t : workers) t.join();
}

// ---------------------------------------------------------------------------
// Proposals.  A proposal slot carries family, scale kind, location and scale
// storage exactly as src/proposals.jl:158-189 and the packed histories /
// banks (src/adaptive_gaussian.jl:46-60, src/proposal_banks.jl:143-163).
enum Kind { kSpherical = 0, kDiagonal = 1, kFactor = 2 };

struct Prop {
  int family;     // 0 Gaussian, 1 Student-t
  double dof;
  int kind;       // Kind
  int d;
  const double* mu;     // d
  const double* scale;  // spherical: [1]; diagonal: [d]; factor: d*d column-major lower
  double lognorm;
};

inline double factor_at(const Prop& p, int r, int c) { return p.scale[r + (int64_t)c * p.d]; }
inline double diag_scale(const Prop& p, int c) {
  return p.kind == kSpherical ? p.scale[0] : p.scale[c];
}

// _standardized_difference: src/student_t_range.jl:76-82.
inline double standardized_difference(double sample, double location, double scale) {
  double difference = sample - location;
  if (std::isfinite(difference) || !std::isfinite(sample) || !std::isfinite(location))
    return difference / scale;
  return sample / scale - location / scale;
}

// _radial_logdensity: src/proposals.jl:24-55 for the ordinary range; the range-risk
// branch and the extended-range log radius (src/student_t_range.jl) are not restated
// here: the engine's extended-range densities are pinned by closed forms evaluated in
// mpmath (the reference's BigFloat known answers, test/native_proposals.jl:205-306).
inline double radial_logdensity(const Prop& p, double r2, const double* coords) {
  const int d = p.d;
  if (p.family == 0) {
    double half = 0.5 * r2;
    if (!std::isfinite(r2)) {
      half = 0.0;
      for (int c = 0; c < d; ++c) half += (0.5 * coords[c]) * coords[c];
    }
    return p.lognorm - half;
  }
  const double dof = p.dof;
  const double ratio = r2 / dof;
  double radial;
  if (r2 < DBL_MIN) {
    const double s = std::sqrt(dof);
    double scaled = 0.0;
    for (int c = 0; c < d; ++c) {
      double v = coords[c] / s;
      scaled += v * v;
    }
    radial = std::log1p(scaled);
  } else if (std::isfinite(ratio)) {
    radial = std::log1p(ratio);
  } else if (std::isfinite(r2)) {
    // _student_t_overflow_logradius: src/proposals.jl:58-70.
    radial = std::log(r2) - std::log(dof);
  } else {
    double s = 0.0;
    for (int c = 0; c < d; ++c) s = std::max(s, std::fabs(coords[c]));
    if (!std::isfinite(s)) return kNaN;  // the extended-range path (not restated)
    double scaled = 0.0;
    for (int c = 0; c < d; ++c) {
      double v = coords[c] / s;
      scaled += v * v;
    }
    radial = log1pexp(2.0 * std::log(s) + std::log(scaled) - std::log(dof));
  }
  return p.lognorm - (dof + (double)d) / 2.0 * radial;
}

// Unpacked proposal log density at x: src/proposals.jl:603-697.
double plain_logdensity(const Prop& p, const double* x, double* work) {
  double r2 = 0.0;
  if (p.kind == kFactor) {
    for (int r = 0; r < p.d; ++r) {
      double s = x[r] - p.mu[r];
      for (int c = 0; c < r; ++c) s -= factor_at(p, r, c) * work[c];
      s /= factor_at(p, r, r);
      work[r] = s;
      r2 += s * s;
    }
  } else {
    for (int c = 0; c < p.d; ++c) {
      double s = standardized_difference(x[c], p.mu[c], diag_scale(p, c));
      work[c] = s;
      r2 += s * s;
    }
  }
  return radial_logdensity(p, r2, work);
}

// Packed (history / bank) log density: src/native_execution.jl:346-384
// (factor) and :447-469 (diagonal / scalar).
double packed_logdensity(const Prop& p, const double* x, double* work) {
  double r2 = 0.0;
  if (p.kind == kFactor) {
    for (int r = 0; r < p.d; ++r) {
      double s = x[r] - p.mu[r];
      for (int c = 0; c < r; ++c) s -= factor_at(p, r, c) * work[c];
      s /= factor_at(p, r, r);
      work[r] = s;
      r2 += s * s;
    }
    return radial_logdensity(p, r2, work);
  }
  for (int c = 0; c < p.d; ++c) {
    double s = (x[c] - p.mu[c]) / diag_scale(p, c);
    r2 += s * s;
  }
  if (!std::isfinite(r2)) {
    for (int c = 0; c < p.d; ++c) work[c] = standardized_difference(x[c], p.mu[c], diag_scale(p, c));
  } else {
    for (int c = 0; c < p.d; ++c) work[c] = (x[c] - p.mu[c]) / diag_scale(p, c);
  }
  return radial_logdensity(p, r2, work);
}

// Unpacked affine draw: src/proposals.jl:543-601.
inline double plain_coordinate(const Prop& p, const double* z, int c) {
  if (p.kind == kFactor) {
    double value = 0.0;
    for (int j = 0; j <= c; ++j) value += factor_at(p, c, j) * z[j];
    return p.mu[c] + value;
  }
  return p.mu[c] + diag_scale(p, c) * z[c];
}

// Packed draw: src/native_execution.jl:226-252, src/adaptive_gaussian_execution.jl:35-62.
inline double packed_coordinate(const Prop& p, const double* z, int c) {
  if (p.kind == kFactor) {
    double value = p.mu[c];
    for (int j = 0; j <= c; ++j) value += factor_at(p, c, j) * z[j];
    return value;
  }
  return p.mu[c] + diag_scale(p, c) * z[c];
}

// _gamma_parameters / _gamma_candidate: src/proposals.jl:738-751;
// _buffered_unit_gamma: src/native_execution.jl:164-180.
bool buffered_unit_gamma(double shape, const double* normals, const double* uniforms, double* out) {
  const double adjusted = shape < 1.0 ? shape + 1.0 : shape;
  const double offset = adjusted - 1.0 / 3.0;
  const double coefficient = (1.0 / std::sqrt(offset)) / 3.0;
  for (int a = 0; a < kGammaAttempts; ++a) {
    const double normal = normals[a];
    const double uniform = uniforms[a];
    const double root = 1.0 + coefficient * normal;
    if (!(root > 0.0)) continue;
    const double cube = root * root * root;
    const double n2 = normal * normal;
    const bool accepted = uniform < 1.0 - 0.0331 * (n2 * n2) ||
                          std::log(uniform) < 0.5 * n2 + offset * (1.0 - cube + std::log(cube));
    if (!accepted) continue;
    double gamma = offset * cube;
    if (shape < 1.0) gamma *= std::pow(uniforms[kGammaAttempts], 1.0 / shape);
    if (!(std::isfinite(gamma) && gamma > 0.0)) break;
    *out = gamma;
    return true;
  }
  return false;
}

// _student_t_radial_multiplier: src/native_execution.jl:147-162.
bool student_t_multiplier(double dof, const double* radial_normals, const double* radial_uniforms,
                          double* out) {
  if (radial_uniforms == nullptr || dof == 1.0) {
    double denominator = std::fabs(radial_normals[0]);
    if (!(std::isfinite(denominator) && denominator > 0.0)) return false;
    *out = 1.0 / denominator;
    return true;
  }
  double gamma;
  if (!buffered_unit_gamma(dof / 2.0, radial_normals, radial_uniforms, &gamma)) return false;
  *out = std::sqrt(dof / (2.0 * gamma));
  return true;
}

inline int radial_normal_stride(int family, double dof) {  // src/proposal_banks.jl:219
  return family == 0 ? 0 : (dof == 1.0 ? 1 : kGammaAttempts);
}
inline int radial_uniform_stride(int family, double dof) {  // src/proposal_banks.jl:220-221
  return family == 0 ? 0 : (dof == 1.0 ? 0 : kGammaAttempts + (dof < 2.0 ? 1 : 0));
}

// ---------------------------------------------------------------------------
// Built-in targets (the engine's own target kernels; the reference's are user
// closures evaluated per sample, src/execution.jl:775-783).
//   0  standard:  -sum(abs2, x)/2                                  (no params)
//   1  gaussian:  lognorm - r^2/2, r = L \ (x - m)                 params [m(d), L(d*d col-major), lognorm]
//   2  logistic:  sum_i y_i*eta_i - log1pexp(eta_i) + prior         params [n, p, prior_var, X(n*p row-major), y(n)]
//   3  linreg:    Gaussian likelihood + N(0, tau^2 I) prior         params [n, p, sigma, tau, X(n*p row-major), y(n)]
//   5  transformed: inner(T(z)) + log|dT/dz| per coordinate       params [inner kind, kinds(d), bounds(2d), inner params]
//   6  layout:     inner(T(z)) + log|J| of a flat selector layout  params [inner kind, B, B x (first, count, kind,
//                  lower, upper), inner params of the logical dimension]
enum TargetKind { kTStandard = 0, kTGaussian = 1, kTLogistic = 2, kTLinreg = 3, kTTransformed = 5, kTLayout = 6 };

// Scalar forward transforms z -> x with log|dx/dz| (src/transforms.jl:160-235, helpers
// :295-333), FP64; returns the reason bits (src/transforms.jl:106-109), 0 on success.
// kind 0 identity, 1 positive, 2 softplus, 3 interval (an infinite bound is open).
inline double tf_softplus(double z) { return z > 0.0 ? z + std::log1p(std::exp(-z)) : std::log1p(std::exp(z)); }
inline double tf_log_positive_difference(double upper, double lower) {
  if (lower < 0.0 && upper > 0.0) {
    const double larger = std::max(-lower, upper), smaller = std::min(-lower, upper);
    return std::log(larger) + std::log1p(smaller / larger);
  }
  return std::log(upper - lower);
}
int tf_forward(int kind, double lo, double hi, double z, double* x, double* lj) {
  *x = z;
  *lj = kind == 0 ? 0.0 : z;
  if (!std::isfinite(z)) return 0x0001;
  if (kind == 0) return 0;
  if (kind == 1) {
    *x = std::exp(z);
    if (!std::isfinite(*x)) return 0x0002;
    return *x > 0.0 ? 0 : 0x0008;
  }
  if (kind == 2) {
    *x = tf_softplus(z);
    *lj = z - *x;
    if (!std::isfinite(*x)) return 0x0002;
    if (!std::isfinite(*lj)) return 0x0004;
    return *x > 0.0 ? 0 : 0x0008;
  }
  const bool hl = std::isfinite(lo), hh = std::isfinite(hi);
  if (hl && !hh) {  // lower + PositiveTransform
    const double dist = std::exp(z);
    *x = dist;
    if (!std::isfinite(dist)) return 0x0002;
    if (!(dist > 0.0)) return 0x0008;
    *x = lo + dist;
    if (!std::isfinite(*x)) return 0x0002;
    return *x > lo ? 0 : 0x0008;
  }
  if (!hl && hh) {  // upper - PositiveTransform
    const double dist = std::exp(z);
    *x = dist;
    if (!std::isfinite(dist)) return 0x0002;
    if (!(dist > 0.0)) return 0x0008;
    *x = hi - dist;
    if (!std::isfinite(*x)) return 0x0002;
    return *x < hi ? 0 : 0x0008;
  }
  if (!hl) return 0;
  // _logistic, _bounded_interval_value, and the log Jacobian
  const double p = z >= 0.0 ? 1.0 / (1.0 + std::exp(-z)) : std::exp(z) / (1.0 + std::exp(z));
  if (lo < 0.0 && hi > 0.0)
    *x = (1.0 - p) * lo + p * hi;
  else
    *x = p <= 0.5 ? lo + p * (hi - lo) : hi - (1.0 - p) * (hi - lo);
  *lj = (tf_log_positive_difference(hi, lo) + -tf_softplus(-z)) + -tf_softplus(z);
  if (!std::isfinite(*x)) return 0x0002;
  if (!std::isfinite(*lj)) return 0x0004;
  return (lo < *x && *x < hi) ? 0 : 0x0008;
}

// _compensated_add (src/transforms.jl:347-352)
inline void compensated_add(double& total, double& correction, double value) {
  const double corrected = value - correction;
  const double updated = total + corrected;
  correction = (updated - total) - corrected;
  total = updated;
}

// _native_simplex_transform! (src/native_execution.jl:556-626) with the embedding constants
// of src/transforms.jl:341-345, FP64: K - 1 coordinates z to K weights x and the forward
// log Jacobian; the checks in the reference's order, the first failure's reason bit.
uint16_t simplex_forward(int km1, const double* z, double* x, double* lj) {
  const int K = km1 + 1;
  double sum = 0.0, sum_c = 0.0;
  *lj = 0.0;
  for (int i = 0; i < km1; ++i) {
    if (!std::isfinite(z[i])) return 0x0001;
    x[i] = z[i];
    compensated_add(sum, sum_c, z[i]);
  }
  if (!std::isfinite(sum)) return 0x0002;
  const double inv_root = 1.0 / std::sqrt((double)K);
  const double shared = (1.0 + inv_root) / (double)(K - 1);
  double maxl = inv_root * sum;
  x[km1] = maxl;
  for (int i = 0; i < km1; ++i) {
    const double logit = x[i] - shared * sum;
    if (!std::isfinite(logit)) return 0x0002;
    x[i] = logit;
    maxl = std::max(maxl, logit);
  }
  double es = 0.0, es_c = 0.0;
  for (int i = 0; i < K; ++i) {
    x[i] = std::exp(x[i] - maxl);
    compensated_add(es, es_c, x[i]);
  }
  if (!std::isfinite(es)) return 0x0002;
  double acc = 0.5 * std::log((double)K), acc_c = 0.0;
  for (int i = 0; i < K; ++i) {
    const double w = x[i] / es;
    if (!(w > 0.0)) return 0x0008;
    x[i] = w;
    compensated_add(acc, acc_c, std::log(w));
  }
  *lj = acc;
  if (!std::isfinite(acc)) return 0x0004;
  return 0;
}

// _native_transform_flat_blocks! (src/native_execution.jl:628-712): blocks (first, count,
// kind, lower, upper) in name order map z to the logical x block by block; the blocks' log
// Jacobians are summed as the reference's recursion does, block 1 + (block 2 + (... + 0)).
// Returns the first failing block (1-based) with its reason in *reason, else 0.
int layout_forward(int nb, const double* blocks, const double* z, double* x, double* lj, uint16_t* reason) {
  std::vector<double> ljs(nb, 0.0);
  int failed = 0;
  *reason = 0;
  int row = 0;
  for (int b = 0; b < nb && !failed; ++b) {
    const double* B = blocks + 5 * b;
    const int first = (int)B[0], count = (int)B[1], kind = (int)B[2];
    uint16_t r = 0;
    if (kind == 4) {
      r = simplex_forward(count, z + first, x + row, &ljs[b]);
      row += count + 1;
    } else if (kind == 0) {
      for (int c = 0; c < count; ++c) {
        x[row + c] = z[first + c];
        if (!std::isfinite(z[first + c]) && !r) r = 0x0001;
      }
      row += count;
    } else {
      r = (uint16_t)tf_forward(kind, B[3], B[4], z[first], &x[row], &ljs[b]);
      row += 1;
    }
    if (r) {
      failed = b + 1;
      *reason = r;
    }
  }
  double acc = 0.0;
  for (int b = nb - 1; b >= 0; --b) acc = ljs[b] + acc;
  *lj = acc;
  return failed;
}

int layout_dim(int nb, const double* blocks) {
  int dl = 0;
  for (int b = 0; b < nb; ++b) dl += (int)blocks[5 * b + 1] + ((int)blocks[5 * b + 2] == 4 ? 1 : 0);
  return dl;
}

double target_eval(int kind, const double* tp, int d, const double* x, double* work) {
  switch (kind) {
    case kTLayout: {
      const int inner = (int)tp[0];
      const int nb = (int)tp[1];
      const double* blocks = tp + 2;
      const int dl = layout_dim(nb, blocks);
      std::vector<double> xl(dl), w(std::max(dl, 1));
      double lj;
      uint16_t reason;
      if (layout_forward(nb, blocks, x, xl.data(), &lj, &reason)) return kNaN;
      return target_eval(inner, tp + 2 + 5 * nb, dl, xl.data(), w.data()) + lj;
    }
    case kTTransformed: {
      const int inner = (int)tp[0];
      const double* kinds = tp + 1;
      const double* bounds = tp + 1 + d;
      std::vector<double> xl(d);
      double lj = 0.0;
      for (int c = 0; c < d; ++c) {
        double ljc;
        if (tf_forward((int)kinds[c], bounds[2 * c], bounds[2 * c + 1], x[c], &xl[c], &ljc)) return kNaN;
        lj += ljc;
      }
      return target_eval(inner, tp + 1 + 3 * d, d, xl.data(), work) + lj;
    }
    case kTStandard: {
      double s = 0.0;
      for (int c = 0; c < d; ++c) s += x[c] * x[c];
      return -s / 2.0;
    }
    case kTGaussian: {
      const double* m = tp;
      const double* L = tp + d;
      const double lognorm = tp[d + (int64_t)d * d];
      double r2 = 0.0;
      for (int r = 0; r < d; ++r) {
        double v = x[r] - m[r];
        for (int c = 0; c < r; ++c) v -= L[r + (int64_t)c * d] * work[c];
        work[r] = v / L[r + (int64_t)r * d];
      }
      for (int r = 0; r < d; ++r) r2 += work[r] * work[r];
      return lognorm - 0.5 * r2;
    }
    case kTLogistic: {
      const int64_t n = (int64_t)tp[0];
      const int p = (int)tp[1];
      const double prior_var = tp[2];
      const double* X = tp + 3;
      const double* y = X + n * p;
      double ll = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        double eta = 0.0;
        for (int c = 0; c < p; ++c) eta += X[i * p + c] * x[c];
        ll += y[i] * eta - log1pexp(eta);
      }
      double b2 = 0.0;
      for (int c = 0; c < p; ++c) b2 += x[c] * x[c];
      return ll - 0.5 * b2 / prior_var - 0.5 * p * std::log(2.0 * M_PI * prior_var);
    }
    case kTLinreg: {
      const int64_t n = (int64_t)tp[0];
      const int p = (int)tp[1];
      const double sigma = tp[2];
      const double tau = tp[3];
      const double* X = tp + 4;
      const double* y = X + n * p;
      double rss = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        double eta = 0.0;
        for (int c = 0; c < p; ++c) eta += X[i * p + c] * x[c];
        double r = y[i] - eta;
        rss += r * r;
      }
      double b2 = 0.0;
      for (int c = 0; c < p; ++c) b2 += x[c] * x[c];
      return -0.5 * rss / (sigma * sigma) - (double)n * std::log(sigma) - 0.5 * (double)n * std::log(2.0 * M_PI)
             - 0.5 * b2 / (tau * tau) - (double)p * std::log(tau) - 0.5 * (double)p * std::log(2.0 * M_PI);
    }
  }
  return kNaN;
}

// _native_target_reason: src/execution.jl:769-773.
inline uint16_t target_reason(double v) {
  if (std::isnan(v)) return kTargetNaN;
  if (v == kInf) return kTargetPosInf;
  return 0;
}
// _native_proposal_reason: src/execution.jl:820-823.
inline uint16_t proposal_reason(double v) {
  return (std::isnan(v) || v == -kInf) ? kProposalInvalid : 0;
}
// _subtract_logweight: src/methods/importance_sampling.jl:1261-1266.
inline uint16_t subtract_logweight(double t, double q, double* out) {
  double lw = t - q;
  *out = lw;
  return (std::isnan(lw) || lw == kInf) ? kLogweightInvalid : 0;
}

// ---------------------------------------------------------------------------
// Deterministic exponential used by the resampling CDF.  Cody-Waite
// reduction x = k*ln2 + r (two-part ln2), |r| <= ln2/2, then the degree-13
// Taylor polynomial of e^r by Horner with fused multiply-adds, then scaling by
// 2^k.  The GPU CDF kernel evaluates the identical operation sequence, so the
// CDF is bit-reproducible on both sides.
inline double det_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return kInf;
  if (x < -745.2) return 0.0;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  const double k = std::nearbyint(x * inv_ln2);
  double r = std::fma(-k, ln2_hi, x);
  r = std::fma(-k, ln2_lo, r);
  double p = 1.0 / 6227020800.0;                 // 1/13!
  p = std::fma(p, r, 1.0 / 479001600.0);          // 1/12!
  p = std::fma(p, r, 1.0 / 39916800.0);
  p = std::fma(p, r, 1.0 / 3628800.0);
  p = std::fma(p, r, 1.0 / 362880.0);
  p = std::fma(p, r, 1.0 / 40320.0);
  p = std::fma(p, r, 1.0 / 5040.0);
  p = std::fma(p, r, 1.0 / 720.0);
  p = std::fma(p, r, 1.0 / 120.0);
  p = std::fma(p, r, 1.0 / 24.0);
  p = std::fma(p, r, 1.0 / 6.0);
  p = std::fma(p, r, 0.5);
  p = std::fma(p, r, 1.0);
  p = std::fma(p, r, 1.0);
  int ki = (int)k;
  // two-step scaling keeps subnormal results exact for k < -1022
  if (ki < -1000) return std::ldexp(std::ldexp(p, ki + 600), -600);
  return std::ldexp(p, ki);
}

}  // namespace

extern "C" {

// ---- scalar helpers exported for tests ----------------------------------------------
double or_logaddexp(double a, double b) { return logaddexp(a, b); }
double or_log1pexp(double x) { return log1pexp(x); }
double or_det_exp(double x) { return det_exp(x); }

// _student_t_logconstant: src/proposals.jl:410-437.
double or_student_t_logconstant(double dof, int dimension) {
  double value = -(double)dimension * std::log(2.0 * M_PI) / 2.0;
  if (dimension % 2 == 1) {
    double shifted = dof;
    double correction = 0.0;
    while (shifted < 32.0) {
      correction += shifted < 1.0
                        ? (std::log(shifted) + std::log(shifted + 2.0)) / 2.0 - std::log1p(shifted)
                        : std::log1p(-((1.0 / (shifted + 1.0)) * (1.0 / (shifted + 1.0)))) / 2.0;
      shifted += 2.0;
    }
    double r = 1.0 / shifted;
    double r2 = r * r;
    // evalpoly(r2, (-1/4, 1/24, -1/20, 17/112, -31/36, 691/88)) by Horner.
    double poly = 691.0 / 88.0;
    poly = poly * r2 + (-31.0 / 36.0);
    poly = poly * r2 + (17.0 / 112.0);
    poly = poly * r2 + (-1.0 / 20.0);
    poly = poly * r2 + (1.0 / 24.0);
    poly = poly * r2 + (-1.0 / 4.0);
    correction += r * poly;
    value += correction;
  }
  double acc = 0.0;
  for (int offset = (dimension % 2 == 1) ? 1 : 0; offset <= dimension - 2; offset += 2) {
    double x = (double)offset;
    acc += x <= dof ? std::log1p(x / dof) : std::log(x) + std::log1p(dof / x) - std::log(dof);
  }
  return value + acc;
}

// Proposal log density for a single point, unpacked (packed=0) or packed (packed=1).
double or_proposal_logdensity(int family, double dof, int kind, int d, const double* mu,
                              const double* scale, double lognorm, const double* x, int packed) {
  Prop p{family, dof, kind, d, mu, scale, lognorm};
  std::vector<double> work(d);
  return packed ? packed_logdensity(p, x, work.data()) : plain_logdensity(p, x, work.data());
}

int or_transform_forward(int kind, double lo, double hi, double z, double* x, double* lj) {
  return tf_forward(kind, lo, hi, z, x, lj);
}

int or_simplex_forward(int km1, const double* z, double* x, double* lj) { return simplex_forward(km1, z, x, lj); }

// n samples (d x n column-major z) through a layout: x (dl x n), lj (n), reason (n) and the
// failing block (n, 0 when none).
void or_layout_forward(int nb, const double* blocks, int d, int64_t n, const double* z, double* x, double* lj,
                       int32_t* reason, int32_t* block) {
  const int dl = layout_dim(nb, blocks);
  for (int64_t i = 0; i < n; ++i) {
    uint16_t r;
    block[i] = layout_forward(nb, blocks, z + i * d, x + i * dl, &lj[i], &r);
    reason[i] = r;
  }
}

double or_target(int kind, const double* tp, int d, const double* x) {
  std::vector<double> work(std::max(d, 1));
  return target_eval(kind, tp, d, x, work.data());
}

int or_student_t_multiplier(double dof, const double* radial_normals, const double* radial_uniforms,
                            double* out) {
  return student_t_multiplier(dof, radial_normals, radial_uniforms, out) ? 0 : 1;
}

// ---- plain IS: _native_fused_kernel!, src/native_execution.jl:741-812 -----------------
// normals: stride*n, stride = d (+1 for nu == 1, +8 otherwise for Student-t);
// uniforms: (8 + (nu < 2)) * n for Student-t with nu != 1, else unused.
void or_plain_is(int family, double dof, int kind, int d, const double* mu, const double* scale,
                 double lognorm, int tkind, const double* tp, int64_t n, const double* normals,
                 const double* uniforms, int nthreads, double* samples, double* logweights,
                 uint64_t* failure) {
  Prop p{family, dof, kind, d, mu, scale, lognorm};
  const int nstride = d + radial_normal_stride(family, dof);
  const int ustride = radial_uniform_stride(family, dof);
  std::vector<FailureRecord> records(std::max(nthreads, 1));
  parallel_for(n, nthreads, [&](int64_t first, int64_t last, int worker) {
    std::vector<double> z(d), work(std::max(d, 1));
    FailureRecord& rec = records[worker];
    for (int64_t i = first; i < last; ++i) {
      const double* nz = normals + i * nstride;
      double m = 1.0;
      if (family == 1) {
        bool ok = student_t_multiplier(dof, nz + d, ustride ? uniforms + i * ustride : nullptr, &m);
        if (!ok) {
          rec.record(i + 1, 0, kDrawExhausted);
          continue;
        }
      }
      for (int c = 0; c < d; ++c) z[c] = nz[c] * m;
      double* x = samples + i * d;
      bool valid = true;
      for (int c = 0; c < d; ++c) {
        x[c] = plain_coordinate(p, z.data(), c);
        valid &= std::isfinite(x[c]);
      }
      if (!valid) {
        rec.record(i + 1, 0, kGeneratedNonfinite);
        continue;
      }
      double t = target_eval(tkind, tp, d, x, work.data());
      uint16_t tr = target_reason(t);
      if (tr) {
        rec.record(i + 1, 0, tr);
        continue;
      }
      double q = plain_logdensity(p, x, work.data());
      uint16_t qr = proposal_reason(q);
      if (qr) {
        rec.record(i + 1, 0, qr);
        continue;
      }
      double lw;
      uint16_t wr = subtract_logweight(t, q, &lw);
      if (wr) {
        rec.record(i + 1, 0, wr);
        continue;
      }
      logweights[i] = lw;
    }
  });
  FailureRecord all;
  for (auto& r : records) all.merge(r);
  std::memcpy(failure, all.w, sizeof(all.w));
}

// ---- log-weight summary: src/execution.jl:103-186 (host branch) ------------------------
void or_logweight_summary(const double* lw, int64_t n, double* ess, double* logz) {
  double mx = -kInf;
  for (int64_t i = 0; i < n; ++i) mx = std::max(mx, lw[i]);
  if (mx == -kInf) {
    *ess = 0.0;
    *logz = -kInf;
    return;
  }
  double s = 0.0, s2 = 0.0;
  for (int64_t i = 0; i < n; ++i) s += std::exp(lw[i] - mx);
  for (int64_t i = 0; i < n; ++i) {
    double e = std::exp(lw[i] - mx);
    s2 += e * e;
  }
  *ess = (s * s) / s2;
  *logz = mx + std::log(s) - std::log((double)n);
}

// _merge_logweight_moments: src/execution.jl:165-173 (device monoid).
void or_merge_logweight_moments(const double* a, const double* b, double* out) {
  if (a[1] == 0.0) {
    out[0] = b[0]; out[1] = b[1]; out[2] = b[2];
    return;
  }
  if (b[1] == 0.0) {
    out[0] = a[0]; out[1] = a[1]; out[2] = a[2];
    return;
  }
  double m = std::max(a[0], b[0]);
  double l = std::exp(a[0] - m), r = std::exp(b[0] - m);
  out[0] = m;
  out[1] = l * a[1] + r * b[1];
  out[2] = (l * l) * a[2] + (r * r) * b[2];
}

// ---- results API: src/results.jl:548-626, src/statistics.jl:11-103,188-237 -----------
// LSE accumulator (max, scaled_sum) with the host branch of _logsumexp_accumulator.
void or_logsumexp_accumulator(const double* lw, int64_t n, double* out) {
  double mx = -kInf;
  for (int64_t i = 0; i < n; ++i) mx = std::max(mx, lw[i]);
  if (mx == -kInf) {
    out[0] = -kInf;
    out[1] = 0.0;
    return;
  }
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += std::exp(lw[i] - mx);
  out[0] = mx;
  out[1] = s;
}

double or_lognormalizer(const double* lw, int64_t n) {
  double acc[2];
  or_logsumexp_accumulator(lw, n, acc);
  if (acc[0] == -kInf) return -kInf;
  return (acc[0] + std::log(acc[1])) - std::log((double)n);
}

int or_normalized_weights(const double* lw, int64_t n, double* w) {
  double acc[2];
  or_logsumexp_accumulator(lw, n, acc);
  if (acc[0] == -kInf) return 1;  // AllZeroWeightsError
  double inv = 1.0 / acc[1];
  for (int64_t i = 0; i < n; ++i) w[i] = std::exp(lw[i] - acc[0]) * inv;
  return 0;
}

// mean (d), var (d), cov (d*d col-major) for samples stored d x n column-major.
int or_moments(const double* x, const double* lw, int64_t n, int d, double* mean, double* var,
               double* cov) {
  std::vector<double> w(n);
  if (or_normalized_weights(lw, n, w.data())) return 1;
  for (int c = 0; c < d; ++c) mean[c] = 0.0;
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < d; ++c) mean[c] += x[i * d + c] * w[i];
  for (int c = 0; c < d; ++c) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      double dev = x[i * d + c] - mean[c];
      double contrib = dev * (w[i] * dev);
      if (!std::isfinite(dev)) {
        double sw = std::sqrt(w[i]);
        double sv = sw * x[i * d + c] - sw * mean[c];
        contrib = sv * sv;
      }
      if (w[i] == 0.0) contrib = 0.0;
      s += contrib;
    }
    var[c] = s;
  }
  if (cov) {
    std::vector<double> cen((size_t)n * d);
    for (int64_t i = 0; i < n; ++i) {
      double sw = std::sqrt(w[i]);
      for (int c = 0; c < d; ++c) {
        double dev = x[i * d + c] - mean[c];
        cen[i * d + c] = std::isfinite(dev) ? sw * dev : sw * x[i * d + c] - sw * mean[c];
      }
    }
    for (int r = 0; r < d; ++r)
      for (int c = 0; c < d; ++c) {
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += cen[i * d + r] * cen[i * d + c];
        cov[r + (int64_t)c * d] = s;
      }
  }
  return 0;
}

// ---- Cholesky (column-oriented lower factor, validation/reproducers/amis.jl:120-131) ----
int or_cholesky(const double* A, int d, double* L) {
  for (int64_t i = 0; i < (int64_t)d * d; ++i) L[i] = 0.0;
  for (int c = 0; c < d; ++c) {
    for (int r = c; r < d; ++r) {
      double v = A[r + (int64_t)c * d];
      for (int k = 0; k < c; ++k) v -= L[r + (int64_t)k * d] * L[c + (int64_t)k * d];
      if (r == c) {
        if (!(v > 0.0) || !std::isfinite(v)) return c + 1;  // PosDefException(info)
        L[r + (int64_t)c * d] = std::sqrt(v);
      } else {
        L[r + (int64_t)c * d] = v / L[c + (int64_t)c * d];
      }
    }
  }
  return 0;
}

// ---- resampling ----------------------------------------------------------------------
// Sequential reference CDF: src/resampling.jl:199-225.  Returns 1 on AllZeroWeights.
int or_cdf_sequential(const double* lw, int64_t n, double* cdf) {
  double mx = -kInf;
  for (int64_t i = 0; i < n; ++i) mx = std::max(mx, lw[i]);
  if (mx == -kInf) return 1;
  double total = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    cdf[i] = std::exp(lw[i] - mx);
    total += cdf[i];
  }
  if (!(std::isfinite(total) && total > 0.0)) return 1;
  for (int64_t i = 1; i < n; ++i) cdf[i] = cdf[i - 1] + cdf[i];
  double last = cdf[n - 1];
  for (int64_t i = 0; i < n; ++i) cdf[i] /= last;
  return 0;
}

// Blocked CDF: the engine's fixed summation order (the wavefront-scan CDF kernels of
// the HIP library).  Tiles of threads*per elements; thread j of a tile sums its `per`
// elements in order; the block scan below gives each thread its exclusive prefix and
// the tile total; tile carries come from the same block scan over per-thread runs of
// C = ceil(ntiles / threads) consecutive tile totals; element value =
// (carry + (thread_prefix + running_sum)) / last, where `last` is that value for the
// final element, so the CDF ends at exactly 1 as the reference's does (cdf ./= cdf[end],
// src/resampling.jl:199-225); AllZeroWeights is judged on the grand total of the tile
// totals.  Only the association of the sums is the engine's.

// Exclusive prefix of `tau` and its total in the device's block-scan association:
// per 64-lane wave a Hillis-Steele inclusive scan (step `off`: x_l <- x_{l-off} + x_l
// for l >= off, off = 1, 2, ..., 32), shifted by one lane (lane 0: 0); then the wave
// totals in wave order, the running value added to each lane's prefix.
static void block_excl_scan(const std::vector<double>& tau, std::vector<double>& pre, double& total) {
  const int threads = (int)tau.size();
  const int waves = threads / 64;
  pre.assign(threads, 0.0);
  std::vector<double> wt(waves);
  for (int w = 0; w < waves; ++w) {
    double x[64], y[64];
    for (int l = 0; l < 64; ++l) x[l] = tau[64 * w + l];
    for (int off = 1; off < 64; off <<= 1) {
      for (int l = 0; l < 64; ++l) y[l] = (l >= off) ? x[l - off] + x[l] : x[l];
      for (int l = 0; l < 64; ++l) x[l] = y[l];
    }
    for (int l = 0; l < 64; ++l) pre[64 * w + l] = l ? x[l - 1] : 0.0;
    wt[w] = x[63];
  }
  double run = 0.0;
  for (int w = 0; w < waves; ++w) {
    for (int l = 0; l < 64; ++l) pre[64 * w + l] = run + pre[64 * w + l];
    run = run + wt[w];
  }
  total = run;
}

int or_cdf_blocked(const double* lw, int64_t n, double* cdf, int threads, int per) {
  if (threads % 64 != 0 || threads < 64) return 2;
  double mx = -kInf;
  for (int64_t i = 0; i < n; ++i) mx = std::max(mx, lw[i]);
  if (mx == -kInf) return 1;
  const int64_t tile = (int64_t)threads * per;
  const int64_t ntiles = (n + tile - 1) / tile;
  std::vector<double> tile_total(ntiles), tau(threads), pre;
  std::vector<double> thread_pre((size_t)ntiles * threads);
  for (int64_t t = 0; t < ntiles; ++t) {
    for (int j = 0; j < threads; ++j) {
      double s = 0.0;
      for (int e = 0; e < per; ++e) {
        const int64_t i = t * tile + (int64_t)j * per + e;
        const double w = i < n ? det_exp(lw[i] - mx) : 0.0;
        s = s + w;
        if (i < n) cdf[i] = s;  // the thread's running sum for now
      }
      tau[j] = s;
    }
    block_excl_scan(tau, pre, tile_total[t]);
    for (int j = 0; j < threads; ++j) thread_pre[(size_t)t * threads + j] = pre[j];
  }
  // tile carries: per-thread runs of C tile totals, then the block scan
  const int64_t C = (ntiles + threads - 1) / threads;
  std::vector<double> carry(ntiles);
  for (int j = 0; j < threads; ++j) {
    double s = 0.0;
    for (int64_t b = (int64_t)j * C; b < std::min<int64_t>((int64_t)(j + 1) * C, ntiles); ++b) s = s + tile_total[b];
    tau[j] = s;
  }
  double total;
  block_excl_scan(tau, pre, total);
  for (int j = 0; j < threads; ++j) {
    double run = 0.0;
    for (int64_t b = (int64_t)j * C; b < std::min<int64_t>((int64_t)(j + 1) * C, ntiles); ++b) {
      carry[b] = pre[j] + run;
      run = run + tile_total[b];
    }
  }
  if (!(std::isfinite(total) && total > 0.0)) return 1;
  for (int64_t t = 0; t < ntiles; ++t)
    for (int j = 0; j < threads; ++j)
      for (int e = 0; e < per; ++e) {
        const int64_t i = t * tile + (int64_t)j * per + e;
        if (i < n) cdf[i] = carry[t] + (thread_pre[(size_t)t * threads + j] + cdf[i]);
      }
  const double last = cdf[n - 1];
  for (int64_t i = 0; i < n; ++i) cdf[i] = cdf[i] / last;
  return 0;
}

// LocalResampling ancestor of each proposal from its own group [group[k], group[k+1])
// of the round's log-weights, 1-based round indices; returns 1 when a group has no
// positive finite weight (AllZeroWeightsError).  lanes == 0: the CPU kernel's order
// (src/dm_pmc_execution.jl:80-125: the group maximum, the sum of exp(lw - max), the
// first running sum above u * total, else the group's last sample).  lanes > 0: the
// cooperative device order (:173-246) with `lanes` chunks, restated as the engine
// runs it, exponentials by det_exp.
int or_local_select(const double* lw, const int64_t* group, int K, const double* u, int64_t* anc, int lanes) {
  for (int k = 0; k < K; ++k) {
    const int64_t g0 = group[k], g1 = group[k + 1], count = g1 - g0;
    double m = -kInf;
    for (int64_t i = g0; i < g1; ++i) m = std::max(m, lw[i]);
    anc[k] = 0;
    if (lanes == 0) {
      if (m == -kInf) return 1;
      double total = 0.0;
      for (int64_t i = g0; i < g1; ++i) total += std::exp(lw[i] - m);
      if (!(std::isfinite(total) && total > 0.0)) return 1;
      const double threshold = u[k] * total;
      double cumulative = 0.0;
      anc[k] = g1;
      for (int64_t i = g0; i < g1; ++i) {
        cumulative += std::exp(lw[i] - m);
        if (threshold < cumulative) {
          anc[k] = i + 1;
          break;
        }
      }
      continue;
    }
    const int64_t cs = (count + lanes - 1) / lanes;
    std::vector<double> chunk(lanes, 0.0);
    for (int l = 0; l < lanes; ++l) {
      const int64_t c0 = g0 + (int64_t)l * cs, c1 = std::min(c0 + cs, g1);
      double tot = 0.0;
      if (m != -kInf)
        for (int64_t i = c0; i < c1; ++i) tot = tot + det_exp(lw[i] - m);
      chunk[l] = tot;
    }
    double total = 0.0;
    for (int l = 0; l < lanes; ++l) total = total + chunk[l];
    if (!(std::isfinite(total) && total > 0.0)) return 1;
    const double draw = u[k] * total;
    int chosen = (int)std::min<int64_t>(count, lanes) - 1;
    double rem = chunk[chosen];
    double prefix = 0.0;
    for (int l = 0; l < lanes; ++l) {
      const double next = prefix + chunk[l];
      if (draw < next) {
        chosen = l;
        rem = draw - prefix;
        break;
      }
      prefix = next;
    }
    const int64_t c0 = g0 + (int64_t)chosen * cs, c1 = std::min(c0 + cs, g1);
    double cum = 0.0;
    int64_t a = c1 - 1;
    for (int64_t i = c0; i < c1; ++i) {
      cum = cum + det_exp(lw[i] - m);
      if (rem < cum) {
        a = i;
        break;
      }
    }
    anc[k] = a + 1;
  }
  return 0;
}

// Ancestor selection: src/resampling.jl:92-118 (binary search + zero-mass plateau fix-up).
// Returns 1-based ancestors.
void or_select_ancestors(const double* cdf, const double* lw, int64_t n, const double* u,
                         int64_t count, int64_t* anc) {
  for (int64_t k = 0; k < count; ++k) {
    int64_t first = 1, last = n;
    const double uniform = u[k];
    while (first < last) {
      int64_t middle = first + ((last - first) >> 1);
      if (uniform < cdf[middle - 1])
        last = middle;
      else
        first = middle + 1;
    }
    if (lw[first - 1] == -kInf) {
      while (first > 1 && lw[first - 1] == -kInf) --first;
      while (first < n && lw[first - 1] == -kInf) ++first;
    }
    anc[k] = first;
  }
}

// Systematic points (new method; the reference has only multinomial):
// p_j = (j + u) / count for j = 0..count-1 with one uniform u in [0, 1), then
// the same search and plateau fix-up as multinomial selection.
void or_systematic_points(double u, int64_t count, double* points) {
  for (int64_t j = 0; j < count; ++j) points[j] = ((double)j + u) / (double)count;
}

// ---- AMIS / NPMC: src/adaptive_gaussian_execution.jl:897-1112 ----------------------------
// method 0 = AMIS (src/amis_execution.jl:261-357), 1 = NPMC (src/npmc_execution.jl:11-84).
// scalar=1 selects the scalar history (d must be 1): scale0 = [sigma].
// Otherwise the factor history: scale0 = d*d column-major lower factor.
// normals: per round d*maxN (round t uses the first d*N_t entries).
// radial_normals / radial_uniforms: per round rs*maxN / us*maxN (Student-t only).
// Returns 0 on success; otherwise fail_info = {round, phase} with phase
// 1 sampling, 2 target, 3 denominator, 4 weight, 5 moment, 6 factorization.
int or_adaptive(int method, int family, double dof, int scalar, int d, const double* mu0,
                const double* scale0, double lognorm0, int T, const int64_t* schedule,
                const double* normals, const double* radial_normals, const double* radial_uniforms,
                int tkind, const double* tp, int nthreads, double* samples, double* logtargets,
                double* lognums, double* logweights, int64_t* round_ids, double* round_ess,
                double* round_logz, double* adaptation_ess, double* hist_means,
                double* hist_factors, double* hist_lognorm, uint64_t* failure, int* fail_info,
                int replay, double* fit_means, double* fit_factors, double* fit_lognorm) {
  // replay = 1: hist_* are inputs holding every slot 0..T (for example a device
  // run's fitted history); each round's fit is written to fit_* instead, so a
  // long adaptive trajectory is checked one round at a time.
  const double logconstant = family == 1 ? or_student_t_logconstant(dof, d) : 0.0;
  const double multiplier = family == 1 ? dof / (dof - 2.0) : 1.0;  // src/proposals.jl:73-75
  int64_t maxN = 0, total = 0;
  for (int t = 0; t < T; ++t) {
    maxN = std::max<int64_t>(maxN, schedule[t]);
    total += schedule[t];
  }
  std::vector<int64_t> offsets(T + 1);
  offsets[0] = 0;
  for (int t = 0; t < T; ++t) offsets[t + 1] = offsets[t] + schedule[t];
  std::vector<double> logcounts(T);
  for (int t = 0; t < T; ++t) logcounts[t] = std::log((double)schedule[t]);
  const int64_t dd = (int64_t)d * d;
  // history slot 0 = initial proposal
  if (!replay) {
    for (int c = 0; c < d; ++c) hist_means[c] = mu0[c];
    if (scalar) {
      hist_factors[0] = scale0[0];
    } else {
      for (int64_t i = 0; i < dd; ++i) hist_factors[i] = scale0[i];
    }
    hist_lognorm[0] = lognorm0;
  }
  double* out_means = replay ? fit_means : hist_means;
  double* out_factors = replay ? fit_factors : hist_factors;
  double* out_lognorm = replay ? fit_lognorm : hist_lognorm;
  auto slot_prop = [&](int s) {
    return Prop{family, dof, scalar ? (int)kSpherical : (int)kFactor, d, hist_means + (int64_t)s * d,
                hist_factors + (int64_t)s * (scalar ? 1 : dd), hist_lognorm[s]};
  };
  const int rs = radial_normal_stride(family, dof);
  const int us = radial_uniform_stride(family, dof);
  for (int64_t i = 0; i < total; ++i) {
    round_ids[i] = 0;
    lognums[i] = -kInf;
  }
  std::vector<double> zbuf((size_t)d * maxN);
  std::vector<FailureRecord> records(std::max(nthreads, 1));
  for (int i = 0; i < 3; ++i) failure[i] = 0;
  fail_info[0] = 0;
  fail_info[1] = 0;
  for (int t = 0; t < T; ++t) {
    const int64_t N = schedule[t];
    const int64_t first = offsets[t];
    const int64_t last = offsets[t + 1];  // exclusive
    const double logtotal = method == 0 ? std::log((double)last) : 0.0;
    for (auto& r : records) r = FailureRecord();
    const Prop cur = slot_prop(t);
    // Student-t normal scaling: _scale_mis_normals_kernel!, src/native_execution.jl:184-198.
    const double* nz = normals + (int64_t)t * d * maxN;
    for (int64_t j = 0; j < (int64_t)d * N; ++j) zbuf[j] = nz[j];
    if (family == 1) {
      const double* rn = radial_normals + (int64_t)t * rs * maxN;
      const double* ru = us ? radial_uniforms + (int64_t)t * us * maxN : nullptr;
      for (int64_t j = 0; j < N; ++j) {
        double m;
        if (!student_t_multiplier(dof, rn + j * rs, ru ? ru + j * us : nullptr, &m)) {
          records[0].record(j + 1, 0, kDrawExhausted);
          m = kNaN;
        }
        for (int c = 0; c < d; ++c) zbuf[j * d + c] *= m;
      }
    }
    if (method == 0 && first > 0) {
      // append pass: _append_logmixture_kernel!, src/amis_execution.jl:37-65.
      parallel_for(first, nthreads, [&](int64_t a, int64_t b, int worker) {
        std::vector<double> work(d);
        for (int64_t i = a; i < b; ++i) {
          double q = packed_logdensity(cur, samples + i * d, work.data());
          uint16_t reason = std::isnan(q) ? kProposalInvalid : 0;
          if (!reason) {
            double value = logaddexp(lognums[i], logcounts[t] + q);
            reason = proposal_reason(value);
            if (!reason) lognums[i] = value;
          }
          if (reason) records[worker].record(i + 1, 0, reason);
        }
      });
    }
    // new samples: _gaussian_round_launch_kernel!, src/mis_execution.jl:722-771
    // with _mis_round_values! (:601-672) and _mis_logdenominator_core (:30-65).
    parallel_for(N, nthreads, [&](int64_t a, int64_t b, int worker) {
      std::vector<double> work(std::max(d, 1));
      FailureRecord& rec = records[worker];
      for (int64_t j = a; j < b; ++j) {
        const int64_t i = first + j;
        logtargets[i] = -kInf;
        lognums[i] = 0.0;
        logweights[i] = -kInf;
        round_ids[i] = 0;
        double* x = samples + i * d;
        bool valid = true;
        for (int c = 0; c < d; ++c) {
          x[c] = packed_coordinate(cur, zbuf.data() + j * d, c);
          valid &= std::isfinite(x[c]);
        }
        if (!valid) {
          rec.record(j + 1, 0, kGeneratedNonfinite);
          continue;
        }
        double tv = target_eval(tkind, tp, d, x, work.data());
        uint16_t tr = target_reason(tv);
        if (tr) {
          rec.record(j + 1, 0, tr);
          continue;
        }
        double value = -kInf, generating = 0.0;
        const int k0 = method == 0 ? 0 : t;
        for (int k = k0; k <= t; ++k) {
          double q = packed_logdensity(slot_prop(k), x, work.data());
          if (k == t) generating = q;
          const double logmass = method == 0 ? logcounts[k] : 0.0;
          value = logaddexp(value, logmass + q);
        }
        if (std::isnan(generating) || generating == -kInf || std::isnan(value) || value == -kInf) {
          rec.record(j + 1, 0, kProposalInvalid);
          continue;
        }
        double lw;
        uint16_t wr = subtract_logweight(tv, value - logtotal, &lw);  // src/mis_execution.jl:94-95
        if (wr) {
          rec.record(j + 1, 0, wr);
          continue;
        }
        logtargets[i] = tv;
        lognums[i] = value;
        logweights[i] = lw;
        round_ids[i] = t + 1;
      }
    });
    if (method == 0) {
      // weight pass: _form_amis_logweights_kernel!, src/amis_execution.jl:107-132.
      parallel_for(last, nthreads, [&](int64_t a, int64_t b, int worker) {
        for (int64_t i = a; i < b; ++i) {
          if (round_ids[i] == 0) continue;
          double lw;
          uint16_t wr = subtract_logweight(logtargets[i], lognums[i] - logtotal, &lw);
          if (wr)
            records[worker].record(i + 1, 0, wr);
          else
            logweights[i] = lw;
        }
      });
    }
    FailureRecord all;
    for (auto& r : records) all.merge(r);
    if (all.w[0] != 0) {
      std::memcpy(failure, all.w, sizeof(all.w));
      uint16_t bits = (uint16_t)(all.w[1] & 0xffff);
      int phase = 3;
      if (all.w[2] != 0 || (bits & kDrawReasons)) phase = 1;
      else if (bits & (kTargetNaN | kTargetPosInf)) phase = 2;
      else if (bits & kLogweightInvalid) phase = 4;
      fail_info[0] = t + 1;
      fail_info[1] = phase;
      return 1;
    }
    // adaptation workspace (src/amis_execution.jl:8; src/npmc_execution.jl:21-33)
    const int64_t fit_first = method == 0 ? 0 : first;
    const int64_t fit_n = last - fit_first;
    std::vector<double> fitlw(logweights + fit_first, logweights + last);
    if (method == 1) {
      std::vector<double> sorted(fitlw);
      int64_t isq = (int64_t)std::floor(std::sqrt((double)fit_n));
      while (isq * isq > fit_n) --isq;
      while ((isq + 1) * (isq + 1) <= fit_n) ++isq;
      int64_t threshold_index = fit_n - isq;  // 0-based of (n - isqrt(n) + 1)
      std::nth_element(sorted.begin(), sorted.begin() + threshold_index, sorted.end());
      double threshold = sorted[threshold_index];
      for (auto& v : fitlw) v = std::min(v, threshold);
    }
    // fit: _normalize_moment_weights!, src/adaptive_gaussian_execution.jl:186-205, then
    // _fit_moment_proposal! (:303-425).
    double mx = -kInf;
    for (double v : fitlw) mx = std::max(mx, v);
    if (mx == -kInf) {
      fail_info[0] = t + 1;
      fail_info[1] = 5;
      return 2;
    }
    std::vector<double> w(fit_n);
    double tot = 0.0;
    for (int64_t i = 0; i < fit_n; ++i) {
      w[i] = std::exp(fitlw[i] - mx);
      tot += w[i];
    }
    if (!(std::isfinite(tot) && tot > 0.0)) {
      fail_info[0] = t + 1;
      fail_info[1] = 5;
      return 2;
    }
    double inv_total = 1.0 / tot;
    for (auto& v : w) v *= inv_total;
    double* mean_out = out_means + (int64_t)(t + 1) * d;
    double* fac_out = out_factors + (int64_t)(t + 1) * (scalar ? 1 : dd);
    const double* prev_fac = hist_factors + (int64_t)t * (scalar ? 1 : dd);
    if (scalar) {
      double mean = 0.0;
      for (int64_t i = 0; i < fit_n; ++i) mean += w[i] * samples[fit_first + i];
      double variance = 0.0;
      for (int64_t i = 0; i < fit_n; ++i) {
        double cen = (samples[fit_first + i] - mean) * std::sqrt(w[i]);
        variance += cen * cen;
      }
      // _scalar_scale_aware_ridge: src/adaptive_gaussian_execution.jl:275-281.
      const double reg = std::sqrt(DBL_EPSILON);
      double ridge = reg * (prev_fac[0] * prev_fac[0] * multiplier) / 1.0;
      if (!std::isfinite(ridge)) {
        double s = prev_fac[0] * (std::sqrt(reg) * std::sqrt(multiplier));
        ridge = s * s;
      }
      variance += ridge;
      if (!(std::isfinite(variance) && variance > 0.0)) {
        fail_info[0] = t + 1;
        fail_info[1] = 6;
        return 3;
      }
      double sc = std::sqrt(variance / multiplier);
      double ln = family == 1 ? logconstant - std::log(sc)
                              : -0.5 * 1.0 * std::log(2.0 * M_PI) - std::log(sc);
      if (!(sc > 0.0 && std::isfinite(ln))) {
        fail_info[0] = t + 1;
        fail_info[1] = 6;
        return 3;
      }
      mean_out[0] = mean;
      fac_out[0] = sc;
      out_lognorm[t + 1] = ln;
    } else {
      std::vector<double> mean(d, 0.0), cov(dd, 0.0), cen((size_t)d * fit_n);
      for (int64_t i = 0; i < fit_n; ++i)
        for (int c = 0; c < d; ++c) mean[c] += w[i] * samples[(fit_first + i) * d + c];
      for (int64_t i = 0; i < fit_n; ++i) {
        double sw = std::sqrt(w[i]);
        for (int c = 0; c < d; ++c) cen[i * d + c] = (samples[(fit_first + i) * d + c] - mean[c]) * sw;
      }
      for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) {
          double s = 0.0;
          for (int64_t i = 0; i < fit_n; ++i) s += cen[i * d + r] * cen[i * d + c];
          cov[r + (int64_t)c * d] = s;
        }
      // _factor_scale_aware_ridge: src/adaptive_gaussian_execution.jl:283-301.
      const double reg = std::sqrt(DBL_EPSILON);
      double trace = 0.0;
      for (int c = 0; c < d; ++c)
        for (int r = c; r < d; ++r) trace += prev_fac[r + (int64_t)c * d] * prev_fac[r + (int64_t)c * d];
      double ridge = reg * (trace * multiplier) / (double)d;
      if (!std::isfinite(ridge)) {
        double s = std::sqrt(reg) * std::sqrt(multiplier / (double)d);
        ridge = 0.0;
        for (int c = 0; c < d; ++c)
          for (int r = c; r < d; ++r) {
            double v = prev_fac[r + (int64_t)c * d] * s;
            ridge += v * v;
          }
      }
      for (int c = 0; c < d; ++c) cov[c + (int64_t)c * d] += ridge;
      std::vector<double> L(dd);
      if (or_cholesky(cov.data(), d, L.data()) != 0) {
        fail_info[0] = t + 1;
        fail_info[1] = 6;
        return 3;
      }
      if (family == 1) {
        double f = std::sqrt(1.0 / multiplier);  // _scale_covariance_factor!: src/proposals.jl:77-80
        for (auto& v : L) v *= f;
      }
      double logabsdet = 0.0;
      for (int c = 0; c < d; ++c) logabsdet += std::log(std::fabs(L[c + (int64_t)c * d]));
      double ln = family == 1 ? logconstant - logabsdet
                              : -0.5 * (double)d * std::log(2.0 * M_PI) - logabsdet;
      bool ok = std::isfinite(ln);
      for (int c = 0; c < d; ++c) ok &= std::isfinite(mean[c]);
      for (int c = 0; c < d; ++c)
        for (int r = c; r < d; ++r) {
          double v = L[r + (int64_t)c * d];
          ok &= std::isfinite(v);
          ok &= r != c || v > 0.0;
        }
      if (!ok) {
        fail_info[0] = t + 1;
        fail_info[1] = 6;
        return 3;
      }
      for (int c = 0; c < d; ++c) mean_out[c] = mean[c];
      for (int64_t i = 0; i < dd; ++i) fac_out[i] = L[i];
      out_lognorm[t + 1] = ln;
    }
    // round summary (AMIS: all retained; NPMC: current round raw weights).
    const int64_t s_first = method == 0 ? 0 : first;
    or_logweight_summary(logweights + s_first, last - s_first, round_ess + t, round_logz + t);
    if (method == 1 && adaptation_ess) {
      double s2 = 0.0;
      for (double v : w) s2 += v * v;
      adaptation_ess[t] = 1.0 / s2;
    }
  }
  return 0;
}

// ---- DM-PMC: src/population_execution.jl:407-567 -------------------------------------
// Bank of K active slots in allocation order (ascending mass, src/proposal_banks.jl:301-326).
// kind: 0 spherical (scales K), 1 diagonal (scales d*K), 2 factor (factors d*d*K).
// counts: K*T allocation plan (column per round), logcoef: K*T.
// uniforms: per round K uniforms (systematic uses the first of each round).
// resampling: 0 multinomial (sequential CDF as the reference), 1 systematic,
//             2 multinomial with the engine's blocked CDF, 3 systematic with the blocked CDF.
int or_dmpmc(int family, double dof, int kind, int d, int K, const double* locs0,
             const double* scales, const double* lognorms, const int64_t* proposal_ids, int T,
             const int64_t* schedule, const int64_t* counts, const double* logcoef, int resampling,
             int cdf_threads, int cdf_per, const double* normals, const double* radial_normals,
             const double* radial_uniforms, const double* uniforms, int tkind, const double* tp,
             int nthreads, double* samples, double* logweights, int64_t* round_ids,
             int64_t* out_proposal_ids, int64_t* ancestors_out, double* round_ess,
             double* round_logz, double* final_locs, uint64_t* failure, int* fail_info) {
  int64_t maxN = 0;
  for (int t = 0; t < T; ++t) maxN = std::max<int64_t>(maxN, schedule[t]);
  std::vector<double> locs(locs0, locs0 + (int64_t)d * K);
  const int64_t dd = (int64_t)d * d;
  const int rs = radial_normal_stride(family, dof);
  const int us = radial_uniform_stride(family, dof);
  std::vector<double> zbuf((size_t)d * maxN), cdf(maxN);
  std::vector<FailureRecord> records(std::max(nthreads, 1));
  int64_t offset = 0;
  fail_info[0] = fail_info[1] = 0;
  for (int i = 0; i < 3; ++i) failure[i] = 0;
  for (int t = 0; t < T; ++t) {
    const int64_t N = schedule[t];
    std::vector<int> assign(N);
    {
      int64_t a = 0;
      for (int k = 0; k < K; ++k)
        for (int64_t c = 0; c < counts[k + (int64_t)t * K]; ++c) assign[a++] = k;
    }
    auto slot = [&](int k) {
      const double* sc = kind == kSpherical ? scales + k : kind == kDiagonal ? scales + (int64_t)k * d
                                                                               : scales + (int64_t)k * dd;
      return Prop{family, dof, kind, d, locs.data() + (int64_t)k * d, sc, lognorms[k]};
    };
    for (auto& r : records) r = FailureRecord();
    const double* nz = normals + (int64_t)t * d * maxN;
    for (int64_t j = 0; j < (int64_t)d * N; ++j) zbuf[j] = nz[j];
    if (family == 1) {
      const double* rn = radial_normals + (int64_t)t * rs * maxN;
      const double* ru = us ? radial_uniforms + (int64_t)t * us * maxN : nullptr;
      for (int64_t j = 0; j < N; ++j) {
        double m;
        if (!student_t_multiplier(dof, rn + j * rs, ru ? ru + j * us : nullptr, &m)) {
          records[0].record(j + 1, 0, kDrawExhausted);
          m = kNaN;
        }
        for (int c = 0; c < d; ++c) zbuf[j * d + c] *= m;
      }
    }
    // _mis_round_launch_kernel!: src/mis_execution.jl:674-720.
    parallel_for(N, nthreads, [&](int64_t a, int64_t b, int worker) {
      std::vector<double> work(std::max(d, 1));
      FailureRecord& rec = records[worker];
      for (int64_t j = a; j < b; ++j) {
        const int64_t i = offset + j;
        logweights[i] = -kInf;
        out_proposal_ids[i] = 0;
        round_ids[i] = t + 1;
        const Prop gen = slot(assign[j]);
        double* x = samples + i * d;
        bool valid = true;
        for (int c = 0; c < d; ++c) {
          x[c] = packed_coordinate(gen, zbuf.data() + j * d, c);
          valid &= std::isfinite(x[c]);
        }
        if (!valid) {
          rec.record(j + 1, 0, kGeneratedNonfinite);
          continue;
        }
        double tv = target_eval(tkind, tp, d, x, work.data());
        uint16_t tr = target_reason(tv);
        if (tr) {
          rec.record(j + 1, 0, tr);
          continue;
        }
        double value = -kInf, generating = 0.0;
        for (int k = 0; k < K; ++k) {
          double q = packed_logdensity(slot(k), x, work.data());
          if (k == assign[j]) generating = q;
          value = logaddexp(value, logcoef[k + (int64_t)t * K] + q);
        }
        if (std::isnan(generating) || generating == -kInf || std::isnan(value) || value == -kInf) {
          rec.record(j + 1, 0, kProposalInvalid);
          continue;
        }
        double lw;
        uint16_t wr = subtract_logweight(tv, value, &lw);
        if (wr) {
          rec.record(j + 1, 0, wr);
          continue;
        }
        logweights[i] = lw;
        out_proposal_ids[i] = proposal_ids[assign[j]];
      }
    });
    FailureRecord all;
    for (auto& r : records) all.merge(r);
    if (all.w[0] != 0) {
      std::memcpy(failure, all.w, sizeof(all.w));
      fail_info[0] = t + 1;
      fail_info[1] = 1;
      return 1;
    }
    const double* lw = logweights + offset;
    if (resampling >= 4) {
      // local resampling: src/dm_pmc_execution.jl:297-374, each proposal's ancestor from
      // its own group (codes 4: the engine's cooperative order, 5: the CPU kernel's)
      std::vector<int64_t> group(K + 1, 0), anc(K);
      for (int k = 0; k < K; ++k) group[k + 1] = group[k] + counts[k + (int64_t)t * K];
      if (or_local_select(lw, group.data(), K, uniforms + (int64_t)t * K, anc.data(), resampling == 4 ? 256 : 0)) {
        fail_info[0] = t + 1;
        fail_info[1] = 7;
        return 2;
      }
      for (int k = 0; k < K; ++k) {
        if (ancestors_out) ancestors_out[(int64_t)t * K + k] = anc[k];
        for (int c = 0; c < d; ++c) locs[(int64_t)k * d + c] = samples[(offset + anc[k] - 1) * d + c];
      }
      or_logweight_summary(lw, N, round_ess + t, round_logz + t);
      offset += N;
      continue;
    }
    // global resampling: src/dm_pmc_execution.jl:267-295.
    int bad = (resampling >= 2) ? or_cdf_blocked(lw, N, cdf.data(), cdf_threads, cdf_per)
                                : or_cdf_sequential(lw, N, cdf.data());
    if (bad) {
      fail_info[0] = t + 1;
      fail_info[1] = 7;
      return 2;
    }
    std::vector<double> pts(K);
    const double* u = uniforms + (int64_t)t * K;
    if (resampling == 1 || resampling == 3)
      or_systematic_points(u[0], K, pts.data());
    else
      for (int k = 0; k < K; ++k) pts[k] = u[k];
    std::vector<int64_t> anc(K);
    or_select_ancestors(cdf.data(), lw, N, pts.data(), K, anc.data());
    for (int k = 0; k < K; ++k) {
      if (ancestors_out) ancestors_out[(int64_t)t * K + k] = anc[k];
      for (int c = 0; c < d; ++c) locs[(int64_t)k * d + c] = samples[(offset + anc[k] - 1) * d + c];
    }
    or_logweight_summary(lw, N, round_ess + t, round_logz + t);
    offset += N;
  }
  std::memcpy(final_locs, locs.data(), sizeof(double) * d * K);
  return 0;
}

// Scale-aware ridges exported for the bit-exact characterization test
// (src/adaptive_gaussian_execution.jl:275-301; test/amis_kernel.jl:282-329).
double or_ridge_scalar(double scale, int family, double dof) {
  const double reg = std::sqrt(DBL_EPSILON);
  const double multiplier = family == 1 ? dof / (dof - 2.0) : 1.0;
  double ridge = reg * ((scale * scale) * multiplier) / 1.0;
  if (std::isfinite(ridge)) return ridge;
  double s = scale * (std::sqrt(reg) * std::sqrt(multiplier));
  return s * s;
}

double or_ridge_factor(const double* factor, int d, int family, double dof) {
  const double reg = std::sqrt(DBL_EPSILON);
  const double multiplier = family == 1 ? dof / (dof - 2.0) : 1.0;
  double trace = 0.0;
  for (int c = 0; c < d; ++c)
    for (int r = c; r < d; ++r) trace += factor[r + (int64_t)c * d] * factor[r + (int64_t)c * d];
  double ridge = reg * (trace * multiplier) / (double)d;
  if (std::isfinite(ridge)) return ridge;
  double s = std::sqrt(reg) * std::sqrt(multiplier / (double)d);
  ridge = 0.0;
  for (int c = 0; c < d; ++c)
    for (int r = c; r < d; ++r) {
      double v = factor[r + (int64_t)c * d] * s;
      ridge += v * v;
    }
  return ridge;
}

}  // extern "C"

