// Host-side helpers shared by the engine translation units: error reporting,
// device buffers, the context, and proposal construction / slot packing.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/issampler.h"
#include "device.h"

struct is_ctx {
  int device = 0;
  int cus = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  std::mutex mu;
};

namespace issampler {
namespace host {

inline thread_local std::string g_last_error;

inline int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
inline int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess) return fail(IS_EDEVICE, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

// ---------------------------------------------------------------- host math
// Student-t log normalizing constant: src/proposals.jl:410-437 (gamma
// recurrence for odd d shifted to nu >= 32, six Stirling terms; pairwise
// log1p terms for the remaining gamma-ratio factors).
inline double student_t_logconstant(double dof, int dimension) {
  double value = -(double)dimension * std::log(2.0 * M_PI) / 2.0;
  if (dimension % 2 == 1) {
    double shifted = dof, correction = 0.0;
    while (shifted < 32.0) {
      correction += shifted < 1.0
                        ? (std::log(shifted) + std::log(shifted + 2.0)) / 2.0 - std::log1p(shifted)
                        : std::log1p(-((1.0 / (shifted + 1.0)) * (1.0 / (shifted + 1.0)))) / 2.0;
      shifted += 2.0;
    }
    const double r = 1.0 / shifted, r2 = r * r;
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
    const double x = (double)offset;
    acc += x <= dof ? std::log1p(x / dof) : std::log(x) + std::log1p(dof / x) - std::log(dof);
  }
  return value + acc;
}

inline double gaussian_lognormalizer(int d, double logabsdet) {  // src/proposals.jl:237-240
  return -0.5 * (double)d * std::log(2.0 * M_PI) - logabsdet;
}

// _factor_exponent_bound (src/student_t_range.jl:21-31).
inline int factor_exponent_bound(const std::vector<double>& Lcol, int d) {
  int bound = 0;
  for (int c = 0; c < d; ++c)
    for (int r = c; r < d; ++r) {
      const double v = Lcol[r + (size_t)c * d];
      if (v == 0.0) continue;
      if (!std::isfinite(v)) return 1 << 30;
      int e;
      std::frexp(v, &e);
      bound = std::max(bound, std::max(std::abs(e), std::abs(e - 1)));
    }
  return bound;
}

// A proposal in host form: location and a full lower factor (column-major).
struct HostProp {
  std::vector<double> mu;
  std::vector<double> L;  // d*d column-major lower
  double lognorm = 0.0;
};

inline void pack_slot(const HostProp& p, int d, double logmass, double* slot) {
  const int stride = slot_stride(d);
  std::fill(slot, slot + stride, 0.0);
  for (int r = 0; r < d; ++r) {
    slot[r] = p.mu[r];
    slot[d + r] = slot_inv_entry<double>(p.L[r + (size_t)r * d]);
    for (int c = 0; c <= r; ++c) slot[2 * d + r * (r + 1) / 2 + c] = p.L[r + (size_t)c * d];
  }
  slot[slot_lognorm_off(d)] = p.lognorm;
  slot[slot_logmass_off(d)] = logmass;
}

inline void unpack_slot(const double* slot, int d, HostProp& p) {
  p.mu.assign(slot, slot + d);
  p.L.assign((size_t)d * d, 0.0);
  for (int r = 0; r < d; ++r)
    for (int c = 0; c <= r; ++c) p.L[r + (size_t)c * d] = slot[2 * d + r * (r + 1) / 2 + c];
  p.lognorm = slot[slot_lognorm_off(d)];
}

// Builds one proposal from (kind, location, scale storage); returns an error string or "".
inline std::string make_prop(int family, double dof, int kind, int d, const double* loc, const double* scale,
                      HostProp& out) {
  out.mu.assign(loc, loc + d);
  out.L.assign((size_t)d * d, 0.0);
  double logabsdet = 0.0;
  for (int c = 0; c < d; ++c)
    if (!std::isfinite(loc[c])) return "native parameter must contain only finite values";
  if (kind == IS_SPHERICAL) {
    const double s = scale[0];
    if (!(std::isfinite(s) && s > 0.0)) return "scale must be finite and positive";
    for (int c = 0; c < d; ++c) out.L[c + (size_t)c * d] = s;
    logabsdet = (double)d * std::log(s);
  } else if (kind == IS_DIAGONAL) {
    for (int c = 0; c < d; ++c) {
      const double s = scale[c];
      if (!(std::isfinite(s) && s > 0.0)) return "diagonal scales must be finite and positive";
      out.L[c + (size_t)c * d] = s;
      logabsdet += std::log(s);
    }
  } else if (kind == IS_FACTOR) {
    for (int c = 0; c < d; ++c)
      for (int r = 0; r < d; ++r) {
        const double v = scale[r + (size_t)c * d];
        if (!std::isfinite(v)) return "factor must contain only finite values";
        if (r < c && v != 0.0) return "factor must be lower triangular";
        out.L[r + (size_t)c * d] = v;
      }
    for (int c = 0; c < d; ++c) {
      const double v = out.L[c + (size_t)c * d];
      if (!(v > 0.0)) return "factor diagonal must be positive";
      logabsdet += std::log(v);
    }
  } else {
    return "unknown scale kind";
  }
  if (family == IS_STUDENT_T) {
    if (!(std::isfinite(dof) && dof > 0.0)) return "degrees of freedom must be finite and positive";
    out.lognorm = student_t_logconstant(dof, d) - logabsdet;
  } else {
    out.lognorm = gaussian_lognormalizer(d, logabsdet);
  }
  return "";
}


inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Round-to-nearest-even FP32 -> BF16 (the hardware conversion's rounding), and back.
inline uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
inline double bf16_value(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return (double)f;
}

// ---------------------------------------------------------------- capture lock
// A stream capture on one thread must not overlap a device-wide runtime call on
// another: hipMalloc, hipFree and the blocking copies and fills may synchronise
// every stream of the device, which invalidates a capture in progress.  Every
// capture section and every such call takes this one process-wide lock (it is
// recursive, so a capture section may reach an allocation on its own thread).
inline std::recursive_mutex& capture_lock() {
  static std::recursive_mutex m;
  return m;
}
using CaptureGuard = std::lock_guard<std::recursive_mutex>;

inline hipError_t locked_memcpy(void* dst, const void* src, size_t n, hipMemcpyKind kind) {
  CaptureGuard g(capture_lock());
  return hipMemcpy(dst, src, n, kind);
}
inline hipError_t locked_memset(void* dst, int v, size_t n) {
  CaptureGuard g(capture_lock());
  return hipMemset(dst, v, n);
}

// ---------------------------------------------------------------- device buffer
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) {
      CaptureGuard g(capture_lock());
      (void)hipFree(p);
    }
  }
  int alloc(size_t n) {
    CaptureGuard g(capture_lock());
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = n;
    if (n == 0) return IS_OK;
    if (hipMalloc(&p, n) != hipSuccess) {
      p = nullptr;
      return fail(IS_ENOMEM, "hipMalloc(%zu) failed", n);
    }
    return IS_OK;
  }
  template <class U>
  U* as() const {
    return reinterpret_cast<U*>(p);
  }
};

// A part of a DevBuf (no ownership; same accessors).
struct DevView {
  void* p = nullptr;
  size_t bytes = 0;
  template <class U>
  U* as() const {
    return reinterpret_cast<U*>(p);
  }
};

// Page-locked host memory from hipHostMalloc: copies to and from it are
// asynchronous on the issuing stream, with no staging through pageable memory.
struct HostPinned {
  void* p = nullptr;
  size_t bytes = 0;
  HostPinned() = default;
  HostPinned(const HostPinned&) = delete;
  HostPinned& operator=(const HostPinned&) = delete;
  ~HostPinned() {
    if (p) {
      CaptureGuard g(capture_lock());
      (void)hipHostFree(p);
    }
  }
  int alloc(size_t n) {
    CaptureGuard g(capture_lock());
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = n;
    if (n == 0) return IS_OK;
    if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return fail(IS_ENOMEM, "hipHostMalloc(%zu) failed", n);
    }
    std::memset(p, 0, n);
    return IS_OK;
  }
  template <class U>
  U* as() const {
    return reinterpret_cast<U*>(p);
  }
};


}  // namespace host
}  // namespace issampler
