// Per-dimension kernel launchers.  Each dimension instantiation D (d = 1..16
// exactly, then the buckets 24, 32, 48, 64) is compiled in its own
// translation unit (dim.hip built with -DIS_DIM=D, in parallel) and linked
// into the library; the engine dispatches on the job's D.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdint.h>

#include "kernels.h"

namespace issampler {

#define IS_DIMS(X) \
  X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(24) X(32) X(48) X(64)

inline int dim_bucket(int d) {
  if (d <= kExactDims) return d;
  if (d <= 24) return 24;
  if (d <= 32) return 32;
  if (d <= 48) return 48;
  return 64;
}

// Moment tiles per side for the augmented vector (x - c, 1) of length d + 1.
inline int sweep_nt(int d, bool moments) { return moments ? (d + 1 + 15) / 16 : 0; }
inline int sweep_waves(int nt) { return nt <= 1 ? 4 : nt == 2 ? 3 : nt == 3 ? 2 : 1; }
inline size_t sweep_lds(int nt, int nw) {
  const size_t per_wave = nt > 0 ? (size_t)64 * sweep_row_stride(nt) * sizeof(double) : 0;
  const size_t staging = nw * per_wave;
  const size_t P = (size_t)nt * (nt + 1) / 2;
  const size_t reduce = (16 + (size_t)nw * P * 256 + P * 256) * sizeof(double);
  return staging > reduce ? staging : reduce;
}

constexpr size_t kNewSlotBudget = 48 * 1024;  // LDS bytes for staged slots per new-sample block

// Fused round launches (k_round, k_round_group; D <= 15): the mixture slots staged per
// chunk and the dynamic LDS bytes, for nt moment tiles and ncomp mixture components.
constexpr size_t kRoundLds = 40 * 1024;
inline int round_slot_chunk(int D, int nt, int stride, int tsize, bool gaussian_target, int ncomp, size_t* lds) {
  const size_t slot_bytes = (size_t)stride * tsize;
  const size_t tslot_bytes = gaussian_target ? slot_bytes : 0;
  const size_t fixed = new_lds_fixed(D, 4, tsize) + tslot_bytes;
  const size_t cap = std::max(sweep_lds(nt, 4), kRoundLds);
  const size_t room = cap > fixed ? cap - fixed : 0;
  const size_t comp_bytes = slot_bytes + tsize;  // a staged slot and its log mixture mass
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, ncomp), room / comp_bytes));
  *lds = std::max(sweep_lds(nt, 4), fixed + (size_t)chunk * comp_bytes);
  return chunk;
}

template <int D>
struct DimKernels {
  // k_new in mode `mode` (kDrawFused / kDrawOnly / kPreTarget), 64 samples per block of W waves
  static void new_samples(int mode, const RoundArgs<float>& a, int ncomp, int blocks, int W, hipStream_t s);
  static void new_samples(int mode, const RoundArgs<double>& a, int ncomp, int blocks, int W, hipStream_t s);
  // k_sweep with `nt` moment tiles per side (0: summary only)
  static void sweep(int nt, const SweepArgs<float>& a, int blocks, int nw, size_t lds, hipStream_t s);
  static void sweep(int nt, const SweepArgs<double>& a, int blocks, int nw, size_t lds, hipStream_t s);
  // k_round: one round's new samples beside the sweep of the retained ones, with
  // nt = 1 moment tiles (AMIS) or nt = 0 (summary only); false when this
  // instantiation has no fused round (d > 15)
  static bool round(int nt, const RoundArgs<float>& a, const SweepArgs<float>& w, int ncomp, int nb_new,
                    int nrec_old, hipStream_t s);
  static bool round(int nt, const RoundArgs<double>& a, const SweepArgs<double>& w, int ncomp, int nb_new,
                    int nrec_old, hipStream_t s);
  // k_round_group: the same round for `jobs` same-shape jobs (device argument tables, the
  // slot chunk already set by round_slot_chunk); false without a fused round
  static bool round_group(int nt, bool student, const RoundArgs<float>* ga, const SweepArgs<float>* gw, int jobs,
                          int nb_new, int nrec_old, size_t lds, hipStream_t s);
  static bool round_group(int nt, bool student, const RoundArgs<double>* ga, const SweepArgs<double>* gw, int jobs,
                          int nb_new, int nrec_old, size_t lds, hipStream_t s);
  // k_target_gemm over `splits` row slices
  static void target(const TargetArgs<float>& t, int64_t splits, hipStream_t s);
  static void target(const TargetArgs<double>& t, int64_t splits, hipStream_t s);
};

// The new-sample and sweep kernels with the extended-range densities
// (slot_logdensity<..., RANGE = true>), 64-coordinate instantiation serving any d; a
// run whose densities left the floating-point range is replayed on these.
constexpr int kRangeDims = 1000;  // pseudo dimension instantiation selecting them
struct RangeKernels {
  static void new_samples(int mode, const RoundArgs<float>& a, int ncomp, int blocks, int W, hipStream_t s);
  static void new_samples(int mode, const RoundArgs<double>& a, int ncomp, int blocks, int W, hipStream_t s);
  static void sweep(int nt, const SweepArgs<float>& a, int blocks, int nw, size_t lds, hipStream_t s);
  static void sweep(int nt, const SweepArgs<double>& a, int blocks, int nw, size_t lds, hipStream_t s);
};

}  // namespace issampler
