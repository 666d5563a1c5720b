// Definitions of DimKernels<D> (see launch.h); included only by dim.hip.
#pragma once

#include <algorithm>

#include "launch.h"

namespace issampler {

template <class T, int D, int MODE, bool RG = false>
void launch_new_mode(RoundArgs<T> a, int ncomp, int blocks, int W, hipStream_t s) {
  const bool st = a.fam.student != 0;
  W = std::min(W, new_block_limit(D, (int)sizeof(T), st) / 64);
  const size_t slot_bytes = (size_t)a.stride * sizeof(T);
  const size_t tslot_bytes = a.tkind == kTGaussian ? slot_bytes : 0;
  const size_t avail = kNewSlotBudget > tslot_bytes ? kNewSlotBudget - tslot_bytes : 0;
  // each staged component: its slot and its log mixture mass
  const size_t comp_bytes = slot_bytes + sizeof(T);
  a.slot_chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, ncomp), avail / comp_bytes));
  const size_t lds = new_lds_fixed(D, W, (int)sizeof(T)) + tslot_bytes + (size_t)a.slot_chunk * comp_bytes;
  if (MODE == kPreTarget)
    hipLaunchKernelGGL((k_new<T, D, MODE, false, RG>), dim3(blocks), dim3(64 * W), lds, s, a);
  else if (st)
    hipLaunchKernelGGL((k_new<T, D, MODE, true, RG>), dim3(blocks), dim3(64 * W), lds, s, a);
  else
    hipLaunchKernelGGL((k_new<T, D, MODE, false, RG>), dim3(blocks), dim3(64 * W), lds, s, a);
}

template <class T, int D, bool RG = false>
void launch_new_any(int mode, const RoundArgs<T>& a, int ncomp, int blocks, int W, hipStream_t s) {
  if (mode == kDrawFused)
    launch_new_mode<T, D, kDrawFused, RG>(a, ncomp, blocks, W, s);
  else if (mode == kDrawOnly)
    launch_new_mode<T, D, kDrawOnly, RG>(a, ncomp, blocks, W, s);
  else
    launch_new_mode<T, D, kPreTarget, RG>(a, ncomp, blocks, W, s);
}

// The extended-range sweep: the 64-coordinate instantiation for any d <= 64, so
// every moment-tile count 0..5 is instantiated.
template <class T>
void launch_sweep_range(int nt, const SweepArgs<T>& a, int blocks, int nw, size_t lds, hipStream_t s) {
  switch (nt) {
#define IS_RANGE_SWEEP(v)                                                                             \
  case v:                                                                                           \
    hipLaunchKernelGGL((k_sweep<T, kMaxDim, v, true>), dim3(blocks), dim3(64 * nw), lds, s, a); \
    break;
    IS_RANGE_SWEEP(0) IS_RANGE_SWEEP(1) IS_RANGE_SWEEP(2) IS_RANGE_SWEEP(3) IS_RANGE_SWEEP(4) IS_RANGE_SWEEP(5)
#undef IS_RANGE_SWEEP
    default:
      break;
  }
}



// moment tile counts an instantiation serves: d in the bucket of D, d + 1 values
constexpr int nt_lo(int D) { return D <= kExactDims ? (D + 16) / 16 : (D == 24 ? 2 : D == 32 ? 2 : D == 48 ? 3 : 4); }
constexpr int nt_hi(int D) { return (D + 16) / 16; }

template <class T, int D>
void launch_sweep_any(int nt, const SweepArgs<T>& a, int blocks, int nw, size_t lds, hipStream_t s) {
  if (nt == 0)
    hipLaunchKernelGGL((k_sweep<T, D, 0>), dim3(blocks), dim3(64 * nw), lds, s, a);
  else if (nt == nt_lo(D))
    hipLaunchKernelGGL((k_sweep<T, D, nt_lo(D)>), dim3(blocks), dim3(64 * nw), lds, s, a);
  else
    hipLaunchKernelGGL((k_sweep<T, D, nt_hi(D)>), dim3(blocks), dim3(64 * nw), lds, s, a);
}

// LDS of a fused round block: the larger of the sweep staging and the new-sample
// layout, whose slot chunk may grow into kRoundLds.

template <class T, int D>
bool launch_round_any(int nt, const RoundArgs<T>& a0, const SweepArgs<T>& w, int ncomp, int nb_new, int nrec_old,
                      hipStream_t s) {
  if constexpr (D > kExactDims || nt_lo(D) != 1) {
    return false;
  } else {
    if (nt != 0 && nt != 1) return false;
    RoundArgs<T> a = a0;
    const bool st = a.fam.student != 0;
    size_t lds = 0;
    a.slot_chunk = round_slot_chunk(D, nt, a.stride, (int)sizeof(T), a.tkind == kTGaussian, ncomp, &lds);
    const int blocks = nb_new + nrec_old;
    if (nt == 1) {
      if (st)
        hipLaunchKernelGGL((k_round<T, D, 1, true>), dim3(blocks), dim3(256), lds, s, a, w, nb_new, nrec_old);
      else
        hipLaunchKernelGGL((k_round<T, D, 1, false>), dim3(blocks), dim3(256), lds, s, a, w, nb_new, nrec_old);
    } else {
      if (st)
        hipLaunchKernelGGL((k_round<T, D, 0, true>), dim3(blocks), dim3(256), lds, s, a, w, nb_new, nrec_old);
      else
        hipLaunchKernelGGL((k_round<T, D, 0, false>), dim3(blocks), dim3(256), lds, s, a, w, nb_new, nrec_old);
    }
    return true;
  }
}

template <class T, int D>
bool launch_round_group_any(int nt, bool st, const RoundArgs<T>* ga, const SweepArgs<T>* gw, int jobs, int nb_new,
                            int nrec_old, size_t lds, hipStream_t s) {
  if constexpr (D > kExactDims || nt_lo(D) != 1) {
    return false;
  } else {
    if (nt != 0 && nt != 1) return false;
    const dim3 grid((unsigned)(nb_new + nrec_old), (unsigned)jobs);
    if (nt == 1) {
      if (st)
        hipLaunchKernelGGL((k_round_group<T, D, 1, true>), grid, dim3(256), lds, s, ga, gw, nb_new, nrec_old);
      else
        hipLaunchKernelGGL((k_round_group<T, D, 1, false>), grid, dim3(256), lds, s, ga, gw, nb_new, nrec_old);
    } else {
      if (st)
        hipLaunchKernelGGL((k_round_group<T, D, 0, true>), grid, dim3(256), lds, s, ga, gw, nb_new, nrec_old);
      else
        hipLaunchKernelGGL((k_round_group<T, D, 0, false>), grid, dim3(256), lds, s, ga, gw, nb_new, nrec_old);
    }
    return true;
  }
}

template <class T, int D>
void launch_target_any(const TargetArgs<T>& t, int64_t splits, hipStream_t s) {
  hipLaunchKernelGGL((k_target_gemm<T, D>), dim3((unsigned)((t.n_new + 63) / 64), (unsigned)splits), dim3(256),
                     target_lds_bytes((D + 3) / 4, (int)sizeof(T)), s, t);
}

template <int D>
void DimKernels<D>::new_samples(int mode, const RoundArgs<float>& a, int ncomp, int blocks, int W, hipStream_t s) {
  launch_new_any<float, D>(mode, a, ncomp, blocks, W, s);
}
template <int D>
void DimKernels<D>::new_samples(int mode, const RoundArgs<double>& a, int ncomp, int blocks, int W, hipStream_t s) {
  launch_new_any<double, D>(mode, a, ncomp, blocks, W, s);
}
template <int D>
void DimKernels<D>::sweep(int nt, const SweepArgs<float>& a, int blocks, int nw, size_t lds, hipStream_t s) {
  launch_sweep_any<float, D>(nt, a, blocks, nw, lds, s);
}
template <int D>
void DimKernels<D>::sweep(int nt, const SweepArgs<double>& a, int blocks, int nw, size_t lds, hipStream_t s) {
  launch_sweep_any<double, D>(nt, a, blocks, nw, lds, s);
}
template <int D>
bool DimKernels<D>::round(int nt, const RoundArgs<float>& a, const SweepArgs<float>& w, int ncomp, int nb_new,
                          int nrec_old, hipStream_t s) {
  return launch_round_any<float, D>(nt, a, w, ncomp, nb_new, nrec_old, s);
}
template <int D>
bool DimKernels<D>::round(int nt, const RoundArgs<double>& a, const SweepArgs<double>& w, int ncomp, int nb_new,
                          int nrec_old, hipStream_t s) {
  return launch_round_any<double, D>(nt, a, w, ncomp, nb_new, nrec_old, s);
}
template <int D>
bool DimKernels<D>::round_group(int nt, bool student, const RoundArgs<float>* ga, const SweepArgs<float>* gw, int jobs,
                                int nb_new, int nrec_old, size_t lds, hipStream_t s) {
  return launch_round_group_any<float, D>(nt, student, ga, gw, jobs, nb_new, nrec_old, lds, s);
}
template <int D>
bool DimKernels<D>::round_group(int nt, bool student, const RoundArgs<double>* ga, const SweepArgs<double>* gw,
                                int jobs, int nb_new, int nrec_old, size_t lds, hipStream_t s) {
  return launch_round_group_any<double, D>(nt, student, ga, gw, jobs, nb_new, nrec_old, lds, s);
}
template <int D>
void DimKernels<D>::target(const TargetArgs<float>& t, int64_t splits, hipStream_t s) {
  launch_target_any<float, D>(t, splits, s);
}
template <int D>
void DimKernels<D>::target(const TargetArgs<double>& t, int64_t splits, hipStream_t s) {
  launch_target_any<double, D>(t, splits, s);
}

}  // namespace issampler
