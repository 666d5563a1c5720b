// One part of one dimension instantiation of the engine's kernels; compiled once
// per IS_DIM value listed in IS_DIMS (launch.h) and per IS_DIM_PART:
//   0: new-sample kernels, FP32     1: new-sample kernels, FP64
//   2: sweep, fused round and target kernels, FP32
//   3: sweep, fused round and target kernels, FP64
//   4, 5 (IS_DIM 64 only): the extended-range new-sample and sweep kernels
//        (RangeKernels), FP32 and FP64
//   6, 7 (IS_DIM <= 15): the grouped fused rounds (k_round_group), FP32 and FP64
// so that the wide instantiations compile in parallel.
#include <type_traits>

#include "dim_impl.h"

#ifndef IS_DIM
#error "compile dim.hip with -DIS_DIM=<dimension instantiation>"
#endif
#ifndef IS_DIM_PART
#error "compile dim.hip with -DIS_DIM_PART=<0..7>"
#endif

namespace issampler {
#if IS_DIM_PART >= 6
using T_ = std::conditional_t<IS_DIM_PART == 6, float, double>;
template bool DimKernels<IS_DIM>::round_group(int, bool, const RoundArgs<T_>*, const SweepArgs<T_>*, int, int, int,
                                              size_t, hipStream_t);
#elif IS_DIM_PART >= 4
using T_ = std::conditional_t<IS_DIM_PART == 4, float, double>;
void RangeKernels::new_samples(int mode, const RoundArgs<T_>& a, int ncomp, int blocks, int W, hipStream_t s) {
  launch_new_any<T_, kMaxDim, true>(mode, a, ncomp, blocks, W, s);
}
void RangeKernels::sweep(int nt, const SweepArgs<T_>& a, int blocks, int nw, size_t lds, hipStream_t s) {
  launch_sweep_range<T_>(nt, a, blocks, nw, lds, s);
}
#else
#if IS_DIM_PART == 0 || IS_DIM_PART == 2
using T_ = float;
#else
using T_ = double;
#endif
#if IS_DIM_PART < 2
template void DimKernels<IS_DIM>::new_samples(int, const RoundArgs<T_>&, int, int, int, hipStream_t);
#else
template void DimKernels<IS_DIM>::sweep(int, const SweepArgs<T_>&, int, int, size_t, hipStream_t);
template bool DimKernels<IS_DIM>::round(int, const RoundArgs<T_>&, const SweepArgs<T_>&, int, int, int, hipStream_t);
template void DimKernels<IS_DIM>::target(const TargetArgs<T_>&, int64_t, hipStream_t);
#if IS_DIM > 15
// no fused round at this dimension: the grouped launcher only reports that
template bool DimKernels<IS_DIM>::round_group(int, bool, const RoundArgs<T_>*, const SweepArgs<T_>*, int, int, int,
                                              size_t, hipStream_t);
#endif
#endif
#endif
}  // namespace issampler
