// Kernel translation unit of the batched AMIS engine for one (precision, waves)
// pair: compiled with -DIS_BATCH_FP=32|64 and -DIS_BATCH_W=2|4|8.
#include "batch_kernel.h"

#if !defined(IS_BATCH_FP) || !defined(IS_BATCH_W)
#error "compile batch_k.hip with -DIS_BATCH_FP=<32|64> -DIS_BATCH_W=<waves>"
#endif

namespace issampler {

#if IS_BATCH_FP == 32
typedef float BatchT;
#else
typedef double BatchT;
#endif

template <class T, int W>
void launch_batch_amis(const BatchArgs<T>& a, int blocks, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_amis<T, W>), dim3(blocks), dim3(64 * W), lds, s, a);
}
template void launch_batch_amis<BatchT, IS_BATCH_W>(const BatchArgs<BatchT>&, int, size_t, hipStream_t);

#if IS_BATCH_W == 4
template <class T>
void launch_batch_begin(const BatchJob<T>* jobs, int njobs, unsigned int* queue, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_begin<T>), dim3(njobs), dim3(64), 0, s, jobs, njobs, queue);
}
template <class T>
void launch_batch_end(const BatchJob<T>* jobs, int njobs, int commit, int* status, unsigned long long* run_word,
                      hipStream_t s) {
  hipLaunchKernelGGL((k_batch_end<T>), dim3(njobs), dim3(64), 0, s, jobs, njobs, commit, status, run_word);
}
template void launch_batch_begin<BatchT>(const BatchJob<BatchT>*, int, unsigned int*, hipStream_t);
template void launch_batch_end<BatchT>(const BatchJob<BatchT>*, int, int, int*, unsigned long long*, hipStream_t);
#endif

}  // namespace issampler
