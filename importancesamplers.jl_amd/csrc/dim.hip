// One dimension instantiation of the engine's kernels; compiled once per
// IS_DIM value listed in IS_DIMS (launch.h).
#include "dim_impl.h"

#ifndef IS_DIM
#error "compile dim.hip with -DIS_DIM=<dimension instantiation>"
#endif

namespace issampler {
template struct DimKernels<IS_DIM>;
}
