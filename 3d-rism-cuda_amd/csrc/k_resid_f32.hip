// k_resid_f32.hip -- FP32 (mixed precision) instantiation of the inverse-x /
// residual / Anderson inner-product kernel (k_iterate_impl.h).
#include "k_iterate_impl.h"

namespace rism {

template hipError_t launch_fx_resid<float, double>(const IterArgs<float, double> &, const LaunchGeom &, int, hipStream_t);

}  // namespace rism
