// k_resid_f64.hip -- FP64 instantiation of the inverse-x / residual /
// Anderson inner-product kernel (k_iterate_impl.h).
#include "k_iterate_impl.h"

namespace rism {

template hipError_t launch_fx_resid<double, double>(const IterArgs<double, double> &, const LaunchGeom &, int, hipStream_t);

}  // namespace rism
