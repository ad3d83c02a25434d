// k_iterate_f32.hip -- FP32 (mixed precision) instantiation of the
// closure/update kernel (k_iterate_impl.h).
#include "k_iterate_impl.h"

namespace rism {

template hipError_t launch_closure_fx<float, double>(const IterArgs<float, double> &, const LaunchGeom &, bool, int, hipStream_t);

}  // namespace rism
