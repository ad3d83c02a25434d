// k_mixed_f32.hip -- the FP32 chain's closure and residual passes of the
// mixed-precision iteration (k_mixed_impl.h).
#include "k_mixed_impl.h"

namespace rism {

template <>
hipError_t launch_mx_closure_fx<float>(const IterArgs<float, double> &a, const LaunchGeom &g, int hm, hipStream_t st) {
  return launch_mx_closure_fx_impl<float>(a, g, hm, st);
}
template <>
hipError_t launch_mx_resid<float>(const IterArgs<float, double> &a, const LaunchGeom &g, int hm, hipStream_t st) {
  return launch_mx_resid_impl<float>(a, g, hm, st);
}

}  // namespace rism
