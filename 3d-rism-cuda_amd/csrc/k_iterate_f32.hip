// k_iterate_f32.hip -- FP32 (mixed precision) instantiations of the iteration kernels
// (k_iterate_impl.h), and the FP32 FFT seam.
#include "k_iterate_impl.h"

namespace rism {

template hipError_t launch_closure_fx<float, double>(const IterArgs<float, double> &, const LaunchGeom &, bool, int, hipStream_t);
template hipError_t launch_fft_y<float, double>(const IterArgs<float, double> &, const LaunchGeom &, bool, hipStream_t);
template hipError_t launch_fft_z_oz<float, double>(const IterArgs<float, double> &, const LaunchGeom &, hipStream_t);
template hipError_t launch_fx_resid<float, double>(const IterArgs<float, double> &, const LaunchGeom &, int, hipStream_t);
template hipError_t launch_fft3d<float>(float2 *, const AxisPlan &, const AxisPlan &, const AxisPlan &,
                                         const LaunchGeom &, int, int, int, double, double, int, hipStream_t);

}  // namespace rism
