// k_fft_f32.hip -- FP32 instantiations of the y pass, the z pass with the
// k-space Ornstein-Zernike step, and the standalone 3-D transform
// (k_iterate_impl.h).
#include "k_iterate_impl.h"

namespace rism {

template hipError_t launch_fft_y<float, double>(const IterArgs<float, double> &, const LaunchGeom &, bool, hipStream_t);
template hipError_t launch_fft_z_oz<float, double>(const IterArgs<float, double> &, const LaunchGeom &, hipStream_t);
template hipError_t launch_fft3d<float>(float2 *, const AxisPlan &, const AxisPlan &, const AxisPlan &,
                                        const LaunchGeom &, int, int, int, double, double, int, hipStream_t);

}  // namespace rism
