// k_iterate_f64.hip -- FP64 instantiations of the iteration kernels
// (k_iterate_impl.h), the Anderson solve launcher and the FP64 FFT seam.
#include "k_iterate_impl.h"

namespace rism {

hipError_t launch_anderson_solve(Ctl *ctl, const double *partials, int nblk, int nv, int hm,
                                 double *rms_hist, int iter, hipStream_t st) {
  return launch_anderson_solve_impl(ctl, partials, nblk, nv, hm, rms_hist, iter, st);
}

template hipError_t launch_closure_fx<double, double>(const IterArgs<double, double> &, const LaunchGeom &, bool, int, hipStream_t);
template hipError_t launch_fft_y<double, double>(const IterArgs<double, double> &, const LaunchGeom &, bool, hipStream_t);
template hipError_t launch_fft_z_oz<double, double>(const IterArgs<double, double> &, const LaunchGeom &, hipStream_t);
template hipError_t launch_fx_resid<double, double>(const IterArgs<double, double> &, const LaunchGeom &, int, hipStream_t);
template hipError_t launch_fft3d<double>(double2 *, const AxisPlan &, const AxisPlan &, const AxisPlan &,
                                         const LaunchGeom &, int, int, int, double, double, int, hipStream_t);

}  // namespace rism
