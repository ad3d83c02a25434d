// k_iterate_f64.hip -- FP64 instantiation of the closure/update kernel
// (k_iterate_impl.h) and the Anderson solve launcher.
#include "k_iterate_impl.h"

namespace rism {

hipError_t launch_anderson_solve(Ctl *ctl, const double *partials, int nblk, int nv, int hm,
                                 double *rms_hist, hipStream_t st, int gate, long astride, int nsol, int *rlist) {
  return launch_anderson_solve_impl(ctl, partials, nblk, nv, hm, rms_hist, st, gate, astride, nsol, rlist);
}

template hipError_t launch_closure_fx<double, double>(const IterArgs<double, double> &, const LaunchGeom &, bool, int, hipStream_t);

}  // namespace rism
