// k_mixed_f64.hip -- the FP64 refresh chain's closure and residual passes of
// the mixed-precision iteration (k_mixed_impl.h), and the conversions
// between its state (A, dl) and the short-range state t_s.
#include "k_mixed_impl.h"

namespace rism {

template <>
hipError_t launch_mx_closure_fx<double>(const IterArgs<double, double> &a, const LaunchGeom &g, int hm,
                                        hipStream_t st) {
  return launch_mx_closure_fx_impl<double>(a, g, hm, st);
}
template <>
hipError_t launch_mx_resid<double>(const IterArgs<double, double> &a, const LaunchGeom &g, int hm, hipStream_t st) {
  return launch_mx_resid_impl<double>(a, g, hm, st);
}

// dir 0: T = ts, A = ts - u, dl = 0 (a new reference at the state);
// dir 1: ts = T + dl
__global__ __launch_bounds__(kThreads) void k_mx_convert(double *ts, double *T, float *A, float *dl,
                                                         const double *u, long n, int dir) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  if (dir == 0) {
    T[i] = ts[i];
    A[i] = ref_f32(ts[i] - u[i]);
    dl[i] = 0.0f;
  } else {
    ts[i] = T[i] + (double)dl[i];
  }
}

hipError_t launch_mx_convert(double *ts, double *T, float *A, float *dl, const double *u, long n, int dir,
                             hipStream_t st) {
  const long nb = (n + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_mx_convert, dim3((unsigned)nb), dim3(kThreads), 0, st, ts, T, A, dl, u, n, dir);
  return hipGetLastError();
}

}  // namespace rism
