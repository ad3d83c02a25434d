// k_rmdft.hip -- the RMDFT free-energy functional of the converged h
// (src-cuda/cal_rmdft.cu:5-282, with the tables of solvent_spline2.cu:7-92),
// FP64, one-time after the solve.  h comes packed two sites per complex
// value (the FP64 evaluation's spectrum buffer); every k-space factor is a
// natural spline evaluated at |k| of the half-shifted grid, with the
// interval the reference's bisection (splint.cc:4-20) selects.
#include "rism_kernels.h"

namespace rism {

__device__ __forceinline__ double h_site(const double2 *Z, long N, int v, long ip) {
  const double2 z = Z[(long)(v >> 1) * N + ip];
  return (v & 1) ? z.y : z.x;
}

__device__ __forceinline__ double kmag(long ip, const RmdftGrid &g) {
  const int x = (int)(ip % g.nx), y = (int)((ip / g.nx) % g.ny), z = (int)(ip / ((long)g.nx * g.ny));
  const double kx = 2.0 * M_PI * ((double)x - g.nx / 2.0 + 0.5) / g.box[0];
  const double ky = 2.0 * M_PI * ((double)y - g.ny / 2.0 + 0.5) / g.box[1];
  const double kz = 2.0 * M_PI * ((double)z - g.nz / 2.0 + 0.5) / g.box[2];
  return sqrt(kx * kx + ky * ky + kz * kz);
}

// splint.cc:9-16: the largest knot index lo with x[lo] <= k (0 below the
// table, n-2 at or beyond its end)
__device__ __forceinline__ int bisect(const double *x, int n, double k) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int m = (hi + lo) >> 1;
    if (x[m] > k) hi = m;
    else lo = m;
  }
  return lo;
}

__device__ __forceinline__ double seval(const double4 &c, double t) {
  return ((c.w * t + c.z) * t + c.y) * t + c.x;  // (y, c0, c1, c2)
}

// h_v -> complex buffer (imaginary part 0), and per-block sums of h_v (xmu1)
__global__ __launch_bounds__(kThreads) void k_rm_unpack(const double2 *Z, long N, int v, double2 *out,
                                                        double *partials) {
  double s = 0.0;
  for (long ip = (long)blockIdx.x * kThreads + threadIdx.x; ip < N; ip += (long)gridDim.x * kThreads) {
    const double hv = h_site(Z, N, v, ip);
    out[ip] = make_double2(hv, 0.0);
    s += hv;
  }
  __shared__ double red[kThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    partials[blockIdx.x] = t;
  }
}

// cal_dwork and rho0wfkhk (cal_rmdft.cu:167-201): for output site v,
// w1(k) = sum_v' c~_vv'(|k|) rho_v' H_v'(k) with c~_vv = c_vv - chs_v rho_0 / rho_v
// (solvent_spline2.cu:64-69), w2(k) = rho_0 w_v(|k|) H_v(k)
__global__ __launch_bounds__(kThreads) void k_rm_kmix(RmdftGrid g, RmdftTabs t, int v, const double2 *const *H,
                                                      double2 *w1, double2 *w2) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= g.N) return;
  const double kk = kmag(ip, g);
  const int nseg = t.ntab - 1;
  int lo = (int)(kk / t.dk) - 1;  // uniform knots k_i = dk (i + 1)
  lo = max(0, min(lo, nseg - 1));
  if (lo < nseg - 1 && t.dk * (lo + 2) <= kk) ++lo;
  if (lo > 0 && t.dk * (lo + 1) > kk) --lo;
  const double tc = kk - t.dk * (lo + 1);
  const int lh = bisect(t.khs, t.nhs, kk);
  const double th = kk - t.khs[lh];
  const int V = t.V;
  double re = 0.0, im = 0.0;
  for (int v1 = 0; v1 < V; ++v1) {
    double c = seval(t.cvk[(long)(v * V + v1) * nseg + lo], tc);
    if (v1 == v) c -= seval(t.chs[(long)v * (t.nhs - 1) + lh], th) * t.rho0 / t.rho[v];
    const double2 hk = H[v1][ip];
    re += hk.x * c * t.rho[v1];
    im += hk.y * c * t.rho[v1];
  }
  w1[ip] = make_double2(re, im);
  const double w = t.rho0 * seval(t.whs[(long)v * (t.nhs - 1) + lh], th);
  const double2 hv = H[v][ip];
  w2[ip] = make_double2(w * hv.x, w * hv.y);
}

// rho0 + rho0guvfmbex (cal_rmdft.cu:203-228): from the weighted density
// n = rho_0 + rho_0 (w * h), f = rho_0 ((h + 1) fd(n) - fd_bulk)
__global__ __launch_bounds__(kThreads) void k_rm_fd(const double2 *Z, long N, int v, double a, double bex,
                                                    double rho0, double2 *w2) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= N) return;
  const double dens = w2[ip].x + rho0;
  const double p = a * dens;
  const double fd = (4.0 - 2.0 * p) / ((1.0 - p) * (1.0 - p) * (1.0 - p)) * (p / dens);
  w2[ip] = make_double2(rho0 * ((h_site(Z, N, v, ip) + 1.0) * fd - bex), 0.0);
}

// wfka (cal_rmdft.cu:230-240): multiply by w_v(|k|)
__global__ __launch_bounds__(kThreads) void k_rm_wmul(RmdftGrid g, RmdftTabs t, int v, double2 *w2) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= g.N) return;
  const double kk = kmag(ip, g);
  const int lh = bisect(t.khs, t.nhs, kk);
  const double w = seval(t.whs[(long)v * (t.nhs - 1) + lh], kk - t.khs[lh]);
  w2[ip] = make_double2(w2[ip].x * w, w2[ip].y * w);
}

// per-block sums: mode 0: (sum Re a, sum Re a * h_v) (xmu2, xmu3);
// mode 1: sum (Re a + e0)(h_v + 1) - e0 (cal_eda, cal_rmdft.cu:242-282)
__global__ __launch_bounds__(kThreads) void k_rm_reduce(const double2 *Z, long N, int v, const double2 *a,
                                                        int mode, double e0, double *partials) {
  double s0 = 0.0, s1 = 0.0;
  for (long ip = (long)blockIdx.x * kThreads + threadIdx.x; ip < N; ip += (long)gridDim.x * kThreads) {
    const double ar = a[ip].x, hv = h_site(Z, N, v, ip);
    if (mode == 0) {
      s0 += ar;
      s1 += ar * hv;
    } else {
      s0 += (ar + e0) * (hv + 1.0) - e0;
    }
  }
  __shared__ double red[2][kThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o, 64);
    s1 += __shfl_xor(s1, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = s0;
    red[1][threadIdx.x >> 6] = s1;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double u = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) u += red[threadIdx.x][w];
    partials[2 * blockIdx.x + threadIdx.x] = u;
  }
}

static unsigned blocks(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

hipError_t launch_rm_unpack(const double2 *Z, long N, int v, double2 *out, double *partials, int nblk,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_rm_unpack, dim3(nblk), dim3(kThreads), 0, st, Z, N, v, out, partials);
  return hipGetLastError();
}
hipError_t launch_rm_kmix(const RmdftGrid &g, const RmdftTabs &t, int v, const double2 *const *H, double2 *w1,
                          double2 *w2, hipStream_t st) {
  hipLaunchKernelGGL(k_rm_kmix, dim3(blocks(g.N)), dim3(kThreads), 0, st, g, t, v, H, w1, w2);
  return hipGetLastError();
}
hipError_t launch_rm_fd(const double2 *Z, long N, int v, double a, double bex, double rho0, double2 *w2,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_rm_fd, dim3(blocks(N)), dim3(kThreads), 0, st, Z, N, v, a, bex, rho0, w2);
  return hipGetLastError();
}
hipError_t launch_rm_wmul(const RmdftGrid &g, const RmdftTabs &t, int v, double2 *w2, hipStream_t st) {
  hipLaunchKernelGGL(k_rm_wmul, dim3(blocks(g.N)), dim3(kThreads), 0, st, g, t, v, w2);
  return hipGetLastError();
}
hipError_t launch_rm_reduce(const double2 *Z, long N, int v, const double2 *a, int mode, double e0,
                            double *partials, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_rm_reduce, dim3(nblk), dim3(kThreads), 0, st, Z, N, v, a, mode, e0, partials);
  return hipGetLastError();
}

}  // namespace rism
