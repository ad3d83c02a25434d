// k_adjoint.hip -- adjoint of the converged 3D-RISM fixed point.  All
// arithmetic is FP64; the spectrum of the iterations' K^T applications is
// complex FP64, or complex FP32 in the mixed path (T2 = float2: FP32
// transforms, everything else FP64).
//
// The solve finds the short-range state s with
//   R(s) = h(s) + 1 - g(s) = 0,  g = closure(s - u),  c_s = g - 1 - s,
//   h = K c_s - K_fk,  K = IFFT chi(|k|) FFT (site-mixing OZ operator),
// and Delta-mu = F(h, t) = kT sum_v rho_v dv sum_r phi(h_v, c_v) with
// t = s + q_v cf fr, c = h - t (cal_exchem.cu integrands).  With
// L = F + lambda.R, dDelta-mu/dp = dL/dp once lambda solves
//   lambda = (g' - 1)(K^T lambda - lambda) + b,
//   b = (g' - 1) K^T alpha + beta,  alpha = kT rho dv (phi_h + phi_c),
//   beta = -kT rho dv phi_c.
// K^T is the same convolution with the site matrix transposed (chi(|k|) is
// real and even, so the real-space kernel is symmetric).  The field
// adjoints are then, with mu = lambda + alpha,
//   ubar_v = -g'_v (K^T mu - lambda)_v          (u = beta u_LJ + q_v cf (de - fr))
//   debar  = cf sum_v q_v ubar_v,  frbar = cf sum_v q_v (beta_v - ubar_v)
//   fkbar(k) = -(cf / (vol dv)) Y(k),  Y = sum_w conj(M_w) sum_v chi_wv q_v,  M = FFT(mu)
// and the per-atom derivatives are grid sums of these against the
// derivatives of the potentials of cal_LJ.cu / cal_Coulomb.cu.
#include "rism_kernels.h"

namespace rism {

static inline dim3 agrid(long n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

__device__ __forceinline__ double adj_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// g', alpha, beta at the converged state; h from the packed FP64 spectrum
// buffer that holds real-space h after the evaluation pass.
__global__ __launch_bounds__(kThreads) void k_adj_fields(AdjArgs a) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= a.N) return;
  const long N = a.N;
  for (int v = 0; v < a.V; ++v) {
    const long iv = (long)v * N + ip;
    const double d = a.ts[iv] - a.u[iv];
    const double gp = (a.closure == 0 && d >= 0.0) ? 1.0 : exp(d);
    const double2 z = a.hz[(long)(v >> 1) * N + ip];
    const double h = (v & 1) ? z.y : z.x;
    const double t = a.ts[iv] + a.qcf[v] * a.fr[ip];
    const double c = h - t;
    // phi: KH -c - hc/2 + h^2/2 Theta(-h); HNC h^2/2 - c - hc/2 (cal_exchem.cu:55-154)
    const double phic = -1.0 - 0.5 * h;
    const double phih = -0.5 * c + ((a.closure != 0 || h <= 0.0) ? h : 0.0);
    const double w = a.kt * a.rho[v] * a.dv;
    a.gp[iv] = gp;
    a.alpha[iv] = w * (phih + phic);
    a.beta[iv] = -w * phic;
  }
}

template <typename T2> __device__ __forceinline__ T2 mk2(double a, double b) {
  T2 r;
  r.x = (decltype(r.x))a;
  r.y = (decltype(r.y))b;
  return r;
}

// Z_p = x_{2p} + i x_{2p+1} (+ y): two real site fields per complex field
template <typename T2>
__global__ __launch_bounds__(kThreads) void k_adj_pack(long N, int V, const double *x, const double *y, T2 *Z) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= N) return;
  for (int p = 0; p < (V + 1) / 2; ++p) {
    const long i0 = (long)(2 * p) * N + ip, i1 = (long)(2 * p + 1) * N + ip;
    double re = x[i0], im = (2 * p + 1 < V) ? x[i1] : 0.0;
    if (y) {
      re += y[i0];
      if (2 * p + 1 < V) im += y[i1];
    }
    Z[(long)p * N + ip] = mk2<T2>(re, im);
  }
}

// k-space: out_v(k) = sum_w chi_wv(|k|) X_w(k) on a full spectrum (pairs
// ip and N-1-ip are k and -k on the half-shifted grid); optionally
// Y(k) = sum_w conj(X_w(k)) sum_v chi_wv(k) q_v.
template <typename T2>
__global__ __launch_bounds__(kThreads) void k_adj_kspace(AdjArgs a, T2 *Z, double2 *Y) {
  const long N = a.N;
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= N / 2) return;
  const long im = N - 1 - ip;
  const int nx = a.nx, ny = a.ny, nz = a.nz, V = a.V, P = (a.V + 1) / 2;
  const int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
  const double kx = 2.0 * M_PI / a.box[0] * (x - nx / 2.0 + 0.5);
  const double ky = 2.0 * M_PI / a.box[1] * (y - ny / 2.0 + 0.5);
  const double kz = 2.0 * M_PI / a.box[2] * (z - nz / 2.0 + 0.5);
  const double kk = sqrt(kx * kx + ky * ky + kz * kz);
  const int nseg = a.ntab - 1;
  int lo = (int)floor(kk / a.dk) - 1;  // splint.cc:4-20 on k_i = dk (i + 1)
  if (lo < 0) lo = 0;
  if (lo > nseg - 1) lo = nseg - 1;
  if (lo < nseg - 1 && a.dk * (lo + 2) <= kk) ++lo;
  if (lo > 0 && a.dk * (lo + 1) > kk) --lo;
  const double tt = kk - a.dk * (lo + 1);
  double chi[kMaxV][kMaxV];
  for (int w = 0; w < V; ++w)
    for (int v = 0; v < V; ++v) {
      const double4 q4 = a.chi[(long)(w * V + v) * nseg + lo];
      chi[w][v] = ((q4.w * tt + q4.z) * tt + q4.y) * tt + q4.x;
    }
  double2 X[kMaxV];  // X_w(k)
  for (int p = 0; p < P; ++p) {
    const T2 za = Z[(long)p * N + ip], zb = Z[(long)p * N + im];
    X[2 * p] = make_double2(0.5 * (za.x + zb.x), 0.5 * (za.y - zb.y));
    if (2 * p + 1 < V) X[2 * p + 1] = make_double2(0.5 * (za.y + zb.y), -0.5 * (za.x - zb.x));
  }
  if (Y) {
    double yr = 0.0, yi = 0.0;
    for (int w = 0; w < V; ++w) {
      double s = 0.0;
      for (int v = 0; v < V; ++v) s += chi[w][v] * a.qv[v];
      yr += X[w].x * s;   // conj(X_w) * s
      yi -= X[w].y * s;
    }
    Y[ip] = make_double2(yr, yi);
    Y[im] = make_double2(yr, -yi);
  }
  double2 O[kMaxV + 1];
  for (int v = 0; v < V; ++v) {
    double re = 0.0, imv = 0.0;
    for (int w = 0; w < V; ++w) {
      re += chi[w][v] * X[w].x;
      imv += chi[w][v] * X[w].y;
    }
    O[v] = make_double2(re, imv);
  }
  O[V] = make_double2(0.0, 0.0);
  for (int p = 0; p < P; ++p) {
    const double2 o0 = O[2 * p], o1 = O[2 * p + 1];
    // Z(k) = O0 + i O1 ; Z(-k) = conj(O0) + i conj(O1)
    Z[(long)p * N + ip] = mk2<T2>(o0.x - o1.y, o0.y + o1.x);
    Z[(long)p * N + im] = mk2<T2>(o0.x + o1.y, o1.x - o0.y);
  }
}

// b = (g' - 1) (K^T alpha) + beta, with K^T alpha unpacked from Z
__global__ __launch_bounds__(kThreads) void k_adj_rhs(AdjArgs a, const double2 *Z) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= a.N) return;
  for (int v = 0; v < a.V; ++v) {
    const long iv = (long)v * a.N + ip;
    const double2 z = Z[(long)(v >> 1) * a.N + ip];
    const double ka = (v & 1) ? z.y : z.x;
    a.b[iv] = (a.gp[iv] - 1.0) * ka + a.beta[iv];
  }
}

// pending Anderson update of lambda (absolute history form, as the FP64
// forward pass: anderson_cuda.cu:1-23)
template <int HM>
__global__ __launch_bounds__(kThreads) void k_adj_update(AdjArgs a) {
  const Ctl *ctl = a.ctl;
  if (ctl->upd_iter != ctl->cur) return;
  const long VN = (long)a.V * a.N;
  const long iv = (long)blockIdx.x * kThreads + threadIdx.x;
  if (iv >= VN) return;
  const int nf = ctl->upd_nf, w = ctl->upd_w;
  const double mix = ctl->upd_mix;
  const double l = a.lam[iv], r = a.res[iv];
  double lb = l, rb = r;
#pragma unroll
  for (int s = 0; s < HM; ++s)
    if (s < nf) {
      const double th = ctl->upd_theta[s];
      lb += th * (a.hist_l[(long)s * VN + iv] - l);
      rb += th * (a.hist_r[(long)s * VN + iv] - r);
    }
  a.hist_l[(long)w * VN + iv] = l;
  a.hist_r[(long)w * VN + iv] = r;
  a.lam[iv] = lb + mix * rb;
}

// residual r = (g' - 1)(K^T lambda - lambda) + b - lambda and the Anderson
// inner products (same partial layout as the forward residual pass)
template <int HM, typename T2>
__global__ __launch_bounds__(kThreads) void k_adj_resid(AdjArgs a, const T2 *Z) {
  constexpr int NV = 1 + HM + HM * (HM + 1) / 2;
  const Ctl *ctl = a.ctl;
  if (ctl->run_iter != ctl->cur) return;
  const long N = a.N, VN = (long)a.V * N;
  const int nf = ctl->filled < HM ? ctl->filled : HM;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (long ip = (long)blockIdx.x * kThreads + threadIdx.x; ip < N; ip += (long)gridDim.x * kThreads) {
    for (int v = 0; v < a.V; ++v) {
      const long iv = (long)v * N + ip;
      const T2 z = Z[(long)(v >> 1) * N + ip];
      const double kl = (double)((v & 1) ? z.y : z.x);
      const double l = a.lam[iv];
      const double r = (a.gp[iv] - 1.0) * (kl - l) + a.b[iv] - l;
      a.res[iv] = r;
      const double wgt = a.rho_inv[v];
      acc[0] += r * r;
      double d[HM];
#pragma unroll
      for (int s = 0; s < HM; ++s) d[s] = (s < nf) ? r - a.hist_r[(long)s * VN + iv] : 0.0;
#pragma unroll
      for (int s = 0; s < HM; ++s) acc[1 + s] += wgt * r * d[s];
      int kk = 1 + HM;
#pragma unroll
      for (int s = 0; s < HM; ++s)
#pragma unroll
        for (int s2 = s; s2 < HM; ++s2) acc[kk++] += wgt * d[s] * d[s2];
    }
  }
  __shared__ double red[kThreads / 64][NV];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double sw = adj_wave_sum(acc[k]);
    if (lane == 0) red[wave][k] = sw;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NV; k += kThreads) {
    double sm = 0.0;
    for (int wv = 0; wv < kThreads / 64; ++wv) sm += red[wv][k];
    a.partials[(long)blockIdx.x * NV + k] = sm;
  }
}

// field adjoints: ubar_v = -g'_v (K^T mu - lambda)_v, debar, frbar
__global__ __launch_bounds__(kThreads) void k_adj_bar(AdjArgs a, const double2 *Z) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= a.N) return;
  double de = 0.0, fr = 0.0;
  for (int v = 0; v < a.V; ++v) {
    const long iv = (long)v * a.N + ip;
    const double2 z = Z[(long)(v >> 1) * a.N + ip];
    const double km = (v & 1) ? z.y : z.x;
    const double ub = -a.gp[iv] * (km - a.lam[iv]);
    a.ubar[iv] = ub;
    de += a.qcf[v] * ub;
    fr += a.qcf[v] * (a.beta[iv] - ub);
  }
  a.debar[ip] = de;
  a.frbar[ip] = fr;
}

// Per (chunk of points, atom): sum over the chunk of the field adjoints
// times the derivatives of the potentials of cal_LJ.cu:47-72 and
// cal_Coulomb.cu:100-124 with respect to the atom's position, charge,
// sigma and epsilon (Lorentz-Berthelot mixing, cal_LJ.cu:24-31).
constexpr int kAdjPPT = 16;

__global__ __launch_bounds__(kThreads) void k_adj_grad_partial(AdjArgs a, double *partials) {
  const long N = a.N;
  const int nx = a.nx, ny = a.ny, nz = a.nz, V = a.V, nu = a.natu;
  const long cbase = (long)blockIdx.x * kThreads * kAdjPPT;
  const double cut2 = 1.0e-2 * 1.0e-2;
  __shared__ double red[kThreads / 64][6];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int u = 0; u < nu; ++u) {
    const double4 at = a.atoms[u];
    double g6[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < kAdjPPT; ++k) {
      const long ip = cbase + threadIdx.x + (long)k * kThreads;
      if (ip >= N) break;
      const int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
      const double dx = (x - nx / 2) * a.dr[0] - at.x;
      const double dy = (y - ny / 2) * a.dr[1] - at.y;
      const double dz = (z - nz / 2) * a.dr[2] - at.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double rinv = rsqrt_f64(fmax(r2, 1.0e-300));
      const double r = r2 * rinv;
      double fpos = 0.0;  // d/dR_u = fpos * (dx, dy, dz)
      // LJ (beta u_LJ = ikbt sum 4 eps (s^12 - s^6)), floor r >= 0.01
      const bool clamp = r2 < cut2;
      const double ir2 = clamp ? 1.0 / cut2 : rinv * rinv;
      for (int v = 0; v < V; ++v) {
        const double ub = a.ubar[(long)v * N + ip] * a.ikbt;
        const double2 p = a.lj[v * nu + u];   // (sigma_uv^2, 4 eps_uv)
        const double2 q = a.ljd[v * nu + u];  // (1 / (2 sigma_uv), 4 eps_uv / (2 eps_u))
        const double s2 = p.x * ir2, s6 = s2 * s2 * s2, s12 = s6 * s6;
        if (!clamp) fpos += ub * p.y * (12.0 * s12 - 6.0 * s6) * ir2;
        g6[4] += ub * p.y * (12.0 * s12 - 6.0 * s6) * q.x;
        g6[5] += ub * q.y * (s12 - s6);
      }
      // Coulomb: de = ubeta q/r (r >= 1e-5); fr = ubeta q (1 - e^-r)/r, -> ubeta q at r < 1e-5
      const double deb = a.debar[ip] * a.ubeta, frb = a.frbar[ip] * a.ubeta;
      if (r >= 1.0e-5) {
        const double ex = r < 38.0 ? exp(-r) : 0.0;
        const double f = (1.0 - ex) * rinv;
        g6[3] += deb * rinv + frb * f;
        // d(1/r)/dR = d/r^3 ; d f/dR = ((1 - e^-r)/r^2 - e^-r/r) d/r
        fpos += at.w * (deb * rinv * rinv * rinv + frb * (f * rinv - ex * rinv) * rinv);
      } else {
        g6[3] += frb;
      }
      g6[0] += fpos * dx;
      g6[1] += fpos * dy;
      g6[2] += fpos * dz;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double s = adj_wave_sum(g6[c]);
      if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
      double s = 0.0;
      for (int w = 0; w < kThreads / 64; ++w) s += red[w][threadIdx.x];
      partials[((long)blockIdx.x * nu + u) * 6 + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

// k-space part: per (chunk of k points, atom) the sums
//   S_u = sum_k Y(k) G(k) e^{-i k.R_u},  G = 4 pi / (k^2 (k^2 + 1))
// as (Re S, sum_k k Im(Y G e^{-ik.R}))  -> d/dq_u and d/dR_u of the fk term.
__global__ __launch_bounds__(kThreads) void k_adj_kgrad_partial(AdjArgs a, const double2 *Y,
                                                                double *partials) {
  const long N = a.N;
  const int nx = a.nx, ny = a.ny, nz = a.nz, nu = a.natu;
  const int nt = nx + ny + nz;
  const long cbase = (long)blockIdx.x * kThreads * kAdjPPT;
  __shared__ double red[kThreads / 64][4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int u = 0; u < nu; ++u) {
    const double2 *ph = a.phase + (long)u * nt;
    double g4[4] = {0, 0, 0, 0};
    for (int k = 0; k < kAdjPPT; ++k) {
      const long ip = cbase + threadIdx.x + (long)k * kThreads;
      if (ip >= N) break;
      const int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
      const double kx = 2.0 * M_PI / a.box[0] * (x - nx / 2.0 + 0.5);
      const double ky = 2.0 * M_PI / a.box[1] * (y - ny / 2.0 + 0.5);
      const double kz = 2.0 * M_PI / a.box[2] * (z - nz / 2.0 + 0.5);
      const double k2 = kx * kx + ky * ky + kz * kz;
      const double G = 4.0 * M_PI / (k2 * (k2 + 1.0));
      const double2 yv = Y[ip];
      const double2 ex = ph[x], ey = ph[nx + y], ez = ph[nx + ny + z];
      // e = ex ey ez
      const double exy_r = ex.x * ey.x - ex.y * ey.y, exy_i = ex.x * ey.y + ex.y * ey.x;
      const double e_r = exy_r * ez.x - exy_i * ez.y, e_i = exy_r * ez.y + exy_i * ez.x;
      const double zr = G * (yv.x * e_r - yv.y * e_i), zi = G * (yv.x * e_i + yv.y * e_r);
      g4[0] += zr;
      g4[1] += kx * zi;
      g4[2] += ky * zi;
      g4[3] += kz * zi;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double s = adj_wave_sum(g4[c]);
      if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      double s = 0.0;
      for (int w = 0; w < kThreads / 64; ++w) s += red[w][threadIdx.x];
      partials[((long)blockIdx.x * nu + u) * 4 + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

// fixed-order column sums of [nblk][n] partials
__global__ __launch_bounds__(kThreads) void k_adj_colsum(const double *partials, int nblk, int n, double *out) {
  const long id = (long)blockIdx.x * kThreads + threadIdx.x;
  if (id >= n) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partials[(long)b * n + id];
  out[id] = s;
}

// ---- launchers ----------------------------------------------------------
hipError_t launch_adj_fields(const AdjArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_fields, agrid(a.N), dim3(kThreads), 0, st, a);
  return hipGetLastError();
}
template <typename T2>
hipError_t launch_adj_pack(const AdjArgs &a, const double *x, const double *y, T2 *Z, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_pack<T2>, agrid(a.N), dim3(kThreads), 0, st, a.N, a.V, x, y, Z);
  return hipGetLastError();
}
template <typename T2>
hipError_t launch_adj_kspace(const AdjArgs &a, T2 *Z, double2 *Y, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_kspace<T2>, agrid(a.N / 2), dim3(kThreads), 0, st, a, Z, Y);
  return hipGetLastError();
}
hipError_t launch_adj_rhs(const AdjArgs &a, const double2 *Z, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_rhs, agrid(a.N), dim3(kThreads), 0, st, a, Z);
  return hipGetLastError();
}
hipError_t launch_adj_update(const AdjArgs &a, int hm, hipStream_t st) {
  const dim3 g = agrid((long)a.V * a.N);
  switch (hm) {
    case 2: hipLaunchKernelGGL(k_adj_update<2>, g, dim3(kThreads), 0, st, a); break;
    case 5: hipLaunchKernelGGL(k_adj_update<5>, g, dim3(kThreads), 0, st, a); break;
    case 10: hipLaunchKernelGGL(k_adj_update<10>, g, dim3(kThreads), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <typename T2>
hipError_t launch_adj_resid(const AdjArgs &a, const T2 *Z, int hm, int nblk, hipStream_t st) {
  switch (hm) {
    case 2: hipLaunchKernelGGL((k_adj_resid<2, T2>), dim3(nblk), dim3(kThreads), 0, st, a, Z); break;
    case 5: hipLaunchKernelGGL((k_adj_resid<5, T2>), dim3(nblk), dim3(kThreads), 0, st, a, Z); break;
    case 10: hipLaunchKernelGGL((k_adj_resid<10, T2>), dim3(nblk), dim3(kThreads), 0, st, a, Z); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template hipError_t launch_adj_pack<double2>(const AdjArgs &, const double *, const double *, double2 *, hipStream_t);
template hipError_t launch_adj_pack<float2>(const AdjArgs &, const double *, const double *, float2 *, hipStream_t);
template hipError_t launch_adj_kspace<double2>(const AdjArgs &, double2 *, double2 *, hipStream_t);
template hipError_t launch_adj_kspace<float2>(const AdjArgs &, float2 *, double2 *, hipStream_t);
template hipError_t launch_adj_resid<double2>(const AdjArgs &, const double2 *, int, int, hipStream_t);
template hipError_t launch_adj_resid<float2>(const AdjArgs &, const float2 *, int, int, hipStream_t);
hipError_t launch_adj_bar(const AdjArgs &a, const double2 *Z, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_bar, agrid(a.N), dim3(kThreads), 0, st, a, Z);
  return hipGetLastError();
}
int adj_grad_blocks(long N) { return (int)((N + (long)kThreads * kAdjPPT - 1) / ((long)kThreads * kAdjPPT)); }
hipError_t launch_adj_grad(const AdjArgs &a, const double2 *Y, double *part6, double *part4, double *out6,
                           double *out4, hipStream_t st) {
  const int nblk = adj_grad_blocks(a.N);
  hipLaunchKernelGGL(k_adj_grad_partial, dim3(nblk), dim3(kThreads), 0, st, a, part6);
  hipLaunchKernelGGL(k_adj_colsum, agrid((long)a.natu * 6), dim3(kThreads), 0, st, part6, nblk, a.natu * 6, out6);
  hipLaunchKernelGGL(k_adj_kgrad_partial, dim3(nblk), dim3(kThreads), 0, st, a, Y, part4);
  hipLaunchKernelGGL(k_adj_colsum, agrid((long)a.natu * 4), dim3(kThreads), 0, st, part4, nblk, a.natu * 4, out4);
  return hipGetLastError();
}

}  // namespace rism
