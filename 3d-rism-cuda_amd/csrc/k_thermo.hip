// k_thermo.hip -- converged-field reductions on gfx950 (one-time, after the
// solve): excess chemical potential densities (src-cuda/cal_exchem.cu:55-154),
// the PMV sum of c = h - t (cal_pmv.cu:25-55), solute-solvent energy SE
// (cal_euv.cu:35-76; total from the potentials pass's atom sums, per atom
// for the .euv file) and the cal_grad forces (cal_grad.cu:51-145).  All sums
// are FP64 with a fixed block order, so they are reproducible run to run.
#include "rism_kernels.h"

namespace rism {

template <typename R> __device__ __forceinline__ R closure_gt(R d, int clos) {
  if (clos == 0) return d >= (R)0 ? (R)1 + d : exp(d);
  return exp(d);
}

__device__ __forceinline__ double wave_sum_t(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// g from the state: g_v = closure(ts_v - u_v), stored FP64 [V][N]
template <typename R, typename S>
__global__ __launch_bounds__(kThreads) void k_gfield(IterArgs<R, S> a, double *g) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= a.N) return;
  for (int v = 0; v < a.V; ++v) {
    const long iv = (long)v * a.N + ip;
    g[iv] = (double)closure_gt<R>((R)a.ts[iv] - a.u[iv], a.closure);
  }
}

// per site: [0] SC (KH or HNC integrand), [1] HNC integrand, [2] GF, [3] c
template <typename R, typename S>
__global__ __launch_bounds__(kThreads) void k_exchem(IterArgs<R, S> a, const double *frd,
                                                     double *partials) {
  const int V = a.V;
  const long N = a.N;
  double acc[4 * kMaxV];
#pragma unroll
  for (int k = 0; k < 4 * kMaxV; ++k) acc[k] = 0.0;
  for (long ip = (long)blockIdx.x * kThreads + threadIdx.x; ip < N; ip += (long)gridDim.x * kThreads) {
    const double f = frd[ip];
#pragma unroll
    for (int v = 0; v < kMaxV; ++v) {
      if (v < V) {
        const typename C2<R>::T z = a.Z[(long)(v >> 1) * N + ip];
        const double h = (double)((v & 1) ? z.y : z.x);
        const double t = (double)a.ts[(long)v * N + ip] + a.qcf[v] * f;
        const double c = h - t;
        const double hnc = h * 0.5 * t - c;
        const double gf = -(h * 0.5 + 1.0) * c;
        const double kh = (h > 0.0) ? gf : hnc;
        acc[4 * v + 0] += (a.closure == 0) ? kh : hnc;
        acc[4 * v + 1] += hnc;
        acc[4 * v + 2] += gf;
        acc[4 * v + 3] += c;
      }
    }
  }
  __shared__ double red[kThreads / 64][4 * kMaxV];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 4 * kMaxV; ++k) {
    const double s = wave_sum_t(acc[k]);
    if (lane == 0) red[wave][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < 4 * V) {
    double s = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) s += red[w][threadIdx.x];
    partials[(long)blockIdx.x * 4 * V + threadIdx.x] = s;
  }
}

// SE totals (cal_euv.cu:35-76): rho_v sum_r g_v sum_u 4 eps_uv (s^12 - s^6)
// and rho_v q_v cc sum_r g_v sum_u q_u / r, both over r >= sigma_uv / 2.
// The atom sums per point and site (sel, seq) come from the potentials
// pass, so this is one weighted sum over the grid.  Points are walked in
// block-uniform grid strides and reduced in a fixed order.
__global__ __launch_bounds__(kThreads) void k_se_sum(long N, int V, const double *rho, const double *qv,
                                                     const double *g, const double *sel, const double *seq,
                                                     double *partials) {
  const double cc = 4.359744722207185e-18 * 0.5291771500001 * 6.02214076e23;
  double elj = 0.0, ees = 0.0;
  for (long b0 = (long)blockIdx.x * kThreads; b0 < N; b0 += (long)gridDim.x * kThreads) {
    const long ip = b0 + threadIdx.x;
    if (ip < N) {
      for (int v = 0; v < V; ++v) {
        const long iv = (long)v * N + ip;
        const double w = rho[v] * g[iv];
        elj += w * sel[iv];
        ees += w * qv[v] * cc * seq[iv];
      }
    }
  }
  __shared__ double red[kThreads / 64][2];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  elj = wave_sum_t(elj);
  ees = wave_sum_t(ees);
  if (lane == 0) { red[wave][0] = elj; red[wave][1] = ees; }
  __syncthreads();
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) s += red[w][threadIdx.x];
    partials[(long)blockIdx.x * 2 + threadIdx.x] = s;
  }
}

// cal_grad: per atom, sum_v rho_v sum_r g dU/dR terms; a workgroup owns a
// chunk of grid points and walks the atoms u0 .. u0+nb-1, writing
// per-(chunk, atom) partials [nblk][nb][6] that a second kernel sums in chunk
// order (each atom's sums do not depend on the batch it is in).
constexpr int kGradPPT = 16;

__global__ __launch_bounds__(kThreads) void k_grad_partial(int nx, int ny, int nz, double drx,
                                                           double dry, double drz, int V, int natu,
                                                           const double4 *atoms, const double2 *lj,
                                                           const double *sigh, const double *qv,
                                                           const double *rho, const double *g, int u0,
                                                           int nb, double *partials) {
  const long N = (long)nx * ny * nz;
  const double cc = 4.359744722207185e-18 * 0.5291771500001 * 6.02214076e23;
  const long cbase = (long)blockIdx.x * kThreads * kGradPPT;
  __shared__ double red[kThreads / 64][6];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int u = u0; u < u0 + nb; ++u) {
    const double4 at = atoms[u];
    double a6[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < kGradPPT; ++k) {
      const long ip = cbase + threadIdx.x + (long)k * kThreads;
      if (ip >= N) break;
      const int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
      const double dx = (x - nx / 2) * drx - at.x;
      const double dy = (y - ny / 2) * dry - at.y;
      const double dz = (z - nz / 2) * drz - at.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double ir1 = rsqrt_f64(fmax(r2, 1.0e-300));
      const double r1 = r2 * ir1, ir2 = ir1 * ir1;
      for (int v = 0; v < V; ++v) {
        if (r1 < sigh[v * natu + u]) continue;  // r < sigma_uv / 2 excluded
        const double2 p = lj[v * natu + u];
        const double gg = g[(long)v * N + ip];
        const double rs2i = p.x * ir2, rs6i = rs2i * rs2i * rs2i;
        const double flj = rho[v] * (p.y * 6.0) * rs6i * ir2 * (2.0 * rs6i - 1.0) * gg;
        const double fes = rho[v] * at.w * qv[v] * (ir2 * ir1) * cc * gg;
        a6[0] += flj * dx; a6[1] += flj * dy; a6[2] += flj * dz;
        a6[3] += fes * dx; a6[4] += fes * dy; a6[5] += fes * dz;
      }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double s = wave_sum_t(a6[c]);
      if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
      double s = 0.0;
      for (int w = 0; w < kThreads / 64; ++w) s += red[w][threadIdx.x];
      partials[((long)blockIdx.x * nb + (u - u0)) * 6 + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void k_grad_sum(const double *partials, int nblk, int nb,
                                                       double scale, double *out) {
  const long id = (long)blockIdx.x * kThreads + threadIdx.x;
  if (id >= (long)nb * 6) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partials[(long)b * nb * 6 + id];
  out[id] = s * scale;
}

// cal_euv per (atom, site), as output_euv.cu:14-27 writes it: a workgroup owns
// a chunk of points and walks the atoms, writing per-(chunk, atom) partials
// [nblk][natu][2V] (LJ for each site, then ES for each site) that
// k_euv_sum adds in chunk order.
__global__ __launch_bounds__(kThreads) void k_euv_partial(int nx, int ny, int nz, double drx, double dry,
                                                          double drz, int V, int natu, const double4 *atoms,
                                                          const double2 *lj, const double *sigh,
                                                          const double *qv, const double *rho,
                                                          const double *g, double *partials) {
  const long N = (long)nx * ny * nz;
  const double cc = 4.359744722207185e-18 * 0.5291771500001 * 6.02214076e23;
  const long cbase = (long)blockIdx.x * kThreads * kGradPPT;
  __shared__ double red[kThreads / 64][2 * kMaxV];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int u = 0; u < natu; ++u) {
    const double4 at = atoms[u];
    double acc[2 * kMaxV];
#pragma unroll
    for (int c = 0; c < 2 * kMaxV; ++c) acc[c] = 0.0;
    for (int k = 0; k < kGradPPT; ++k) {
      const long ip = cbase + threadIdx.x + (long)k * kThreads;
      if (ip >= N) break;
      const int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
      const double dx = (x - nx / 2) * drx - at.x;
      const double dy = (y - ny / 2) * dry - at.y;
      const double dz = (z - nz / 2) * drz - at.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double ir1 = rsqrt_f64(fmax(r2, 1.0e-300));
      const double r1 = r2 * ir1, ir2 = ir1 * ir1;
#pragma unroll
      for (int v = 0; v < kMaxV; ++v) {
        if (v < V && r1 >= sigh[v * natu + u]) {  // r < sigma_uv / 2 excluded (cal_euv.cu:61-64)
          const double2 p = lj[v * natu + u];
          const double w = rho[v] * g[(long)v * N + ip];
          const double s2 = p.x * ir2, s6 = s2 * s2 * s2;
          acc[v] += w * p.y * s6 * (s6 - 1.0);
          acc[kMaxV + v] += w * at.w * qv[v] * cc * ir1;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 2 * kMaxV; ++c) {
      const double s = wave_sum_t(acc[c]);
      if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < 2 * V) {
      const int c = threadIdx.x < V ? threadIdx.x : kMaxV + threadIdx.x - V;
      double s = 0.0;
      for (int w = 0; w < kThreads / 64; ++w) s += red[w][c];
      partials[((long)blockIdx.x * natu + u) * 2 * V + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

// out[LJ/ES][v][u] = dv * sum over chunks (the euv array of cal_euv.cu:7-33, x dv)
__global__ __launch_bounds__(kThreads) void k_euv_sum(const double *partials, int nblk, int natu, int V,
                                                      double dv, double *out) {
  const long id = (long)blockIdx.x * kThreads + threadIdx.x;  // (u, c) with c = kind * V + v
  if (id >= (long)natu * 2 * V) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partials[(long)b * natu * 2 * V + id];
  const int u = (int)(id / (2 * V)), c = (int)(id % (2 * V));
  out[(long)c * natu + u] = s * dv;
}

// qv = sum_v rho_v q_v g_v (cal_qv.cu:20-27)
__global__ __launch_bounds__(kThreads) void k_qv(long N, int V, const double *rho, const double *qv,
                                                 const double *g, double *out) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= N) return;
  double s = 0.0;
  for (int v = 0; v < V; ++v) s += rho[v] * qv[v] * g[(long)v * N + ip];
  out[ip] = s;
}

hipError_t launch_euv_atoms(int nx, int ny, int nz, const double *dr, int V, int natu, const double4 *atoms,
                            const double2 *lj, const double *sigh, const double *qv, const double *rho,
                            const double *g, double *partials, int nblk, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_euv_partial, dim3(nblk), dim3(kThreads), 0, st, nx, ny, nz, dr[0], dr[1], dr[2], V, natu,
                     atoms, lj, sigh, qv, rho, g, partials);
  const double dv = dr[0] * dr[1] * dr[2];
  hipLaunchKernelGGL(k_euv_sum, dim3((unsigned)((natu * 2 * V + kThreads - 1) / kThreads)), dim3(kThreads), 0, st,
                     partials, nblk, natu, V, dv, out);
  return hipGetLastError();
}

hipError_t launch_qv(long N, int V, const double *rho, const double *qv, const double *g, double *out,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_qv, dim3((unsigned)((N + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, N, V, rho, qv,
                     g, out);
  return hipGetLastError();
}

template <typename R, typename S>
hipError_t launch_gfield(const IterArgs<R, S> &a, double *g, hipStream_t st) {
  hipLaunchKernelGGL((k_gfield<R, S>), dim3((unsigned)((a.N + kThreads - 1) / kThreads)), dim3(kThreads),
                     0, st, a, g);
  return hipGetLastError();
}

template <typename R, typename S>
hipError_t launch_exchem(const IterArgs<R, S> &a, const double *frd, double *partials, int nblk,
                         hipStream_t st) {
  hipLaunchKernelGGL((k_exchem<R, S>), dim3(nblk), dim3(kThreads), 0, st, a, frd, partials);
  return hipGetLastError();
}

hipError_t launch_se_sum(long N, int V, const double *rho, const double *qv, const double *g, const double *sel,
                         const double *seq, double *partials, int nblk, hipStream_t st) {
  hipLaunchKernelGGL(k_se_sum, dim3(nblk), dim3(kThreads), 0, st, N, V, rho, qv, g, sel, seq, partials);
  return hipGetLastError();
}

hipError_t launch_grad(int nx, int ny, int nz, const double *dr, int V, int natu,
                       const double4 *atoms, const double2 *lj, const double *sigh,
                       const double *qv, const double *rho, const double *g, int u0, int nb,
                       double *partials, int nblk, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_grad_partial, dim3(nblk), dim3(kThreads), 0, st, nx, ny, nz, dr[0], dr[1], dr[2],
                     V, natu, atoms, lj, sigh, qv, rho, g, u0, nb, partials);
  const double dv = dr[0] * dr[1] * dr[2];
  hipLaunchKernelGGL(k_grad_sum, dim3((unsigned)((nb * 6 + kThreads - 1) / kThreads)), dim3(kThreads),
                     0, st, partials, nblk, nb, dv, out + (long)u0 * 6);
  return hipGetLastError();
}

template hipError_t launch_gfield<double, double>(const IterArgs<double, double> &, double *, hipStream_t);
template hipError_t launch_exchem<double, double>(const IterArgs<double, double> &, const double *,
                                                  double *, int, hipStream_t);

}  // namespace rism
