// k_potential.hip -- solute-solvent grid potentials on gfx950.
//
// Restates src-cuda/cal_LJ.cu:47-72 and cal_Coulomb.cu:100-155 as one pass
// over the grid (every atom for every point, VALU-bound):
//   ulj_v = 1/(N_A k_B T) * sum_u 4 eps_uv (s^12 - s^6),  s = sigma_uv / max(r, 0.01)
//   de    = ubeta * sum_{r >= 1e-5} q_u / r
//   fr    = ubeta * [sum_{r >= 1e-5} q_u (1 - exp(-r)) / r + sum_{r < 1e-5} q_u]
//   sr    = de - fr, summed directly as the short-range remainder
// and, per site, the two sums the solute-solvent energy SE needs
// (cal_euv.cu:35-76): 4 eps_uv (s^12 - s^6) and q_u / r over the atoms with
// r >= sigma_uv / 2 (exact r, no distance floor).  SE is then one weighted
// sum over the converged g instead of a second pass over every atom-point
// pair.
// Atom records are staged through LDS in chunks.
// exp(-r) is used only for r < 38 Angstrom: beyond that 1 - exp(-r) rounds
// to 1 in FP64, so fr is unchanged and the dropped sr terms are below 1e-16
// of q/r.
//
// fk (cal_Coulomb.cu:125-140) = ubeta 4 pi / (k^2 (k^2 + 1)) sum_u q_u
// exp(-i k.R_u) uses the separability exp(-i k.R) = ex(kx) ey(ky) ez(kz):
// per-atom 1D phase tables replace the per-pair sincos.  Only the half grid
// y < ny/2 is stored; fk(-k) = conj(fk(k)).
#include "rism_kernels.h"

namespace rism {

// One grid point per thread; 1/r comes from rsqrt_f64 and gives r = r^2/r
// and 1/r^2 without a division or a square root.  The atoms are staged
// through LDS in chunks (position and charge, then (sigma_uv^2, 4 eps_uv,
// sigma_uv / 2) for VC sites; beyond V a padding record that contributes
// nothing), so the inner loop reads broadcast LDS values instead of waiting
// on a scalar load per atom and site.
//
// The SE sums over the atoms with r >= sigma_uv / 2 are the full sums the
// potential already forms (the LJ terms, and the Coulomb sum de) minus the
// terms of the excluded atoms.  Those are gathered in a branch taken only
// where a point lies within xr_u of atom u, xr_u = max(max_v sigma_uv / 2,
// 0.01 A) (record (u, site 0), fourth component): beyond xr_u every term of
// the full sums is the SE's own (exact r, no floor, r >= 1e-5), so the
// exclusion costs one compare per pair instead of a compare and two 64-bit
// selects per pair and site.  Points with no atom within xr_u get the same
// SE sums, added in the same order, as a per-pair selection would give.
template <int VC>
__global__ __launch_bounds__(kThreads) void k_potentials(PotArgs a) {
  __shared__ double4 s_at[kAtomChunk];
  __shared__ double4 s_lj[kAtomChunk * VC];
  const long ip0 = (long)blockIdx.x * kThreads + threadIdx.x;
  const bool live = ip0 < a.N;
  const long ip = live ? ip0 : a.N - 1;  // every thread stages atoms; only live ones store
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int x = (int)(ip % nx);
  const int y = (int)((ip / nx) % ny);
  const int z = (int)(ip / ((long)nx * ny));
  const double rx = (x - nx / 2) * a.dr[0];
  const double ry = (y - ny / 2) * a.dr[1];
  const double rz = (z - nz / 2) * a.dr[2];
  const double cut2 = 1.0e-2 * 1.0e-2;  // LJ distance floor 0.01 A (cal_LJ.cu:63)
  double lj[VC], xl[VC], xq[VC];         // xl, xq: full-sum terms the SE excludes
#pragma unroll
  for (int v = 0; v < VC; ++v) lj[v] = xl[v] = xq[v] = 0.0;
  double e = 0.0, f = 0.0, s = 0.0;
  const int V = a.V, nu = a.natu;
  for (int c0 = 0; c0 < nu; c0 += kAtomChunk) {
    const int nc = min(kAtomChunk, nu - c0);
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += kThreads) s_at[i] = a.atoms[c0 + i];
    for (int i = threadIdx.x; i < nc * VC; i += kThreads) {
      const int u = i / VC, v = i - u * VC;
      double xr = 0.0;
      if (v == 0) {  // the atom's exclusion radius rides on its first site record
        xr = 1.0e-2;
        for (int w = 0; w < V; ++w) xr = fmax(xr, a.sigh[w * nu + c0 + u]);
      }
      if (v < V) {
        const double2 p = a.lj[v * nu + c0 + u];
        s_lj[i] = make_double4(p.x, p.y, a.sigh[v * nu + c0 + u], xr);
      } else {  // padding site: r >= 1e300 never holds
        s_lj[i] = make_double4(0.0, 0.0, 1.0e300, xr);
      }
    }
    __syncthreads();
    // Per atom: distance, the LJ term per site (tl, distance floored at
    // 0.01 A), then the Coulomb terms behind their branch on r (an exp(-r)
    // only below 38 A); every sum adds the atoms in order.
    for (int k = 0; k < nc; ++k) {
      const double4 at = s_at[k];
      const double dx = rx - at.x, dy = ry - at.y, dz = rz - at.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double rinv = rsqrt_f64(fmax(r2, 1.0e-300));
      const double ra = r2 * rinv;
      const double ir2 = rinv * rinv;
      double tl[VC];
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        const double4 p = s_lj[k * VC + v];
        const double s2 = p.x * ir2;
        const double s6 = s2 * s2 * s2;
        tl[v] = p.y * s6 * (s6 - 1.0);
      }
      if (r2 < cut2) {  // within 0.01 A of the atom (rare): the potential's floor
#pragma unroll
        for (int v = 0; v < VC; ++v) {
          const double4 p = s_lj[k * VC + v];
          const double s2 = p.x * (1.0 / cut2);
          const double s6 = s2 * s2 * s2;
          tl[v] = p.y * s6 * (s6 - 1.0);
        }
      }
#pragma unroll
      for (int v = 0; v < VC; ++v) lj[v] += tl[v];
      const double qa = at.w;
      const double qr = qa * rinv;
      if (ra >= 1.0e-5) {
        e += qr;
        if (ra < 38.0) {
          const double ex = exp(-ra);
          f += qr * (1.0 - ex);
          s += qr * ex;
        } else {
          f += qr;
        }
      } else {
        f += qa;
        s -= qa;
      }
      if (ra < s_lj[k * VC].w) {  // within the atom's exclusion radius (rare)
#pragma unroll
        for (int v = 0; v < VC; ++v) {
          const double4 p = s_lj[k * VC + v];
          const bool in = ra >= p.z;  // the SE keeps this pair (exact r)
          const double s2 = p.x * ir2;
          const double s6 = s2 * s2 * s2;
          const double te = p.y * s6 * (s6 - 1.0);
          xl[v] += tl[v] - (in ? te : 0.0);
          xq[v] += (ra >= 1.0e-5 ? qr : 0.0) - (in ? qr : 0.0);
        }
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int v = 0; v < VC; ++v)
    if (v < V) {
      a.ulj[(long)v * a.N + ip] = lj[v] * a.ikbt;
      a.sel[(long)v * a.N + ip] = lj[v] - xl[v];
      a.seq[(long)v * a.N + ip] = e - xq[v];
    }
  a.de[ip] = e * a.ubeta;
  a.fr[ip] = f * a.ubeta;
  a.sr[ip] = s * a.ubeta;
}

// phase[u][0..nx) = exp(-i kx(j) X_u), then [nx..nx+ny) for y, then z
__global__ __launch_bounds__(kThreads) void k_phase(int nx, int ny, int nz, int natu,
                                                    const double4 *atoms, double bx, double by,
                                                    double bz, double2 *phase) {
  const int nt = nx + ny + nz;
  const long id = (long)blockIdx.x * kThreads + threadIdx.x;
  if (id >= (long)natu * nt) return;
  const int u = (int)(id / nt);
  const int j = (int)(id - (long)u * nt);
  double k, r;
  if (j < nx) {
    k = 2.0 * M_PI * ((double)j - nx / 2.0 + 0.5) / bx;
    r = atoms[u].x;
  } else if (j < nx + ny) {
    k = 2.0 * M_PI * ((double)(j - nx) - ny / 2.0 + 0.5) / by;
    r = atoms[u].y;
  } else {
    k = 2.0 * M_PI * ((double)(j - nx - ny) - nz / 2.0 + 0.5) / bz;
    r = atoms[u].z;
  }
  double sn, cs;
  sincos(k * r, &sn, &cs);
  phase[id] = make_double2(cs, -sn);
}

// FKY rows y = yb*FKY .. yb*FKY+FKY-1 per thread at one (x, z): the x phase
// of an atom (one per lane) is loaded once for FKY points, and the y and z
// phases are the same across a wave of consecutive x.
template <int FKY>
__global__ __launch_bounds__(kThreads) void k_fk_half(int nx, int ny, int nz, int natu,
                                                      const double4 *atoms, const double2 *phase,
                                                      double bx, double by, double bz, double ubeta,
                                                      double2 *fkh) {
  const int nyh = ny / 2, nyb = nyh / FKY;
  const long id = (long)blockIdx.x * kThreads + threadIdx.x;
  if (id >= (long)nx * nyb * nz) return;
  const int x = (int)(id % nx);
  const int yb = (int)((id / nx) % nyb);
  const int z = (int)(id / ((long)nx * nyb));
  const int nt = nx + ny + nz;
  double re[FKY], im[FKY];
#pragma unroll
  for (int j = 0; j < FKY; ++j) re[j] = im[j] = 0.0;
  for (int u = 0; u < natu; ++u) {
    const double2 px = phase[(long)u * nt + x];
    const double2 pz = phase[(long)u * nt + nx + ny + z];
    const double q = atoms[u].w;
#pragma unroll
    for (int j = 0; j < FKY; ++j) {
      const double2 py = phase[(long)u * nt + nx + yb * FKY + j];
      const double ar = px.x * py.x - px.y * py.y, ai = px.x * py.y + px.y * py.x;
      const double br = ar * pz.x - ai * pz.y, bi = ar * pz.y + ai * pz.x;
      re[j] += q * br;
      im[j] += q * bi;
    }
  }
  const double kx = 2.0 * M_PI * ((double)x - nx / 2.0 + 0.5) / bx;
  const double kz = 2.0 * M_PI * ((double)z - nz / 2.0 + 0.5) / bz;
#pragma unroll
  for (int j = 0; j < FKY; ++j) {
    const int y = yb * FKY + j;
    const double ky = 2.0 * M_PI * ((double)y - ny / 2.0 + 0.5) / by;
    const double rk2 = kx * kx + ky * ky + kz * kz;
    const double sc = 4.0 * M_PI / (rk2 * (rk2 + 1.0)) * ubeta;
    fkh[((long)z * nyh + y) * nx + x] = make_double2(re[j] * sc, im[j] * sc);
  }
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_combine_u(const double *ulj, const double *sr, double q0,
                                                        double q1, double q2, double q3, double q4,
                                                        double q5, double q6, double q7, int V,
                                                        long N, R *u) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= N) return;
  const double q[kMaxV] = {q0, q1, q2, q3, q4, q5, q6, q7};
  const double s = sr[ip];
#pragma unroll
  for (int v = 0; v < kMaxV; ++v)
    if (v < V) u[(long)v * N + ip] = (R)(ulj[(long)v * N + ip] + q[v] * s);
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_convert(const double *src, R *dst, long n) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) dst[i] = (R)src[i];
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_convert_c(const double2 *src, typename C2<R>::T *dst,
                                                        long n) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    typename C2<R>::T v;
    v.x = (R)src[i].x;
    v.y = (R)src[i].y;
    dst[i] = v;
  }
}

template <typename S>
__global__ __launch_bounds__(kThreads) void k_axpy_fr(S *x, const double *fr, double c0, double c1,
                                                      double c2, double c3, double c4, double c5,
                                                      double c6, double c7, int V, long N,
                                                      int nslots) {
  const long ip = (long)blockIdx.x * kThreads + threadIdx.x;
  if (ip >= N) return;
  const double c[kMaxV] = {c0, c1, c2, c3, c4, c5, c6, c7};
  const double f = fr[ip];
  for (int s = 0; s < nslots; ++s)
#pragma unroll
    for (int v = 0; v < kMaxV; ++v)
      if (v < V) x[((long)s * V + v) * N + ip] += (S)(c[v] * f);
}

static inline dim3 grid1(long n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

hipError_t launch_potentials(const PotArgs &a, hipStream_t st) {
  if (a.V <= 2)
    hipLaunchKernelGGL(k_potentials<2>, grid1(a.N), dim3(kThreads), 0, st, a);
  else if (a.V <= 4)
    hipLaunchKernelGGL(k_potentials<4>, grid1(a.N), dim3(kThreads), 0, st, a);
  else
    hipLaunchKernelGGL(k_potentials<8>, grid1(a.N), dim3(kThreads), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_fk_half(int nx, int ny, int nz, int natu, const double4 *atoms, const double *box,
                          double ubeta, double2 *phase, double2 *fkh, hipStream_t st) {
  const long nt = (long)natu * (nx + ny + nz);
  if (nt > 0)
    hipLaunchKernelGGL(k_phase, grid1(nt), dim3(kThreads), 0, st, nx, ny, nz, natu, atoms, box[0],
                       box[1], box[2], phase);
  const int nyh = ny / 2;
  if (nyh % 4 == 0)
    hipLaunchKernelGGL(k_fk_half<4>, grid1((long)nx * (nyh / 4) * nz), dim3(kThreads), 0, st, nx, ny, nz, natu,
                       atoms, phase, box[0], box[1], box[2], ubeta, fkh);
  else if (nyh % 2 == 0)
    hipLaunchKernelGGL(k_fk_half<2>, grid1((long)nx * (nyh / 2) * nz), dim3(kThreads), 0, st, nx, ny, nz, natu,
                       atoms, phase, box[0], box[1], box[2], ubeta, fkh);
  else
    hipLaunchKernelGGL(k_fk_half<1>, grid1((long)nx * nyh * nz), dim3(kThreads), 0, st, nx, ny, nz, natu,
                       atoms, phase, box[0], box[1], box[2], ubeta, fkh);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_combine_u(const double *ulj, const double *sr, const double *q, int V, long N, R *u,
                            hipStream_t st) {
  double c[kMaxV] = {0};
  for (int v = 0; v < V && v < kMaxV; ++v) c[v] = q[v];
  hipLaunchKernelGGL((k_combine_u<R>), grid1(N), dim3(kThreads), 0, st, ulj, sr, c[0], c[1], c[2], c[3],
                     c[4], c[5], c[6], c[7], V, N, u);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_convert(const double *src, R *dst, long n, hipStream_t st) {
  hipLaunchKernelGGL((k_convert<R>), grid1(n), dim3(kThreads), 0, st, src, dst, n);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_convert_c(const double2 *src, typename C2<R>::T *dst, long n, hipStream_t st) {
  hipLaunchKernelGGL((k_convert_c<R>), grid1(n), dim3(kThreads), 0, st, src, dst, n);
  return hipGetLastError();
}

template <typename S>
hipError_t launch_axpy_fr(S *x, const double *fr, const double *coef, int V, long N, int nslots,
                          hipStream_t st) {
  double c[kMaxV] = {0};
  for (int v = 0; v < V && v < kMaxV; ++v) c[v] = coef[v];
  hipLaunchKernelGGL((k_axpy_fr<S>), grid1(N), dim3(kThreads), 0, st, x, fr, c[0], c[1], c[2], c[3], c[4],
                     c[5], c[6], c[7], V, N, nslots);
  return hipGetLastError();
}

template hipError_t launch_combine_u<double>(const double *, const double *, const double *, int, long,
                                             double *, hipStream_t);
template hipError_t launch_combine_u<float>(const double *, const double *, const double *, int, long,
                                            float *, hipStream_t);
template hipError_t launch_convert<double>(const double *, double *, long, hipStream_t);
template hipError_t launch_convert<float>(const double *, float *, long, hipStream_t);
template hipError_t launch_convert_c<double>(const double2 *, double2 *, long, hipStream_t);
template hipError_t launch_convert_c<float>(const double2 *, float2 *, long, hipStream_t);
template hipError_t launch_axpy_fr<double>(double *, const double *, const double *, int, long, int,
                                           hipStream_t);
template hipError_t launch_axpy_fr<float>(float *, const double *, const double *, int, long, int,
                                          hipStream_t);

__global__ __launch_bounds__(kThreads) void k_widen(const float *src, double *dst, long n) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) dst[i] = (double)src[i];
}

hipError_t launch_widen(const float *src, double *dst, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_widen, grid1(n), dim3(kThreads), 0, st, src, dst, n);
  return hipGetLastError();
}

}  // namespace rism
