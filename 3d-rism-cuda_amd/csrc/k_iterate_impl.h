// k_iterate_impl.h -- the fused fixed-point iteration of 3D-RISM on gfx950.
//
// One iteration (src-cuda/iterate.cu:51-73 with calculate.cu:4-58 and
// anderson.cu:35-84) is six launches on one stream:
//   1. k_closure_fx   [pending Anderson update of ts] + KH/HNC closure +
//                     c_s = g - 1 - ts for all sites, two sites packed per
//                     complex value, forward FFT along x   (kh/hnc/trm1mt, timekf)
//   2. k_fft_y        forward FFT along y
//   3. k_fft_z_oz     forward FFT along z, unpack the two real sites from the
//                     (k, -k) pair, c(k) = c_s(k) - q_v cf fk(k), h_v(k) =
//                     sum_v' chi_vv'(|k|) c_v'(k), repack, inverse FFT along z
//                     (mqvfk, oz, timeirvol)
//   4. k_fft_y        inverse FFT along y
//   5. k_fx_resid     inverse FFT along x, g, R = h + 1 - g (tr), and the
//                     Anderson inner products <R,R-R_i>, <R-R_i,R-R_j> plus
//                     sum R^2 (theta30, cal_rms) as per-workgroup partials
//   6. k_anderson_solve  one workgroup: reduce partials, rms, convergence
//                     test, Picard/Anderson coefficients (leg), control block
// The control block gates every kernel so the host can enqueue iterations
// ahead without a synchronisation per step.
#pragma once
#include "rism_kernels.h"

namespace rism {

template <typename R> __device__ __forceinline__ R closure_g(R d, int clos) {
  if (clos == 0) return d >= (R)0 ? (R)1 + d : exp(d);  // kh, calculate_cuda.cu:1-12
  return exp(d);                                         // hnc, calculate_cuda.cu:14-20
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------
// 1. update + closure + forward x
// ---------------------------------------------------------------------
template <typename R, typename S, int EPT, bool UPDATE, int VC>
__global__ __launch_bounds__(kThreads) void k_closure_fx(IterArgs<R, S> a) {
  using T2 = typename C2<R>::T;
  const Ctl *ctl = a.ctl;
  const bool do_upd = UPDATE && (ctl->upd_iter == a.iter);
  const bool do_run = (ctl->run_iter == a.iter);
  if (!do_upd && !do_run) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int nx = a.nx, L = a.Lx, ls = line_stride(nx), V = a.V, P = a.P;
  const long N = a.N;
  const long base = (long)blockIdx.x * L * nx;
  const int nel = L * nx;
  const T2 *pre = (const T2 *)a.px.pre;
  int nf = 0, w = 0;
  double mix = 0.0;
  if (do_upd) {
    nf = ctl->upd_nf;
    w = ctl->upd_w;
    mix = ctl->upd_mix;
  }
  const long VN = (long)V * N;
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    const long ip = base + e;
    const int i = e / nx;
    const int x = e - i * nx;
    R cs[VC];
#pragma unroll
    for (int v = 0; v < VC; ++v) {
      if (v < V) {
        const long iv = (long)v * N + ip;
        S t = a.ts[iv];
        if (do_upd) {  // newdt / newdt0 (anderson_cuda.cu:1-23)
          const S r = a.res[iv];
          S tb = t, rb = r;
          for (int s = 0; s < nf; ++s) {
            const double th = ctl->upd_theta[s];
            tb += (S)th * (a.hist_t[s * VN + iv] - t);
            rb += (S)th * (a.hist_r[s * VN + iv] - r);
          }
          a.hist_t[w * VN + iv] = t;
          a.hist_r[w * VN + iv] = r;
          t = tb + (S)mix * rb;
          a.ts[iv] = t;
        }
        if (do_run) {
          if (a.src) {
            cs[v] = (R)a.cref[iv];
          } else {
            // closure in FP64 in both precisions: near a charged solute its
            // argument ts - u is the small difference of large terms
            const double td = (double)t;
            const double g = closure_g<double>(td - (double)a.u[iv], a.closure);
            double c = g - 1.0 - td;
            if (a.refresh) {  // new reference; the rounding remainder goes to the FP32 passes
              const float cr = (float)c;
              a.cref_out[iv] = cr;
              c -= (double)cr;
            } else if (a.delta) {
              c -= (double)a.cref[iv];
            }
            cs[v] = (R)c;
          }
        }
      } else {
        cs[v] = (R)0;
      }
    }
    if (do_run) {
      const T2 pw = pre[x];
#pragma unroll
      for (int p = 0; p < (VC + 1) / 2; ++p)
        if (p < P) lds[(p * L + i) * ls + lpad(x)] = cmul(mk<T2>(cs[2 * p], 2 * p + 1 < VC ? cs[2 * p + 1] : (R)0), pw);
    }
  }
  if (!do_run) return;
  __syncthreads();
  fft_lines<EPT, false>(lds, P * L, ls, a.px);
  const R dv = (R)a.dv;
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    const int i = e / nx;
    const int j = e - i * nx;
    const R sc = dv * (R)jsign(j, a.px.fixsign);
    for (int p = 0; p < P; ++p) a.Z[p * N + base + e] = cscale(lds[(p * L + i) * ls + lpad(j)], sc);
  }
}

// ---------------------------------------------------------------------
// 2/4. FFT along y on tiles of Ty consecutive x
// ---------------------------------------------------------------------
template <typename R, int EPT, bool INV>
__global__ __launch_bounds__(kThreads) void k_fft_y(typename C2<R>::T *Z, const Ctl *ctl, int iter,
                                                     int nx, int ny, long N, int T, AxisPlan py) {
  using T2 = typename C2<R>::T;
  if (ctl && ctl->run_iter != iter) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int ls = line_stride(ny);
  const int x0 = blockIdx.x * T;
  const long zoff = (long)blockIdx.y * nx * ny;
  T2 *Zp = Z + (long)blockIdx.z * N + zoff + x0;
  const T2 *pre = (const T2 *)py.pre;
  const int nel = T * ny;
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    const int y = e / T;
    const int i = e - y * T;
    T2 v = Zp[(long)y * nx + i];
    v = INV ? cscale(v, (decltype(v.x))jsign(y, py.fixsign)) : cmul(v, pre[y]);
    lds[i * ls + lpad(y)] = v;
  }
  __syncthreads();
  fft_lines<EPT, INV>(lds, T, ls, py);
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    const int y = e / T;
    const int i = e - y * T;
    T2 v = lds[i * ls + lpad(y)];
    v = INV ? cmulc(v, pre[y]) : cscale(v, (decltype(v.x))jsign(y, py.fixsign));
    Zp[(long)y * nx + i] = v;
  }
}

// ---------------------------------------------------------------------
// 3. FFT along z + k-space Ornstein-Zernike mixing + inverse FFT along z.
// A workgroup owns tile A = (x in [xa, xa+T), y) and its mirror tile
// B = (x in [nx-xa-T, nx-xa), ny-1-y), so every k and -k are local.
// ---------------------------------------------------------------------
template <typename R, typename S, int EPT, int VC>
__global__ __launch_bounds__(kThreads) void k_fft_z_oz(IterArgs<R, S> a) {
  using T2 = typename C2<R>::T;
  using T4 = typename R4T<R>::T;
  if (a.ctl->run_iter != a.iter) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int nx = a.nx, ny = a.ny, nz = a.nz, T = a.Tz, P = a.P, V = a.V;
  const long N = a.N, nxy = (long)nx * ny;
  const int ls = line_stride(nz);
  const int xa = blockIdx.x * T, ya = blockIdx.y;
  const int xb = nx - xa - T, yb = ny - 1 - ya;
  const T2 *pre = (const T2 *)a.pz.pre;
  const int nel = 2 * P * T * nz;
  // load both tiles of every pack; line = (p*2 + tile)*T + i
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    int r = e;
    const int i = r % T; r /= T;
    const int z = r % nz; r /= nz;
    const int tile = r & 1, p = r >> 1;
    const long off = (long)p * N + (long)z * nxy + (tile ? (long)yb * nx + xb : (long)ya * nx + xa) + i;
    lds[((p * 2 + tile) * T + i) * ls + lpad(z)] = cmul(a.Z[off], pre[z]);
  }
  __syncthreads();
  fft_lines<EPT, false>(lds, 2 * P * T, ls, a.pz);
  // k-space: pair (i, j) of tile A with (T-1-i, nz-1-j) of tile B
  const double tpx = 2.0 * M_PI / a.box[0], tpy = 2.0 * M_PI / a.box[1], tpz = 2.0 * M_PI / a.box[2];
  const double ky = tpy * ((double)ya - ny / 2.0 + 0.5);
  const int nseg = a.ntab - 1;
  for (int e = threadIdx.x; e < T * nz; e += kThreads) {
    const int j = e / T;
    const int i = e - j * T;
    const int ib = T - 1 - i, jb = nz - 1 - j;
    const int x = xa + i;
    const double kx = tpx * ((double)x - nx / 2.0 + 0.5);
    const double kz = tpz * ((double)j - nz / 2.0 + 0.5);
    const double kk = sqrt(kx * kx + ky * ky + kz * kz);
    // spline interval on the uniform table k_i = dk (i + 1) (splint.cc:4-20)
    int lo = (int)floor(kk / a.dk) - 1;
    if (lo < 0) lo = 0;
    if (lo > nseg - 1) lo = nseg - 1;
    if (lo < nseg - 1 && a.dk * (lo + 2) <= kk) ++lo;
    if (lo > 0 && a.dk * (lo + 1) > kk) --lo;
    const R tt = (R)(kk - a.dk * (lo + 1));
    const R sa = (R)jsign(j, a.pz.fixsign), sb = (R)jsign(jb, a.pz.fixsign);
    const T2 fk = a.fkh[((long)j * (ny / 2) + ya) * nx + x];
    T2 c[VC + 1];
#pragma unroll
    for (int p = 0; p < (VC + 1) / 2; ++p) {
      if (p < P) {
        const T2 za = cscale(lds[((p * 2 + 0) * T + i) * ls + lpad(j)], sa);
        const T2 zb = cscale(lds[((p * 2 + 1) * T + ib) * ls + lpad(jb)], sb);
        // C_{2p} = (za + conj zb)/2 ; C_{2p+1} = -i (za - conj zb)/2
        c[2 * p] = mk<T2>((R)0.5 * (za.x + zb.x), (R)0.5 * (za.y - zb.y));
        c[2 * p + 1] = mk<T2>((R)0.5 * (za.y + zb.y), (R)-0.5 * (za.x - zb.x));
      }
    }
    if (a.use_fk) {
#pragma unroll
      for (int v = 0; v < VC; ++v)
        if (v < V) {
          const R q = (R)a.qcf[v];
          c[v] = mk<T2>(c[v].x - q * fk.x, c[v].y - q * fk.y);  // mqvfk
        }
    }
    T2 h[VC + 1];
#pragma unroll
    for (int v = 0; v < VC + 1; ++v) {
      h[v] = mk<T2>((R)0, (R)0);
      if (v < V) {
        R hr = 0, hi = 0;
#pragma unroll
        for (int v1 = 0; v1 < VC; ++v1)
          if (v1 < V) {  // oz: h_v = sum_v' chi[v][v'] c_v'
            const T4 cf = a.chi[(long)(v * V + v1) * nseg + lo];
            const R xv = ((cf.w * tt + cf.z) * tt + cf.y) * tt + cf.x;
            hr += c[v1].x * xv;
            hi += c[v1].y * xv;
          }
        h[v] = mk<T2>(hr, hi);
      }
    }
#pragma unroll
    for (int p = 0; p < (VC + 1) / 2; ++p) {
      if (p < P) {
        const T2 h0 = h[2 * p], h1 = h[2 * p + 1];
        // H(k) = h0 + i h1 ; H(-k) = conj(h0) + i conj(h1)
        const T2 ha = mk<T2>(h0.x - h1.y, h0.y + h1.x);
        const T2 hb = mk<T2>(h0.x + h1.y, h1.x - h0.y);
        lds[((p * 2 + 0) * T + i) * ls + lpad(j)] = cscale(ha, sa);
        lds[((p * 2 + 1) * T + ib) * ls + lpad(jb)] = cscale(hb, sb);
      }
    }
  }
  __syncthreads();
  fft_lines<EPT, true>(lds, 2 * P * T, ls, a.pz);
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    int r = e;
    const int i = r % T; r /= T;
    const int z = r % nz; r /= nz;
    const int tile = r & 1, p = r >> 1;
    const long off = (long)p * N + (long)z * nxy + (tile ? (long)yb * nx + xb : (long)ya * nx + xa) + i;
    a.Z[off] = cmulc(lds[((p * 2 + tile) * T + i) * ls + lpad(z)], pre[z]);
  }
}

// ---------------------------------------------------------------------
// 5. inverse x + residual + Anderson inner products
// ---------------------------------------------------------------------
template <typename R, typename S, int EPT, int HM, int VC>
__global__ __launch_bounds__(kThreads) void k_fx_resid(IterArgs<R, S> a) {
  using T2 = typename C2<R>::T;
  constexpr int NV = 1 + HM + HM * (HM + 1) / 2;
  const Ctl *ctl = a.ctl;
  if (ctl->run_iter != a.iter) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int nx = a.nx, L = a.Lx, ls = line_stride(nx), V = a.V, P = a.P;
  const long N = a.N, VN = (long)V * N;
  const int nel = L * nx;
  const int groups = (int)((long)a.ny * a.nz / L);
  const int nf = ctl->filled < HM ? ctl->filled : HM;
  const T2 *pre = (const T2 *)a.px.pre;
  const R ivol = (R)a.ivol;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int grp = blockIdx.x; grp < groups; grp += gridDim.x) {
    const long base = (long)grp * nel;
    __syncthreads();
    for (int e = threadIdx.x; e < nel; e += kThreads) {
      const int i = e / nx;
      const int j = e - i * nx;
      const R sg = (R)jsign(j, a.px.fixsign);
      for (int p = 0; p < P; ++p) lds[(p * L + i) * ls + lpad(j)] = cscale(a.Z[p * N + base + e], sg);
    }
    __syncthreads();
    fft_lines<EPT, true>(lds, P * L, ls, a.px);
    for (int e = threadIdx.x; e < nel; e += kThreads) {
      const long ip = base + e;
      const int i = e / nx;
      const int x = e - i * nx;
      const T2 pw = pre[x];
      double hv[VC + 1];
#pragma unroll
      for (int p = 0; p < (VC + 1) / 2; ++p) {
        if (p < P) {
          const T2 z = cscale(cmulc(lds[(p * L + i) * ls + lpad(x)], pw), ivol);
          hv[2 * p] = (double)z.x;
          hv[2 * p + 1] = (double)z.y;
          if (a.delta) {
            const double2 hr = a.href[p * N + ip];
            hv[2 * p] += hr.x;
            hv[2 * p + 1] += hr.y;
          }
          if (a.store_h) a.Z[p * N + ip] = z;
        }
      }
      if (a.href_only) continue;
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        if (v < V) {
          const long iv = (long)v * N + ip;
          const double t = (double)a.ts[iv];
          const double g = closure_g<double>(t - (double)a.u[iv], a.closure);
          const S rv = (S)(hv[v] + 1.0 - g);  // tr: R = h + 1 - g
          a.res[iv] = rv;
          const double rd = (double)rv, wgt = a.rho_inv[v];
          acc[0] += rd * rd;
          double d[HM];
#pragma unroll
          for (int s = 0; s < HM; ++s) d[s] = (s < nf) ? rd - (double)a.hist_r[s * VN + iv] : 0.0;
#pragma unroll
          for (int s = 0; s < HM; ++s) acc[1 + s] += wgt * rd * d[s];
          int k = 1 + HM;
#pragma unroll
          for (int s = 0; s < HM; ++s)
#pragma unroll
            for (int s2 = s; s2 < HM; ++s2) acc[k++] += wgt * d[s] * d[s2];
        }
      }
    }
  }
  // workgroup reduction of the accumulators
  __shared__ double red[kThreads / 64][NV];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) red[wave][k] = s;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NV; k += kThreads) {
    double s = 0.0;
    for (int wv = 0; wv < kThreads / 64; ++wv) s += red[wv][k];
    a.partials[(long)blockIdx.x * NV + k] = s;
  }
}

// ---------------------------------------------------------------------
// 6. solve: reduction, rms, convergence, Anderson coefficients
// ---------------------------------------------------------------------
static __device__ void leg_solve_dev(double *A, double *x, double *b, int n) {
  // Gaussian elimination without pivoting (anderson.cu:87-112)
  for (int k = 0; k < n - 1; ++k) {
    const double p = A[k * n + k];
    for (int j = k + 1; j < n; ++j) A[k * n + j] /= p;
    const double r = b[k] /= p;
    for (int i = k + 1; i < n; ++i) {
      const double q = A[i * n + k];
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= q * A[k * n + j];
      b[i] -= q * r;
    }
  }
  x[n - 1] = b[n - 1] / A[(n - 1) * n + (n - 1)];
  for (int k = n - 2; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * x[j];
    x[k] = s;
  }
}

static __global__ __launch_bounds__(kThreads) void k_anderson_solve(Ctl *ctl, const double *partials,
                                                              int nblk, int nv, int hm,
                                                              double *rms_hist, int iter) {
  if (ctl->run_iter != iter) return;
  __shared__ double tot[1 + kHMax + kHMax * (kHMax + 1) / 2];
  __shared__ double wred[kThreads / 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = 0; k < nv; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kThreads) s += partials[(long)b * nv + k];
    s = wave_sum(s);
    if (lane == 0) wred[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int wv = 0; wv < kThreads / 64; ++wv) t += wred[wv];
      tot[k] = t;
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double rms = sqrt(tot[0] / ctl->npts);  // cal_rms.cu:18-20
  ctl->step += 1;
  ctl->total += 1;
  rms_hist[ctl->total - 1] = rms;
  ctl->rms = rms;
  if (!isfinite(rms)) {  // iterate.cu:54-57
    ctl->status = 1;
    ctl->run_iter = -1;
    ctl->upd_iter = -1;
    return;
  }
  if (rms <= ctl->tol) {  // iterate.cu:58-59
    ctl->status = 0;
    ctl->run_iter = -1;
    ctl->upd_iter = -1;
    return;
  }
  const int H = ctl->nhist;
  const int w = ctl->write;
  if (ctl->count > 0) {  // Picard warm-up, anderson.cu:43-49
    ctl->count -= 1;
    ctl->upd_nf = 0;
    ctl->upd_mix = 1.0;
  } else {
    const int nf = ctl->filled < H ? ctl->filled : H;
    double A[kHMax * kHMax], c[kHMax], th[kHMax];
    for (int s = 0; s < nf; ++s) c[s] = tot[1 + s];
    int k = 1 + hm;
    for (int s = 0; s < hm; ++s)
      for (int s2 = s; s2 < hm; ++s2) {
        if (s < nf && s2 < nf) {
          A[s * nf + s2] = tot[k];
          A[s2 * nf + s] = tot[k];
        }
        ++k;
      }
    if (nf > 0) leg_solve_dev(A, th, c, nf);
    const int recent = (w - 1 + H) % H;  // mp on the newest slot (anderson.cu:54-55)
    for (int s = 0; s < nf; ++s) ctl->upd_theta[s] = th[s] + (s == recent ? ctl->mp : 0.0);
    ctl->upd_nf = nf;
    ctl->upd_mix = ctl->m;
  }
  ctl->upd_w = w;
  ctl->write = (w + 1) % H;
  if (ctl->filled < H) ctl->filled += 1;
  ctl->upd_iter = iter + 1;
  if (ctl->step >= ctl->maxstep) {  // iterate.cu:76-79 (the update still applies)
    ctl->status = 2;
    ctl->run_iter = -1;
  } else {
    ctl->run_iter = iter + 1;
  }
}

// ---------------------------------------------------------------------
// standalone transforms for the FFT3D::execute seam
// ---------------------------------------------------------------------
template <typename R, int EPT, bool INV>
__global__ __launch_bounds__(kThreads) void k_fft_x_pure(typename C2<R>::T *Z, int nx, int L,
                                                          AxisPlan px, R scale) {
  using T2 = typename C2<R>::T;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int ls = line_stride(nx), nel = L * nx;
  const long base = (long)blockIdx.x * nel;
  const T2 *pre = (const T2 *)px.pre;
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    const int i = e / nx, x = e - i * nx;
    T2 v = Z[base + e];
    v = INV ? cscale(v, (R)jsign(x, px.fixsign)) : cmul(v, pre[x]);
    lds[i * ls + lpad(x)] = v;
  }
  __syncthreads();
  fft_lines<EPT, INV>(lds, L, ls, px);
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    const int i = e / nx, x = e - i * nx;
    T2 v = lds[i * ls + lpad(x)];
    v = INV ? cscale(cmulc(v, pre[x]), scale) : cscale(v, scale * (R)jsign(x, px.fixsign));
    Z[base + e] = v;
  }
}

template <typename R, int EPT, bool INV>
__global__ __launch_bounds__(kThreads) void k_fft_z_pure(typename C2<R>::T *Z, int nx, int ny, int nz,
                                                          int T, AxisPlan pz) {
  using T2 = typename C2<R>::T;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int ls = line_stride(nz), nel = T * nz;
  const long nxy = (long)nx * ny;
  T2 *Zp = Z + (long)blockIdx.y * nx + blockIdx.x * T;
  const T2 *pre = (const T2 *)pz.pre;
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    const int z = e / T, i = e - z * T;
    T2 v = Zp[(long)z * nxy + i];
    v = INV ? cscale(v, (decltype(v.x))jsign(z, pz.fixsign)) : cmul(v, pre[z]);
    lds[i * ls + lpad(z)] = v;
  }
  __syncthreads();
  fft_lines<EPT, INV>(lds, T, ls, pz);
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    const int z = e / T, i = e - z * T;
    T2 v = lds[i * ls + lpad(z)];
    v = INV ? cmulc(v, pre[z]) : cscale(v, (decltype(v.x))jsign(z, pz.fixsign));
    Zp[(long)z * nxy + i] = v;
  }
}

// ---------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------
#define RISM_EPT_SWITCH(ept, ...)                 \
  switch (ept) {                                  \
    case 4: { constexpr int E = 4; __VA_ARGS__; } break;   \
    case 8: { constexpr int E = 8; __VA_ARGS__; } break;   \
    case 16: { constexpr int E = 16; __VA_ARGS__; } break; \
    case 32: { constexpr int E = 32; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;         \
  }

#define RISM_VC_SWITCH(V, ...)                                  \
  switch ((V) <= 2 ? 2 : ((V) <= 4 ? 4 : 8)) {                 \
    case 2: { constexpr int VC = 2; __VA_ARGS__; } break;       \
    case 4: { constexpr int VC = 4; __VA_ARGS__; } break;       \
    default: { constexpr int VC = 8; __VA_ARGS__; } break;      \
  }

template <typename R, typename S>
hipError_t launch_closure_fx(const IterArgs<R, S> &a, const LaunchGeom &g, bool update, int,
                             hipStream_t st) {
  const dim3 grid(g.xgroups), block(kThreads);
  if (update) {
    RISM_VC_SWITCH(a.V, RISM_EPT_SWITCH(g.ept_x, hipLaunchKernelGGL((k_closure_fx<R, S, E, true, VC>), grid, block, g.lds_x, st, a)))
  } else {
    RISM_VC_SWITCH(a.V, RISM_EPT_SWITCH(g.ept_x, hipLaunchKernelGGL((k_closure_fx<R, S, E, false, VC>), grid, block, g.lds_x, st, a)))
  }
  return hipGetLastError();
}

template <typename R, typename S>
hipError_t launch_fft_y(const IterArgs<R, S> &a, const LaunchGeom &g, bool inverse, hipStream_t st) {
  const dim3 grid(a.nx / g.Ty, a.nz, a.P), block(kThreads);
  if (inverse) {
    RISM_EPT_SWITCH(g.ept_y, hipLaunchKernelGGL((k_fft_y<R, E, true>), grid, block, g.lds_y, st, a.Z,
                                                a.ctl, a.iter, a.nx, a.ny, a.N, g.Ty, a.py))
  } else {
    RISM_EPT_SWITCH(g.ept_y, hipLaunchKernelGGL((k_fft_y<R, E, false>), grid, block, g.lds_y, st, a.Z,
                                                a.ctl, a.iter, a.nx, a.ny, a.N, g.Ty, a.py))
  }
  return hipGetLastError();
}

template <typename R, typename S>
hipError_t launch_fft_z_oz(const IterArgs<R, S> &a, const LaunchGeom &g, hipStream_t st) {
  const dim3 grid(a.nx / g.Tz, a.ny / 2), block(kThreads);
  RISM_VC_SWITCH(a.V, RISM_EPT_SWITCH(g.ept_z, hipLaunchKernelGGL((k_fft_z_oz<R, S, E, VC>), grid, block, g.lds_z, st, a)))
  return hipGetLastError();
}

template <typename R, typename S>
hipError_t launch_fx_resid(const IterArgs<R, S> &a, const LaunchGeom &g, int hm, hipStream_t st) {
  const dim3 grid(g.xblocks), block(kThreads);
#define RISM_RESID(HMV)                                                                       \
  RISM_VC_SWITCH(a.V, RISM_EPT_SWITCH(g.ept_x, hipLaunchKernelGGL((k_fx_resid<R, S, E, HMV, VC>), grid, block, g.lds_x, st, a)))
  switch (hm) {
    case 2: RISM_RESID(2); break;
    case 5: RISM_RESID(5); break;
    case 10: RISM_RESID(10); break;
    default: return hipErrorInvalidValue;
  }
#undef RISM_RESID
  return hipGetLastError();
}

inline hipError_t launch_anderson_solve_impl(Ctl *ctl, const double *partials, int nblk, int nv, int hm,
                                 double *rms_hist, int iter, hipStream_t st) {
  hipLaunchKernelGGL(k_anderson_solve, dim3(1), dim3(kThreads), 0, st, ctl, partials, nblk, nv, hm,
                     rms_hist, iter);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_fft3d(typename C2<R>::T *data, const AxisPlan &px, const AxisPlan &py,
                        const AxisPlan &pz, const LaunchGeom &g, int nx, int ny, int nz, double dv,
                        double ivol, int dir, hipStream_t st) {
  const long N = (long)nx * ny * nz;
  const int L = g.Lx;
  const dim3 block(kThreads);
  const dim3 gx((unsigned)((long)ny * nz / L)), gy(nx / g.Ty, nz, 1), gz(nx / g.Tz, ny);
  const size_t sz = sizeof(typename C2<R>::T);
  const size_t ldsx = (size_t)L * line_stride(nx) * sz, ldsy = (size_t)g.Ty * line_stride(ny) * sz,
               ldsz = (size_t)g.Tz * line_stride(nz) * sz;
  // pure z tiles hold T lines: EPT class from T*nz
  const int ez = (g.Tz * nz <= 4 * kThreads) ? 4 : (g.Tz * nz <= 8 * kThreads) ? 8 : (g.Tz * nz <= 16 * kThreads) ? 16 : 32;
  if (dir < 0) {
    RISM_EPT_SWITCH(g.ept_x, hipLaunchKernelGGL((k_fft_x_pure<R, E, false>), gx, block, ldsx, st, data, nx, L, px, (R)dv))
    RISM_EPT_SWITCH(g.ept_y, hipLaunchKernelGGL((k_fft_y<R, E, false>), gy, block, ldsy, st, data, (const Ctl *)nullptr, 0, nx, ny, N, g.Ty, py))
    RISM_EPT_SWITCH(ez, hipLaunchKernelGGL((k_fft_z_pure<R, E, false>), gz, block, ldsz, st, data, nx, ny, nz, g.Tz, pz))
  } else {
    RISM_EPT_SWITCH(ez, hipLaunchKernelGGL((k_fft_z_pure<R, E, true>), gz, block, ldsz, st, data, nx, ny, nz, g.Tz, pz))
    RISM_EPT_SWITCH(g.ept_y, hipLaunchKernelGGL((k_fft_y<R, E, true>), gy, block, ldsy, st, data, (const Ctl *)nullptr, 0, nx, ny, N, g.Ty, py))
    RISM_EPT_SWITCH(g.ept_x, hipLaunchKernelGGL((k_fft_x_pure<R, E, true>), gx, block, ldsx, st, data, nx, L, px, (R)ivol))
  }
  return hipGetLastError();
}

}  // namespace rism
