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

// Element access relative to a wave-uniform base with a 32-bit lane offset:
// the compiler keeps the base in scalar registers and issues the
// base+offset addressing form (one VGPR per offset instead of a 64-bit
// address per load).  e * sizeof(T) must stay below 4 GiB.
template <typename T> __device__ __forceinline__ const T &lane_at(const T *base, unsigned e) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + e * (unsigned)sizeof(T));
}
template <typename T> __device__ __forceinline__ T &lane_at(T *base, unsigned e) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + e * (unsigned)sizeof(T));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------
// 1. update + closure + forward x.  The points of a thread are processed
// in batches of CH; every global load of a batch is issued (at a clamped
// index, so without branches) before the first use, and only the stores
// are guarded by the true index.
// ---------------------------------------------------------------------
#ifndef RISM_CLOS_CH
#define RISM_CLOS_CH 2
#endif
// The mixed path's vectorised closure pass takes its line groups from the
// last to the first.  The residual pass before it writes the current
// residual in group order, so the last groups' residuals are still in the
// 256 MB Infinity Cache when the closure pass starts on them, and the
// spectrum it writes last (the first groups) is what the forward y pass
// reads first.  Each group's work is the same in either order.  (The FP64
// closure pass measured slower in that order, and keeps the forward one.)
__device__ __forceinline__ long closure_group() { return (long)(gridDim.x - 1 - blockIdx.x); }
template <typename R, typename S, int EPT, int VC, int HM, bool UPD>
__device__ __forceinline__ void closure_points(const IterArgs<R, S> &a, typename C2<R>::T *lds, bool do_run,
                                               const double *th, const long *off, double mix, int w, int nf) {
  const long so = (long)blockIdx.y * a.astride;  // this launch's solute (batched launches)
  using T2 = typename C2<R>::T;
  using Q = typename HistT<R>::T;
  constexpr int CH = EPT < RISM_CLOS_CH ? EPT : RISM_CLOS_CH;
  const int nx = a.nx, L = a.Lc, ls = line_stride(nx), V = a.V;
  const long N = a.N, VN = (long)V * N;
  const long base = (long)blockIdx.x * L * nx;
  const int nel = L * nx;
  const T2 *__restrict__ pre = (const T2 *)a.px.pre;
  S *__restrict__ ts = soff(a.ts, so);
  const Q *__restrict__ res = soff(a.res, so);
  Q *__restrict__ ht = soff(a.hist_t, so);
  Q *__restrict__ hr = soff(a.hist_r, so);
  const R *__restrict__ uu = soff(a.u, so);
#pragma unroll
  for (int k0 = 0; k0 < EPT; k0 += CH) {
    double t[CH][VC], u[CH][VC];
    Q r[CH][VC], hti[CH][VC][UPD ? HM : 1], hri[CH][VC][UPD ? HM : 1];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const unsigned e = (unsigned)min((int)threadIdx.x + (k0 + c) * kThreads, nel - 1);
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        if (v < V) {
          const long vb = (long)v * N + base;
          t[c][v] = (double)lane_at(ts + vb, e);
          u[c][v] = (double)lane_at(uu + vb, e);
          if (UPD) {
            r[c][v] = lane_at(res + vb, e);
#pragma unroll
            for (int s = 0; s < (UPD ? HM : 0); ++s) {
              hti[c][v][s] = lane_at(ht + off[s] + vb, e);
              hri[c][v][s] = lane_at(hr + off[s] + vb, e);
            }
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const unsigned e = threadIdx.x + (k0 + c) * kThreads;
      const bool valid = e < (unsigned)nel;
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        if (v < V) {
          const long vb = (long)v * N + base;
          if (UPD) {  // newdt / newdt0 (anderson_cuda.cu:1-23)
            const double rv = (double)r[c][v];
            double tb = t[c][v], rb = rv;
#pragma unroll
            for (int s = 0; s < (UPD ? HM : 0); ++s) {
              // slots beyond the filled ones: valid loads, values discarded
              const double hts = s < nf ? (double)hti[c][v][UPD ? s : 0] : 0.0;
              const double hrs = s < nf ? (double)hri[c][v][UPD ? s : 0] : rv;
              tb += th[s] * (hts - t[c][v]);
              rb += th[s] * (hrs - rv);
            }
            const double tn = tb + mix * rb;
            if (valid) {
              lane_at(ht + (long)w * VN + vb, e) = (Q)t[c][v];
              lane_at(hr + (long)w * VN + vb, e) = r[c][v];
              lane_at(ts + vb, e) = (S)tn;
            }
            t[c][v] = tn;
          }
        }
      }
      if (do_run && valid) {
        const int i = e / nx, x = e - i * nx;
        R cs[VC + 1];
#pragma unroll
        for (int v = 0; v < VC; ++v) {
          cs[v] = (R)0;
          if (v < V) cs[v] = (R)(closure_g<double>(t[c][v] - u[c][v], a.closure) - 1.0 - t[c][v]);
        }
        cs[VC] = (R)0;
        const T2 pw = pre[x];
#pragma unroll
        for (int p = 0; p < (VC + 1) / 2; ++p)
          if (2 * p < V) lds[(p * L + i) * ls + lpad(x)] = cmul(mk<T2>(cs[2 * p], cs[2 * p + 1]), pw);
      }
    }
    // keep the next batch's loads out of this one (register pressure)
    __builtin_amdgcn_sched_barrier(0);
  }
}

// the FP64 path's first pass (the mixed path's is k_mx_closure_fx)
template <typename R, typename S, int EPT, bool UPDATE, int VC, int HM>
__global__ __launch_bounds__(kThreads) void k_closure_fx(IterArgs<R, S> a) {
  const long so = (long)blockIdx.y * a.astride;  // this launch's solute (batched launches)
  using T2 = typename C2<R>::T;
  const Ctl *ctl = soff(a.ctl, so);
  const int iter = ctl->cur;
  const bool do_upd = UPDATE && (ctl->upd_iter == iter);
  const bool do_run = (ctl->run_iter == iter);
  if (!do_upd && !do_run) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int nx = a.nx, L = a.Lc, ls = line_stride(nx), P = a.P;
  const long N = a.N, VN = (long)a.V * N;
  const long base = (long)blockIdx.x * L * nx;
  const int nel = L * nx;
  // update coefficients by slot; slots beyond the filled ones get a zero
  // coefficient and a valid address
  int w = 0, nf = 0;
  double mix = 0.0;
  double th[HM];
  long off[HM];
#pragma unroll
  for (int k = 0; k < HM; ++k) {
    th[k] = 0.0;
    off[k] = 0;
  }
  if (do_upd) {
    nf = ctl->upd_nf;
    w = ctl->upd_w;
    mix = ctl->upd_mix;
    const int nh = ctl->nhist;
#pragma unroll
    for (int k = 0; k < HM; ++k) {
      const int slot = k < nh ? k : 0;
      th[k] = (k < nf) ? ctl->upd_theta[slot] : 0.0;
      off[k] = (long)slot * VN;
    }
  }
  if (UPDATE && do_upd)
    closure_points<R, S, EPT, VC, HM, true>(a, lds, do_run, th, off, mix, w, nf);
  else
    closure_points<R, S, EPT, VC, HM, false>(a, lds, do_run, th, off, mix, w, nf);
  if (!do_run) return;
  T2 *ltw = lds + P * L * ls;
  stage_tables(ltw, a.px);
  __syncthreads();
  // radix-4 plan: at about 1024 elements per workgroup a radix-16 stage has
  // work for one wave in four; four radix-4 stages keep every wave busy
  // (measured 0-1% faster, and no radix-8/16 stage is compiled in)
  fft_lines<EPT, false, 5>(lds, P * L, ls, a.pxr, ltw);
  const R dv = (R)a.dv;
  T2 *__restrict__ Z = soff(a.Z, so);
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int f = threadIdx.x + k * kThreads;  // (pack, element) flattened
    if (f < P * nel) {
      const int p = f / nel, e = f - p * nel;
      const int i = e / nx, j = e - i * nx;
      const R sc = dv * (R)jsign(j, a.px.fixsign);
      Z[p * N + base + e] = cscale(lds[(p * L + i) * ls + lpad(j)], sc);
    }
  }
}

// ---------------------------------------------------------------------
// 2/4. FFT along y on tiles of T consecutive x (T x 8 bytes = one 128-byte
// line for FP32 at T = 16).  One tile per workgroup; the twiddle and phase
// tables are staged in LDS behind the tile.
// ---------------------------------------------------------------------
template <typename R, int EPT, bool INV, int NL = 0>
__global__ __launch_bounds__(kThreads) void k_fft_y(typename C2<R>::T *Z, const Ctl *ctl, int nx,
                                                     int ny_rt, int nz, long N, int T_rt, int P, AxisPlan py,
                                                     int gate, long astride, const int *rlist) {
  // NL > 0: compile-time line length (power of two) with T = EPT * kThreads / NL
  using T2 = typename C2<R>::T;
  const int sid = batch_solute(gate, rlist);  // this launch's solute (batched launches)
  if (sid < 0) return;
  const long so = (long)sid * astride;
  Z = soff(Z, so);
  if (ctl) ctl = soff(ctl, so);
  if (ctl && !gate_open(ctl, gate)) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int ny = NL > 0 ? NL : ny_rt;
  const int T = NL > 0 ? EPT * kThreads / (NL > 0 ? NL : 1) : T_rt;
  const int ls = line_stride(ny);
  const int nxt = nx / T;
  const int nel = T * ny;
  const int t = blockIdx.x, xt = t % nxt, r = t / nxt;
  T2 *__restrict__ Zp = Z + (long)(r / nz) * N + (long)(r % nz) * nx * ny + xt * T;
  T2 buf[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {  // clamped index: one straight-line batch of loads
    const int e = min((int)threadIdx.x + k * kThreads, nel - 1);
    const int y = e / T, i = e - y * T;
    buf[k] = Zp[(long)y * nx + i];
  }
  T2 *ltw = lds + T * ls, *lpre = ltw + ny;
  stage_tables(ltw, py);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = threadIdx.x + k * kThreads;
    if (e < nel) {
      const int y = e / T, i = e - y * T;
      lds[i * ls + lpad(y)] = INV ? cscale(buf[k], (decltype(buf[k].x))jsign(y, py.fixsign)) : cmul(buf[k], lpre[y]);
    }
  }
  __syncthreads();
  if constexpr (NL > 0)
    fft_lines_ct<NL, EPT, INV>(lds, (const T2 *)ltw);
  else
    fft_lines<EPT, INV>(lds, T, ls, py, ltw);
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = threadIdx.x + k * kThreads;
    if (e < nel) {
      const int y = e / T, i = e - y * T;
      T2 v = lds[i * ls + lpad(y)];
      v = INV ? cmulc(v, lpre[y]) : cscale(v, (decltype(v.x))jsign(y, py.fixsign));
      Zp[(long)y * nx + i] = v;
    }
  }
}

// ---------------------------------------------------------------------
// 3. FFT along z + k-space Ornstein-Zernike mixing + inverse FFT along z.
// A tile pair is A = (x in [xa, xa+T), y) and its mirror
// B = (x in [nx-xa-T, nx-xa), ny-1-y), so every k and -k are local.
// Tiles narrower than a 128-byte line share lines with their x neighbours:
// workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2), so the tile order is remapped to give each XCD a contiguous run of
// tiles, and neighbours meet in the same L2 instead of both reading HBM.
// ---------------------------------------------------------------------
#ifndef RISM_OZ_GROUP
#define RISM_OZ_GROUP 2
#endif
constexpr int kXcds = 8;
template <typename R, typename S, int EPT, int VC>
__global__ __launch_bounds__(kThreads) void k_fft_z_oz(IterArgs<R, S> a) {
  const int sid = batch_solute(a.gate, a.rlist);  // this launch's solute (batched launches)
  if (sid < 0) return;
  const long so = (long)sid * a.astride;
  using T2 = typename C2<R>::T;
  using T4 = typename R4T<R>::T;
  if (!gate_open(soff(a.ctl, so), a.gate)) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int nx = a.nx, ny = a.ny, V = a.V;
  const int nz = a.nz;
  const int T = a.Tz;
  const int P = a.P;
  const long N = a.N, nxy = (long)nx * ny;
  const int ls = line_stride(nz);
  const int nxt = nx / T, ntile = nxt * (ny / 2);
  const int b = blockIdx.x;
  const int t = (ntile % kXcds) ? b : (b % kXcds) * (ntile / kXcds) + b / kXcds;
  const int xa = (t % nxt) * T, ya = t / nxt;
  const int xb = nx - xa - T, yb = ny - 1 - ya;
  T2 *__restrict__ Z = soff(a.Z, so);
  const int nel = 2 * P * T * nz;
  // element e of the tile pair: line (p*2 + tile)*T + i, position z
  auto gofs = [&](int e) -> long {
    const int i = e % T;
    const int r = e / T;
    const int z = r % nz, q = r / nz;
    const int tile = q & 1, p = q >> 1;
    return (long)p * N + (long)z * nxy + (tile ? (long)yb * nx + xb : (long)ya * nx + xa) + i;
  };
  T2 buf[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) buf[k] = Z[gofs(min((int)threadIdx.x + k * kThreads, nel - 1))];
  T2 *ltw = lds + 2 * P * T * ls, *lpre = ltw + nz;
  stage_tables(ltw, a.pz);
  __syncthreads();
  {
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int e = tid + k * kThreads;
      if (e < nel) {
        const int i = e % T, z = (e / T) % nz, line = (e / T / nz) * T + i;
        lds[line * ls + lpad(z)] = cmul(buf[k], lpre[z]);
      }
    }
    __syncthreads();
    // Each phase below takes its own opaque copy of the lane index: the
    // forward and inverse transforms (and the first and last LDS passes)
    // share their index arithmetic, and the compiler would otherwise keep
    // it live across the k-space phase, at the cost of occupancy.
    int tidf = tid;
    asm volatile("" : "+v"(tidf));
    fft_lines<EPT, false>(lds, 2 * P * T, ls, a.pz, ltw, tidf);
    // k-space: pair (i, j) of tile A with (T-1-i, nz-1-j) of tile B.  |k|
    // and the spline interval are evaluated in FP64 in both precisions.
    const double tpx = 2.0 * M_PI / a.box[0], tpz = 2.0 * M_PI / a.box[2];
    const double ky = 2.0 * M_PI / a.box[1] * ((double)ya - ny / 2.0 + 0.5);
    const double dk = a.dk, idk = 1.0 / a.dk;
    const double hx = 0.5 - nx / 2, hz = 0.5 - nz / 2;
    const int nseg = a.ntab - 1;
    const T4 *__restrict__ chi = a.chi;
    const T2 *__restrict__ fkh = soff(a.fkh, so);
    const int npair = T * nz;
    int tido = tid;
    asm volatile("" : "+v"(tido));
#pragma unroll
    for (int k = 0; k < (EPT + 1) / 2; ++k) {
      // clamped index: loads and arithmetic are unconditional, the LDS stores
      // below are guarded by the true index
      const int e0 = tido + k * kThreads;
      const int e = min(e0, npair - 1);
      {
        const int j = e / T;
        const int i = e - j * T;
        const int ib = T - 1 - i, jb = nz - 1 - j;
        const int x = xa + i;
        const double kx = tpx * ((double)x + hx);
        const double kz = tpz * ((double)j + hz);
        const double kk = sqrt(kx * kx + ky * ky + kz * kz);
        // spline interval on the uniform table k_i = dk (i + 1) (splint.cc:4-20)
        int lo = (int)(kk * idk) - 1;
        lo = max(0, min(lo, nseg - 1));
        if (lo < nseg - 1 && dk * (lo + 2) <= kk) ++lo;
        if (lo > 0 && dk * (lo + 1) > kk) --lo;
        const R tt = (R)(kk - dk * (lo + 1));
        T4 cf[VC][VC];
#pragma unroll
        for (int v = 0; v < VC; ++v)
#pragma unroll
          for (int v1 = 0; v1 < VC; ++v1)
            if (v < V && v1 < V) cf[v][v1] = lane_at(chi + (long)(v * V + v1) * nseg, (unsigned)lo);
        T2 fk = mk<T2>((R)0, (R)0);
        if (a.use_fk) fk = fkh[((long)j * (ny / 2) + ya) * nx + x];
        const R sa = (R)jsign(j, a.pz.fixsign), sb = (R)jsign(jb, a.pz.fixsign);
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
#pragma unroll
        for (int v = 0; v < VC; ++v)
          if (v < V) {
            const R q = (R)a.qcf[v];
            c[v] = mk<T2>(c[v].x - q * fk.x, c[v].y - q * fk.y);  // mqvfk (fk = 0 in delta mode)
          }
        T2 h[VC + 1];
#pragma unroll
        for (int v = 0; v < VC + 1; ++v) {
          h[v] = mk<T2>((R)0, (R)0);
          if (v < V && v < VC) {
            R hr_ = 0, hi_ = 0;
#pragma unroll
            for (int v1 = 0; v1 < VC; ++v1)
              if (v1 < V) {  // oz: h_v = sum_v' chi[v][v'] c_v'
                const T4 q4 = cf[v][v1];
                const R xv = ((q4.w * tt + q4.z) * tt + q4.y) * tt + q4.x;
                hr_ += c[v1].x * xv;
                hi_ += c[v1].y * xv;
              }
            h[v] = mk<T2>(hr_, hi_);
          }
        }
#pragma unroll
        for (int p = 0; p < (VC + 1) / 2; ++p) {
          if (p < P && e0 < npair) {
            const T2 h0 = h[2 * p], h1 = h[2 * p + 1];
            // H(k) = h0 + i h1 ; H(-k) = conj(h0) + i conj(h1)
            const T2 ha = mk<T2>(h0.x - h1.y, h0.y + h1.x);
            const T2 hb = mk<T2>(h0.x + h1.y, h1.x - h0.y);
            lds[((p * 2 + 0) * T + i) * ls + lpad(j)] = cscale(ha, sa);
            lds[((p * 2 + 1) * T + ib) * ls + lpad(jb)] = cscale(hb, sb);
          }
        }
      }
      // bound the pairs in flight (table loads are hoisted within a group)
      if ((k + 1) % RISM_OZ_GROUP == 0) __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    int tidi = tid;
    asm volatile("" : "+v"(tidi));
    fft_lines<EPT, true>(lds, 2 * P * T, ls, a.pz, ltw, tidi);
    int tids = tid;
    asm volatile("" : "+v"(tids));
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int e = tids + k * kThreads;
      if (e < nel) {
        const int i = e % T, z = (e / T) % nz, line = (e / T / nz) * T + i;
        const T2 v = cmulc(lds[line * ls + lpad(z)], lpre[z]);
        Z[gofs(e)] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------
// 3'. the z pass for one pack (V <= 2) on a power-of-two z length NL whose
// tile pair fills the workgroup exactly (EPT elements per thread).  The
// first forward stage runs on the values as loaded and the last inverse
// stage stores to global memory, so the spectrum passes through LDS only
// between stages.  In the k-space step thread tid holds column i0 = tid % T
// of both tiles at rows z0 + k ZR (z0 = tid / T, ZR = 2 NL / EPT): every
// LDS position, phase factor and global offset is a per-thread base plus a
// compile-time step, and pair k joins row z0 + k ZR of tile A with row
// NL - 1 - z0 - k ZR of tile B.  |k| and the spline interval are evaluated
// in FP64 in both precisions.
// ---------------------------------------------------------------------
template <typename R, typename S, int EPT, int NL>
__global__ __launch_bounds__(kThreads) void k_fft_z_oz_ct(IterArgs<R, S> a) {
  const int sid = batch_solute(a.gate, a.rlist);  // this launch's solute (batched launches)
  if (sid < 0) return;
  const long so = (long)sid * a.astride;
  using T2 = typename C2<R>::T;
  using T4 = typename R4T<R>::T;
  if (!gate_open(soff(a.ctl, so), a.gate)) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  constexpr int T = EPT * kThreads / (2 * NL);  // columns per tile
  constexpr int ZR = kThreads / T;              // rows between a thread's elements
  constexpr int H = EPT / 2;                    // elements per tile and thread
  constexpr unsigned LS = line_stride(NL);
  static_assert(T >= 1 && ZR * H == NL && ZR % 2 == 0, "tile shape");
  const int nx = a.nx, ny = a.ny;
  const long N = a.N, nxy = (long)nx * ny;
  const int nxt = nx / T, ntile = nxt * (ny / 2);
  const int b = blockIdx.x;
  const int t = (ntile % kXcds) ? b : (b % kXcds) * (ntile / kXcds) + b / kXcds;
  const int xa = (t % nxt) * T, ya = t / nxt;
  const int xb = nx - xa - T, yb = ny - 1 - ya;
  T2 *__restrict__ Z = soff(a.Z, so);
  const long offa = (long)ya * nx + xa, offb = (long)yb * nx + xb;
  // k-space pairs of a thread: column i0 = tid % T of both tiles at rows
  // z0 + k ZR, z0 = tid / T; global offsets are a wave-uniform part plus a
  // 32-bit lane part below NL * nxy <= 2^27 elements
  const unsigned ut = threadIdx.x;
  const unsigned i0 = ut % T, z0 = ut / T;
  const unsigned la = i0 * LS + z0 + (z0 >> 4);  // LDS position of row z0 of tile A
  T2 *ltw = lds + 2 * T * LS, *lpre = ltw + NL;
  // The first forward stage runs on the loaded values and the last inverse
  // stage writes global memory: butterfly bf of the first stage (radix F0,
  // M0 = NL/F0 per line) is column bf % T of tile bf / (T M0) at rows
  // j + M0 a, j = (bf / T) % M0, so a thread loads exactly its butterfly's
  // inputs (lanes take consecutive columns: each load covers a few whole
  // row segments), and the last stage's outputs q0 + SL b go straight out.
  constexpr int NST = ct_stages(NL);
  constexpr int F0 = ct_radix(NL, 0), M0 = NL / F0;
  constexpr int FL = ct_radix(NL, NST - 1), SL = NL / FL;
  constexpr int NB0 = 2 * T * M0, KB0 = (NB0 + kThreads - 1) / kThreads;
  constexpr int NBL = 2 * T * SL, KBL = (NBL + kThreads - 1) / kThreads;
  static_assert(NST >= 2, "first and last stage are distinct");
  T2 v0[KB0][F0];
#pragma unroll
  for (int kb = 0; kb < KB0; ++kb) {
    const unsigned bf = ut + kb * kThreads;
    const unsigned i = bf % T, rest = bf / T, j = rest % M0, tile = rest / M0;
    if (NB0 % kThreads == 0 || bf < NB0)
#pragma unroll
      for (int q = 0; q < F0; ++q)
        v0[kb][q] = lane_at((const T2 *)Z + (long)(M0 * q) * nxy + (tile ? offb : offa), j * (unsigned)nxy + i);
  }
  stage_tables(ltw, a.pz);
  __syncthreads();
#pragma unroll
  for (int kb = 0; kb < KB0; ++kb) {
    const unsigned bf = ut + kb * kThreads;
    const unsigned i = bf % T, rest = bf / T, j = rest % M0, tile = rest / M0;
    if (NB0 % kThreads == 0 || bf < NB0) {
#pragma unroll
      for (int q = 0; q < F0; ++q) v0[kb][q] = cmul(v0[kb][q], lpre[j + M0 * q]);
      Dft<F0, false, T2>::run(v0[kb]);
      // stage 0 (S = 1): outputs F0 j + b, twiddle tw[j b]
      const unsigned wb = (tile * T + i) * LS + F0 * j + ((F0 * j) >> 4);
      lds[wb] = v0[kb][0];
#pragma unroll
      for (int q = 1; q < F0; ++q) lds[wb + q + (q >> 4)] = cmul(v0[kb][q], ltw[j * q]);
    }
  }
  __syncthreads();
  int tidf = threadIdx.x;
  asm volatile("" : "+v"(tidf));
  fft_lines_ct_range<NL, EPT, false, 4, 1, NST>(lds, (const T2 *)ltw, tidf);
  {
    // k-space.  Pair k: A at row z0 + k ZR of line i0, B at row NL - 1 - z0 - k ZR of line
    // 2T - 1 - i0 (its padded position steps down by the same amount as A's
    // steps up: NL - 1 - z0 leaves at least k ZR mod 16 below its 16-block's
    // end).  (-1)^(j + NL/2) is the same for every k (ZR is even), and B's
    // sign is its negative (NL is even).
    const unsigned w0 = NL - 1 - z0;
    const unsigned lb = (2 * T - 1 - i0) * LS + w0 + (w0 >> 4);
    const R sa = (R)jsign((int)z0, a.pz.fixsign), sb = -sa;
    const double tpx = 2.0 * M_PI / a.box[0], tpz = 2.0 * M_PI / a.box[2];
    const double kxd = tpx * ((double)(xa + (int)i0) + (0.5 - nx / 2));
    const double kyd = 2.0 * M_PI / a.box[1] * ((double)ya - ny / 2.0 + 0.5);
    const double hz = 0.5 - NL / 2;
    const int nseg = a.ntab - 1;
    const T4 *__restrict__ chi = a.chi;
    const T2 *__restrict__ fkh = soff(a.fkh, so);
    const double kxy2 = kxd * kxd + kyd * kyd, dk = a.dk, idk = 1.0 / a.dk;
    int tido = threadIdx.x;
    asm volatile("" : "+v"(tido));
    const unsigned z0o = (unsigned)tido / T;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const int j = (int)z0o + k * ZR;
      R tt;
      int lo;
      {
        const double kz = tpz * ((double)j + hz);
        const double kk = sqrt(kxy2 + kz * kz);
        // spline interval on the uniform table k_i = dk (i + 1) (splint.cc:4-20)
        lo = (int)(kk * idk) - 1;
        lo = max(0, min(lo, nseg - 1));
        if (lo < nseg - 1 && dk * (lo + 2) <= kk) ++lo;
        if (lo > 0 && dk * (lo + 1) > kk) --lo;
        tt = (R)(kk - dk * (lo + 1));
      }
      T4 cf[2][2];
      const int V = a.V;
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int v1 = 0; v1 < 2; ++v1)
          if (v < V && v1 < V) cf[v][v1] = lane_at(chi + (long)(v * V + v1) * nseg, (unsigned)lo);
      T2 fk = mk<T2>((R)0, (R)0);
      if (a.use_fk)  // row j = z0 + k ZR, column xa + i0
        fk = lane_at(fkh + ((long)k * ZR * (ny / 2) + ya) * nx + xa, (unsigned)(z0o * (ny / 2) * nx + (unsigned)tido % T));
      const unsigned pa = la + (unsigned)(k * ZR + ((k * ZR) >> 4)), pb = lb - (unsigned)(k * ZR + ((k * ZR) >> 4));
      const T2 za = cscale(lds[pa], sa);
      const T2 zb = cscale(lds[pb], sb);
      // C_0 = (za + conj zb)/2 ; C_1 = -i (za - conj zb)/2
      T2 c[2];
      c[0] = mk<T2>((R)0.5 * (za.x + zb.x), (R)0.5 * (za.y - zb.y));
      c[1] = mk<T2>((R)0.5 * (za.y + zb.y), (R)-0.5 * (za.x - zb.x));
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const R q = (R)a.qcf[v];
        c[v] = mk<T2>(c[v].x - q * fk.x, c[v].y - q * fk.y);  // mqvfk (fk = 0 in delta mode)
      }
      T2 h[2];
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        R hr_ = 0, hi_ = 0;
        if (v < V) {
#pragma unroll
          for (int v1 = 0; v1 < 2; ++v1)
            if (v1 < V) {  // oz: h_v = sum_v' chi[v][v'] c_v'
              const T4 q4 = cf[v][v1];
              const R xv = ((q4.w * tt + q4.z) * tt + q4.y) * tt + q4.x;
              hr_ += c[v1].x * xv;
              hi_ += c[v1].y * xv;
            }
        }
        h[v] = mk<T2>(hr_, hi_);
      }
      // H(k) = h0 + i h1 ; H(-k) = conj(h0) + i conj(h1)
      const T2 ha = mk<T2>(h[0].x - h[1].y, h[0].y + h[1].x);
      const T2 hb = mk<T2>(h[0].x + h[1].y, h[1].x - h[0].y);
      lds[pa] = cscale(ha, sa);
      lds[pb] = cscale(hb, sb);
      if ((k + 1) % RISM_OZ_GROUP == 0) __builtin_amdgcn_sched_barrier(0);
    }
  }
  __syncthreads();
  int tidi = threadIdx.x;
  asm volatile("" : "+v"(tidi));
  fft_lines_ct_range<NL, EPT, true, 4, 0, NST - 1>(lds, (const T2 *)ltw, tidi);
  // last inverse stage (S = SL, no twiddle): butterfly q0 of line i + tile T
  // reads q0 + SL c and writes rows q0 + SL b of global memory
  int tids = threadIdx.x;
  asm volatile("" : "+v"(tids));
#pragma unroll
  for (int kb = 0; kb < KBL; ++kb) {
    const unsigned bf = (unsigned)tids + kb * kThreads;
    const unsigned i = bf % T, rest = bf / T, q0 = rest % SL, tile = rest / SL;
    if (NBL % kThreads == 0 || bf < NBL) {
      T2 u[FL];
      const unsigned rb = (tile * T + i) * LS + q0 + (q0 >> 4);
#pragma unroll
      for (int c = 0; c < FL; ++c) u[c] = lds[rb + SL * c + ((SL * c) >> 4)];
      Dft<FL, true, T2>::run(u);
#pragma unroll
      for (int c = 0; c < FL; ++c)
        lane_at(Z + (long)(SL * c) * nxy + (tile ? offb : offa), q0 * (unsigned)nxy + i) =
            cmulc(u[c], lpre[q0 + SL * c]);
    }
  }
}

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

// One workgroup: reduce the per-workgroup partials (fixed order), rms,
// convergence test, Picard / Anderson coefficients (leg), control block.
template <int HM>
static __device__ void anderson_solve_block(Ctl *gctl, const double *partials, int nblk, double *rms_hist,
                                            int *rlist, int sid) {
  constexpr int NV = 1 + HM + HM * (HM + 1) / 2;
  const int hm = HM;
  __shared__ double tot[NV];
  __shared__ double wred[kThreads / 64][NV];
  // thread 0's small system lives in LDS: indexed by the filled slot count,
  // it would otherwise sit in scratch (private memory), one global-memory
  // round trip per element access
  __shared__ double sA[HM * HM], sc[HM], sth[HM];
  // The control block is staged in LDS by all threads at once, thread 0
  // works on the staged copy (its many dependent reads then wait on LDS, not
  // on global memory) and all threads write it back whole.
  __shared__ Ctl sctl;
  static_assert(sizeof(Ctl) % 4 == 0, "control block copied in 4-byte words");
  constexpr int kCtlWords = (int)(sizeof(Ctl) / 4);
  for (int i = threadIdx.x; i < kCtlWords; i += kThreads)
    reinterpret_cast<unsigned *>(&sctl)[i] = reinterpret_cast<const unsigned *>(gctl)[i];
  Ctl *ctl = &sctl;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // thread t sums rows t, t + 256, ... (fixed order); four rows in flight
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  int b = threadIdx.x;
  for (; b + 3 * kThreads < nblk; b += 4 * kThreads) {
    double r[4][NV];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int k = 0; k < NV; ++k) r[j][k] = partials[(long)(b + j * kThreads) * NV + k];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int k = 0; k < NV; ++k) s[k] += r[j][k];
  }
  for (; b < nblk; b += kThreads)
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] += partials[(long)b * NV + k];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double sw = wave_sum(s[k]);
    if (lane == 0) wred[wave][k] = sw;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NV; k += kThreads) {
    double t = 0.0;
    for (int wv = 0; wv < kThreads / 64; ++wv) t += wred[wv][k];
    tot[k] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) do {  // one pass; break leaves it
    const int iter = ctl->cur;
    ctl->cur = iter + 1;
    if (ctl->ring) ctl->rlast = ctl->rcur;
    const double rms = sqrt(tot[0] / ctl->npts);  // cal_rms.cu:18-20
    ctl->step += 1;
    ctl->total += 1;
    rms_hist[ctl->total - 1] = rms;
    ctl->rms_recent[(ctl->total - 1) % kRmsRecent] = rms;
    ctl->rms = rms;
    if (!isfinite(rms)) {  // iterate.cu:54-57
      ctl->status = 1;
      ctl->run_iter = -1;
      ctl->upd_iter = -1;
      break;
    }
    if (rms <= ctl->tol) {  // iterate.cu:58-59
      ctl->status = 0;
      ctl->run_iter = -1;
      ctl->upd_iter = -1;
      break;
    }
    const int H = ctl->nhist;
    const int w = ctl->write;
    if (ctl->ring) {  // mixed precision: refinement schedule (see Ctl)
      if (ctl->refresh_iter == iter) {
        ctl->ref_rms = rms;
        ctl->best_rms = rms;
        ctl->best_at = ctl->total;
        ctl->since_ref = 0;
      } else {
        ctl->since_ref += 1;
        if (rms < ctl->best_rms) {
          ctl->best_rms = rms;
          ctl->best_at = ctl->total;
        }
        // (a refresh step itself never asks for the next one: its rms is the
        // new reference and since_ref is 0)
        if (ctl->step < ctl->maxstep &&
            (rms < 0.1 * ctl->ref_rms || (ctl->since_ref >= 24 && ctl->total - ctl->best_at >= 24))) {
          bool filed = true;
          if (rlist) {  // batched launches: a place on the refresh list, else again next step
            const int slot = atomicAdd(rlist, 1);
            filed = slot < rlist[1];
            if (filed) rlist[2 + slot] = sid;
          }
          if (filed) ctl->refresh_iter = iter + 1;
        }
      }
    }
    if (ctl->count > 0) {  // Picard warm-up, anderson.cu:43-49
      ctl->count -= 1;
      ctl->upd_nf = 0;
      ctl->upd_mix = 1.0;
    } else {
      const int nf = ctl->filled < H ? ctl->filled : H;
      double *A = sA, *c = sc, *th = sth;
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
    if (ctl->ring) {  // file R into slot w by relabelling; the update reads the old map
      ctl->upd_rcur = ctl->rcur;
      for (int s = 0; s < H; ++s) ctl->upd_rmap[s] = ctl->rmap[s];
      const int freed = ctl->rmap[w];
      ctl->rmap[w] = ctl->rcur;
      ctl->rcur = freed;
    }
    ctl->write = (w + 1) % H;
    if (ctl->filled < H) ctl->filled += 1;
    ctl->upd_iter = iter + 1;
    if (ctl->step >= ctl->maxstep) {  // iterate.cu:76-79 (the update still applies, once, from the host)
      ctl->status = 2;
      ctl->run_iter = -1;
      ctl->upd_hold = iter + 1;
      ctl->upd_iter = -1;
    } else {
      ctl->run_iter = iter + 1;
    }
  } while (0);
  __syncthreads();
  for (int i = threadIdx.x; i < kCtlWords; i += kThreads)
    reinterpret_cast<unsigned *>(gctl)[i] = reinterpret_cast<const unsigned *>(&sctl)[i];
}

template <int HM>
__global__ __launch_bounds__(kThreads) void k_anderson_solve(Ctl *ctl, const double *partials, int nblk,
                                                             double *rms_hist, int gate, long astride,
                                                             int *rlist) {
  const int sid = batch_solute(gate, rlist);  // this launch's solute (batched launches)
  if (sid < 0) return;
  const long so = (long)sid * astride;
  ctl = soff(ctl, so);
  if (!gate_open(ctl, gate)) return;
  // the FP32 chain's solve files refreshes on the list; the refresh chain's reads it
  anderson_solve_block<HM>(ctl, soff(partials, so), nblk, soff(rms_hist, so), gate == 1 ? rlist : nullptr, sid);
}

// ---------------------------------------------------------------------
// 5. inverse x + residual + Anderson inner products.  After the transform
// the points are processed in batches whose loads are all issued first.
// ---------------------------------------------------------------------
template <typename R, typename S, int EPT, int HM, int VC>
__global__ __launch_bounds__(kThreads) void k_fx_resid(IterArgs<R, S> a) {
  const long so = (long)blockIdx.y * a.astride;  // this launch's solute (batched launches)
  using T2 = typename C2<R>::T;
  using Q = typename HistT<R>::T;
  constexpr int NV = 1 + HM + HM * (HM + 1) / 2;
  const Ctl *ctl = soff(a.ctl, so);
  if (ctl->run_iter != ctl->cur) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int nx = a.nx, L = a.Lx, ls = line_stride(nx), V = a.V, P = a.P;
  const long N = a.N, VN = (long)V * N;
  const int nel = L * nx;
  const int groups = (int)((long)a.ny * a.nz / L);
  const int nf = ctl->filled < HM ? ctl->filled : HM;
  T2 *ltw = lds + P * L * ls;
  const T2 *pre = ltw + nx;
  stage_tables(ltw, a.px);  // visible after the first barrier of the group loop
  const R ivol = (R)a.ivol;
  T2 *__restrict__ Z = soff(a.Z, so);
  const S *__restrict__ ts = soff(a.ts, so);
  const R *__restrict__ uu = soff(a.u, so);
  Q *__restrict__ res = soff(a.res, so);
  const Q *__restrict__ hr = soff(a.hist_r, so);
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
#ifndef RISM_RESID_ZPF
#define RISM_RESID_ZPF 1
#endif
  // spectrum of a line group into registers (clamped index: one batch of loads)
  T2 zb[EPT];
  auto load_z = [&](int g, int lane) {
    const long gb = (long)g * nel;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int f = min(lane + k * kThreads, P * nel - 1);
      const int p = f / nel, e = f - p * nel;
      zb[k] = Z[p * N + gb + e];
    }
  };
  if (RISM_RESID_ZPF) load_z(blockIdx.x, threadIdx.x);
  for (int grp = blockIdx.x; grp < groups; grp += gridDim.x) {
    const long base = (long)grp * nel;
    // the lane's index is made opaque per group: otherwise the compiler
    // hoists every per-element index and twiddle out of this loop and keeps
    // them live, which costs more registers than recomputing them
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    __syncthreads();
    if (!RISM_RESID_ZPF) load_z(grp, tid);
    // points per thread: nel <= EPT * kThreads / P and P >= PMIN in this site class
    constexpr int PMIN = VC <= 2 ? 1 : (VC <= 4 ? 2 : 3);
    constexpr int KPT = (EPT + PMIN - 1) / PMIN;
#ifndef RISM_RESID_CH
#define RISM_RESID_CH 1
#endif
    // Operands of a point (t_s, u, history residuals) are loaded after
    // the line transform in batches of CH points.  Loading every point's
    // operands before the transform instead (to overlap their latency with
    // the LDS work) needs so many registers that occupancy halves; measured
    // slower at 256^3.
    constexpr int CH = KPT < RISM_RESID_CH ? KPT : RISM_RESID_CH;
    constexpr bool PREH = CH * VC * HM <= 16;  // history residuals join the batch
    auto load_ops = [&](unsigned e, double *tvo, double *uvo, Q *hqo) {
      if (!a.href_only) {
#pragma unroll
        for (int v = 0; v < VC; ++v)
          if (v < V) {
            const long vb = (long)v * N + base;
            tvo[v] = (double)lane_at(ts + vb, e);
            uvo[v] = (double)lane_at(uu + vb, e);
#pragma unroll
            for (int s = 0; s < (PREH ? HM : 0); ++s) hqo[v * (PREH ? HM : 1) + s] = lane_at(hr + (s < nf ? s : 0) * VN + vb, e);
          }
      }
    };
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int f = tid + k * kThreads;
      if (f < P * nel) {
        const int p = f / nel, e = f - p * nel;
        const int i = e / nx, j = e - i * nx;
        lds[(p * L + i) * ls + lpad(j)] = cscale(zb[k], (R)jsign(j, a.px.fixsign));
      }
    }
    __syncthreads();
    // the transform and the point loop take their own opaque lane index, so
    // the LDS indices of the first pass are not kept live through them
    int tidf = tid;
    asm volatile("" : "+v"(tidf));
    // FP64: the radix-4 plan (no radix-8/16 stage compiled in) lowers the
    // register peak from the transform enough for one more wave per SIMD;
    // FP32 fits its radix-16 plan at the same occupancy as the point loop
    if constexpr (sizeof(R) == sizeof(double))
      fft_lines<EPT, true, 5>(lds, P * L, ls, a.pxr, ltw, tidf);
    else
      fft_lines<EPT, true>(lds, P * L, ls, a.px, ltw, tidf);
    int tidr = tid;
    asm volatile("" : "+v"(tidr));
    // The next group's spectrum is loaded now, so its latency overlaps this
    // group's point loop instead of stalling the next transform (a uniform
    // branch: after a workgroup's last group nothing is loaded).  Registers:
    // the point loop stays below the transform's peak with it.
    if (RISM_RESID_ZPF && grp + (int)gridDim.x < groups) load_z(grp + (int)gridDim.x, tidr);
#pragma unroll
    for (int k0 = 0; k0 < KPT; k0 += CH) {
      // batch of CH points: every load first (clamped index), then the arithmetic
      double hv[CH][VC + 1];
      double tv[CH][VC], uv[CH][VC];
      Q hq[CH][VC][PREH ? HM : 1];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const unsigned e = (unsigned)min(tidr + (k0 + c) * kThreads, nel - 1);
        const int i = e / nx, x = e - i * nx;
        const T2 pw = pre[x];
        load_ops(e, tv[c], uv[c], &hq[c][0][0]);
#pragma unroll
        for (int p = 0; p < (VC + 1) / 2; ++p) {
          if (p < P) {
            const T2 z = cscale(cmulc(lds[(p * L + i) * ls + lpad(x)], pw), ivol);
            hv[c][2 * p] = (double)z.x;
            hv[c][2 * p + 1] = (double)z.y;
          }
        }
      }
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const unsigned e = tidr + (k0 + c) * kThreads;
        if (k0 + c < KPT && e < (unsigned)nel) {
          if (a.store_h) {
            const int i = e / nx, x = e - i * nx;
#pragma unroll
            for (int p = 0; p < (VC + 1) / 2; ++p)
              if (p < P) lane_at(Z + p * N + base, e) = cscale(cmulc(lds[(p * L + i) * ls + lpad(x)], pre[x]), ivol);
          }
          if (a.href_only) continue;
#pragma unroll
          for (int v = 0; v < VC; ++v) {
            if (v < V) {
              const long vb = (long)v * N + base;
              const double g = closure_g<double>(tv[c][v] - uv[c][v], a.closure);
              const double rd = hv[c][v] + 1.0 - g;  // tr: R = h + 1 - g
              lane_at(res + vb, e) = (Q)rd;
              const double wgt = a.rho_inv[v];
              acc[0] += rd * rd;
              double d[HM];
#pragma unroll
              for (int s = 0; s < HM; ++s)
                d[s] = (s < nf) ? rd - (double)(PREH ? hq[c][v][PREH ? s : 0] : lane_at(hr + s * VN + vb, e)) : 0.0;
#pragma unroll
              for (int s = 0; s < HM; ++s) acc[1 + s] += wgt * rd * d[s];
              int kk = 1 + HM;
#pragma unroll
              for (int s = 0; s < HM; ++s)
#pragma unroll
                for (int s2 = s; s2 < HM; ++s2) acc[kk++] += wgt * d[s] * d[s2];
            }
          }
        }
      }
      // keep the next batch's loads from being hoisted into this one
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // workgroup reduction of the accumulators
  __shared__ double red[kThreads / 64][NV];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double sw = wave_sum(acc[k]);
    if (lane == 0) red[wave][k] = sw;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NV; k += kThreads) {
    double sm = 0.0;
    for (int wv = 0; wv < kThreads / 64; ++wv) sm += red[wv][k];
    soff(a.partials, so)[(long)blockIdx.x * NV + k] = sm;
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
  T2 *ltw = lds + L * ls;
  const T2 *pre = ltw + nx;
  stage_tables(ltw, px);
  __syncthreads();
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    const int i = e / nx, x = e - i * nx;
    T2 v = Z[base + e];
    v = INV ? cscale(v, (R)jsign(x, px.fixsign)) : cmul(v, pre[x]);
    lds[i * ls + lpad(x)] = v;
  }
  __syncthreads();
  fft_lines<EPT, INV>(lds, L, ls, px, ltw);
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
  T2 *ltw = lds + T * ls;
  const T2 *pre = ltw + nz;
  stage_tables(ltw, pz);
  __syncthreads();
  for (int e = threadIdx.x; e < nel; e += kThreads) {
    const int z = e / T, i = e - z * T;
    T2 v = Zp[(long)z * nxy + i];
    v = INV ? cscale(v, (decltype(v.x))jsign(z, pz.fixsign)) : cmul(v, pre[z]);
    lds[i * ls + lpad(z)] = v;
  }
  __syncthreads();
  fft_lines<EPT, INV>(lds, T, ls, pz, ltw);
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
// x-line kernels: at most 16 elements per thread (P * Lx * nx <= 4096)
#define RISM_EPTX_SWITCH(ept, ...)                \
  switch (ept) {                                  \
    case 4: { constexpr int E = 4; __VA_ARGS__; } break;   \
    case 8: { constexpr int E = 8; __VA_ARGS__; } break;   \
    case 16: { constexpr int E = 16; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;         \
  }

#define RISM_VC_SWITCH(V, ...)                                  \
  switch ((V) <= 2 ? 2 : ((V) <= 4 ? 4 : 8)) {                 \
    case 2: { constexpr int VC = 2; __VA_ARGS__; } break;       \
    case 4: { constexpr int VC = 4; __VA_ARGS__; } break;       \
    default: { constexpr int VC = 8; __VA_ARGS__; } break;      \
  }

template <typename R, typename S>
hipError_t launch_closure_fx(const IterArgs<R, S> &a, const LaunchGeom &g, bool update, int hm,
                             hipStream_t st) {
  const dim3 grid(g.cgroups, a.nsol > 1 ? a.nsol : 1), block(kThreads);
  if (!update) return hipErrorInvalidValue;  // every launch carries the history class
#define RISM_CLOS(HMV) \
  RISM_VC_SWITCH(a.V, RISM_EPTX_SWITCH(g.ept_c, hipLaunchKernelGGL((k_closure_fx<R, S, E, true, VC, HMV>), grid, block, g.lds_c, st, a)))
  switch (hm) {
    case 2: RISM_CLOS(2); break;
    case 5: RISM_CLOS(5); break;
    case 10: RISM_CLOS(10); break;
    default: return hipErrorInvalidValue;
  }
#undef RISM_CLOS
  return hipGetLastError();
}

// compile-time line length for the y / z passes: a power of two in
// 64..512 on a tile of exactly EPT (8 or 16) elements per thread
inline int ct_length(int n, long tile_elems, int ept) {
  if ((ept != 8 && ept != 16) || tile_elems != (long)ept * kThreads) return 0;
  return (n == 64 || n == 128 || n == 256 || n == 512) ? n : 0;
}
#define RISM_CT_SWITCH(nl, ept, GENERIC, ...)                                     \
  switch ((nl) * 100 + (ept)) {                                                   \
    case 6408: { constexpr int E = 8, NLV = 64; __VA_ARGS__; } break;             \
    case 6416: { constexpr int E = 16, NLV = 64; __VA_ARGS__; } break;            \
    case 12808: { constexpr int E = 8, NLV = 128; __VA_ARGS__; } break;           \
    case 12816: { constexpr int E = 16, NLV = 128; __VA_ARGS__; } break;          \
    case 25608: { constexpr int E = 8, NLV = 256; __VA_ARGS__; } break;           \
    case 25616: { constexpr int E = 16, NLV = 256; __VA_ARGS__; } break;          \
    case 51208: { constexpr int E = 8, NLV = 512; __VA_ARGS__; } break;           \
    case 51216: { constexpr int E = 16, NLV = 512; __VA_ARGS__; } break;          \
    default: GENERIC;                                                             \
  }

template <typename R, typename S>
hipError_t launch_fft_y(const IterArgs<R, S> &a, const LaunchGeom &g, bool inverse, hipStream_t st) {
  const dim3 grid(a.nx / g.Ty * a.nz * a.P, a.nsol > 1 ? a.nsol : 1), block(kThreads);
  const int nl = ct_length(a.ny, (long)g.Ty * a.ny, g.ept_y);
  if (inverse) {
    RISM_CT_SWITCH(nl, g.ept_y,
                   RISM_EPT_SWITCH(g.ept_y, hipLaunchKernelGGL((k_fft_y<R, E, true>), grid, block, g.lds_y, st, a.Z,
                                                               a.ctl, a.nx, a.ny, a.nz, a.N, g.Ty, a.P, a.py, a.gate, a.astride, a.rlist)),
                   hipLaunchKernelGGL((k_fft_y<R, E, true, NLV>), grid, block, g.lds_y, st, a.Z, a.ctl, a.nx,
                                      a.ny, a.nz, a.N, g.Ty, a.P, a.py, a.gate, a.astride, a.rlist))
  } else {
    RISM_CT_SWITCH(nl, g.ept_y,
                   RISM_EPT_SWITCH(g.ept_y, hipLaunchKernelGGL((k_fft_y<R, E, false>), grid, block, g.lds_y, st, a.Z,
                                                               a.ctl, a.nx, a.ny, a.nz, a.N, g.Ty, a.P, a.py, a.gate, a.astride, a.rlist)),
                   hipLaunchKernelGGL((k_fft_y<R, E, false, NLV>), grid, block, g.lds_y, st, a.Z, a.ctl, a.nx,
                                      a.ny, a.nz, a.N, g.Ty, a.P, a.py, a.gate, a.astride, a.rlist))
  }
  return hipGetLastError();
}

template <typename R, typename S>
hipError_t launch_fft_z_oz(const IterArgs<R, S> &a, const LaunchGeom &g, hipStream_t st) {
  const dim3 grid(a.nx / g.Tz * (a.ny / 2), a.nsol > 1 ? a.nsol : 1), block(kThreads);
  const int nl = a.V <= 2 ? ct_length(a.nz, 2L * g.Tz * a.nz, g.ept_z) : 0;
  RISM_CT_SWITCH(nl, g.ept_z,
                 RISM_VC_SWITCH(a.V, RISM_EPT_SWITCH(g.ept_z, hipLaunchKernelGGL((k_fft_z_oz<R, S, E, VC>), grid, block,
                                                                                 g.lds_z, st, a))),
                 hipLaunchKernelGGL((k_fft_z_oz_ct<R, S, E, NLV>), grid, block, g.lds_z, st, a))
  return hipGetLastError();
}

template <typename R, typename S>
hipError_t launch_fx_resid(const IterArgs<R, S> &a, const LaunchGeom &g, int hm, hipStream_t st) {
  const dim3 grid(g.xblocks, a.nsol > 1 ? a.nsol : 1), block(kThreads);
#define RISM_RESID(HMV)                                                                       \
  RISM_VC_SWITCH(a.V, RISM_EPTX_SWITCH(g.ept_x, hipLaunchKernelGGL((k_fx_resid<R, S, E, HMV, VC>), grid, block, g.lds_x, st, a)))
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
                                 double *rms_hist, hipStream_t st, int gate, long astride, int nsol,
                                 int *rlist) {
  const dim3 grid(1, nsol > 1 ? nsol : 1);
  switch (hm) {
    case 2: hipLaunchKernelGGL(k_anderson_solve<2>, grid, dim3(kThreads), 0, st, ctl, partials, nblk, rms_hist, gate, astride, rlist); break;
    case 5: hipLaunchKernelGGL(k_anderson_solve<5>, grid, dim3(kThreads), 0, st, ctl, partials, nblk, rms_hist, gate, astride, rlist); break;
    case 10: hipLaunchKernelGGL(k_anderson_solve<10>, grid, dim3(kThreads), 0, st, ctl, partials, nblk, rms_hist, gate, astride, rlist); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_fft3d(typename C2<R>::T *data, const AxisPlan &px, const AxisPlan &py,
                        const AxisPlan &pz, const LaunchGeom &g, int nx, int ny, int nz, double dv,
                        double ivol, int dir, hipStream_t st) {
  const long N = (long)nx * ny * nz;
  const int L = g.Lx;
  const dim3 block(kThreads);
  const dim3 gx((unsigned)((long)ny * nz / L)), gy(nx / g.Ty * nz), gz(nx / g.Tz, ny);
  const size_t sz = sizeof(typename C2<R>::T);
  const size_t ldsx = ((size_t)L * line_stride(nx) + 2 * nx) * sz, ldsy = ((size_t)g.Ty * line_stride(ny) + 2 * ny) * sz,
               ldsz = ((size_t)g.Tz * line_stride(nz) + 2 * nz) * sz;
  // pure z tiles hold T lines: EPT class from T*nz
  const int ez = (g.Tz * nz <= 4 * kThreads) ? 4 : (g.Tz * nz <= 8 * kThreads) ? 8 : (g.Tz * nz <= 16 * kThreads) ? 16 : 32;
  if (dir < 0) {
    RISM_EPTX_SWITCH(g.ept_x, hipLaunchKernelGGL((k_fft_x_pure<R, E, false>), gx, block, ldsx, st, data, nx, L, px, (R)dv))
    RISM_EPT_SWITCH(g.ept_y, hipLaunchKernelGGL((k_fft_y<R, E, false>), gy, block, ldsy, st, data, (const Ctl *)nullptr, nx, ny, nz, N, g.Ty, 1, py, 0, 0L, (const int *)nullptr))
    RISM_EPT_SWITCH(ez, hipLaunchKernelGGL((k_fft_z_pure<R, E, false>), gz, block, ldsz, st, data, nx, ny, nz, g.Tz, pz))
  } else {
    RISM_EPT_SWITCH(ez, hipLaunchKernelGGL((k_fft_z_pure<R, E, true>), gz, block, ldsz, st, data, nx, ny, nz, g.Tz, pz))
    RISM_EPT_SWITCH(g.ept_y, hipLaunchKernelGGL((k_fft_y<R, E, true>), gy, block, ldsy, st, data, (const Ctl *)nullptr, nx, ny, nz, N, g.Ty, 1, py, 0, 0L, (const int *)nullptr))
    RISM_EPTX_SWITCH(g.ept_x, hipLaunchKernelGGL((k_fft_x_pure<R, E, true>), gx, block, ldsx, st, data, nx, L, px, (R)ivol))
  }
  return hipGetLastError();
}

}  // namespace rism
