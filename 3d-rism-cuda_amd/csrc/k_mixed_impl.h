// k_mixed_impl.h -- the closure and residual passes of the mixed-precision
// iteration on gfx950: FP32 fields relative to an FP64 reference.
//
// The fixed point is the one of the FP64 path (src-cuda/calculate.cu:4-58,
// anderson.cu:35-84); only the representation differs.  With ts the
// short-range state and u the closure potential,
//   c_s = G(ts - u) - 1 - ts,   h = K c (FFT, chi, inverse FFT),
//   R = h + 1 - G(ts - u).
// The state is ts = T + dl: T (FP64) is ts at the last refresh and dl (FP32)
// the change since then; A = T - u (stored in FP32) is the closure argument
// at the refresh.  (ts itself is kept: near the solute's core u reaches 1e29, and
// ts - u cannot carry ts there, while c_s = -1 - ts needs it.)  A refresh
// step (R = double, gate 2) folds dl into T and evaluates the map in FP64,
// storing R_ref = R (FP32).  Every other step (R = float, gate 1)
// transforms only the change of c_s since the refresh,
//   dc = G(A + dl) - G(A) - dl,
// whose long-range part is zero, in FP32; the residual is then
//   R = K dc + B,   B = R_ref - (G(A + dl) - G(A)),
// with G, dc and B formed in FP64 and B handed to the residual pass in
// FP32.  Rounding thus scales with |dc| and |B|, which shrink as the solve
// converges, not with |c_s| or |h|.
//
// Anderson state (FP32 in both chains): hx [H][V][N] holds the step
// increments of ts by slot (age order as in the FP64 update, written by the
// pending update), rring [H+1][V][N] the residuals, addressed through the
// control block's ring map (Ctl::rcur, rmap): the residual pass writes the
// current R into ring slot rcur and the solve kernel files it into history
// by relabelling, so no pass copies a residual.
#pragma once
#include "k_iterate_impl.h"

#ifndef RISM_MXR_PLAN4
#define RISM_MXR_PLAN4 0  // FP32 residual pass on the radix-4 x plan
#endif
#ifndef RISM_MXR_SB
#define RISM_MXR_SB 1     // residual pass: each batch's loads stay in the batch
#endif
#ifndef RISM_MXR_ZPF
#define RISM_MXR_ZPF 1    // residual pass loads the next group's spectrum during the point loop
#endif
#ifndef RISM_MXC_VEC
#define RISM_MXC_VEC 1    // FP32 closure pass with consecutive points per thread (k_mx_closure_vec)
#endif
#ifndef RISM_MXR_VEC
#define RISM_MXR_VEC 1    // FP32 residual pass with consecutive points per thread (k_mx_resid_vec)
#endif

namespace rism {

// ---------------------------------------------------------------------
// 1. pending Anderson update of x + closure + c_s (FP64) or dc (FP32),
// packed two sites per complex value, forward FFT along x
// ---------------------------------------------------------------------
template <typename R, int EPT, int VC, int HM, bool UPD>
__device__ __forceinline__ void mx_points(const IterArgs<R, double> &a, typename C2<R>::T *lds, bool do_run,
                                          const double *th, const long *toff, const long *roff, long rcur_off,
                                          double mix, int w, int nf) {
  const long so = (long)batch_solute(a.gate, a.rlist) * a.astride;  // this launch's solute (batched launches)
  using T2 = typename C2<R>::T;
  constexpr bool F64 = sizeof(R) == sizeof(double);  // the refresh chain
  constexpr int CH = EPT < RISM_CLOS_CH ? EPT : RISM_CLOS_CH;
  const int nx = a.nx, L = a.Lc, ls = line_stride(nx), V = a.V;
  const long N = a.N, VN = (long)V * N;
  const long base = (long)blockIdx.x * L * nx;
  const int nel = L * nx;
  const T2 *__restrict__ pre = (const T2 *)a.px.pre;
  float *__restrict__ Ab = soff(a.A, so);
  double *__restrict__ Tb = soff(a.ts, so);  // the refresh chain's T
  float *__restrict__ dlb = soff(a.dl, so);
  const float *__restrict__ hx = soff(a.hx, so);
  float *__restrict__ hxw = soff(a.hx, so);
  const float *__restrict__ rr = soff(a.rring, so);
  const float *__restrict__ rref = soff(a.rref, so);
  const double *__restrict__ uu = (const double *)soff(a.u, so);
#pragma unroll
  for (int k0 = 0; k0 < EPT; k0 += CH) {
    double A[CH][VC], d0[CH][VC], u[CH][VC], rf[CH][VC];  // A: T in the refresh chain
    float r[CH][VC], hti[CH][VC][UPD ? HM : 1], hri[CH][VC][UPD ? HM : 1];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const unsigned e = (unsigned)min((int)threadIdx.x + (k0 + c) * kThreads, nel - 1);
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        if (v < V) {
          const long vb = (long)v * N + base;
          A[c][v] = F64 ? lane_at(Tb + vb, e) : (double)lane_at(Ab + vb, e);
          d0[c][v] = (double)lane_at(dlb + vb, e);
          if constexpr (F64) u[c][v] = lane_at(uu + vb, e);
          else rf[c][v] = (double)lane_at(rref + vb, e);
          if (UPD) {
            r[c][v] = lane_at(rr + rcur_off + vb, e);
#pragma unroll
            for (int s = 0; s < (UPD ? HM : 0); ++s) {
              hti[c][v][s] = lane_at(hx + toff[s] + vb, e);
              hri[c][v][s] = lane_at(rr + roff[s] + vb, e);
            }
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const unsigned e = threadIdx.x + (k0 + c) * kThreads;
      const bool valid = e < (unsigned)nel;
      R cs[VC + 1];
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        cs[v] = (R)0;
        if (v < V) {
          const long vb = (long)v * N + base;
          // the update as a change of ts: ts_new - ts = sum theta_s (ts_s - ts) + m Rbar
          double d = d0[c][v];
          if (UPD) {  // newdt / newdt0 (anderson_cuda.cu:1-23) on the increment history
            const double rv = (double)r[c][v];
            double tb = 0.0, rb = rv, back = 0.0;
#pragma unroll
            for (int s = 0; s < (UPD ? HM : 0); ++s) {
              // slots beyond the filled ones: valid loads, values discarded
              const double hts = s < nf ? (double)hti[c][v][UPD ? s : 0] : 0.0;
              const double hrs = s < nf ? (double)hri[c][v][UPD ? s : 0] : rv;
              back -= hts;  // age s+1: ts_(s+1) - ts = -(increments of ages 1..s+1)
              tb += th[s] * back;
              rb += th[s] * (hrs - rv);
            }
            d += tb + mix * rb;
          }
          if constexpr (F64) {
            // refresh: T takes the change, A = T - u is the new reference
            const double Tn = A[c][v] + d, An = Tn - u[c][v];
            if (UPD && valid) lane_at(hxw + (long)w * VN + vb, e) = (float)(d - d0[c][v]);
            if (valid && (UPD || do_run)) {
              lane_at(Tb + vb, e) = Tn;
              lane_at(Ab + vb, e) = ref_f32(An);
              lane_at(dlb + vb, e) = 0.0f;
            }
            if (do_run) cs[v] = (R)(closure_g<double>(An, a.closure) - 1.0 - Tn);
          } else {
            // the change since the refresh is stored in FP32: round first, then evaluate
            const float dn = (float)d;
            if (UPD && valid) {
              lane_at(hxw + (long)w * VN + vb, e) = (float)((double)dn - d0[c][v]);
              lane_at(dlb + vb, e) = dn;
            }
            if (do_run) {
              const double dg = closure_g<double>(A[c][v] + (double)dn, a.closure) - closure_g<double>(A[c][v], a.closure);
              if (valid) lane_at(soff(a.bq, so) + vb, e) = (float)(rf[c][v] - dg);
              cs[v] = (R)(dg - (double)dn);
            }
          }
        }
      }
      if (do_run && valid) {
        const int i = e / nx, x = e - i * nx;
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

// W consecutive floats at base + e (e a multiple of W; 16-byte aligned when
// W is a multiple of 4): one 16-byte access per four
template <int W> __device__ __forceinline__ void ldv(float (&d)[W], const float *base, unsigned e) {
  const char *p = reinterpret_cast<const char *>(base) + e * 4u;
  if constexpr (W % 4 == 0) {
#pragma unroll
    for (int j = 0; j < W; j += 4) {
      const float4 t = *reinterpret_cast<const float4 *>(p + 4 * j);
      d[j] = t.x; d[j + 1] = t.y; d[j + 2] = t.z; d[j + 3] = t.w;
    }
  } else if constexpr (W == 2) {
    const float2 t = *reinterpret_cast<const float2 *>(p);
    d[0] = t.x; d[1] = t.y;
  } else {
    d[0] = *reinterpret_cast<const float *>(p);
  }
}
template <int W> __device__ __forceinline__ void ldv(double (&d)[W], const double *base, unsigned e) {
  const char *p = reinterpret_cast<const char *>(base) + e * 8u;
  if constexpr (W % 2 == 0) {
#pragma unroll
    for (int j = 0; j < W; j += 2) {
      const double2 t = *reinterpret_cast<const double2 *>(p + 8 * j);
      d[j] = t.x; d[j + 1] = t.y;
    }
  } else {
    d[0] = *reinterpret_cast<const double *>(p);
  }
}
template <int W> __device__ __forceinline__ void stv(float *base, unsigned e, const float (&d)[W]) {
  char *p = reinterpret_cast<char *>(base) + e * 4u;
  if constexpr (W % 4 == 0) {
#pragma unroll
    for (int j = 0; j < W; j += 4) *reinterpret_cast<float4 *>(p + 4 * j) = make_float4(d[j], d[j + 1], d[j + 2], d[j + 3]);
  } else if constexpr (W == 2) {
    *reinterpret_cast<float2 *>(p) = make_float2(d[0], d[1]);
  } else {
    *reinterpret_cast<float *>(p) = d[0];
  }
}

// ---------------------------------------------------------------------
// 1'. the FP32 chain's closure pass with W consecutive points per thread
// (16-byte operand loads and stores), otherwise k_mx_closure_fx's
// arithmetic: the pending update of dl, dG = G(A + dl) - G(A), B = R_ref -
// dG and dc = dG - dl, packed two sites per complex value, forward x.
// ---------------------------------------------------------------------
template <int EPT, int VC, int HM, bool UPD, int NX>
__device__ __forceinline__ void mx_points_vec(const IterArgs<float, double> &a, float2 *lds, bool do_run, long so,
                                              const double *th, const long *toff, const long *roff, long rcur_off,
                                              double mix, int w, int nf) {
  constexpr int W = VC <= 2 ? EPT : EPT / 2;
  const int nx = NX > 0 ? NX : a.nx, L = NX > 0 ? EPT * kThreads / (VC / 2 * NX) : a.Lc, ls = line_stride(nx);
  const int V = a.V;
  const long N = a.N, VN = (long)V * N;
  const long base = closure_group() * L * nx;
  const int nel = L * nx;
  const float2 *__restrict__ pre = (const float2 *)a.px.pre;
  const float *__restrict__ Ab = soff(a.A, so);
  float *__restrict__ dlb = soff(a.dl, so);
  const float *__restrict__ hx = soff(a.hx, so);
  float *__restrict__ hxw = soff(a.hx, so);
  const float *__restrict__ rr = soff(a.rring, so);
  const float *__restrict__ rref = soff(a.rref, so);
  float *__restrict__ bq = soff(a.bq, so);
  const bool act = (int)threadIdx.x * W < nel;
  const unsigned e0 = (unsigned)(act ? (int)threadIdx.x * W : nel - W);
  float cs[VC + 1][W];
#pragma unroll
  for (int v = 0; v < VC; ++v) {
    if (v < V) {
      const long vb = (long)v * N + base;
      float A[W];
      float d0[W], rf[W], r[W], hti[UPD ? HM : 1][W], hri[UPD ? HM : 1][W];
      ldv<W>(A, Ab + vb, e0);
      ldv<W>(d0, dlb + vb, e0);
      ldv<W>(rf, rref + vb, e0);
      if (UPD) {
        ldv<W>(r, rr + rcur_off + vb, e0);
#pragma unroll
        for (int s = 0; s < (UPD ? HM : 0); ++s) {
          ldv<W>(hti[UPD ? s : 0], hx + toff[s] + vb, e0);
          ldv<W>(hri[UPD ? s : 0], rr + roff[s] + vb, e0);
        }
      }
      float dno[W], hxo[W], bo[W];
#pragma unroll
      for (int j = 0; j < W; ++j) {
        double d = (double)d0[j];
        if (UPD) {  // newdt / newdt0 (anderson_cuda.cu:1-23) on the increment history
          const double rv = (double)r[j];
          double tb = 0.0, rb = rv, back = 0.0;
#pragma unroll
          for (int s = 0; s < (UPD ? HM : 0); ++s) {
            const double hts = s < nf ? (double)hti[UPD ? s : 0][j] : 0.0;
            const double hrs = s < nf ? (double)hri[UPD ? s : 0][j] : rv;
            back -= hts;
            tb += th[s] * back;
            rb += th[s] * (hrs - rv);
          }
          d += tb + mix * rb;
        }
        const float dn = (float)d;
        dno[j] = dn;
        hxo[j] = (float)((double)dn - (double)d0[j]);
        cs[v][j] = 0.0f;
        if (do_run) {
          const double dg = closure_g<double>((double)A[j] + (double)dn, a.closure) - closure_g<double>((double)A[j], a.closure);
          bo[j] = (float)((double)rf[j] - dg);
          cs[v][j] = (float)(dg - (double)dn);
        }
      }
      if (act) {
        if (UPD) {
          stv<W>(hxw + (long)w * VN + vb, e0, hxo);
          stv<W>(dlb + vb, e0, dno);
        }
        if (do_run) stv<W>(bq + vb, e0, bo);
      }
    } else {
#pragma unroll
      for (int j = 0; j < W; ++j) cs[v][j] = 0.0f;
    }
  }
  if (do_run && act) {
#pragma unroll
    for (int j = 0; j < W; ++j) cs[VC][j] = 0.0f;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const int e = (int)e0 + j;
      const int i = e / nx, x = e - i * nx;
      const float2 pw = pre[x];
#pragma unroll
      for (int p = 0; p < (VC + 1) / 2; ++p)
        if (2 * p < V) lds[(p * L + i) * ls + lpad(x)] = cmul(mk<float2>(cs[2 * p][j], cs[2 * p + 1][j]), pw);
    }
  }
}

template <int EPT, int VC, int HM, int NX>
__global__ __launch_bounds__(kThreads) void k_mx_closure_vec(IterArgs<float, double> a) {
  // the refresh list of the next step starts empty (the refresh chain of
  // this step, launched before this pass, has read it)
  if (a.gate == 1 && a.rlist && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) a.rlist[0] = 0;
  const int sid = batch_solute(a.gate, a.rlist);
  if (sid < 0) return;
  const long so = (long)sid * a.astride;
  using T2 = float2;
  const Ctl *ctl = soff(a.ctl, so);
  const int iter = ctl->cur;
  const bool mine = chain_active(ctl, a.gate);
  const bool do_upd = mine && ctl->upd_iter == iter;
  const bool do_run = mine && ctl->run_iter == iter;
  if (!do_upd && !do_run) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  // NX > 0: compile-time line length (a power of two) on a tile of exactly
  // EPT * kThreads elements, VC / 2 packs
  const int nx = NX > 0 ? NX : a.nx, P = NX > 0 ? VC / 2 : a.P;
  const int L = NX > 0 ? EPT * kThreads / (VC / 2 * NX) : a.Lc, ls = line_stride(nx);
  const long N = a.N, VN = (long)a.V * N;
  const long base = closure_group() * L * nx;
  const int nel = L * nx;
  int w = 0, nf = 0;
  double mix = 0.0;
  double th[HM];
  long toff[HM], roff[HM];
  long rcur_off = 0;
#pragma unroll
  for (int k = 0; k < HM; ++k) {
    th[k] = 0.0;
    toff[k] = roff[k] = 0;
  }
  if (do_upd) {
    nf = ctl->upd_nf;
    w = ctl->upd_w;
    mix = ctl->upd_mix;
    const int nh = ctl->nhist;
    rcur_off = (long)ctl->upd_rcur * VN;
#pragma unroll
    for (int k = 0; k < HM; ++k) {
      const int slot = (((w - 1 - k) % nh) + nh) % nh;
      th[k] = (k < nf) ? ctl->upd_theta[slot] : 0.0;
      toff[k] = (long)slot * VN;
      roff[k] = (long)ctl->upd_rmap[slot] * VN;
    }
  }
  if (do_upd)
    mx_points_vec<EPT, VC, HM, true, NX>(a, lds, do_run, so, th, toff, roff, rcur_off, mix, w, nf);
  else
    mx_points_vec<EPT, VC, HM, false, NX>(a, lds, do_run, so, th, toff, roff, rcur_off, mix, w, nf);
  if (!do_run) return;
  T2 *ltw = lds + P * L * ls;
  stage_tables(ltw, a.px);
  __syncthreads();
  // radix-4 stages: at about 1024 elements a radix-16 stage would leave
  // three waves in four idle
  if constexpr (NX > 0)
    fft_lines_ct<NX, EPT, false, 2>(lds, (const T2 *)ltw);
  else
    fft_lines<EPT, false, 5>(lds, P * L, ls, a.pxr, ltw);
  const float dv = (float)a.dv;
  T2 *__restrict__ Z = soff(a.Z, so);
  // EPT consecutive (pack, point) elements per thread, two per 16-byte
  // store (nel is even: a pair never straddles two packs)
#pragma unroll
  for (int k = 0; k < EPT / 2; ++k) {
    const int f = (int)threadIdx.x * EPT + 2 * k;
    if (f < P * nel) {
      const int p = f / nel, e = f - p * nel;
      const int i = e / nx, j = e - i * nx;
      const float sc = dv * (float)jsign(j, a.px.fixsign);
      const T2 z0 = lds[(p * L + i) * ls + lpad(j)], z1 = lds[(p * L + i) * ls + lpad(j + 1)];
      *reinterpret_cast<float4 *>(Z + p * N + base + e) = make_float4(z0.x * sc, z0.y * sc, -z1.x * sc, -z1.y * sc);
    }
  }
}

template <typename R, int EPT, int VC, int HM>
__global__ __launch_bounds__(kThreads) void k_mx_closure_fx(IterArgs<R, double> a) {
  // the refresh list of the next step starts empty (the refresh chain of
  // this step, launched before this pass, has read it)
  if (a.gate == 1 && a.rlist && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) a.rlist[0] = 0;
  const int sid = batch_solute(a.gate, a.rlist);  // this launch's solute (batched launches)
  if (sid < 0) return;
  const long so = (long)sid * a.astride;
  using T2 = typename C2<R>::T;
  const Ctl *ctl = soff(a.ctl, so);
  const int iter = ctl->cur;
  const bool mine = chain_active(ctl, a.gate);
  const bool do_upd = mine && ctl->upd_iter == iter;
  const bool do_run = mine && ctl->run_iter == iter;
  if (!do_upd && !do_run) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int nx = a.nx, L = a.Lc, ls = line_stride(nx), P = a.P;
  const long N = a.N, VN = (long)a.V * N;
  const long base = (long)blockIdx.x * L * nx;
  const int nel = L * nx;
  // update coefficients by age; the residual of each slot through the ring
  // map of this step; slots beyond the filled ones get a zero coefficient
  // and a valid address
  int w = 0, nf = 0;
  double mix = 0.0;
  double th[HM];
  long toff[HM], roff[HM];
  long rcur_off = 0;
#pragma unroll
  for (int k = 0; k < HM; ++k) {
    th[k] = 0.0;
    toff[k] = roff[k] = 0;
  }
  if (do_upd) {
    nf = ctl->upd_nf;
    w = ctl->upd_w;
    mix = ctl->upd_mix;
    const int nh = ctl->nhist;
    rcur_off = (long)ctl->upd_rcur * VN;
#pragma unroll
    for (int k = 0; k < HM; ++k) {
      const int slot = (((w - 1 - k) % nh) + nh) % nh;
      th[k] = (k < nf) ? ctl->upd_theta[slot] : 0.0;
      toff[k] = (long)slot * VN;
      roff[k] = (long)ctl->upd_rmap[slot] * VN;
    }
  }
  if (do_upd)
    mx_points<R, EPT, VC, HM, true>(a, lds, do_run, th, toff, roff, rcur_off, mix, w, nf);
  else
    mx_points<R, EPT, VC, HM, false>(a, lds, do_run, th, toff, roff, rcur_off, mix, w, nf);
  if (!do_run) return;
  T2 *ltw = lds + P * L * ls;
  stage_tables(ltw, a.px);
  __syncthreads();
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
// 5. inverse x + residual into the ring + Anderson inner products
// ---------------------------------------------------------------------
template <typename R, int EPT, int HM, int VC>
__global__ __launch_bounds__(kThreads) void k_mx_resid(IterArgs<R, double> a) {
  const int sid = batch_solute(a.gate, a.rlist);  // this launch's solute (batched launches)
  if (sid < 0) return;
  const long so = (long)sid * a.astride;
  using T2 = typename C2<R>::T;
  constexpr bool F64 = sizeof(R) == sizeof(double);
  constexpr int NV = 1 + HM + HM * (HM + 1) / 2;
  const Ctl *ctl = soff(a.ctl, so);
  if (!gate_open(ctl, a.gate)) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  const int nx = a.nx, L = a.Lx, ls = line_stride(nx), V = a.V, P = a.P;
  const long N = a.N, VN = (long)V * N;
  const int nel = L * nx;
  const int groups = (int)((long)a.ny * a.nz / L);
  const int nf = ctl->filled < HM ? ctl->filled : HM;
  long hoff[HM];
#pragma unroll
  for (int s = 0; s < HM; ++s) hoff[s] = (long)ctl->rmap[s < nf ? s : 0] * VN;
  float *__restrict__ rw = soff(a.rring, so) + (long)ctl->rcur * VN;
  const float *__restrict__ rr = soff(a.rring, so);
  T2 *ltw = lds + P * L * ls;
  const T2 *pre = ltw + nx;
  stage_tables(ltw, a.px);  // visible after the first barrier of the group loop
  const R ivol = (R)a.ivol;
  T2 *__restrict__ Z = soff(a.Z, so);
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
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
  if (RISM_MXR_ZPF) load_z(blockIdx.x, threadIdx.x);
  for (int grp = blockIdx.x; grp < groups; grp += gridDim.x) {
    const long base = (long)grp * nel;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    __syncthreads();
    if (!RISM_MXR_ZPF) load_z(grp, tid);
    constexpr int PMIN = VC <= 2 ? 1 : (VC <= 4 ? 2 : 3);
    constexpr int KPT = (EPT + PMIN - 1) / PMIN;
    constexpr int CH = KPT < RISM_RESID_CH ? KPT : RISM_RESID_CH;
    constexpr bool PREH = CH * VC * HM <= 16;
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
    int tidf = tid;
    asm volatile("" : "+v"(tidf));
    if constexpr (F64 || RISM_MXR_PLAN4)
      fft_lines<EPT, true, 5>(lds, P * L, ls, a.pxr, ltw, tidf);
    else
      fft_lines<EPT, true>(lds, P * L, ls, a.px, ltw, tidf);
    int tidr = tid;
    asm volatile("" : "+v"(tidr));
    if (RISM_MXR_ZPF && grp + (int)gridDim.x < groups) load_z(grp + (int)gridDim.x, tidr);
#pragma unroll
    for (int k0 = 0; k0 < KPT; k0 += CH) {
      double hv[CH][VC + 1];
      double ov[CH][VC];  // FP64: A = T - u (g = G(A)); FP32: B
      float hq[CH][VC][PREH ? HM : 1];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const unsigned e = (unsigned)min(tidr + (k0 + c) * kThreads, nel - 1);
        const int i = e / nx, x = e - i * nx;
        const T2 pw = pre[x];
#pragma unroll
        for (int v = 0; v < VC; ++v)
          if (v < V) {
            const long vb = (long)v * N + base;
            if constexpr (F64) ov[c][v] = lane_at(soff(a.ts, so) + vb, e) - (double)lane_at(soff(a.u, so) + vb, e);
            else ov[c][v] = (double)lane_at(soff(a.bq, so) + vb, e);
#pragma unroll
            for (int s = 0; s < (PREH ? HM : 0); ++s)
              hq[c][v][PREH ? s : 0] = lane_at(rr + hoff[s] + vb, e);
          }
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
#pragma unroll
          for (int v = 0; v < VC; ++v) {
            if (v < V) {
              const long vb = (long)v * N + base;
              double rd;
              if constexpr (F64) {
                rd = hv[c][v] + 1.0 - closure_g<double>(ov[c][v], a.closure);  // tr: R = h + 1 - g
                lane_at(soff(a.rref, so) + vb, e) = (float)rd;
              } else {
                rd = hv[c][v] + ov[c][v];  // R = K dc + B
              }
              const float rq = (float)rd;
              lane_at(rw + vb, e) = rq;
              // the dot products see the stored residual
              rd = (double)rq;
              const double wgt = a.rho_inv[v];
              acc[0] += rd * rd;
              double d[HM];
#pragma unroll
              for (int s = 0; s < HM; ++s)
                d[s] = (s < nf) ? rd - (double)(PREH ? hq[c][v][PREH ? s : 0] : lane_at(rr + hoff[s] + vb, e)) : 0.0;
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
      if (RISM_MXR_SB) __builtin_amdgcn_sched_barrier(0);
    }
  }
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
// 5'. the FP32 chain's residual pass with W consecutive points per thread:
// the operands of a point (B, the history residuals) and the residual it
// writes move as 16-byte vectors, and every operand of the group is in
// flight while the inverse transform runs.  Same arithmetic as k_mx_resid;
// the inner products are summed in another order.
// ---------------------------------------------------------------------
template <int EPT, int HM, int VC, int NX>
__global__ __launch_bounds__(kThreads) void k_mx_resid_vec(IterArgs<float, double> a) {
  const int sid = batch_solute(a.gate, a.rlist);
  if (sid < 0) return;
  const long so = (long)sid * a.astride;
  using T2 = float2;
  constexpr int NV = 1 + HM + HM * (HM + 1) / 2;
  // W points of every site per thread: one pack (V <= 2) or two (V <= 4)
  // share the EPT elements of the transform
  static_assert(VC == 2 || VC == 4, "one or two packs");
  constexpr int W = VC <= 2 ? EPT : EPT / 2;
  const Ctl *ctl = soff(a.ctl, so);
  if (!gate_open(ctl, a.gate)) return;
  extern __shared__ unsigned char smem_raw[];
  T2 *lds = reinterpret_cast<T2 *>(smem_raw);
  // NX > 0: compile-time line length (a power of two) on a tile of exactly
  // EPT * kThreads elements, VC / 2 packs
  const int nx = NX > 0 ? NX : a.nx, P = NX > 0 ? VC / 2 : a.P;
  const int L = NX > 0 ? EPT * kThreads / (VC / 2 * NX) : a.Lx, ls = line_stride(nx);
  const int V = a.V;
  const long N = a.N, VN = (long)V * N;
  const int nel = L * nx;  // a multiple of 4 (nx is)
  const int groups = (int)((long)a.ny * a.nz / L);
  const int nf = ctl->filled < HM ? ctl->filled : HM;
  long hoff[HM];
#pragma unroll
  for (int s = 0; s < HM; ++s) hoff[s] = (long)ctl->rmap[s < nf ? s : 0] * VN;
  float *__restrict__ rw = soff(a.rring, so) + (long)ctl->rcur * VN;
  const float *__restrict__ rr = soff(a.rring, so);
  const float *__restrict__ bq = soff(a.bq, so);
  T2 *ltw = lds + P * L * ls;
  const T2 *pre = ltw + nx;
  stage_tables(ltw, a.px);  // visible after the first barrier of the group loop
  const float ivol = (float)a.ivol;
  T2 *__restrict__ Z = soff(a.Z, so);
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  // spectrum: EPT consecutive (pack, point) elements per thread, two per
  // 16-byte load (nel is even, so a pair never straddles two packs)
  float4 zb[EPT / 2];
  auto load_z = [&](int g, int lane) {
    const long gb = (long)g * nel;
#pragma unroll
    for (int k = 0; k < EPT / 2; ++k) {
      const int f = min(lane * EPT + 2 * k, P * nel - 2);
      const int p = f / nel, e = f - p * nel;
      zb[k] = *reinterpret_cast<const float4 *>(Z + p * N + gb + e);
    }
  };
  load_z(blockIdx.x, threadIdx.x);
  for (int grp = blockIdx.x; grp < groups; grp += gridDim.x) {
    const long base = (long)grp * nel;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    // this thread's points: e0 .. e0 + W - 1 (clamped copies past the end)
    const bool act = tid * W < nel;
    const unsigned e0 = (unsigned)(act ? tid * W : nel - W);
    // operands of the group, in flight during the transform
    float ob[VC][W], oh[VC][HM][W];
#pragma unroll
    for (int v = 0; v < VC; ++v)
      if (v < V) {
        const long vb = (long)v * N + base;
        ldv<W>(ob[v], bq + vb, e0);
#pragma unroll
        for (int s = 0; s < HM; ++s) ldv<W>(oh[v][s], rr + hoff[s] + vb, e0);
      }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < EPT / 2; ++k) {
      const int f = tid * EPT + 2 * k;
      if (f < P * nel) {
        const int p = f / nel, e = f - p * nel;
        const int i = e / nx, j = e - i * nx;  // j even: j and j + 1 share the line
        const float sg = (float)jsign(j, a.px.fixsign);
        lds[(p * L + i) * ls + lpad(j)] = mk<T2>(zb[k].x * sg, zb[k].y * sg);
        lds[(p * L + i) * ls + lpad(j + 1)] = mk<T2>(-zb[k].z * sg, -zb[k].w * sg);
      }
    }
    __syncthreads();
    int tidf = tid;
    asm volatile("" : "+v"(tidf));
    // at 1024 elements a radix-16 stage would leave three waves in four
    // idle: the radix-4 plan then (its twiddle table is the same)
    if constexpr (NX > 0)
      fft_lines_ct<NX, EPT, true, 2>(lds, (const T2 *)ltw, tidf);
    else if constexpr (EPT <= 4)
      fft_lines<EPT, true, 5>(lds, P * L, ls, a.pxr, ltw, tidf);
    else
      fft_lines<EPT, true>(lds, P * L, ls, a.px, ltw, tidf);
    int tidr = tid;
    asm volatile("" : "+v"(tidr));
    if (grp + (int)gridDim.x < groups) load_z(grp + (int)gridDim.x, tidr);
    float rq[VC][W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const int e = (int)e0 + j;
      const int i = e / nx, x = e - i * nx;
      const T2 pw = pre[x];
      float hv[VC + 1];
#pragma unroll
      for (int p = 0; p < (VC + 1) / 2; ++p)
        if (p < P) {
          const T2 z = cscale(cmulc(lds[(p * L + i) * ls + lpad(x)], pw), ivol);
          hv[2 * p] = z.x;
          hv[2 * p + 1] = z.y;
        }
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        if (v < V) {
          // R = K dc + B; the dot products see the stored residual
          const float q = (float)((double)hv[v] + (double)ob[v][j]);
          rq[v][j] = q;
          if (act) {
            const double rd = (double)q;
            const double wgt = a.rho_inv[v];
            acc[0] += rd * rd;
            double d[HM];
#pragma unroll
            for (int s = 0; s < HM; ++s) d[s] = (s < nf) ? rd - (double)oh[v][s][j] : 0.0;
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
    if (act)
#pragma unroll
      for (int v = 0; v < VC; ++v)
        if (v < V) stv<W>(rw + (long)v * N + base, e0, rq[v]);
  }
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

// compile-time x length for the FP32 chain's x passes: a power of two in
// 64..512 on a tile of exactly 4 * kThreads elements (else 0: runtime plan)
inline int ct_xlen(int nx, long tile_elems, int ept) {
  if (ept != 4 || tile_elems != 4L * kThreads) return 0;
  return (nx == 64 || nx == 128 || nx == 256 || nx == 512) ? nx : 0;
}
#define RISM_NX_SWITCH(nxc, CT, ...)                            \
  switch (nxc) {                                                \
    case 64: { constexpr int NXV = 64; CT; } break;             \
    case 128: { constexpr int NXV = 128; CT; } break;           \
    case 256: { constexpr int NXV = 256; CT; } break;           \
    case 512: { constexpr int NXV = 512; CT; } break;           \
    default: { __VA_ARGS__; } break;                            \
  }

template <typename R>
hipError_t launch_mx_closure_fx_impl(const IterArgs<R, double> &a, const LaunchGeom &g, int hm, hipStream_t st) {
  const dim3 grid(g.cgroups, a.nsol > 1 ? a.nsol : 1), block(kThreads);
#if RISM_MXC_VEC
  // k_mx_closure_vec when a thread's W points of every site tile the group
  // exactly (one or two packs, up to 16 elements per thread)
  if constexpr (sizeof(R) == sizeof(float)) {
    const int W = a.V <= 2 ? g.ept_c : g.ept_c / 2;
    if (a.V <= 4 && (a.Lc * a.nx) % W == 0 && W <= 8) {
      const int nxc = ct_xlen(a.nx, (long)a.P * a.Lc * a.nx, g.ept_c);
#define RISM_MXCV(HMV)                                                                                              \
  if (a.V <= 2) {                                                                                                   \
    constexpr int VC = 2;                                                                                           \
    RISM_NX_SWITCH(nxc, hipLaunchKernelGGL((k_mx_closure_vec<4, VC, HMV, NXV>), grid, block, g.lds_c, st, a),      \
                   RISM_EPTX_SWITCH(g.ept_c, hipLaunchKernelGGL((k_mx_closure_vec<E, VC, HMV, 0>), grid, block, \
                                                                g.lds_c, st, a)));                                  \
  } else {                                                                                                          \
    constexpr int VC = 4;                                                                                           \
    RISM_NX_SWITCH(nxc, hipLaunchKernelGGL((k_mx_closure_vec<4, VC, HMV, NXV>), grid, block, g.lds_c, st, a),      \
                   RISM_EPTX_SWITCH(g.ept_c, hipLaunchKernelGGL((k_mx_closure_vec<E, VC, HMV, 0>), grid, block, \
                                                                g.lds_c, st, a)));                                  \
  }
      switch (hm) {
        case 2: RISM_MXCV(2); break;
        case 5: RISM_MXCV(5); break;
        case 10: RISM_MXCV(10); break;
        default: return hipErrorInvalidValue;
      }
#undef RISM_MXCV
      return hipGetLastError();
    }
  }
#endif
#define RISM_MXC(HMV) \
  RISM_VC_SWITCH(a.V, RISM_EPTX_SWITCH(g.ept_c, hipLaunchKernelGGL((k_mx_closure_fx<R, E, VC, HMV>), grid, block, g.lds_c, st, a)))
  switch (hm) {
    case 2: RISM_MXC(2); break;
    case 5: RISM_MXC(5); break;
    case 10: RISM_MXC(10); break;
    default: return hipErrorInvalidValue;
  }
#undef RISM_MXC
  return hipGetLastError();
}

template <typename R>
hipError_t launch_mx_resid_impl(const IterArgs<R, double> &a, const LaunchGeom &g, int hm, hipStream_t st) {
  const dim3 grid(g.xblocks, a.nsol > 1 ? a.nsol : 1), block(kThreads);
#if RISM_MXR_VEC
  // k_mx_resid_vec when a thread's W points of every site tile the group
  // exactly and its operands (V (1 + H) W floats) fit in 48 registers
  if constexpr (sizeof(R) == sizeof(float)) {
    const int W = a.V <= 2 ? g.ept_x : g.ept_x / 2;
    if (a.V <= 4 && (a.Lx * a.nx) % W == 0 && a.V * (1 + hm) * W <= 48) {
      const int nxc = ct_xlen(a.nx, (long)a.P * a.Lx * a.nx, g.ept_x);
#define RISM_MXRV(HMV)                                                                                            \
  if (a.V <= 2) {                                                                                                 \
    constexpr int VC = 2;                                                                                         \
    RISM_NX_SWITCH(nxc, hipLaunchKernelGGL((k_mx_resid_vec<4, HMV, VC, NXV>), grid, block, g.lds_x, st, a),      \
                   RISM_EPTX_SWITCH(g.ept_x, hipLaunchKernelGGL((k_mx_resid_vec<E, HMV, VC, 0>), grid, block, \
                                                                g.lds_x, st, a)));                                \
  } else {                                                                                                        \
    constexpr int VC = 4;                                                                                         \
    RISM_NX_SWITCH(nxc, hipLaunchKernelGGL((k_mx_resid_vec<4, HMV, VC, NXV>), grid, block, g.lds_x, st, a),      \
                   RISM_EPTX_SWITCH(g.ept_x, hipLaunchKernelGGL((k_mx_resid_vec<E, HMV, VC, 0>), grid, block, \
                                                                g.lds_x, st, a)));                                \
  }
      switch (hm) {
        case 2: RISM_MXRV(2); break;
        case 5: RISM_MXRV(5); break;
        case 10: RISM_MXRV(10); break;
        default: return hipErrorInvalidValue;
      }
#undef RISM_MXRV
      return hipGetLastError();
    }
  }
#endif
#define RISM_MXR(HMV) \
  RISM_VC_SWITCH(a.V, RISM_EPTX_SWITCH(g.ept_x, hipLaunchKernelGGL((k_mx_resid<R, E, HMV, VC>), grid, block, g.lds_x, st, a)))
  switch (hm) {
    case 2: RISM_MXR(2); break;
    case 5: RISM_MXR(5); break;
    case 10: RISM_MXR(10); break;
    default: return hipErrorInvalidValue;
  }
#undef RISM_MXR
  return hipGetLastError();
}

}  // namespace rism
