// rism_kernels.h -- device-side data layout and host launchers for the
// 3D-RISM iterate path on gfx950.
//
// HBM layout (N = nx*ny*nz, x fastest: ip = x + nx*(y + ny*z), V sites,
// P = ceil(V/2) complex packs, H history slots):
//   ts   S[V][N]      short-range state  ts_v = t_v - q_v*cf*fr
//   res  S[V][N]      residual R = h + 1 - g  (iterate.cu's dtr)
//   hist S[H][V][N]   Anderson history of ts and of R (two arrays)
//   u    R[V][N]      beta*u_LJ + q_v*cf*(de - fr): the closure exponent
//                     is ts - u = t - beta*u_LJ - q_v*cf*de (calculate_cuda.cu:6)
//   fr   R[N]         long-range part of the solute field (cal_Coulomb.cu:117)
//   fkh  C[nz][ny/2][nx]  its transform on the half grid y < ny/2
//   Z    C[P][N]      work spectrum: two real sites per complex value
//   chi  R4[V*V][ntab-1]  natural-spline coefficients of chi_vv'(k)
// S is the state precision (double), R the field/FFT precision (double for
// RISM_PREC_FP64, float for the mixed path's FP32 chain; its state is the
// FP64 reference plus FP32 change of IterArgs::A, dl).
#pragma once

#include <hip/hip_runtime.h>

#include "fft_core.h"

namespace rism {

constexpr int kMaxV = 8;    // reduced solvent sites
constexpr int kAtomChunk = 128;  // solute atoms staged in LDS per pass of the grid x atom kernels
constexpr int kHMax = 16;   // Anderson history slots
constexpr int kRmsRecent = 16;  // rms values the control block carries (Ctl::rms_recent)

// 1/sqrt(x) in FP64 for x > 0: the hardware estimate refined by two Newton
// steps y <- y + y (1 - x y^2) / 2 (each roughly doubles the correct bits),
// within a few ulp of 1/sqrt(x) and far cheaper than sqrt plus a division.
__device__ __forceinline__ double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// Solver control block in device memory.  Written by the single-workgroup
// solve kernel, read by the field kernels to decide whether they run.
struct Ctl {
  int cur;        // current iteration index; advanced by the solve kernel, so
                  // one iteration's launches are identical from step to step
                  // (replayable as a HIP graph)
  int run_iter;   // iteration index whose closure/FFT/residual pass runs (-1 none)
  int upd_iter;   // iteration whose first pass applies the pending update (-1 none)
  int status;     // -1 running, else RISM_SOLVE_*
  int step;       // steps done in this charge-up stage
  int total;      // steps over all stages
  int maxstep;
  int count;      // remaining Picard warm-up steps (anderson.cu:43-49)
  int write;      // next history slot to write
  int filled;     // history slots holding data
  int nhist;
  int upd_nf;     // slots 0..upd_nf-1 enter the pending update
  int upd_w;      // slot the pending update writes
  int upd_hold;   // step of the update a step-limit stop leaves pending: the
                  // host applies it once (update-only launch); no first pass
                  // applies it on its own, so launches queued beyond the
                  // stop leave the state alone
  double upd_theta[kHMax];
  double upd_mix;
  double tol, m, mp;
  double rms;
  double npts;    // V*N
  // ---- mixed precision (ring = 1) ----
  // The residuals live in a ring of nhist + 1 slots: the residual pass
  // writes the current R into slot rcur, and the history slot s (0..nhist-1)
  // is ring slot rmap[s].  When the solve kernel files R into history slot
  // w it only relabels (rmap[w] <- rcur, and rcur takes the freed slot); the
  // pending update reads through the map of its own step (upd_rcur,
  // upd_rmap).
  int ring;
  int rcur, upd_rcur;
  int rlast;      // ring slot of the last residual pass's R
  int rmap[kHMax + 1], upd_rmap[kHMax + 1];
  // Refinement schedule, decided on the device after every step: the step
  // refresh_iter runs the FP64 chain (a new reference A = x, and R_ref),
  // every other step the FP32 chain on differences to that reference.  A
  // refresh follows when the residual has fallen 10x below the one at the
  // last refresh, or has not improved for 24 steps (the FP32 floor).
  int refresh_iter;
  int since_ref, best_at;
  double ref_rms, best_rms;
  // rms of the last kRmsRecent steps, step k (over all stages, from 0) in
  // slot k % kRmsRecent: the host reads them with the control block after
  // each chunk of at most kRmsRecent steps, with no copy of its own
  double rms_recent[kRmsRecent];
};

// Which chain of kernels runs at the current step.  gate 0: the FP64 path
// (every step); 1: the FP32 chain of the mixed path (steps that are no
// refresh); 2: its FP64 refresh chain.
__device__ __forceinline__ bool chain_active(const Ctl *c, int gate) {
  return gate == 0 || ((c->refresh_iter == c->cur) == (gate == 2));
}
__device__ __forceinline__ bool gate_open(const Ctl *c, int gate) {
  return c->run_iter == c->cur && chain_active(c, gate);
}

// The solute of a workgroup of a batched launch: blockIdx.y; in the refresh
// chain's compacted launches (gate 2 with a refresh list
// rl = {count, capacity, solute ...}) entry blockIdx.y of the list, and -1
// (no solute: the workgroup exits) past its end.
__device__ __forceinline__ int batch_solute(int gate, const int *rl) {
  if (gate != 2 || !rl) return (int)blockIdx.y;
  const int n = min(rl[0], rl[1]);
  return (int)blockIdx.y < n ? rl[2 + blockIdx.y] : -1;
}

// a per-solute pointer of a batched launch (o = solute * astride bytes)
template <typename P> __device__ __forceinline__ P soff(P p, long o) {
  return (P)((const char *)p + o);
}

// The mixed path's closure argument at the reference, A = T - u, rounded to
// FP32 for storage.  The FP32 chain only forms G(A + dl) - G(A) from it, whose
// error from rounding A is second order (G''(A) dA dl); the refresh chain
// evaluates T - u in FP64 itself.  Inside the solute's cores u reaches 1e30,
// so the value is clamped to the FP32 range (G is 0 there either way).
__device__ __forceinline__ float ref_f32(double a) {
  return (float)fmin(fmax(a, -3.0e38), 3.0e38);
}

// Residual / history storage type of the FP64 path's passes (the mixed
// path's state is described with IterArgs::A below).
template <typename R> struct HistT;
template <> struct HistT<double> { using T = double; };
template <> struct HistT<float> { using T = float; };

template <typename R> struct R4T;
template <> struct R4T<double> { using T = double4; };
template <> struct R4T<float> { using T = float4; };

template <typename R, typename S>
struct IterArgs {
  int nx, ny, nz;
  long N;
  int V, P;
  S *ts;
  typename HistT<R>::T *res, *hist_t, *hist_r;
  const R *u, *fr;
  typename C2<R>::T *Z;
  const typename C2<R>::T *fkh;
  const typename R4T<R>::T *chi;
  int ntab;
  double dk;
  double box[3];
  double qcf[kMaxV];
  double rho_inv[kMaxV];
  double dv, ivol;
  int closure;
  Ctl *ctl;
  double *partials;
  int nv;           // accumulators per block in the residual pass
  AxisPlan px, py, pz;
  AxisPlan pxr;     // x plan with radices <= 5 (closure pass, FP64 residual pass)
  int Lx;           // x lines per workgroup of the residual pass
  int Lc;           // x lines per workgroup of the closure/update pass
  int Ty, Tz;       // tile widths (consecutive x) of the y / z passes
  int xblocks;      // workgroups of the x passes (grid-stride over line groups)
  int store_h;      // residual pass also writes real-space h into Z
  int href_only;        // residual pass: store h only (the final evaluation)
  int use_fk;           // z pass subtracts q_v cf fk(k)
  double *rms_hist;
  // ---- mixed precision on differences to an FP64 reference (gate 1, 2) ----
  // ts = T + dl: T [V][N] (FP64, in the ts field) is ts at the last refresh,
  // dl [V][N] (FP32) the change since; A [V][N] (FP32, ref_f32) = T - u, the
  // closure argument at the refresh; rref [V][N] (FP32) the refresh step's FP64
  // residual, hx [H][V][N] (FP32) the step increments of ts, rring
  // [H+1][V][N] (FP32) the residual ring (Ctl::rmap).  The FP32 chain
  // transforms dc = G(A + dl) - G(A) - dl (the change of c_s, whose
  // long-range part is constant) and the residual is R = K dc + B with
  // B = R_ref - (G(A + dl) - G(A)) from the closure pass.
  int gate;             // chain_active() selector
  // ---- batched launches: nsol solutes of one grid class in one launch ----
  // blockIdx.y selects the solute; every per-solute field (state, history,
  // potentials, spectrum, control block, partials) of solute s sits
  // s * astride bytes after solute 0's; the plans and the chi table are
  // the same for all and are read from solute 0.
  long astride;
  int nsol;
  // Refresh list of a batched mixed solve, {count, capacity, solute ...}:
  // the FP32 chain's solve kernel files the solutes whose next step is a
  // refresh (up to capacity; the others try again at their next step), the
  // refresh chain launches capacity rows of workgroups mapped through it
  // (batch_solute), and the FP32 chain's closure pass empties it again.
  // nullptr: blockIdx.y is the solute.
  int *rlist;
  float *A;
  float *dl;
  float *rref;
  float *bq;            // B, closure pass -> residual pass
  float *hx, *rring;
};

// launch configuration chosen on the host for a grid
struct LaunchGeom {
  int Lx, Ty, Tz;
  int ept_x, ept_y, ept_z;  // elements-per-thread class (4, 8, 16, 32)
  int xgroups;              // number of x line groups (ny*nz / Lx)
  int xblocks;              // residual-pass grid size
  size_t lds_x, lds_y, lds_z;  // bytes, including any staged tables
  // closure/update pass: its own x line group (one workgroup per group);
  // it reduces nothing, so its shape never changes a result
  int Lc, ept_c, cgroups;
  size_t lds_c;
};

// ---- iteration launchers (k_iterate.hip) ----
template <typename R, typename S>
hipError_t launch_closure_fx(const IterArgs<R, S> &a, const LaunchGeom &g, bool update, int hm,
                             hipStream_t st);
template <typename R, typename S>
hipError_t launch_fft_y(const IterArgs<R, S> &a, const LaunchGeom &g, bool inverse, hipStream_t st);
template <typename R, typename S>
hipError_t launch_fft_z_oz(const IterArgs<R, S> &a, const LaunchGeom &g, hipStream_t st);
template <typename R, typename S>
hipError_t launch_fx_resid(const IterArgs<R, S> &a, const LaunchGeom &g, int hm, hipStream_t st);
hipError_t launch_anderson_solve(Ctl *ctl, const double *partials, int nblk, int nv, int hm,
                                 double *rms_hist, hipStream_t st, int gate = 0, long astride = 0, int nsol = 1,
                                 int *rlist = nullptr);
// mixed precision: closure / residual passes of the FP32 chain (R = float)
// and of the FP64 refresh chain (R = double), k_mixed_f32.hip / k_mixed_f64.hip
template <typename R>
hipError_t launch_mx_closure_fx(const IterArgs<R, double> &a, const LaunchGeom &g, int hm, hipStream_t st);
template <typename R>
hipError_t launch_mx_resid(const IterArgs<R, double> &a, const LaunchGeom &g, int hm, hipStream_t st);
// t_s <-> state conversions of the mixed path: dir 0: T = ts, A = ts - u,
// dl = 0; dir 1: ts = T + dl
hipError_t launch_mx_convert(double *ts, double *T, float *A, float *dl, const double *u, long n, int dir,
                             hipStream_t st);
// ---- adjoint of the converged fixed point (k_adjoint.hip) ----
struct AdjArgs {
  long N;
  int nx, ny, nz, V, closure, natu, ntab;
  double box[3], dr[3], dv, dk, kt, ikbt, ubeta;
  double qcf[kMaxV], qv[kMaxV], rho[kMaxV], rho_inv[kMaxV];
  const double *ts, *u, *fr;   // converged state, FP64 potential, long-range field
  const double2 *hz;           // real-space h, packed two sites per complex
  const double4 *chi;          // spline table [V*V][ntab-1]
  double *gp, *alpha, *beta, *b;       // [V][N]
  double *lam, *res, *hist_l, *hist_r; // adjoint state, residual, history [H][V][N]
  double *ubar, *debar, *frbar;        // field adjoints
  Ctl *ctl;
  double *partials;
  const double4 *atoms;        // x, y, z, q
  const double2 *lj;           // [V][natu] (sigma_uv^2, 4 eps_uv)
  const double2 *ljd;          // [V][natu] (1 / (2 sigma_uv), 4 eps_uv / (2 eps_u))
  const double2 *phase;        // [natu][nx+ny+nz] exp(-i k R) per axis
};
hipError_t launch_adj_fields(const AdjArgs &a, hipStream_t st);
// T2: the spectrum of a K^T application, double2, or float2 in the mixed path
template <typename T2>
hipError_t launch_adj_pack(const AdjArgs &a, const double *x, const double *y, T2 *Z, hipStream_t st);
template <typename T2>
hipError_t launch_adj_kspace(const AdjArgs &a, T2 *Z, double2 *Y, hipStream_t st);
hipError_t launch_adj_rhs(const AdjArgs &a, const double2 *Z, hipStream_t st);
hipError_t launch_adj_update(const AdjArgs &a, int hm, hipStream_t st);
template <typename T2>
hipError_t launch_adj_resid(const AdjArgs &a, const T2 *Z, int hm, int nblk, hipStream_t st);
hipError_t launch_adj_bar(const AdjArgs &a, const double2 *Z, hipStream_t st);
int adj_grad_blocks(long N);
hipError_t launch_adj_grad(const AdjArgs &a, const double2 *Y, double *part6, double *part4, double *out6,
                           double *out4, hipStream_t st);

// standalone 3D FFT on a caller buffer (reference FFT3D::execute seam)
template <typename R>
hipError_t launch_fft3d(typename C2<R>::T *data, const AxisPlan &px, const AxisPlan &py,
                        const AxisPlan &pz, const LaunchGeom &g, int nx, int ny, int nz, double dv,
                        double ivol, int dir, hipStream_t st);

// ---- potentials (k_potential.hip) ----
struct PotArgs {
  int nx, ny, nz;
  long N;
  int V, natu;
  double dr[3];
  double box[3];
  const double4 *atoms;  // x, y, z, q (centred coordinates)
  const double2 *lj;     // [V][natu]: (sigma_uv^2, 4*eps_uv)
  const double *sigh;    // [V][natu]: sigma_uv / 2 (the SE's exclusion radius)
  double ikbt, ubeta;
  double *ulj, *de, *fr, *sr;  // sr = de - fr
  // [V][N] per-site sums over the atoms for the SE (cal_euv.cu:35-76):
  // 4 eps (s^12 - s^6) and q_u / r, both over r >= sigma_uv / 2
  double *sel, *seq;
};
hipError_t launch_potentials(const PotArgs &a, hipStream_t st);
hipError_t launch_fk_half(int nx, int ny, int nz, int natu, const double4 *atoms, const double *box,
                          double ubeta, double2 *phase /*[natu*(nx+ny+nz)]*/, double2 *fkh,
                          hipStream_t st);
template <typename R>
hipError_t launch_combine_u(const double *ulj, const double *sr, const double *qcf, int V, long N,
                            R *u, hipStream_t st);
template <typename R>
hipError_t launch_convert(const double *src, R *dst, long n, hipStream_t st);
template <typename R>
hipError_t launch_convert_c(const double2 *src, typename C2<R>::T *dst, long n, hipStream_t st);
template <typename S>
hipError_t launch_axpy_fr(S *x, const double *fr, const double *coef, int V, long N, int nslots,
                          hipStream_t st);  // x[s][v] += coef[v] * fr
hipError_t launch_widen(const float *src, double *dst, long n, hipStream_t st);

// ---- thermodynamics (k_thermo.hip) ----
template <typename R, typename S>
hipError_t launch_gfield(const IterArgs<R, S> &a, double *g, hipStream_t st);
template <typename R, typename S>
hipError_t launch_exchem(const IterArgs<R, S> &a, const double *frd, double *partials, int nblk,
                         hipStream_t st);
hipError_t launch_se_sum(long N, int V, const double *rho, const double *qv, const double *g, const double *sel,
                         const double *seq, double *partials, int nblk, hipStream_t st);
// RMDFT (cal_rmdft.cu): grid and spline tables (coefficients (y, c0, c1, c2)
// per segment)
struct RmdftGrid {
  int nx, ny, nz;
  long N;
  double box[3];
};
struct RmdftTabs {
  int V, ntab, nhs;
  double dk, rho0;
  double rho[kMaxV];
  const double4 *cvk;  // [V*V][ntab-1] on k_i = dk (i + 1)
  const double *khs;   // [nhs] knots of the hard-sphere tables
  const double4 *chs;  // [V][nhs-1]
  const double4 *whs;  // [V][nhs-1]
};
hipError_t launch_rm_unpack(const double2 *Z, long N, int v, double2 *out, double *partials, int nblk,
                            hipStream_t st);
hipError_t launch_rm_kmix(const RmdftGrid &g, const RmdftTabs &t, int v, const double2 *const *H, double2 *w1,
                          double2 *w2, hipStream_t st);
hipError_t launch_rm_fd(const double2 *Z, long N, int v, double a, double bex, double rho0, double2 *w2,
                        hipStream_t st);
hipError_t launch_rm_wmul(const RmdftGrid &g, const RmdftTabs &t, int v, double2 *w2, hipStream_t st);
hipError_t launch_rm_reduce(const double2 *Z, long N, int v, const double2 *a, int mode, double e0,
                            double *partials, int nblk, hipStream_t st);
hipError_t launch_euv_atoms(int nx, int ny, int nz, const double *dr, int V, int natu, const double4 *atoms,
                            const double2 *lj, const double *sigh, const double *qv, const double *rho,
                            const double *g, double *partials, int nblk, double *out, hipStream_t st);
hipError_t launch_qv(long N, int V, const double *rho, const double *qv, const double *g, double *out,
                     hipStream_t st);
// cal_grad sums for the atoms u0 .. u0+nb-1: partials [nblk][nb][6], results
// into out[u0 .. u0+nb-1][6]
hipError_t launch_grad(int nx, int ny, int nz, const double *dr, int V, int natu,
                       const double4 *atoms, const double2 *lj, const double *sigh,
                       const double *qv, const double *rho, const double *g, int u0, int nb,
                       double *partials, int nblk, double *out, hipStream_t st);

}  // namespace rism
