// fft_core.h -- LDS-resident batched Stockham FFT building blocks for gfx950.
//
// A workgroup owns a set of whole lines (one axis of the 3D grid) in LDS.
// Each radix stage reads every element once from LDS into registers, runs
// radix-F butterflies (F in {2,3,4,5,8,16}), applies the Stockham twiddle and
// writes the results back in autosorted order.  Stage parameters are
// runtime values (one kernel serves every supported length); the butterfly
// width and the per-thread butterfly count are compile-time so the data
// stays in VGPRs.  Twiddles come from a per-axis table tw[k] = exp(-2 pi i
// k / n) computed on the host in FP64.
//
// The 3D transform with the reference's half-shifted k grid
// (src-cuda/fft3d.cu:22-94) is separable: along each axis
//   F(j) = sum_x f(x) exp(-i k_j r_x),  r_x = (x - n/2) dr,
//   k_j = 2 pi (j - n/2 + 1/2) / L
//        = (-1)^(j + n/2) * DFT_n[ f(x) (-1)^x exp(-i pi (x - n/2) / n) ](j)
// so every pass multiplies by its axis's pre-twiddle before and a sign
// after (forward), or the sign before and the conjugate pre-twiddle after
// (inverse).  Those multiplies are folded into the passes' load/store.
#pragma once

#include <hip/hip_runtime.h>

namespace rism {

constexpr int kThreads = 256;
constexpr int kMaxStages = 12;

template <typename R> struct C2;
template <> struct C2<double> { using T = double2; };
template <> struct C2<float> { using T = float2; };

template <typename T2> __device__ __forceinline__ T2 mk(decltype(T2::x) a, decltype(T2::x) b) {
  T2 r; r.x = a; r.y = b; return r;
}
template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) { return mk<T2>(a.x + b.x, a.y + b.y); }
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) { return mk<T2>(a.x - b.x, a.y - b.y); }
template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  return mk<T2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T2> __device__ __forceinline__ T2 cmulc(T2 a, T2 b) {  // a * conj(b)
  return mk<T2>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T2> __device__ __forceinline__ T2 cscale(T2 a, decltype(T2::x) s) { return mk<T2>(a.x * s, a.y * s); }
template <typename T2> __device__ __forceinline__ T2 cconj(T2 a) { return mk<T2>(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename T2> __device__ __forceinline__ T2 rot90(T2 a) {
  return INV ? mk<T2>(-a.y, a.x) : mk<T2>(a.y, -a.x);
}

// Per-axis plan passed by value as a kernel argument.
struct AxisPlan {
  int n;
  int nst;
  int rad[kMaxStages];
  int s[kMaxStages];      // Stockham stride before the stage
  int m[kMaxStages];      // current length / radix
  int tstep[kMaxStages];  // n / current length
  int ls2[kMaxStages];    // log2(s) when s is a power of two, else -1
  int lb2[kMaxStages];    // log2(n / radix) when a power of two, else -1
  int fixsign;            // (-1)^(n/2): 1 when n % 4 == 0
  const void *tw;         // complex[n] exp(-2 pi i k/n)
  const void *pre;        // complex[n] (-1)^x exp(-i pi (x - n/2)/n)
};

// ---------------------------------------------------------------------
// small DFT kernels (in registers)
// ---------------------------------------------------------------------
template <int F, bool INV, typename T2> struct Dft;

template <bool INV, typename T2> struct Dft<2, INV, T2> {
  static __device__ __forceinline__ void run(T2 *v) {
    T2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <bool INV, typename T2> struct Dft<4, INV, T2> {
  static __device__ __forceinline__ void run(T2 *v) {
    T2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    T2 t2 = cadd(v[1], v[3]), t3 = rot90<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <bool INV, typename T2> struct Dft<3, INV, T2> {
  static __device__ __forceinline__ void run(T2 *v) {
    using R = decltype(T2::x);
    const R s3 = (R)0.86602540378443864676372317075294;  // sin(2 pi / 3)
    T2 t = cadd(v[1], v[2]);
    T2 mid = csub(v[0], cscale(t, (R)0.5));
    // (v1 - v2) * (sign * i * sqrt(3)/2), sign = -1 forward
    T2 d = cscale(rot90<INV>(csub(v[1], v[2])), s3);
    v[0] = cadd(v[0], t);
    v[1] = cadd(mid, d);
    v[2] = csub(mid, d);
  }
};

template <bool INV, typename T2> struct Dft<5, INV, T2> {
  static __device__ __forceinline__ void run(T2 *v) {
    using R = decltype(T2::x);
    const R c1 = (R)0.30901699437494742410229341718282;   // cos(2pi/5)
    const R c2 = (R)-0.80901699437494742410229341718282;  // cos(4pi/5)
    const R s1 = (R)0.95105651629515357211643933337938;   // sin(2pi/5)
    const R s2 = (R)0.58778525229247312916870595463907;   // sin(4pi/5)
    T2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    T2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    T2 x0 = v[0];
    T2 p1 = mk<T2>(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    T2 p2 = mk<T2>(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    // sign * i * (s1 b1 + s2 b2) and sign * i * (s2 b1 - s1 b2)
    T2 q1 = rot90<INV>(mk<T2>(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    T2 q2 = rot90<INV>(mk<T2>(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    v[0] = mk<T2>(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
  }
};

template <bool INV, typename T2> struct Dft<8, INV, T2> {
  static __device__ __forceinline__ void run(T2 *v) {
    using R = decltype(T2::x);
    const R h = (R)0.70710678118654752440084436210485;
    T2 e[4] = {v[0], v[2], v[4], v[6]};
    T2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, INV, T2>::run(e);
    Dft<4, INV, T2>::run(o);
    // w8^k with sign: w1 = (1 - i)/sqrt2 fwd, (1 + i)/sqrt2 inv
    T2 o1 = INV ? mk<T2>((o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h)
                : mk<T2>((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
    T2 o2 = rot90<INV>(o[2]);
    T2 o3 = INV ? mk<T2>(-(o[3].x + o[3].y) * h, (o[3].x - o[3].y) * h)
                : mk<T2>((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
  }
};

template <bool INV, typename T2> struct Dft<16, INV, T2> {
  static __device__ __forceinline__ void run(T2 *v) {
    using R = decltype(T2::x);
    // cos / sin of pi k / 8, k = 0..7
    const R cs[8] = {(R)1.0, (R)0.92387953251128675612818318939679, (R)0.70710678118654752440084436210485,
                     (R)0.38268343236508977172845998403040, (R)0.0, (R)-0.38268343236508977172845998403040,
                     (R)-0.70710678118654752440084436210485, (R)-0.92387953251128675612818318939679};
    const R sn[8] = {(R)0.0, (R)0.38268343236508977172845998403040, (R)0.70710678118654752440084436210485,
                     (R)0.92387953251128675612818318939679, (R)1.0, (R)0.92387953251128675612818318939679,
                     (R)0.70710678118654752440084436210485, (R)0.38268343236508977172845998403040};
    T2 e[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
    Dft<8, INV, T2>::run(e);
    Dft<8, INV, T2>::run(o);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      // w16^k: k = 4 is a quarter turn, k = 2 and 6 are the eighth turns
      // (a sum and a difference times 1/sqrt 2), the rest general
      T2 t;
      if (k == 0) {
        t = o[0];
      } else if (k == 4) {
        t = rot90<INV>(o[4]);
      } else if (k == 2 || k == 6) {
        const R h = cs[2];
        const T2 q = o[k];
        if (k == 2) t = INV ? mk<T2>((q.x - q.y) * h, (q.x + q.y) * h) : mk<T2>((q.x + q.y) * h, (q.y - q.x) * h);
        else t = INV ? mk<T2>(-(q.x + q.y) * h, (q.x - q.y) * h) : mk<T2>((q.y - q.x) * h, -(q.x + q.y) * h);
      } else {
        t = cmul(o[k], mk<T2>(cs[k], INV ? sn[k] : -sn[k]));
      }
      v[k] = cadd(e[k], t);
      v[k + 8] = csub(e[k], t);
    }
  }
};

// LDS position of element i of a line: one pad slot every 16 elements, so
// the stride-F writes of a Stockham stage and the column loads of the y/z
// passes spread over the banks instead of hitting a few of them.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
// padded line stride (odd, so consecutive lines also start on other banks)
__host__ __device__ constexpr int line_stride(int n) { return n + (n >> 4) + 1; }

// ---------------------------------------------------------------------
// one Stockham stage over all lines held in LDS (in place: load phase,
// barrier, store phase, barrier).  Element (line l, index i) lives at
// lds[l * ls + i].  KMAX = butterflies per thread (compile-time bound).
// ---------------------------------------------------------------------
template <int F, int KMAX, bool INV, typename T2>
__device__ __forceinline__ void fft_stage(T2 *lds, int nlines, int ls, int n, int s, int m,
                                          int tstep, int ls2, int lb2, const T2 *__restrict__ tw, int tid) {
  const int nbl = n / F;
  const int total = nlines * nbl;
  T2 v[KMAX][F];
  int lbase[KMAX], q0[KMAX], ppv[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const int bf = tid + k * kThreads;
    lbase[k] = -1;
    if (bf < total) {
      const int line = lb2 >= 0 ? (bf >> lb2) : bf / nbl;
      const int j = bf - line * nbl;
      const int pp = ls2 >= 0 ? (j >> ls2) : j / s;
      lbase[k] = line * ls;
      q0[k] = j - pp * s;
      ppv[k] = pp;
    }
  }
#pragma unroll
  for (int k = 0; k < KMAX; ++k)
    if (lbase[k] >= 0)
#pragma unroll
      for (int a = 0; a < F; ++a) v[k][a] = lds[lbase[k] + lpad(q0[k] + s * (ppv[k] + a * m))];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    if (lbase[k] >= 0) {
      Dft<F, INV, T2>::run(v[k]);
      const int o = q0[k] + s * (F * ppv[k]);
      lds[lbase[k] + lpad(o)] = v[k][0];
#pragma unroll
      for (int b = 1; b < F; ++b) {
        const T2 w = tw[ppv[k] * b * tstep];  // tw lives in LDS
        lds[lbase[k] + lpad(o + s * b)] = INV ? cmulc(v[k][b], w) : cmul(v[k][b], w);
      }
    }
  }
  __syncthreads();
}

// stage an axis plan's twiddle and phase tables (n complex each) into LDS
template <typename T2>
__device__ __forceinline__ void stage_tables(T2 *ltw, const AxisPlan &p) {
  for (int k = threadIdx.x; k < p.n; k += kThreads) {
    ltw[k] = ((const T2 *)p.tw)[k];
    ltw[p.n + k] = ((const T2 *)p.pre)[k];
  }
}

// Run the full plan on nlines lines of length p.n held in LDS.  EPT = upper
// bound of elements per thread (nlines * n <= kThreads * EPT); tw = the
// plan's twiddle table staged in LDS (stage_tables); tid = the lane's
// thread index (callers in persistent loops pass an opaque copy).  MAXF < 8
// leaves the radix-8 and radix-16 stages out of the kernel (a smaller
// register peak); the plan must then hold no such stage.
template <int EPT, bool INV, int MAXF = 16, typename T2>
__device__ __forceinline__ void fft_lines(T2 *lds, int nlines, int ls, const AxisPlan &p, const T2 *tw,
                                          int tid = threadIdx.x) {
  for (int st = 0; st < p.nst; ++st) {
    const int F = p.rad[st], s = p.s[st], m = p.m[st], ts = p.tstep[st];
    const int a2 = p.ls2[st], b2 = p.lb2[st];
    switch (F) {
      case 16:
        if constexpr (MAXF >= 16) fft_stage<16, (EPT + 15) / 16, INV>(lds, nlines, ls, p.n, s, m, ts, a2, b2, tw, tid);
        break;
      case 8:
        if constexpr (MAXF >= 8) fft_stage<8, (EPT + 7) / 8, INV>(lds, nlines, ls, p.n, s, m, ts, a2, b2, tw, tid);
        break;
      case 4: fft_stage<4, (EPT + 3) / 4, INV>(lds, nlines, ls, p.n, s, m, ts, a2, b2, tw, tid); break;
      case 2: fft_stage<2, (EPT + 1) / 2, INV>(lds, nlines, ls, p.n, s, m, ts, a2, b2, tw, tid); break;
      case 3: fft_stage<3, (EPT + 2) / 3, INV>(lds, nlines, ls, p.n, s, m, ts, a2, b2, tw, tid); break;
      case 5: fft_stage<5, (EPT + 4) / 5, INV>(lds, nlines, ls, p.n, s, m, ts, a2, b2, tw, tid); break;
      default: break;
    }
  }
}

// ---------------------------------------------------------------------
// Compile-time plans for power-of-two lengths N = 2^e on tiles that fill
// every thread (nlines * N == EPT * kThreads).  The stages are those of
// the runtime plan (ceil(e/4) radices of balanced bit counts, so results
// are identical); with every size a constant, the butterfly index
// arithmetic reduces to shifts and masks and no bounds test remains.
// ---------------------------------------------------------------------
constexpr int ct_log2(int n) { return n <= 1 ? 0 : 1 + ct_log2(n / 2); }
// RL = log2 of the largest radix (4: radix-16 stages; 2: radix-4 stages,
// which give a butterfly to every thread of a 1024-element tile)
constexpr int ct_stages(int N, int RL = 4) { return (ct_log2(N) + RL - 1) / RL; }
constexpr int ct_radix(int N, int k, int RL = 4) {  // radix of stage k
  const int e = ct_log2(N), ns = ct_stages(N, RL);
  return 1 << (e / ns + (k < e % ns ? 1 : 0));
}

// With N, S and F powers of two, every padded LDS position of a butterfly
// is its first one plus a compile-time offset: input a sits a*S*M after
// input 0 and output b sits S*b after output 0, and
//   lpad(i0 + d) = lpad(i0) + d + (d >> 4)
// holds for those d (for d a multiple of 16 always; for S*M < 16 because
// i0 < S*M and the sum stays inside one 16-element block; for the outputs
// because o mod 16 < S when S*F >= 16, and o mod 16 + S*(F-1) < 16 when S*F
// divides 16).  The index arithmetic is unsigned, so the divisions by
// powers of two are shifts.
template <int F, bool INV, int N, int S, int EPT, typename T2>
__device__ __forceinline__ void fft_stage_ct(T2 *lds, const T2 *__restrict__ tw, int tid) {
  static_assert((N & (N - 1)) == 0, "power-of-two length");
  constexpr int TOT = EPT * kThreads / F;  // butterflies in the tile
  constexpr int KMAX = (TOT + kThreads - 1) / kThreads;
  constexpr bool FULL = TOT % kThreads == 0;
  constexpr unsigned NBL = N / F;  // butterflies per line
  constexpr unsigned M = N / (S * F);
  constexpr unsigned LS = line_stride(N);
  constexpr unsigned SM = S * M;
  T2 v[KMAX][F];
  unsigned rb[KMAX], wb[KMAX], tb[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const unsigned bf = (unsigned)tid + k * kThreads;
    const unsigned line = bf / NBL, j = bf % NBL;
    const unsigned pp = j / S, q0 = j % S;
    const unsigned i0 = q0 + S * pp, o = q0 + S * F * pp;
    rb[k] = line * LS + i0 + (i0 >> 4);
    wb[k] = line * LS + o + (o >> 4);
    tb[k] = pp * S;  // twiddle step (tstep = S)
  }
  const bool act = FULL || tid < TOT;
#pragma unroll
  for (int k = 0; k < KMAX; ++k)
    if (act)
#pragma unroll
      for (int a = 0; a < F; ++a) v[k][a] = lds[rb[k] + a * SM + ((a * SM) >> 4)];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    if (act) {
      Dft<F, INV, T2>::run(v[k]);
      lds[wb[k]] = v[k][0];
#pragma unroll
      for (int b = 1; b < F; ++b) {
        if constexpr (M == 1) {  // last stage: pp = 0, every twiddle is 1
          lds[wb[k] + S * b + ((S * b) >> 4)] = v[k][b];
        } else {
          const T2 w = tw[tb[k] * b];
          lds[wb[k] + S * b + ((S * b) >> 4)] = INV ? cmulc(v[k][b], w) : cmul(v[k][b], w);
        }
      }
    }
  }
  __syncthreads();
}

// lines of length N at the padded stride line_stride(N); radices up to 2^RL
template <int N, int EPT, bool INV, int RL = 4, int K = 0, int S = 1, typename T2>
__device__ __forceinline__ void fft_lines_ct(T2 *lds, const T2 *tw, int tid = threadIdx.x) {
  if constexpr (K < ct_stages(N, RL)) {
    constexpr int F = ct_radix(N, K, RL);
    fft_stage_ct<F, INV, N, S, EPT>(lds, tw, tid);
    fft_lines_ct<N, EPT, INV, RL, K + 1, S * F>(lds, tw, tid);
  }
}

// stages K..KEND-1 of the compile-time plan (the caller runs the others)
constexpr int ct_sprod(int N, int K, int RL) {  // product of the radices before stage K
  int s = 1;
  for (int k = 0; k < K; ++k) s *= ct_radix(N, k, RL);
  return s;
}
template <int N, int EPT, bool INV, int RL, int K, int KEND, typename T2>
__device__ __forceinline__ void fft_lines_ct_range(T2 *lds, const T2 *tw, int tid) {
  if constexpr (K < KEND) {
    constexpr int F = ct_radix(N, K, RL);
    fft_stage_ct<F, INV, N, ct_sprod(N, K, RL), EPT>(lds, tw, tid);
    fft_lines_ct_range<N, EPT, INV, RL, K + 1, KEND>(lds, tw, tid);
  }
}

// sign (-1)^(j + n/2) applied after a forward / before an inverse pass
__device__ __forceinline__ int jsign(int j, int fixsign) { return (j & 1) ? -fixsign : fixsign; }

}  // namespace rism
