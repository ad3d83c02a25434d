// chi1d.cpp -- the solvent susceptibility chi_vv(k) carried by .xvk files.
//
// The reference reads chi_vv(k) (src-cuda/solvent_read.cc:10-86) and does not
// produce it.  Its own tables (example/h2o_vv.xvk, example_rmdft/h2o_vv.xvk,
// example/naclaq_vv.xvk) are XRISM solutions on a DST-I grid: their last row
// is the bare intramolecular omega (h(k_max) = 0), and their small-k charge
// correlation gives the ideal-dipole dielectric constant 1 + 3y.  This file
// computes such tables:
//
//   r_j = j dr (j = 1..N-1),  k_i = i dk (i = 1..N-1),  dk = pi / (N dr);
//   row N of the table (k = pi/dr) is omega itself.
//   omega_ag(k) = delta_ag + [same molecule, a != g] j0(k l_ag)
//   h(k) = omega c omega + omega c rho h  per k  (XRISM, site densities rho)
//   chi_ag = omega_ag + rho_a h_ag;  reduced table (what solvent_read.cc:51-77
//   expects): xvv[v][v'] = sum_{s' in v'} chi_{s' s}, s the first site of v
//   closures KH (d > 0: h = d, else e^d - 1) and HNC (e^d - 1), d = -beta u + t
//   Coulomb split at eta = 1 Angstrom: beta q q erf(r/eta)/r is handled in k
//   space (4 pi beta q q e^{-k^2 eta^2 / 4} / k^2), so every transformed
//   function is short-ranged: c_s = h - t_s, t_s(k) = h(k) - c_s(k)
//   charge-up: Coulomb terms scaled by 0, 1/n, ..., 1, each stage from the last
//   Anderson acceleration in the DIIS form on t_s, off-diagonal pairs weighted
//   twice (a site-pair matrix is symmetric and stored once)
//   isothermal compressibility: rho kT chi_T = 1 / (1 - sum_ag rho_a rho_g
//   c_ag(k=0) / rho_tot), the xt of the header line (solvent_read.cc:41)
//
// DRISM (Perkyns & Pettitt, Chem. Phys. Lett. 190, 626 (1992)), opts.dielec
// > 0: for the one polar species (density rho_p, dipole mu), its sites in a
// frame centred on the centre of mass with z along mu,
//   zeta_ag(k) = f_a h_c f_g,  f_a = j0(k x_a) j0(k y_a) j1(k z_a),
//   h_c(k) = h_c0 exp(-k^2 a^2 / 4),  h_c0 = ((eps - 1) / y - 3) / rho_p,
//   y = 4 pi beta rho_p mu^2 / 9,
//   omega' = omega + rho zeta,  h = h' + zeta,  h' = omega' c omega' + omega' c rho h'.
// At k -> 0 this gives the longitudinal dielectric constant eps (XRISM's
// 1 + 3y is h_c0 = 0).  The solve reaches it from the XRISM solution,
// scaling zeta up in 10 stages.
//
// This runs on the host: N is 4096-16384 and a solve is 0.1-1 s once per
// solvent, a problem far too small to occupy a 256-CU device.  Physical
// constants are those of src-cuda/physical.h:5-17.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rism_host.h"

namespace rism {
namespace {

// In-place radix-2 complex FFT of length m = 2^p, sign -1, unnormalised.
class Fft2 {
 public:
  explicit Fft2(int m) : m_(m), wr_(m / 2), wi_(m / 2), rev_(m) {
    for (int i = 0; i < m / 2; ++i) {
      wr_[i] = std::cos(2.0 * M_PI * i / m);
      wi_[i] = -std::sin(2.0 * M_PI * i / m);
    }
    int lg = 0;
    while ((1 << lg) < m) ++lg;
    for (int i = 0; i < m; ++i) {
      int r = 0;
      for (int b = 0; b < lg; ++b)
        if ((i >> b) & 1) r |= 1 << (lg - 1 - b);
      rev_[i] = r;
    }
  }
  void run(double *re, double *im) const {
    for (int i = 0; i < m_; ++i)
      if (i < rev_[i]) {
        std::swap(re[i], re[rev_[i]]);
        std::swap(im[i], im[rev_[i]]);
      }
    for (int len = 2; len <= m_; len <<= 1) {
      const int half = len / 2, step = m_ / len;
      for (int i = 0; i < m_; i += len)
        for (int j = 0; j < half; ++j) {
          const double cr = wr_[j * step], ci = wi_[j * step];
          const int a = i + j, b = a + half;
          const double vr = re[b] * cr - im[b] * ci, vi = re[b] * ci + im[b] * cr;
          re[b] = re[a] - vr;
          im[b] = im[a] - vi;
          re[a] += vr;
          im[a] += vi;
        }
    }
  }

 private:
  int m_;
  std::vector<double> wr_, wi_;
  std::vector<int> rev_;
};

// Gaussian elimination with partial pivoting: A[n][n] X = B[n][m], in place
// (X in B).  Returns false on an exactly zero pivot.
bool gauss_solve(double *A, double *B, int n, int m) {
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c])) p = r;
    if (A[p * n + c] == 0.0) return false;
    if (p != c) {
      for (int j = 0; j < n; ++j) std::swap(A[c * n + j], A[p * n + j]);
      for (int j = 0; j < m; ++j) std::swap(B[c * m + j], B[p * m + j]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      if (f == 0.0) continue;
      for (int j = c; j < n; ++j) A[r * n + j] -= f * A[c * n + j];
      for (int j = 0; j < m; ++j) B[r * m + j] -= f * B[c * m + j];
    }
  }
  for (int c = n - 1; c >= 0; --c) {
    for (int j = 0; j < m; ++j) {
      double s = B[c * m + j];
      for (int k = c + 1; k < n; ++k) s -= A[c * n + k] * B[k * m + j];
      B[c * m + j] = s / A[c * n + c];
    }
  }
  return true;
}

double j0(double x) { return std::fabs(x) < 1e-12 ? 1.0 : std::sin(x) / x; }
double j1(double x) {
  return std::fabs(x) < 1e-3 ? x / 3.0 - x * x * x / 30.0 : (std::sin(x) / x - std::cos(x)) / x;
}

// atomic mass (g/mol) of the element a site name starts with (two letters
// when known: Na, Cl), 0 when unknown
double element_mass(const char *name) {
  static const struct { const char *e; double m; } tab[] = {
      {"He", 4.0026}, {"Li", 6.94}, {"Na", 22.990}, {"Mg", 24.305}, {"Cl", 35.45}, {"Ca", 40.078},
      {"Br", 79.904}, {"H", 1.008}, {"C", 12.011}, {"N", 14.007}, {"O", 15.999}, {"F", 18.998},
      {"P", 30.974}, {"S", 32.06}, {"K", 39.098}, {"I", 126.90}};
  for (const auto &t : tab)
    if (std::strncmp(name, t.e, std::strlen(t.e)) == 0) return t.m;
  return 0.0;
}

struct Chi1D {
  int S = 0, N = 0, P = 0, M = 0;  // sites, table length, site pairs, points per function
  int closure = RISM_CLOSURE_KH;
  double dr = 0, dk = 0, lam = 1.0, temper = 0;
  std::vector<int> pa, pg, pidx;
  std::vector<double> wt, rho, r, k;
  std::vector<double> bu_lj, bu_sr, bul_k;  // [P][M]
  std::vector<double> omega;                // [M][S][S]
  std::vector<double> zeta;                 // [M][S][S], DRISM (empty: XRISM)
  double zscale = 0.0;                      // DRISM stage: zeta x zscale in the map
  std::vector<double> xre, xim;             // FFT buffers, length 2N
  Fft2 fft;

  Chi1D(const rism_chi_site *st, int ns, const rism_chi_opts &o)
      : S(ns), N(o.ntab), M(o.ntab - 1), closure(o.closure), dr(o.dr), temper(o.temper), fft(2 * o.ntab) {
    dk = M_PI / (N * dr);
    P = S * (S + 1) / 2;
    pidx.assign(S * S, 0);
    for (int a = 0, p = 0; a < S; ++a)
      for (int g = a; g < S; ++g, ++p) {
        pa.push_back(a);
        pg.push_back(g);
        wt.push_back(a == g ? 1.0 : 2.0);
        pidx[a * S + g] = pidx[g * S + a] = p;
      }
    r.resize(M);
    k.resize(M);
    for (int j = 0; j < M; ++j) {
      r[j] = dr * (j + 1);
      k[j] = dk * (j + 1);
    }
    for (int a = 0; a < S; ++a) rho.push_back(st[a].dens * kAvogadro * 1.0e-27);
    const double beta_lj = 1.0 / (kAvogadro * kBoltzmann * temper);
    const double beta_c = kHartree * kBohr / (kBoltzmann * temper);
    const double eta = 1.0;
    bu_lj.assign((size_t)P * M, 0.0);
    bu_sr.assign((size_t)P * M, 0.0);
    bul_k.assign((size_t)P * M, 0.0);
    for (int p = 0; p < P; ++p) {
      const rism_chi_site &A = st[pa[p]], &G = st[pg[p]];
      const double sg = 0.5 * (A.sig + G.sig), ep = std::sqrt(A.eps * G.eps), qq = A.q * G.q;
      for (int j = 0; j < M; ++j) {
        const double s2 = sg * sg / (r[j] * r[j]), s6 = s2 * s2 * s2;
        bu_lj[(size_t)p * M + j] = beta_lj * 4.0 * ep * (s6 * s6 - s6);
        bu_sr[(size_t)p * M + j] = beta_c * qq * std::erfc(r[j] / eta) / r[j];
        bul_k[(size_t)p * M + j] =
            4.0 * M_PI * beta_c * qq * std::exp(-k[j] * k[j] * eta * eta / 4.0) / (k[j] * k[j]);
      }
    }
    omega.assign((size_t)M * S * S, 0.0);
    for (int i = 0; i < M; ++i) omega_at(st, k[i], &omega[(size_t)i * S * S]);
    xre.assign(2 * N, 0.0);
    xim.assign(2 * N, 0.0);
  }

  // DRISM zeta_ag(k) of the one polar species (file comment); false (with
  // the error set) when the model has no or several polar species or a site
  // of unknown element
  bool set_zeta(const rism_chi_site *st, double eps, double a, double beta_c) {
    std::vector<int> spc;
    for (int s = 0; s < S; ++s)
      if (std::find(spc.begin(), spc.end(), st[s].species) == spc.end()) spc.push_back(st[s].species);
    int polar = -1;
    double mu[3] = {0, 0, 0};
    for (int sp : spc) {
      double m[3] = {0, 0, 0};
      for (int s = 0; s < S; ++s)
        if (st[s].species == sp)
          for (int d = 0; d < 3; ++d) m[d] += st[s].q * st[s].xyz[d];
      if (std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) > 1e-8) {
        if (polar >= 0) {
          set_error("DRISM: more than one polar species");
          return false;
        }
        polar = sp;
        std::copy(m, m + 3, mu);
      }
    }
    if (polar < 0) {
      set_error("DRISM: no polar species");
      return false;
    }
    std::vector<int> idx;
    double com[3] = {0, 0, 0}, mt = 0.0;
    for (int s = 0; s < S; ++s)
      if (st[s].species == polar) {
        const double m = element_mass(st[s].name);
        if (m <= 0.0) {
          set_error("DRISM: a polar site's name must start with its element (for the centre of mass)");
          return false;
        }
        idx.push_back(s);
        for (int d = 0; d < 3; ++d) com[d] += m * st[s].xyz[d];
        mt += m;
      }
    for (double &c : com) c /= mt;
    const double mun = std::sqrt(mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2]);
    const double ez[3] = {mu[0] / mun, mu[1] / mun, mu[2] / mun};
    double ex[3] = {0, 0, 0};
    bool have_x = false;
    for (int s : idx) {  // x axis: the first site off the dipole axis
      double v[3], t[3];
      for (int d = 0; d < 3; ++d) v[d] = st[s].xyz[d] - com[d];
      const double vz = v[0] * ez[0] + v[1] * ez[1] + v[2] * ez[2];
      for (int d = 0; d < 3; ++d) t[d] = v[d] - vz * ez[d];
      const double tn = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      if (tn > 1e-8) {
        for (int d = 0; d < 3; ++d) ex[d] = t[d] / tn;
        have_x = true;
        break;
      }
    }
    if (!have_x) {  // linear molecule: any axis normal to the dipole
      const double u[3] = {std::fabs(ez[0]) < 0.9 ? 1.0 : 0.0, std::fabs(ez[0]) < 0.9 ? 0.0 : 1.0, 0.0};
      ex[0] = ez[1] * u[2] - ez[2] * u[1];
      ex[1] = ez[2] * u[0] - ez[0] * u[2];
      ex[2] = ez[0] * u[1] - ez[1] * u[0];
      const double n = std::sqrt(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
      for (double &e : ex) e /= n;
    }
    const double ey[3] = {ez[1] * ex[2] - ez[2] * ex[1], ez[2] * ex[0] - ez[0] * ex[2], ez[0] * ex[1] - ez[1] * ex[0]};
    const double rp = rho[idx[0]];
    const double y = 4.0 * M_PI * beta_c * rp * mun * mun / 9.0;
    const double hc0 = ((eps - 1.0) / y - 3.0) / rp;
    std::vector<double> fx(S), fy(S), fz(S);
    for (int s : idx) {
      double v[3];
      for (int d = 0; d < 3; ++d) v[d] = st[s].xyz[d] - com[d];
      fx[s] = v[0] * ex[0] + v[1] * ex[1] + v[2] * ex[2];
      fy[s] = v[0] * ey[0] + v[1] * ey[1] + v[2] * ey[2];
      fz[s] = v[0] * ez[0] + v[1] * ez[1] + v[2] * ez[2];
    }
    zeta.assign((size_t)M * S * S, 0.0);
    std::vector<double> f(S);
    for (int i = 0; i < M; ++i) {
      const double kk = k[i], hc = hc0 * std::exp(-kk * kk * a * a / 4.0);
      for (int s = 0; s < S; ++s) f[s] = 0.0;
      for (int s : idx) f[s] = j0(kk * fx[s]) * j0(kk * fy[s]) * j1(kk * fz[s]);
      for (int s : idx)
        for (int g : idx) zeta[(size_t)i * S * S + s * S + g] = f[s] * hc * f[g];
    }
    return true;
  }

  void omega_at(const rism_chi_site *st, double kk, double *w) const {
    for (int a = 0; a < S; ++a)
      for (int g = 0; g < S; ++g) {
        double v = 0.0;
        if (a == g) {
          v = 1.0;
        } else if (st[a].species == st[g].species) {
          const double dx = st[a].xyz[0] - st[g].xyz[0], dy = st[a].xyz[1] - st[g].xyz[1],
                       dz = st[a].xyz[2] - st[g].xyz[2];
          v = j0(kk * std::sqrt(dx * dx + dy * dy + dz * dz));
        }
        w[a * S + g] = v;
      }
  }

  // out_i = sum_{j=1}^{N-1} f_j sin(pi i j / N), i = 1..N-1, for two
  // functions at once (f1 real part, f2 imaginary part of one 2N transform).
  void dst2(const double *f1, const double *f2, double *o1, double *o2) {
    std::fill(xre.begin(), xre.end(), 0.0);
    std::fill(xim.begin(), xim.end(), 0.0);
    for (int j = 1; j < N; ++j) {
      xre[j] = f1[j - 1];
      xre[2 * N - j] = -f1[j - 1];
      if (f2) {
        xim[j] = f2[j - 1];
        xim[2 * N - j] = -f2[j - 1];
      }
    }
    fft.run(xre.data(), xim.data());
    for (int i = 1; i < N; ++i) {
      o1[i - 1] = -0.5 * xim[i];
      if (o2) o2[i - 1] = 0.5 * xre[i];
    }
  }

  // radial transforms of P functions (in/out [P][M]); fwd: 4 pi dr / k DST(r f),
  // inv: dk / (2 pi^2 r) DST(k F)
  void transform(const double *in, double *out, bool fwd) {
    std::vector<double> a(M), b(M);
    const double *w = fwd ? r.data() : k.data();
    for (int p = 0; p < P; p += 2) {
      const bool two = p + 1 < P;
      for (int j = 0; j < M; ++j) {
        a[j] = w[j] * in[(size_t)p * M + j];
        if (two) b[j] = w[j] * in[(size_t)(p + 1) * M + j];
      }
      dst2(a.data(), two ? b.data() : nullptr, &out[(size_t)p * M], two ? &out[(size_t)(p + 1) * M] : nullptr);
      for (int q = p; q < p + (two ? 2 : 1); ++q)
        for (int j = 0; j < M; ++j) {
          double &v = out[(size_t)q * M + j];
          v *= fwd ? 4.0 * M_PI * dr / k[j] : dk / (2.0 * M_PI * M_PI * r[j]);
        }
    }
  }

  // one fixed-point map: t_s -> new t_s; optionally h(k) and c_s(r)
  bool map(const double *ts, double *tnew, double *hk_out, double *cs_out) {
    std::vector<double> cs((size_t)P * M), csk((size_t)P * M), tsk((size_t)P * M);
    for (size_t e = 0; e < cs.size(); ++e) {
      const double d = -bu_lj[e] - lam * bu_sr[e] + ts[e];
      const double h = closure == RISM_CLOSURE_KH ? (d > 0.0 ? d : std::expm1(d)) : std::expm1(d);
      cs[e] = h - ts[e];
    }
    transform(cs.data(), csk.data(), true);
    std::vector<double> C(S * S), WC(S * S), A(S * S), B(S * S), Wz(S * S);
    const bool dr_on = !zeta.empty() && zscale != 0.0;
    for (int i = 0; i < M; ++i) {
      const double *W = &omega[(size_t)i * S * S];
      const double *Z = dr_on ? &zeta[(size_t)i * S * S] : nullptr;
      if (dr_on) {  // omega' = omega + rho zeta
        for (int a = 0; a < S; ++a)
          for (int g = 0; g < S; ++g) Wz[a * S + g] = W[a * S + g] + rho[a] * zscale * Z[a * S + g];
        W = Wz.data();
      }
      for (int a = 0; a < S; ++a)
        for (int g = 0; g < S; ++g) {
          const size_t e = (size_t)pidx[a * S + g] * M + i;
          C[a * S + g] = csk[e] - lam * bul_k[e];
        }
      for (int a = 0; a < S; ++a)
        for (int g = 0; g < S; ++g) {
          double s = 0.0;
          for (int m = 0; m < S; ++m) s += W[a * S + m] * C[m * S + g];
          WC[a * S + g] = s;
        }
      for (int a = 0; a < S; ++a)
        for (int g = 0; g < S; ++g) {
          A[a * S + g] = (a == g ? 1.0 : 0.0) - WC[a * S + g] * rho[g];
          double s = 0.0;
          for (int m = 0; m < S; ++m) s += WC[a * S + m] * W[m * S + g];
          B[a * S + g] = s;
        }
      if (!gauss_solve(A.data(), B.data(), S, S)) return false;
      for (int p = 0; p < P; ++p) {
        const double h = B[pa[p] * S + pg[p]] + (dr_on ? zscale * Z[pa[p] * S + pg[p]] : 0.0);
        tsk[(size_t)p * M + i] = h - csk[(size_t)p * M + i];
        if (hk_out) hk_out[(size_t)p * M + i] = h;
      }
    }
    transform(tsk.data(), tnew, false);
    if (cs_out) std::memcpy(cs_out, cs.data(), sizeof(double) * cs.size());
    return true;
  }

  double wdot(const double *x, const double *y) const {
    double s = 0.0;
    for (int p = 0; p < P; ++p) {
      double q = 0.0;
      for (int j = 0; j < M; ++j) q += x[(size_t)p * M + j] * y[(size_t)p * M + j];
      s += wt[p] * q;
    }
    return s;
  }

  // Anderson (DIIS form) to rms <= tol; returns RISM_SOLVE_* or an error
  int solve(std::vector<double> &ts, double tol, int maxstep, int nhist, double mix, int *steps, double *rms) {
    const size_t L = (size_t)P * M;
    std::vector<std::vector<double>> T, R;
    std::vector<std::vector<double>> G;  // Gram matrix of the stored residuals
    std::vector<double> tn(L), res(L);
    const double denom = (double)S * S * M;
    for (int it = 1; it <= maxstep; ++it) {
      if (!map(ts.data(), tn.data(), nullptr, nullptr)) {
        set_error("1D RISM: singular site matrix");
        return RISM_ERR_ARG;
      }
      for (size_t e = 0; e < L; ++e) res[e] = tn[e] - ts[e];
      const double rr = wdot(res.data(), res.data());
      *rms = std::sqrt(rr / denom);
      *steps = it;
      if (!std::isfinite(*rms)) return RISM_SOLVE_DIVERGED;
      if (*rms <= tol) return RISM_SOLVE_CONVERGED;
      T.push_back(ts);
      R.push_back(res);
      for (size_t a = 0; a + 1 < R.size(); ++a) G[a].push_back(wdot(R[a].data(), res.data()));
      G.emplace_back();
      for (size_t a = 0; a + 1 < R.size(); ++a) G.back().push_back(G[a].back());
      G.back().push_back(rr);
      if ((int)T.size() > nhist) {
        T.erase(T.begin());
        R.erase(R.begin());
        G.erase(G.begin());
        for (auto &row : G) row.erase(row.begin());
      }
      const int n = (int)R.size();
      if (n == 1) {
        for (size_t e = 0; e < L; ++e) ts[e] += mix * res[e];
        continue;
      }
      std::vector<double> Mx(n * n);
      double mmax = 0.0;
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
          Mx[a * n + b] = G[a][b];
          mmax = std::max(mmax, Mx[a * n + b]);
        }
      const int nb = n + 1;
      std::vector<double> Bm(nb * nb, 0.0), rhs(nb, 0.0);
      for (int a = 0; a < n; ++a) {
        for (int b = 0; b < n; ++b) Bm[a * nb + b] = Mx[a * n + b] / mmax;
        Bm[a * nb + n] = Bm[n * nb + a] = -1.0;
      }
      rhs[n] = -1.0;
      if (!gauss_solve(Bm.data(), rhs.data(), nb, 1)) {
        T.erase(T.begin(), T.end() - 1);
        R.erase(R.begin(), R.end() - 1);
        G.assign(1, std::vector<double>(1, rr));
        for (size_t e = 0; e < L; ++e) ts[e] += mix * res[e];
        continue;
      }
      std::fill(ts.begin(), ts.end(), 0.0);
      for (int a = 0; a < n; ++a)
        for (size_t e = 0; e < L; ++e) ts[e] += rhs[a] * (T[a][e] + mix * R[a][e]);
    }
    return RISM_SOLVE_MAXSTEP;
  }
};

}  // namespace
}  // namespace rism

using namespace rism;

extern "C" int rism_chi_nreduced(const rism_chi_site *sites, int nsite) {
  int V = 0;
  for (int a = 0; a < nsite; ++a) V = std::max(V, std::abs(sites[a].sym));
  return V;
}

static int check_sites(const rism_chi_site *st, int ns, int *V) {
  if (!st || ns < 1 || ns > 16) {
    set_error("1D RISM: 1..16 solvent sites are supported");
    return RISM_ERR_ARG;
  }
  *V = rism_chi_nreduced(st, ns);
  if (*V < 1 || *V > 8) {
    set_error("1D RISM: 1..8 reduced sites are supported");
    return RISM_ERR_ARG;
  }
  for (int v = 1; v <= *V; ++v) {
    int first = -1;
    for (int a = 0; a < ns && first < 0; ++a)
      if (std::abs(st[a].sym) == v) first = a;
    if (first < 0 || st[first].sym < 0) {
      set_error("1D RISM: each reduced site needs a first site with a positive sym index");
      return RISM_ERR_FORMAT;
    }
  }
  double qtot = 0.0;
  for (int a = 0; a < ns; ++a) {
    if (!(st[a].dens > 0.0) || !(st[a].sig > 0.0) || st[a].eps < 0.0) {
      set_error("1D RISM: site density and sigma must be > 0, epsilon >= 0");
      return RISM_ERR_ARG;
    }
    qtot += st[a].dens * st[a].q;
  }
  if (std::fabs(qtot) > 1e-6) {
    set_error("1D RISM: the solvent must be electroneutral (sum of density x charge = 0)");
    return RISM_ERR_ARG;
  }
  return RISM_OK;
}

extern "C" void rism_chi_default_opts(rism_chi_opts *o) {
  o->ntab = 16384;
  o->dr = 0.01;
  o->temper = 298.15;
  o->closure = RISM_CLOSURE_KH;
  o->tol = 1e-10;
  o->maxstep = 10000;
  o->nhist = 10;
  o->mix = 0.3;
  o->chargeup = 4;
  o->dielec = 0.0;
  o->adbcor = 0.5;
}

extern "C" int rism_chi_solve(const rism_chi_site *sites, int nsite, const rism_chi_opts *o, double *chi,
                              double *xt, int *steps, double *rms) {
  int V = 0;
  int rc = check_sites(sites, nsite, &V);
  if (rc) return rc;
  if (!o || o->ntab < 16 || (o->ntab & (o->ntab - 1)) || o->ntab > (1 << 22) || !(o->dr > 0.0) ||
      !(o->temper > 0.0) || o->nhist < 1 || !(o->mix > 0.0) || o->maxstep < 1 || o->chargeup < 0 ||
      (o->closure != RISM_CLOSURE_KH && o->closure != RISM_CLOSURE_HNC) || o->dielec < 0.0 ||
      (o->dielec > 0.0 && !(o->adbcor > 0.0))) {
    set_error("1D RISM: bad options (ntab a power of two >= 16, dr, T, mix > 0, nhist, maxstep >= 1, KH/HNC, "
              "dielec >= 0, adbcor > 0)");
    return RISM_ERR_ARG;
  }
  Chi1D s(sites, nsite, *o);
  if (o->dielec > 0.0 &&
      !s.set_zeta(sites, o->dielec, o->adbcor, kHartree * kBohr / (kBoltzmann * o->temper)))
    return RISM_ERR_ARG;
  std::vector<double> ts((size_t)s.P * s.M, 0.0);
  int total = 0, st = 0;
  double r = 0.0;
  const int nst = o->chargeup;
  for (int stage = 0; stage <= nst; ++stage) {
    s.lam = nst == 0 ? 1.0 : (double)stage / nst;
    rc = s.solve(ts, o->tol, o->maxstep, o->nhist, o->mix, &st, &r);
    total += st;
    if (rc == RISM_SOLVE_DIVERGED || rc < 0) break;
  }
  // DRISM: from the XRISM solution, zeta scaled up in 10 stages
  constexpr int kZetaStages = 10;
  for (int stage = 1; !s.zeta.empty() && stage <= kZetaStages && rc >= 0 && rc != RISM_SOLVE_DIVERGED; ++stage) {
    s.zscale = (double)stage / kZetaStages;
    rc = s.solve(ts, o->tol, o->maxstep, o->nhist, o->mix, &st, &r);
    total += st;
  }
  if (steps) *steps = total;
  if (rms) *rms = r;
  if (rc < 0) return rc;
  if (rc == RISM_SOLVE_DIVERGED) {
    set_error("1D RISM diverged");
    return rc;
  }
  std::vector<double> hk((size_t)s.P * s.M), cs((size_t)s.P * s.M), tn((size_t)s.P * s.M);
  s.map(ts.data(), tn.data(), hk.data(), cs.data());
  // first site of each reduced site
  std::vector<int> first(V, -1);
  for (int a = 0; a < nsite; ++a) {
    const int v = std::abs(sites[a].sym) - 1;
    if (first[v] < 0) first[v] = a;
  }
  if (chi) {
    std::vector<double> wN(nsite * nsite);
    s.omega_at(sites, s.N * s.dk, wN.data());
    for (int i = 0; i < s.N; ++i) {
      for (int v = 0; v < V; ++v)
        for (int vp = 0; vp < V; ++vp) {
          const int a = first[v];
          double acc = 0.0;
          for (int sp = 0; sp < nsite; ++sp) {
            if (std::abs(sites[sp].sym) != vp + 1) continue;
            if (i < s.M)
              acc += s.omega[(size_t)i * nsite * nsite + sp * nsite + a] +
                     s.rho[sp] * hk[(size_t)s.pidx[sp * nsite + a] * s.M + i];
            else
              acc += wN[sp * nsite + a];
          }
          chi[((size_t)i * V + v) * V + vp] = acc;
        }
    }
  }
  if (xt) {
    // molecular densities per species, and sum rho_a rho_g c_ag(0)
    std::vector<int> spc;
    double rtot = 0.0;
    for (int a = 0; a < nsite; ++a)
      if (std::find(spc.begin(), spc.end(), sites[a].species) == spc.end()) {
        spc.push_back(sites[a].species);
        rtot += s.rho[a];
      }
    double acc = 0.0;
    for (int p = 0; p < s.P; ++p) {
      double c0 = 0.0;
      for (int j = 0; j < s.M; ++j) c0 += s.r[j] * s.r[j] * cs[(size_t)p * s.M + j];
      c0 *= 4.0 * M_PI * s.dr;
      acc += s.wt[p] * s.rho[s.pa[p]] * s.rho[s.pg[p]] * c0;
    }
    *xt = 1.0 / (1.0 - acc / rtot) / (rtot * 1e30 * kBoltzmann * s.temper);
  }
  return rc;
}

extern "C" int rism_write_xvk(const char *path, const rism_chi_site *sites, int nsite, int ntab, double dr,
                              double temper, double xt, const double *chi) {
  int V = 0;
  int rc = check_sites(sites, nsite, &V);
  if (rc) return rc;
  if (!path || !chi || ntab < 4 || !(dr > 0.0)) {
    set_error("rism_write_xvk: bad arguments");
    return RISM_ERR_ARG;
  }
  FILE *f = std::fopen(path, "w");
  if (!f) {
    set_error(std::string("cannot open ") + path + " for writing");
    return RISM_ERR_IO;
  }
  // header lines as solvent_read.cc:31-48 parses them
  std::vector<int> spc;
  for (int a = 0; a < nsite; ++a)
    if (std::find(spc.begin(), spc.end(), sites[a].species) == spc.end()) spc.push_back(sites[a].species);
  std::fprintf(f, "##  reduced xvv(k) data with solvent parameters\n");
  std::fprintf(f, "## %9d %7d %7d %15.8f %15.8f\n", nsite, V, ntab, dr, M_PI / (dr * ntab));
  std::fprintf(f, "## %4d %16.8f %15.8E\n", (int)spc.size(), temper, xt);
  std::fprintf(f, "## ATOM SPC SYM sig[Angs]   eps[J/mol]  charge[e]    ---X---     ---Y---     ---Z---    density[M]\n");
  for (int a = 0; a < nsite; ++a) {
    const rism_chi_site &s = sites[a];
    char name[16];
    std::snprintf(name, sizeof name, "%.8s", s.name[0] ? s.name : "X");
    std::fprintf(f, "## %-4s %4d %3d %11.5f %11.5f %11.5f %11.5f %11.5f %11.5f %14.8f\n", name, s.species, s.sym,
                 s.sig, s.eps, s.q, s.xyz[0], s.xyz[1], s.xyz[2], s.dens);
  }
  std::fprintf(f, "##\n");
  for (int i = 0; i < ntab; ++i) {
    for (int e = 0; e < V * V; ++e) std::fprintf(f, " %.15E", chi[(size_t)i * V * V + e]);
    std::fputc('\n', f);
  }
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) {
    set_error(std::string("write error on ") + path);
    return RISM_ERR_IO;
  }
  return RISM_OK;
}

extern "C" int rism_xvk_read_sites(const char *path, rism_chi_site *sites, int maxsite, int *nsite, int *ntab,
                                   double *dr, double *temper) {
  // the header of an .xvk file (solvent_read.cc:31-48) as a solvent model
  std::ifstream in(path);
  if (!in) {
    set_error(std::string("Solvent file could not be opened! (") + path + ")");
    return RISM_ERR_IO;
  }
  std::string line, d0, d1;
  int num = 0, natv = 0, nt = 0;
  double r = 0.0, t = 0.0, xt = 0.0;
  for (int ln = 0; std::getline(in, line); ++ln) {
    std::istringstream iss(line);
    if (ln == 1) {
      if (!(iss >> d0 >> num >> natv >> nt >> r)) break;
    } else if (ln == 2) {
      if (!(iss >> d0 >> d1 >> t >> xt)) break;
    } else if (ln > 3 && ln <= num + 3) {
      const int a = ln - 4;
      if (a >= maxsite) {
        set_error("rism_xvk_read_sites: more sites than the caller's array holds");
        return RISM_ERR_ARG;
      }
      rism_chi_site &s = sites[a];
      std::string name;
      if (!(iss >> d0 >> name >> s.species >> s.sym >> s.sig >> s.eps >> s.q >> s.xyz[0] >> s.xyz[1] >> s.xyz[2] >>
            s.dens)) {
        set_error("malformed .xvk site line");
        return RISM_ERR_FORMAT;
      }
      std::snprintf(s.name, sizeof s.name, "%s", name.c_str());
      if (a == num - 1) {
        *nsite = num;
        *ntab = nt;
        *dr = r;
        *temper = t;
        return RISM_OK;
      }
    }
  }
  set_error("incomplete .xvk header");
  return RISM_ERR_FORMAT;
}
