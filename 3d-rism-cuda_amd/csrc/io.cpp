// io.cpp -- the reference's file surface: .inp decks (src-cuda/read_input.cu:21-89),
// solute centring (solute.cu:13-49), .xvk solvent files (solvent_read.cc:10-86),
// .xmu (output_xmu.cu:7-111) and .gra (output_grad.cu:7-30) writers, plus the
// host numerics shared by the solver (natural spline, table lookup).
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rism_host.h"

namespace rism {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

void natural_spline(const double *x, const double *y, int n, double *c0, double *c1, double *c2) {
  // second derivatives M with M_0 = M_{n-1} = 0 from
  // h_{i-1} M_{i-1} + 2 (h_{i-1} + h_i) M_i + h_i M_{i+1} = 6 (s_i - s_{i-1}),
  // s_i the secant slope of interval i; Thomas algorithm.
  std::vector<double> M(n, 0.0), cp(n, 0.0), dp(n, 0.0);
  for (int i = 1; i < n - 1; ++i) {
    const double hl = x[i] - x[i - 1], hr = x[i + 1] - x[i];
    const double a = hl, b = 2.0 * (hl + hr), c = hr;
    const double r = 6.0 * ((y[i + 1] - y[i]) / hr - (y[i] - y[i - 1]) / hl);
    const double den = b - a * cp[i - 1];
    cp[i] = c / den;
    dp[i] = (r - a * dp[i - 1]) / den;
  }
  for (int i = n - 2; i >= 1; --i) M[i] = dp[i] - cp[i] * M[i + 1];
  for (int i = 0; i < n - 1; ++i) {
    const double h = x[i + 1] - x[i];
    c1[i] = 0.5 * M[i];
    c2[i] = (M[i + 1] - M[i]) / (6.0 * h);
    c0[i] = (y[i + 1] - y[i]) / h - h * (2.0 * M[i] + M[i + 1]) / 6.0;
  }
}

int table_interval(double k, double dk, int ntab) {
  const int nseg = ntab - 1;
  int lo = (int)std::floor(k / dk) - 1;
  if (lo < 0) lo = 0;
  if (lo > nseg - 1) lo = nseg - 1;
  if (lo < nseg - 1 && dk * (lo + 2) <= k) ++lo;
  if (lo > 0 && dk * (lo + 1) > k) --lo;
  return lo;
}

bool is_supported_length(int n) {
  if (n < 4 || n > 1024 || (n & 1)) return false;
  int m = n;
  for (int p : {2, 3, 5})
    while (m % p == 0) m /= p;
  return m == 1;
}

}  // namespace rism

using rism::set_error;

// ---------------------------------------------------------------------
// tokenizer over a whole file (the reference reads with operator>>)
// ---------------------------------------------------------------------
namespace {

struct Tokens {
  std::vector<std::string> t;
  size_t i = 0;
  bool next(std::string &s) {
    if (i >= t.size()) return false;
    s = t[i++];
    return true;
  }
  bool dbl(double &v) {
    std::string s;
    if (!next(s)) return false;
    char *e = nullptr;
    v = std::strtod(s.c_str(), &e);
    return e != s.c_str();
  }
  bool integer(int &v) {
    std::string s;
    if (!next(s)) return false;
    char *e = nullptr;
    long x = std::strtol(s.c_str(), &e, 10);
    if (e == s.c_str()) return false;
    v = (int)x;
    return true;
  }
};

bool slurp_tokens(const std::string &path, Tokens &tk) {
  std::ifstream in(path.c_str());
  if (!in) return false;
  std::string w;
  while (in >> w) tk.t.push_back(w);
  return true;
}

bool file_exists(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

std::string dir_of(const std::string &p) {
  size_t k = p.rfind('/');
  return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

double *dup(const std::vector<double> &v) {
  double *p = (double *)std::malloc(sizeof(double) * (v.empty() ? 1 : v.size()));
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(double) * v.size());
  return p;
}

}  // namespace

extern "C" int rism_xvk_read(const char *path, rism_deck *d) {
  // solvent_read.cc:10-86: line 1 "## nsite natv ntab dr dk", line 2
  // "## nspecies T xt", lines 4..nsite+3 site rows, one skipped line, then
  // ntab records of natv*natv values in [v][v'] order.
  std::ifstream in(path);
  if (!in) {
    set_error(std::string("Solvent file could not be opened! (") + path + ")");
    return RISM_ERR_IO;
  }
  std::string line;
  int lineno = 0, num = 0, natv = 0, ntab = 0;
  double dr = 0.0, temper = 0.0, xt = 0.0;
  std::vector<double> sig, eps, q, den;
  std::vector<int> sym;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    std::string d0, d1;
    if (lineno == 1) {
      if (!(iss >> d0 >> num >> natv >> ntab >> dr)) {
        set_error("malformed .xvk size line");
        return RISM_ERR_FORMAT;
      }
      sig.resize(num); eps.resize(num); q.resize(num); den.resize(num); sym.resize(num);
    } else if (lineno == 2) {
      if (!(iss >> d0 >> d1 >> temper >> xt)) {
        set_error("malformed .xvk temperature line");
        return RISM_ERR_FORMAT;
      }
    } else if (lineno > 3 && lineno <= num + 3) {
      const int k = lineno - 4;
      std::string name, spc;
      double x, y, z;
      if (!(iss >> d0 >> name >> spc >> sym[k] >> sig[k] >> eps[k] >> q[k] >> x >> y >> z >> den[k])) {
        set_error("malformed .xvk site line");
        return RISM_ERR_FORMAT;
      }
    }
    ++lineno;
    if (num > 0 && lineno == num + 5) break;
  }
  if (num <= 0 || natv <= 0 || ntab < 4 || dr <= 0.0) {
    set_error("incomplete .xvk header");
    return RISM_ERR_FORMAT;
  }
  if (natv > 8) {
    set_error("more than 8 reduced solvent sites is not supported");
    return RISM_ERR_ARG;
  }
  std::vector<double> chi((size_t)ntab * natv * natv);
  for (size_t i = 0; i < chi.size(); ++i)
    if (!(in >> chi[i])) {
      set_error("chi table truncated in .xvk");
      return RISM_ERR_FORMAT;
    }
  std::vector<double> rho(natv, 0.0), sq(natv, 0.0), ssig(natv, 0.0), seps(natv, 0.0);
  for (int i = 0; i < num; ++i) {
    const int k = std::abs(sym[i]) - 1;
    if (k < 0 || k >= natv) {
      set_error("bad symmetry index in .xvk");
      return RISM_ERR_FORMAT;
    }
    ssig[k] = sig[i];
    seps[k] = eps[i];
    sq[k] = q[i];
    rho[k] += den[i];
  }
  for (auto &r : rho) r *= rism::kAvogadro * 1.0e-27;  // mol/L -> 1/Angstrom^3
  std::free(d->rho); std::free(d->sq); std::free(d->ssig); std::free(d->seps); std::free(d->chi);
  d->nsite = natv;
  d->ntab = ntab;
  d->dk = M_PI / (dr * ntab);
  d->temper = temper;
  d->xt = xt;
  d->rho = dup(rho);
  d->sq = dup(sq);
  d->ssig = dup(ssig);
  d->seps = dup(seps);
  d->chi = dup(chi);
  return RISM_OK;
}

extern "C" int rism_deck_read(const char *control, const char *structure, int centering, int zero,
                              rism_deck **out) {
  *out = nullptr;
  Tokens tk;
  if (!slurp_tokens(control, tk)) {
    set_error(std::string("cannot open input file ") + control);
    return RISM_ERR_IO;
  }
  rism_deck *d = (rism_deck *)std::calloc(1, sizeof(rism_deck));
  auto fail = [&](int code, const std::string &m) {
    set_error(m);
    rism_deck_free(d);
    return code;
  };
  std::string outlist, version, closure, solvent;
  if (!tk.next(outlist) || !tk.integer(d->ksave) || !tk.next(version))
    return fail(RISM_ERR_FORMAT, "truncated header line");
  if (version != "20220301") return fail(RISM_ERR_VERSION, "Error: This input file is for old version.");
  if (!tk.next(closure)) return fail(RISM_ERR_FORMAT, "missing closure");
  if (closure == "KH") d->opts.closure = RISM_CLOSURE_KH;
  else if (closure == "HNC") d->opts.closure = RISM_CLOSURE_HNC;
  else return fail(RISM_ERR_FORMAT, "Error: Unexpected closure switch ");
  if (!tk.next(solvent) || !tk.dbl(d->opts.tol) || !tk.integer(d->opts.maxstep) ||
      !tk.integer(d->opts.count) || !tk.dbl(d->opts.m) || !tk.dbl(d->opts.mp) ||
      !tk.dbl(d->grid.box[0]) || !tk.dbl(d->grid.box[1]) || !tk.dbl(d->grid.box[2]) ||
      !tk.integer(d->grid.n[0]) || !tk.integer(d->grid.n[1]) || !tk.integer(d->grid.n[2]))
    return fail(RISM_ERR_FORMAT, "truncated control block");
  std::snprintf(d->outlist, sizeof d->outlist, "%s", outlist.c_str());
  d->opts.nhist = 2;
  d->opts.precision = RISM_PREC_FP64;
  if (structure && structure[0]) {
    tk = Tokens();
    if (!slurp_tokens(structure, tk)) return fail(RISM_ERR_IO, std::string("cannot open solute file ") + structure);
  }
  // atom count must be a natural-number token (read_input.cu:8-19)
  std::string ns;
  if (!tk.next(ns)) return fail(RISM_ERR_FORMAT, "Error: Number of solute atom is not a natural number.");
  char *e = nullptr;
  long num = std::strtol(ns.c_str(), &e, 10);
  if (e == ns.c_str() || *e != '\0' || num < 0)
    return fail(RISM_ERR_FORMAT, "Error: Number of solute atom is not a natural number.");
  d->natom = (int)num;
  d->sig = (double *)std::calloc(num + 1, sizeof(double));
  d->eps = (double *)std::calloc(num + 1, sizeof(double));
  d->q = (double *)std::calloc(num + 1, sizeof(double));
  d->xyz = (double *)std::calloc(3 * num + 1, sizeof(double));
  for (long u = 0; u < num; ++u)
    if (!tk.dbl(d->sig[u]) || !tk.dbl(d->eps[u]) || !tk.dbl(d->q[u]) || !tk.dbl(d->xyz[3 * u]) ||
        !tk.dbl(d->xyz[3 * u + 1]) || !tk.dbl(d->xyz[3 * u + 2]))
      return fail(RISM_ERR_FORMAT, "truncated solute record");
  if (centering && num > 0) {  // solute.cu:13-43
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) { lo[k] = HUGE_VAL; hi[k] = -HUGE_VAL; }
    for (long u = 0; u < num; ++u)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::fmin(lo[k], d->xyz[3 * u + k]);
        hi[k] = std::fmax(hi[k], d->xyz[3 * u + k]);
      }
    for (int k = 0; k < 3; ++k) d->shift[k] = std::round(-(hi[k] - lo[k]) / 2 - lo[k]);
    for (long u = 0; u < num; ++u)
      for (int k = 0; k < 3; ++k) d->xyz[3 * u + k] += d->shift[k];
  }
  if (zero)
    for (long u = 0; u < num; ++u) d->q[u] = 0.0;
  // output base name (set_fname.cu:5-12)
  std::string fname = (structure && structure[0]) ? structure : control;
  fname = fname.substr(0, fname.rfind('.'));
  std::snprintf(d->fname, sizeof d->fname, "%s", fname.c_str());
  // solvent: as written (relative to the working directory), else next to the deck
  std::string sp = solvent;
  if (!file_exists(sp) && !solvent.empty() && solvent[0] != '/') {
    std::string alt = dir_of(control) + solvent;
    if (file_exists(alt)) sp = alt;
  }
  std::snprintf(d->solvent_path, sizeof d->solvent_path, "%s", sp.c_str());
  int rc = rism_xvk_read(sp.c_str(), d);
  if (rc != RISM_OK) {
    std::string m = rism::last_error();
    return fail(rc, m);
  }
  *out = d;
  return RISM_OK;
}

extern "C" void rism_deck_free(rism_deck *d) {
  if (!d) return;
  std::free(d->sig); std::free(d->eps); std::free(d->q); std::free(d->xyz);
  std::free(d->rho); std::free(d->sq); std::free(d->ssig); std::free(d->seps); std::free(d->chi);
  std::free(d);
}

extern "C" void rism_deck_views(const rism_deck *d, rism_grid *g, rism_solvent *sv, rism_solute *su,
                                rism_opts *o) {
  if (g) *g = d->grid;
  if (sv) {
    sv->nsite = d->nsite;
    sv->ntab = d->ntab;
    sv->dk = d->dk;
    sv->temper = d->temper;
    sv->xt = d->xt;
    sv->rho = d->rho;
    sv->q = d->sq;
    sv->sig = d->ssig;
    sv->eps = d->seps;
    sv->chi = d->chi;
  }
  if (su) {
    su->natom = d->natom;
    su->sig = d->sig;
    su->eps = d->eps;
    su->q = d->q;
    su->xyz = d->xyz;
  }
  if (o) *o = d->opts;
}

extern "C" int rism_write_xmu(const rism_xmu *x, int closure, const char *path) {
  FILE *f = std::fopen(path, "w");
  if (!f) {
    set_error(std::string("cannot write ") + path);
    return RISM_ERR_IO;
  }
  const int V = x->nsite;
  std::fprintf(f, " $RESULT\n");
  if (x->has_rmdft) std::fprintf(f, "SFE_RMDFT= %.5f !(J/mol)\n\n", x->sfe_rmdft);
  std::fprintf(f, "SFE_SC= %.5f !(J/mol)\n", x->sfe_sc);
  for (int v = 0; v < V; ++v) std::fprintf(f, "  SFEC_SC(%d)= %.5f\n", v, x->sfec_sc[v]);
  std::fprintf(f, "\n");
  if (closure == RISM_CLOSURE_KH) {
    std::fprintf(f, "SFE_SC_HNC= %.5f !(J/mol)\n", x->sfe_sc_hnc);
    for (int v = 0; v < V; ++v) std::fprintf(f, "  SFEC_SC_HNC(%d)= %.5f\n", v, x->sfec_sc_hnc[v]);
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "SFE_GF= %.5f !(J/mol)\n", x->sfe_gf);
  for (int v = 0; v < V; ++v) std::fprintf(f, "  SFEC_GF(%d)= %.5f\n", v, x->sfec_gf[v]);
  std::fprintf(f, "\n");
  std::fprintf(f, "SE= %.5f !(J/mol)\n", x->se);
  std::fprintf(f, "  SE_LJ= %.5f\n", x->se_lj);
  std::fprintf(f, "  SE_ES= %.5f\n", x->se_es);
  std::fprintf(f, "\n");
  std::fprintf(f, "PMV= %.5f !(L/mol)\n", x->pmv);
  std::fprintf(f, "Pressure= %.5f !(J/m^3)\n", x->pressure);
  std::fprintf(f, "Correction_Term= %.5f !(J/mol)\n", x->correction);
  std::fprintf(f, " $END\n");
  std::fclose(f);
  return RISM_OK;
}

extern "C" int rism_write_gra(const double *g, int natom, const char *path) {
  FILE *f = std::fopen(path, "w");
  if (!f) {
    set_error(std::string("cannot write ") + path);
    return RISM_ERR_IO;
  }
  for (int u = 0; u < natom; ++u) {
    const double *r = g + 6 * u;
    std::fprintf(f, "%16.8e %16.8e %16.8e %16.8e %16.8e %16.8e\n", r[0], r[1], r[2], r[3], r[4], r[5]);
  }
  std::fclose(f);
  return RISM_OK;
}

// ---- field writers (output_guv.cu, output_huv.cu, output_cuv.cu,
// output_euv.cu, output_qv.cu).  Text values use the reference's stream
// formats: the default (%g, 6 significant digits) for .huv/.cuv, setw(16)
// scientific 8 digits for .euv, fixed 12.4 + scientific 16.8 for .qv.

namespace {

struct FileCloser {
  FILE *f;
  ~FileCloser() {
    if (f) std::fclose(f);
  }
};

int finish(FILE *&f, const char *path) {
  const bool bad = std::ferror(f) != 0;
  const int rc = std::fclose(f);
  f = nullptr;
  if (bad || rc != 0) {
    set_error(std::string("write error on ") + path);
    return RISM_ERR_IO;
  }
  return RISM_OK;
}

}  // namespace

extern "C" int rism_write_guv(const char *path, int natv, const int *n, const double *box, const double *shift,
                              const double *g) {
  // output_guv.cu:11-21: int natv, int grid[3], double box[3], double shift[3],
  // then g [natv][N] as doubles, native byte order
  if (!path || !n || !box || !g || natv < 1) {
    set_error("rism_write_guv: bad arguments");
    return RISM_ERR_ARG;
  }
  FileCloser fc{std::fopen(path, "wb")};
  if (!fc.f) {
    set_error(std::string("cannot write ") + path);
    return RISM_ERR_IO;
  }
  const double zero[3] = {0.0, 0.0, 0.0};
  const size_t N = (size_t)n[0] * n[1] * n[2];
  std::fwrite(&natv, sizeof(int), 1, fc.f);
  std::fwrite(n, sizeof(int), 3, fc.f);
  std::fwrite(box, sizeof(double), 3, fc.f);
  std::fwrite(shift ? shift : zero, sizeof(double), 3, fc.f);
  std::fwrite(g, sizeof(double), N * natv, fc.f);
  return finish(fc.f, path);
}

extern "C" int rism_write_huv(const char *path, const int *n, const double *box, int natom, const double *q,
                              const double *sig, const double *eps, const double *xyz, int natv,
                              const double *f) {
  // output_huv.cu:9-42 (and output_cuv.cu for c = h - t): box, grid, the
  // solute rows q sig eps x y z, natv, then one row of natv values per grid
  // point.  The reference prints r[3 iu .. 3 iu + 2] of its stride-4 array,
  // which are not atom iu's coordinates for iu > 0; the coordinates are
  // written here.
  if (!path || !n || !box || !f || natv < 1 || natom < 0 || (natom > 0 && (!q || !sig || !eps || !xyz))) {
    set_error("rism_write_huv: bad arguments");
    return RISM_ERR_ARG;
  }
  FileCloser fc{std::fopen(path, "w")};
  if (!fc.f) {
    set_error(std::string("cannot write ") + path);
    return RISM_ERR_IO;
  }
  std::fprintf(fc.f, "%g %g %g \n%d %d %d \n%d\n", box[0], box[1], box[2], n[0], n[1], n[2], natom);
  for (int u = 0; u < natom; ++u)
    std::fprintf(fc.f, "%g %g %g %g %g %g\n", q[u], sig[u], eps[u], xyz[3 * u], xyz[3 * u + 1], xyz[3 * u + 2]);
  std::fprintf(fc.f, "%d\n", natv);
  const size_t N = (size_t)n[0] * n[1] * n[2];
  std::string line;
  char buf[32];
  for (size_t ip = 0; ip < N; ++ip) {
    line.clear();
    for (int v = 0; v < natv; ++v) {
      std::snprintf(buf, sizeof buf, "%g ", f[(size_t)v * N + ip]);
      line += buf;
    }
    line += '\n';
    std::fputs(line.c_str(), fc.f);
  }
  return finish(fc.f, path);
}

extern "C" int rism_write_euv(const char *path, int natv, int natom, const double *euv) {
  // output_euv.cu:14-27: per atom, the LJ energy with each site, then the
  // Coulomb energy with each site; euv[(kind * natv + v) * natom + u], J/mol
  if (!path || !euv || natv < 1 || natom < 0) {
    set_error("rism_write_euv: bad arguments");
    return RISM_ERR_ARG;
  }
  FileCloser fc{std::fopen(path, "w")};
  if (!fc.f) {
    set_error(std::string("cannot write ") + path);
    return RISM_ERR_IO;
  }
  for (int u = 0; u < natom; ++u) {
    for (int c = 0; c < 2 * natv; ++c) std::fprintf(fc.f, "%16.8e ", euv[(size_t)c * natom + u]);
    std::fputc('\n', fc.f);
  }
  return finish(fc.f, path);
}

extern "C" int rism_write_qv(const char *path, const int *n, const double *box, const double *qv) {
  // output_qv.cu:15-32: x y z (grid-point coordinates) and qv * dv per point
  if (!path || !n || !box || !qv) {
    set_error("rism_write_qv: bad arguments");
    return RISM_ERR_ARG;
  }
  FileCloser fc{std::fopen(path, "w")};
  if (!fc.f) {
    set_error(std::string("cannot write ") + path);
    return RISM_ERR_IO;
  }
  const double dx = box[0] / n[0], dy = box[1] / n[1], dz = box[2] / n[2], dv = dx * dy * dz;
  const size_t N = (size_t)n[0] * n[1] * n[2];
  for (size_t ip = 0; ip < N; ++ip) {
    const int x = (int)(ip % n[0]), y = (int)((ip / n[0]) % n[1]), z = (int)(ip / ((size_t)n[0] * n[1]));
    std::fprintf(fc.f, "    %12.4f%12.4f%12.4f  %16.8e\n", (x - n[0] / 2) * dx, (y - n[1] / 2) * dy,
                 (z - n[2] / 2) * dz, qv[ip] * dv);
  }
  return finish(fc.f, path);
}

extern "C" int rism_read_cube(const char *path, const int *n, double *out) {
  // The ESP cube of main.cu -e (cal_Coulomb.cu:40-96): two comment lines;
  // natoms and origin (the reference reads a fifth value after the origin;
  // both forms are accepted here); three axis lines "n dx dy dz" that must
  // match the grid; natoms atom lines; then the values with z fastest,
  // stored x-fastest as out[x + nx (y + ny z)].  Values are Hartree/e.
  if (!path || !n || !out) {
    set_error("rism_read_cube: bad arguments");
    return RISM_ERR_ARG;
  }
  std::ifstream in(path);
  if (!in) {
    set_error(std::string("Error during read esp file: ") + path);
    return RISM_ERR_IO;
  }
  std::string line;
  std::getline(in, line);
  std::getline(in, line);
  if (!std::getline(in, line)) {
    set_error("Error reading atom count and origin.");
    return RISM_ERR_FORMAT;
  }
  int natom = 0;
  double o[3];
  {
    std::istringstream iss(line);
    if (!(iss >> natom >> o[0] >> o[1] >> o[2])) {
      set_error("Error reading atom count and origin.");
      return RISM_ERR_FORMAT;
    }
  }
  int m[3];
  for (int a = 0; a < 3; ++a) {
    double d[3];
    if (!std::getline(in, line) || !(std::istringstream(line) >> m[a] >> d[0] >> d[1] >> d[2])) {
      set_error("malformed cube axis line");
      return RISM_ERR_FORMAT;
    }
  }
  if (m[0] != n[0] || m[1] != n[1] || m[2] != n[2]) {
    set_error("Error. cube file doesn't match input.");
    return RISM_ERR_FORMAT;
  }
  for (int a = 0; a < std::abs(natom); ++a)
    if (!std::getline(in, line)) {
      set_error("cube atom lines truncated");
      return RISM_ERR_FORMAT;
    }
  for (int x = 0; x < n[0]; ++x)
    for (int y = 0; y < n[1]; ++y)
      for (int z = 0; z < n[2]; ++z) {
        double v;
        if (!(in >> v)) {
          set_error("cube values truncated");
          return RISM_ERR_FORMAT;
        }
        out[(size_t)x + (size_t)n[0] * ((size_t)y + (size_t)n[1] * z)] = v;
      }
  return RISM_OK;
}

extern "C" int rism_read_cvk(const char *path, int natv, int ntab, double *out) {
  // solvent_read.cc:88-115: a .cvk file has the .xvk layout (header, site
  // rows, then ntab records of natv*natv values); only the table is read
  std::ifstream in(path);
  if (!in) {
    set_error(std::string("cvk file could not be opened! (") + path + ")");
    return RISM_ERR_IO;
  }
  std::string line;
  int num = -1;
  for (int ln = 0; std::getline(in, line); ++ln) {
    if (ln == 1) {
      std::string d0;
      int v = 0, nt = 0;
      if (!(std::istringstream(line) >> d0 >> num >> v >> nt) || v != natv || nt != ntab) {
        set_error("the .cvk size line does not match the solvent (" + line + ")");
        return RISM_ERR_FORMAT;
      }
    }
    if (num >= 0 && ln == num + 4) break;  // the line after the site rows
  }
  if (num < 0) {
    set_error("incomplete .cvk header");
    return RISM_ERR_FORMAT;
  }
  for (size_t i = 0; i < (size_t)ntab * natv * natv; ++i)
    if (!(in >> out[i])) {
      set_error("c(k) table truncated in .cvk");
      return RISM_ERR_FORMAT;
    }
  return RISM_OK;
}

extern "C" int rism_read_xsv(const char *path, int cap, int *nhs, double *k, double *c, double *w, double *pf,
                             double *w0) {
  // solvent_read.cc:118-142: line 1 the table length, line 2 packing
  // fraction and w(0), then rows "k c_hs(k) w(k)" of the reference fluid
  std::ifstream in(path);
  if (!in) {
    set_error(std::string("Hardsphare file could not be opened! (") + path + ")");
    return RISM_ERR_IO;
  }
  int n = 0;
  if (!(in >> n) || n < 4) {
    set_error("malformed .xsv length line");
    return RISM_ERR_FORMAT;
  }
  if (nhs) *nhs = n;
  if (!k) return RISM_OK;  // size query
  if (n > cap) {
    set_error("rism_read_xsv: table longer than the caller's arrays");
    return RISM_ERR_ARG;
  }
  if (!(in >> *pf >> *w0)) {
    set_error("malformed .xsv packing-fraction line");
    return RISM_ERR_FORMAT;
  }
  for (int i = 0; i < n; ++i)
    if (!(in >> k[i] >> c[i] >> w[i])) {
      set_error("hard-sphere table truncated in .xsv");
      return RISM_ERR_FORMAT;
    }
  return RISM_OK;
}
