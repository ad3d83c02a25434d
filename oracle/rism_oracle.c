/*
 * rism_oracle.c -- FP64 host oracle (C11 + OpenMP) for the 3D-RISM
 * calculate/iterate path.  TEST INFRASTRUCTURE ONLY: the checker used by
 * tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py.  The
 * product library never links or calls this code.
 *
 * Each function names the reference file:line it restates.  Reductions use a
 * fixed blocked summation order so results do not depend on the thread count.
 */
#define _GNU_SOURCE
#include "rism_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* physical constants, identical values to src-cuda/physical.h:5-17 (they
   must be identical for the numbers to agree) */
#define ORC_NA 6.02214076e23
#define ORC_BOHR 0.5291771500001
#define ORC_KB 1.380649e-23
#define ORC_HARTREE 4.359744722207185e-18

#define ORC_BLK 4096L

/* ORC_MUTANT (tests only, never in the shipped build): a nonzero value plants
   one known mistake in the SE sums, so the closed-form pins can be shown to
   reject it */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

static __thread char g_err[1024];

static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

const char *orc_last_error(void) { return g_err; }

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu x %zu)\n", n, sz);
    abort();
  }
  return p;
}

orc_sys *orc_new(void) {
  orc_sys *s = (orc_sys *)xcalloc(1, sizeof(orc_sys));
  s->nhist = 2;
  return s;
}

#define FREE0(p) do { free(p); (p) = NULL; } while (0)

static void free_fields(orc_sys *s) {
  FREE0(s->t); FREE0(s->tr); FREE0(s->guv); FREE0(s->huv);
  FREE0(s->tp); FREE0(s->rp);
}

void orc_free(orc_sys *s) {
  if (!s) return;
  FREE0(s->sigu); FREE0(s->epsu); FREE0(s->qu); FREE0(s->ru);
  FREE0(s->rhov); FREE0(s->qv); FREE0(s->sigv); FREE0(s->epsv);
  FREE0(s->xvv); FREE0(s->ttab);
  FREE0(s->ga); FREE0(s->indga); FREE0(s->chik);
  FREE0(s->du); FREE0(s->de); FREE0(s->fr); FREE0(s->fk);
  free_fields(s);
  free(s);
}

/* ------------------------------------------------------------------ */
/* cell: cell.cc:8-22                                                  */
/* ------------------------------------------------------------------ */
static void cell_setup(orc_sys *s) {
  s->volume = s->box[0] * s->box[1] * s->box[2];
  s->ng = (long)s->n[0] * s->n[1] * s->n[2];
  s->dv = s->volume / (double)s->ng;
  for (int i = 0; i < 3; ++i) s->dr[i] = s->box[i] / s->n[i];
  s->shift[0] = s->shift[1] = s->shift[2] = 0.0;
}

/* solute centring: solute.cu:13-43 (integer-rounded shift per axis) */
static void solute_centre(orc_sys *s) {
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) { lo[d] = HUGE_VAL; hi[d] = -HUGE_VAL; }
  for (int u = 0; u < s->natu; ++u)
    for (int d = 0; d < 3; ++d) {
      double v = s->ru[3 * u + d];
      if (v < lo[d]) lo[d] = v;
      if (v > hi[d]) hi[d] = v;
    }
  for (int d = 0; d < 3; ++d) s->shift[d] = round(-(hi[d] - lo[d]) / 2 - lo[d]);
  for (int u = 0; u < s->natu; ++u)
    for (int d = 0; d < 3; ++d) s->ru[3 * u + d] += s->shift[d];
}

/* ------------------------------------------------------------------ */
/* input deck: read_input.cu:21-89                                     */
/* ------------------------------------------------------------------ */
static int read_tok(FILE *f, char *buf, size_t n) {
  char fmt[32];
  snprintf(fmt, sizeof fmt, "%%%zus", n - 1);
  return fscanf(f, fmt, buf) == 1;
}

static int read_dbl(FILE *f, double *v) {
  char b[256];
  if (!read_tok(f, b, sizeof b)) return 0;
  char *e;
  *v = strtod(b, &e);
  return e != b;
}

static int read_int(FILE *f, int *v) {
  char b[256];
  if (!read_tok(f, b, sizeof b)) return 0;
  char *e;
  long x = strtol(b, &e, 10);
  if (e == b) return 0;
  *v = (int)x;
  return 1;
}

int orc_read_inp(orc_sys *s, const char *control, const char *structure,
                 int centering, int zero) {
  FILE *f = fopen(control, "r");
  if (!f) { set_err("cannot open input file %s", control); return ORC_ERR; }
  char check[64], clos[64];
  if (!read_tok(f, s->outlist, sizeof s->outlist) || !read_int(f, &s->ksave) ||
      !read_tok(f, check, sizeof check)) {
    fclose(f); set_err("truncated header in %s", control); return ORC_ERR;
  }
  if (strcmp(check, "20220301") != 0) {  /* version.h:5 */
    fclose(f); set_err("Error: This input file is for old version."); return ORC_ERR;
  }
  if (!read_tok(f, clos, sizeof clos)) { fclose(f); set_err("no closure"); return ORC_ERR; }
  if (strcmp(clos, "KH") == 0) s->clos = 0;
  else if (strcmp(clos, "HNC") == 0) s->clos = 1;
  else { fclose(f); set_err("Error: Unexpected closure switch "); return ORC_ERR; }
  int ok = read_tok(f, s->fsolvent, sizeof s->fsolvent) && read_dbl(f, &s->tol) &&
           read_int(f, &s->maxstep) && read_int(f, &s->count) &&
           read_dbl(f, &s->m) && read_dbl(f, &s->mp) && read_dbl(f, &s->box[0]) &&
           read_dbl(f, &s->box[1]) && read_dbl(f, &s->box[2]) &&
           read_int(f, &s->n[0]) && read_int(f, &s->n[1]) && read_int(f, &s->n[2]);
  if (!ok) { fclose(f); set_err("truncated control block in %s", control); return ORC_ERR; }
  if (structure && structure[0]) {
    fclose(f);
    f = fopen(structure, "r");
    if (!f) { set_err("cannot open solute file %s", structure); return ORC_ERR; }
  }
  /* atom count must be a natural number token (read_input.cu:8-19,57-67) */
  char tmp[256];
  if (!read_tok(f, tmp, sizeof tmp)) { fclose(f); set_err("Error: Number of solute atom is not a natural number."); return ORC_ERR; }
  char *e;
  long num = strtol(tmp, &e, 10);
  if (e == tmp || *e != '\0' || num < 0) {
    fclose(f); set_err("Error: Number of solute atom is not a natural number."); return ORC_ERR;
  }
  cell_setup(s);
  s->natu = (int)num;
  s->sigu = xcalloc(num, sizeof(double));
  s->epsu = xcalloc(num, sizeof(double));
  s->qu = xcalloc(num, sizeof(double));
  s->ru = xcalloc(3 * num, sizeof(double));
  for (long u = 0; u < num; ++u) {
    if (!(read_dbl(f, &s->sigu[u]) && read_dbl(f, &s->epsu[u]) && read_dbl(f, &s->qu[u]) &&
          read_dbl(f, &s->ru[3 * u]) && read_dbl(f, &s->ru[3 * u + 1]) &&
          read_dbl(f, &s->ru[3 * u + 2]))) {
      fclose(f); set_err("truncated solute record %ld", u); return ORC_ERR;
    }
  }
  fclose(f);
  if (centering) solute_centre(s);
  if (zero) for (long u = 0; u < num; ++u) s->qu[u] = 0.0;
  if (s->nhist < 1) s->nhist = 2;
  return 0;
}

int orc_set_deck(orc_sys *s, const int n[3], const double box[3], int clos,
                 double tol, int maxstep, int count, double m, double mp,
                 int nhist, int natu, const double *sig, const double *eps,
                 const double *q, const double *xyz) {
  for (int d = 0; d < 3; ++d) { s->n[d] = n[d]; s->box[d] = box[d]; }
  strcpy(s->outlist, "m");
  s->clos = clos; s->tol = tol; s->maxstep = maxstep; s->count = count;
  s->m = m; s->mp = mp; s->nhist = nhist > 0 ? nhist : 2;
  cell_setup(s);
  s->natu = natu;
  s->sigu = xcalloc(natu, sizeof(double));
  s->epsu = xcalloc(natu, sizeof(double));
  s->qu = xcalloc(natu, sizeof(double));
  s->ru = xcalloc(3 * (size_t)natu, sizeof(double));
  memcpy(s->sigu, sig, natu * sizeof(double));
  memcpy(s->epsu, eps, natu * sizeof(double));
  memcpy(s->qu, q, natu * sizeof(double));
  memcpy(s->ru, xyz, 3 * natu * sizeof(double));
  return 0;
}

/* ------------------------------------------------------------------ */
/* solvent: solvent_read.cc:10-86                                      */
/* ------------------------------------------------------------------ */
int orc_read_xvk(orc_sys *s, const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) { set_err("Solvent file could not be opened!"); return ORC_ERR; }
  char *line = NULL;
  size_t cap = 0;
  int num = 0, lineno = 0;
  double drt = 0.0;
  double *sig = NULL, *eps = NULL, *q = NULL, *den = NULL;
  int *sym = NULL;
  /* header lines 0..num+4; line num+4 is skipped (solvent_read.cc:45-51) */
  while (getline(&line, &cap, f) > 0) {
    if (lineno == 1) {
      char d0[64], d1[64];
      if (sscanf(line, "%63s %d %d %d %lf %63s", d0, &num, &s->natv, &s->ntab, &drt, d1) < 5) {
        free(line); fclose(f); set_err("bad xvk size line"); return ORC_ERR;
      }
      sig = xcalloc(num, sizeof(double)); eps = xcalloc(num, sizeof(double));
      q = xcalloc(num, sizeof(double)); den = xcalloc(num, sizeof(double));
      sym = xcalloc(num, sizeof(int));
    } else if (lineno == 2) {
      char d0[64], d1[64];
      if (sscanf(line, "%63s %63s %lf %lf", d0, d1, &s->temper, &s->xt) != 4) {
        free(line); fclose(f); set_err("bad xvk temperature line"); return ORC_ERR;
      }
    } else if (lineno > 3 && lineno <= num + 3) {
      int i2 = lineno - 4;
      char d0[64], d1[64], d2[64];
      double x, y, z;
      if (sscanf(line, "%63s %63s %63s %d %lf %lf %lf %lf %lf %lf %lf", d0, d1, d2,
                 &sym[i2], &sig[i2], &eps[i2], &q[i2], &x, &y, &z, &den[i2]) != 11) {
        free(line); fclose(f); set_err("bad xvk site line %d", lineno); return ORC_ERR;
      }
    }
    ++lineno;
    if (lineno == num + 5 && num > 0) break;  /* separator consumed */
  }
  free(line);
  if (num <= 0 || s->natv <= 0 || s->ntab <= 1) {
    fclose(f); set_err("xvk header incomplete"); return ORC_ERR;
  }
  int V = s->natv;
  s->xvv = xcalloc((size_t)V * V * s->ntab, sizeof(double));
  for (int i = 0; i < s->ntab; ++i)
    for (int v2 = 0; v2 < V; ++v2)
      for (int v1 = 0; v1 < V; ++v1) {
        double val;
        if (fscanf(f, "%lf", &val) != 1) {
          fclose(f); set_err("xvk table truncated at record %d", i); return ORC_ERR;
        }
        s->xvv[((size_t)v2 * V + v1) * s->ntab + i] = val;
      }
  fclose(f);
  s->nsite = num;
  s->sigv = xcalloc(V, sizeof(double)); s->epsv = xcalloc(V, sizeof(double));
  s->qv = xcalloc(V, sizeof(double)); s->rhov = xcalloc(V, sizeof(double));
  for (int i = 0; i < num; ++i) {
    int i2 = abs(sym[i]) - 1;
    if (i2 < 0 || i2 >= V) { set_err("bad symmetry index"); return ORC_ERR; }
    s->sigv[i2] = sig[i]; s->epsv[i2] = eps[i]; s->qv[i2] = q[i];
    s->rhov[i2] += den[i];
  }
  for (int i = 0; i < V; ++i) s->rhov[i] *= (ORC_NA * 1.0e-27);
  double dk = M_PI / (drt * s->ntab);
  s->ttab = xcalloc(s->ntab, sizeof(double));
  for (int i = 0; i < s->ntab; ++i) s->ttab[i] = dk * (i + 1);
  free(sig); free(eps); free(q); free(den); free(sym);
  return 0;
}

/* ------------------------------------------------------------------ */
/* natural cubic spline: spline.cc:4-43, splint.cc:4-20                */
/* ------------------------------------------------------------------ */
void orc_spline_coef(const double *xp, const double *yp, int np, double *c0,
                     double *c1, double *c2) {
  /* tridiagonal system for the second derivatives with natural ends,
     solved in the same elimination order as the reference */
  double *a = xcalloc(np, sizeof(double)), *b = xcalloc(np, sizeof(double));
  double *c = xcalloc(np, sizeof(double)), *d = xcalloc(np, sizeof(double));
  double *sd = xcalloc(np, sizeof(double));
  int n = np - 2;
  for (int k = 1; k <= n; ++k) {
    double hl = xp[k] - xp[k - 1], hr = xp[k + 1] - xp[k];
    a[k] = hl / (hl + hr);
    b[k] = 1 - a[k];
    double sr = (yp[k + 1] - yp[k]) / hr, sl = (yp[k] - yp[k - 1]) / hl;
    c[k] = 6.0 * ((sr - sl) / (hl + hr));
    d[k] = 2.0;
  }
  for (int k = 1; k < n; ++k) {
    b[k] = b[k] / d[k];
    c[k] = c[k] / d[k];
    d[k + 1] = d[k + 1] - a[k + 1] * b[k];
    c[k + 1] = c[k + 1] - a[k + 1] * c[k];
  }
  if (n >= 1) sd[n] = c[n] / d[n];
  for (int k = n - 1; k >= 1; --k) sd[k] = c[k] - b[k] * sd[k + 1];
  sd[0] = sd[np - 1] = 0.0;
  for (int k = 0; k <= n; ++k) {
    double h = xp[k + 1] - xp[k];
    c1[k] = sd[k] / 2.0;
    c2[k] = (sd[k + 1] - sd[k]) / (6.0 * h);
    c0[k] = (yp[k + 1] - yp[k]) / h - c1[k] * h - c2[k] * h * h;
  }
  free(a); free(b); free(c); free(d); free(sd);
}

double orc_splint(const double *xp, const double *yp, const double *c0,
                  const double *c1, const double *c2, int np, double x) {
  int lo = 0, hi = np - 1;
  while (hi - lo > 1) {
    int k = (hi + lo) >> 1;
    if (xp[k] > x) hi = k; else lo = k;
  }
  double t = x - xp[lo];
  return ((c2[lo] * t + c1[lo]) * t + c0[lo]) * t + yp[lo];
}

/* ------------------------------------------------------------------ */
/* k grid + unique |k|: initialize_g.cu:7-63                           */
/* ------------------------------------------------------------------ */
typedef struct { double g2; long ip; } kkey;

static int kkey_cmp(const void *a, const void *b) {
  const kkey *x = a, *y = b;
  if (x->g2 < y->g2) return -1;
  if (x->g2 > y->g2) return 1;
  return (x->ip > y->ip) - (x->ip < y->ip);
}

static inline double kcomp(int i, int n, double L) {
  return 2.0 * M_PI * ((double)i - n / 2.0 + 0.5) / L;
}

int orc_kgrid(orc_sys *s) {
  long N = s->ng;
  int nx = s->n[0], ny = s->n[1], nz = s->n[2];
  kkey *keys = xcalloc(N, sizeof(kkey));
#pragma omp parallel for
  for (long ip = 0; ip < N; ++ip) {
    int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
    double kx = kcomp(x, nx, s->box[0]), ky = kcomp(y, ny, s->box[1]), kz = kcomp(z, nz, s->box[2]);
    keys[ip].g2 = kx * kx + ky * ky + kz * kz;
    keys[ip].ip = ip;
  }
  qsort(keys, N, sizeof(kkey), kkey_cmp);
  free(s->indga); free(s->ga);
  s->indga = xcalloc(N, sizeof(int));
  s->ga = xcalloc(N, sizeof(double));
  double prev = -1.0;
  int nga = 0;
  for (long i = 0; i < N; ++i) {
    if (keys[i].g2 > prev) {
      s->ga[nga++] = sqrt(keys[i].g2);
      prev = keys[i].g2;
    }
    s->indga[keys[i].ip] = nga - 1;
  }
  s->nga = nga;
  free(keys);
  return 0;
}

/* chi spline on unique |k|: solvent_spline.cu:8-91 (kept on the unique
   list plus the index map, the same values the reference expands to the
   full grid) */
int orc_spline_chi(orc_sys *s) {
  int V = s->natv, nt = s->ntab;
  if (s->ga[s->nga - 1] > s->ttab[nt - 1]) {
    set_err("insufficient maximal T tabulated");
    return ORC_ERR;
  }
  free(s->chik);
  s->chik = xcalloc((size_t)V * V * s->nga, sizeof(double));
  double *c0 = xcalloc(nt, sizeof(double)), *c1 = xcalloc(nt, sizeof(double));
  double *c2 = xcalloc(nt, sizeof(double));
  for (int v2 = 0; v2 < V; ++v2)
    for (int v1 = 0; v1 < V; ++v1) {
      const double *y = s->xvv + ((size_t)v2 * V + v1) * nt;
      orc_spline_coef(s->ttab, y, nt, c0, c1, c2);
      double *out = s->chik + ((size_t)v2 * V + v1) * s->nga;
#pragma omp parallel for
      for (int i = 0; i < s->nga; ++i)
        out[i] = orc_splint(s->ttab, y, c0, c1, c2, nt, s->ga[i]);
    }
  free(c0); free(c1); free(c2);
  return 0;
}

/* ------------------------------------------------------------------ */
/* potentials: cal_LJ.cu:4-72, cal_Coulomb.cu:8-155                    */
/* ------------------------------------------------------------------ */
int orc_potentials(orc_sys *s) {
  long N = s->ng;
  int V = s->natv, nu = s->natu;
  int nx = s->n[0], ny = s->n[1], nz = s->n[2];
  free(s->du); free(s->de); free(s->fr); free(s->fk);
  s->du = xcalloc((size_t)V * N, sizeof(double));
  s->de = xcalloc(N, sizeof(double));
  s->fr = xcalloc(N, sizeof(double));
  s->fk = xcalloc(N, sizeof(double complex));
  double *sig = xcalloc((size_t)nu * V, sizeof(double));
  double *eps = xcalloc((size_t)nu * V, sizeof(double));
  for (int v = 0; v < V; ++v)
    for (int u = 0; u < nu; ++u) {  /* Lorentz-Berthelot, cal_LJ.cu:24-31 */
      sig[u + nu * v] = (s->sigu[u] + s->sigv[v]) * 0.5;
      eps[u + nu * v] = sqrt(s->epsu[u] * s->epsv[v]);
    }
  const double cut2 = 1.0e-2 * 1.0e-2;
  const double ikbt = 1.0 / (ORC_NA * ORC_KB * s->temper);
  const double ubeta = ORC_HARTREE * ORC_BOHR / (ORC_KB * s->temper);
  const long Np = (s->sample > 0 && s->sample < N) ? s->sample : N;
#pragma omp parallel for schedule(static)
  for (long ip = 0; ip < Np; ++ip) {
    int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
    double rx = (x - nx / 2) * s->dr[0];
    double ry = (y - ny / 2) * s->dr[1];
    double rz = (z - nz / 2) * s->dr[2];
    for (int v = 0; v < V; ++v) {
      double acc = 0.0;
      for (int u = 0; u < nu; ++u) {
        int iuv = u + nu * v;
        double dx = rx - s->ru[3 * u], dy = ry - s->ru[3 * u + 1], dz = rz - s->ru[3 * u + 2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < cut2) r2 = cut2;
        double irs2 = sig[iuv] * sig[iuv] / r2;
        double irs6 = irs2 * irs2 * irs2;
        acc += eps[iuv] * 4.0 * irs6 * (irs6 - 1.0);
      }
      s->du[(size_t)v * N + ip] = acc * ikbt;
    }
    double e = 0.0, f = 0.0;
    for (int u = 0; u < nu; ++u) {
      double dx = rx - s->ru[3 * u], dy = ry - s->ru[3 * u + 1], dz = rz - s->ru[3 * u + 2];
      double ra = sqrt(dx * dx + dy * dy + dz * dz);
      if (ra >= 1.0e-5) {
        double qr = s->qu[u] / ra;
        e += qr;
        f += qr * (1 - exp(-ra));
      } else {
        f += s->qu[u];
      }
    }
    s->de[ip] = e * ubeta;
    s->fr[ip] = f * ubeta;
    double kx = kcomp(x, nx, s->box[0]), ky = kcomp(y, ny, s->box[1]), kz = kcomp(z, nz, s->box[2]);
    double rk2 = kx * kx + ky * ky + kz * kz;
    double rk4i = 1.0 / (rk2 * (rk2 + 1.0));
    double fre = 0.0, fim = 0.0;
    for (int u = 0; u < nu; ++u) {
      double ruk = kx * s->ru[3 * u] + ky * s->ru[3 * u + 1] + kz * s->ru[3 * u + 2];
      double tmp = 4.0 * M_PI * s->qu[u] * rk4i;
      fre += tmp * cos(ruk);
      fim -= tmp * sin(ruk);
    }
    s->fk[ip] = (fre * ubeta) + I * (fim * ubeta);
  }
  free(sig); free(eps);
  return 0;
}

int orc_setup(orc_sys *s) {
  for (int d = 0; d < 3; ++d)
    if (s->n[d] < 2 || (s->n[d] & 1)) { set_err("grid sizes must be even"); return ORC_ERR; }
  if (orc_kgrid(s)) return ORC_ERR;
  if (orc_spline_chi(s)) return ORC_ERR;
  if (orc_potentials(s)) return ORC_ERR;
  return 0;
}

/* ------------------------------------------------------------------ */
/* FFT: fft3d.cu:4-94 convention around a plain DFT                    */
/* ------------------------------------------------------------------ */
typedef struct { int n, nf, fac[40]; double complex *w; } fplan;

static void fplan_init(fplan *p, int n) {
  p->n = n; p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  for (int r = 3; m > 1; r += 2)
    while (m % r == 0) { p->fac[p->nf++] = r; m /= r; }
  p->w = xcalloc(n, sizeof(double complex));
  for (int k = 0; k < n; ++k) {
    double a = 2.0 * M_PI * (double)k / (double)n;
    p->w[k] = cos(a) - I * sin(a);  /* exp(-2 pi i k / n) */
  }
}

static void fplan_free(fplan *p) { free(p->w); p->w = NULL; }

static inline double complex cmul(double complex a, double complex b) {
  double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
  return (ar * br - ai * bi) + I * (ar * bi + ai * br);
}

/* Stockham autosort, decimation in frequency, radix = plan factors.
   sign -1: sum x_j exp(-2 pi i jk/n); sign +1: exp(+...).  x in/out, y scratch */
static void fft1_exec(const fplan *p, double complex *x, double complex *y, int sign) {
  int n = p->n;
  double complex *src = x, *dst = y;
  int nn = n, s = 1;
  double complex v[64], V[64];
  for (int f = 0; f < p->nf; ++f) {
    int r = p->fac[f];
    int m = nn / r;
    int tstep = n / nn;  /* omega_nn^k = w[k * tstep] */
    int rstep = n / r;
    for (int pp = 0; pp < m; ++pp) {
      for (int q = 0; q < s; ++q) {
        for (int a = 0; a < r; ++a) v[a] = src[q + s * (pp + a * m)];
        if (r == 2) {
          V[0] = v[0] + v[1]; V[1] = v[0] - v[1];
        } else if (r == 4) {
          double complex t0 = v[0] + v[2], t1 = v[0] - v[2];
          double complex t2 = v[1] + v[3], t3 = v[1] - v[3];
          /* multiply t3 by -i (forward) or +i (inverse) */
          double complex jt3 = (sign < 0) ? (cimag(t3) - I * creal(t3)) : (-cimag(t3) + I * creal(t3));
          V[0] = t0 + t2; V[2] = t0 - t2; V[1] = t1 + jt3; V[3] = t1 - jt3;
        } else {
          for (int b = 0; b < r; ++b) {
            double complex acc = 0;
            for (int a = 0; a < r; ++a) {
              double complex w = p->w[(long)((a * b) % r) * rstep];
              if (sign > 0) w = conj(w);
              acc += cmul(v[a], w);
            }
            V[b] = acc;
          }
        }
        for (int b = 0; b < r; ++b) {
          double complex w = p->w[(long)((pp * b) % nn) * tstep];
          if (sign > 0) w = conj(w);
          dst[q + s * (r * pp + b)] = (b == 0) ? V[0] : cmul(V[b], w);
        }
      }
    }
    nn = m; s *= r;
    double complex *t = src; src = dst; dst = t;
  }
  if (src != x) memcpy(x, src, (size_t)n * sizeof(double complex));
}

void orc_fft1d(double complex *a, int n, int sign) {
  fplan p;
  fplan_init(&p, n);
  double complex *y = xcalloc(n, sizeof(double complex));
  fft1_exec(&p, a, y, sign);
  free(y);
  fplan_free(&p);
}

/* unnormalised 3D DFT of x-fastest data, one axis at a time */
static void dft3(double complex *a, const int n[3], int sign) {
  int nx = n[0], ny = n[1], nz = n[2];
  long nxy = (long)nx * ny;
  fplan px, py, pz;
  fplan_init(&px, nx); fplan_init(&py, ny); fplan_init(&pz, nz);
  const int B = 8;
#pragma omp parallel
  {
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    double complex *buf = xcalloc((size_t)B * nmax, sizeof(double complex));
    double complex *scr = xcalloc(nmax, sizeof(double complex));
#pragma omp for schedule(static)
    for (long l = 0; l < (long)ny * nz; ++l) fft1_exec(&px, a + l * nx, scr, sign);
    /* y lines: batches of B consecutive x at fixed z */
    long nbx = (nx + B - 1) / B;
#pragma omp for schedule(static)
    for (long blk = 0; blk < nbx * nz; ++blk) {
      long z = blk / nbx; int x0 = (int)(blk % nbx) * B;
      int w = (x0 + B <= nx) ? B : nx - x0;
      for (int y = 0; y < ny; ++y)
        for (int i = 0; i < w; ++i) buf[(size_t)i * ny + y] = a[z * nxy + (long)y * nx + x0 + i];
      for (int i = 0; i < w; ++i) fft1_exec(&py, buf + (size_t)i * ny, scr, sign);
      for (int y = 0; y < ny; ++y)
        for (int i = 0; i < w; ++i) a[z * nxy + (long)y * nx + x0 + i] = buf[(size_t)i * ny + y];
    }
    /* z lines: batches of B consecutive x at fixed y */
#pragma omp for schedule(static)
    for (long blk = 0; blk < nbx * ny; ++blk) {
      long y = blk / nbx; int x0 = (int)(blk % nbx) * B;
      int w = (x0 + B <= nx) ? B : nx - x0;
      for (int z = 0; z < nz; ++z)
        for (int i = 0; i < w; ++i) buf[(size_t)i * nz + z] = a[(long)z * nxy + y * nx + x0 + i];
      for (int i = 0; i < w; ++i) fft1_exec(&pz, buf + (size_t)i * nz, scr, sign);
      for (int z = 0; z < nz; ++z)
        for (int i = 0; i < w; ++i) a[(long)z * nxy + y * nx + x0 + i] = buf[(size_t)i * nz + z];
    }
    free(buf); free(scr);
  }
  fplan_free(&px); fplan_free(&py); fplan_free(&pz);
}

/* FFT3D::execute (fft3d.cu:22-38) with timekf/timekb/timeirvol and the
   set_kf/set_ir tables (fft3d.cu:41-94).  dir -1: F(k) = dv sum f(r)
   exp(-i k.r); dir +1: f(r) = (1/volume) sum F(k) exp(+i k.r), with
   r = (i - n/2) dr and k = 2 pi (j - n/2 + 1/2) / L.  The per-axis factor
   (-1)^(n/2) (1 for n % 4 == 0, where the reference is exact) is included. */
void orc_fft3d(double complex *a, const int n[3], const double box[3], int dir) {
  int nx = n[0], ny = n[1], nz = n[2];
  long N = (long)nx * ny * nz;
  double vol = box[0] * box[1] * box[2];
  double dv = vol / (double)N;
  double fix = 1.0;
  for (int d = 0; d < 3; ++d)
    if ((n[d] / 2) % 2) fix = -fix;
  double dkx = M_PI / nx, dky = M_PI / ny, dkz = M_PI / nz;
  if (dir < 0) {
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) {
      int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
      double dkr = dkx * (x - nx / 2) + dky * (y - ny / 2) + dkz * (z - nz / 2);
      double kr = cos(dkr), ki = -sin(dkr);
      double sg = ((x + y + z) % 2 == 0) ? 1.0 : -1.0;
      double tr = creal(a[ip]) * sg, ti = cimag(a[ip]) * sg;
      a[ip] = (tr * kr - ti * ki) + I * (ti * kr + tr * ki);
    }
    dft3(a, n, -1);
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) {
      int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
      double sg = ((x + y + z) % 2 == 0) ? 1.0 : -1.0;
      double f = sg * dv * fix;
      a[ip] = (creal(a[ip]) * f) + I * (cimag(a[ip]) * f);
    }
  } else {
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) {
      int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
      double sg = ((x + y + z) % 2 == 0) ? 1.0 : -1.0;
      a[ip] = (creal(a[ip]) * sg) + I * (cimag(a[ip]) * sg);
    }
    dft3(a, n, +1);
    double volb = 1.0 / vol;
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) {
      int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
      double dkr = dkx * (x - nx / 2) + dky * (y - ny / 2) + dkz * (z - nz / 2);
      double kr = cos(dkr), ki = -sin(dkr);
      double sg = ((x + y + z) % 2 == 0) ? 1.0 : -1.0;
      double tr = creal(a[ip]) * sg * volb * fix, ti = cimag(a[ip]) * sg * volb * fix;
      a[ip] = (tr * kr + ti * ki) + I * (ti * kr - tr * ki);
    }
  }
}

/* ------------------------------------------------------------------ */
/* fixed-point map: calculate.cu:4-58, calculate_cuda.cu:1-66          */
/* ------------------------------------------------------------------ */
static void ensure_fields(orc_sys *s) {
  size_t VN = (size_t)s->natv * s->ng;
  if (!s->t) s->t = xcalloc(VN, sizeof(double));
  if (!s->tr) s->tr = xcalloc(VN, sizeof(double));
  if (!s->guv) s->guv = xcalloc(VN, sizeof(double complex));
  if (!s->huv) s->huv = xcalloc(VN, sizeof(double complex));
  if (!s->tp) {
    s->tp = xcalloc((size_t)s->nhist * VN, sizeof(double));
    s->rp = xcalloc((size_t)s->nhist * VN, sizeof(double));
  }
}

void orc_calculate(orc_sys *s, double cf) {
  ensure_fields(s);
  long N = s->ng;
  int V = s->natv;
  for (int v = 0; v < V; ++v) {
    double q = s->qv[v] * cf;
    const double *t = s->t + (size_t)v * N, *du = s->du + (size_t)v * N;
    double *tr = s->tr + (size_t)v * N;
    double complex *g = s->guv + (size_t)v * N;
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) {
      double earg = -du[ip] - s->de[ip] * q + t[ip];
      double gg;
      if (s->clos == 0) gg = (earg >= 0.0) ? 1.0 + earg : exp(earg);   /* kh */
      else gg = exp(earg);                                             /* hnc */
      tr[ip] = gg;
      g[ip] = (gg - 1.0 - t[ip] + q * s->fr[ip]) + I * 0.0;            /* trm1mt */
    }
  }
  for (int v = 0; v < V; ++v) orc_fft3d(s->guv + (size_t)v * N, s->n, s->box, -1);
  for (int v = 0; v < V; ++v) {
    double q = s->qv[v] * cf;
    double complex *g = s->guv + (size_t)v * N;
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip)                                    /* mqvfk */
      g[ip] = (creal(g[ip]) - q * creal(s->fk[ip])) + I * (cimag(g[ip]) - q * cimag(s->fk[ip]));
  }
  for (int v = 0; v < V; ++v) {                                          /* oz */
    double complex *h = s->huv + (size_t)v * N;
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) {
      double hr = 0.0, hi = 0.0;
      int ig = s->indga[ip];
      for (int v1 = 0; v1 < V; ++v1) {
        double x = s->chik[((size_t)v * V + v1) * s->nga + ig];
        double complex c = s->guv[(size_t)v1 * N + ip];
        hr += creal(c) * x;
        hi += cimag(c) * x;
      }
      h[ip] = hr + I * hi;
    }
  }
  for (int v = 0; v < V; ++v) orc_fft3d(s->huv + (size_t)v * N, s->n, s->box, +1);
  for (int v = 0; v < V; ++v) {                                          /* tr */
    double complex *g = s->guv + (size_t)v * N, *h = s->huv + (size_t)v * N;
    double *tr = s->tr + (size_t)v * N;
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) {
      double gg = tr[ip];
      g[ip] = gg + I * cimag(g[ip]);
      tr[ip] = creal(h[ip]) + 1.0 - gg;
    }
  }
}

/* cal_rms.cu:13-22 */
double orc_rms(const orc_sys *s) {
  long n = (long)s->natv * s->ng;
  long nb = (n + ORC_BLK - 1) / ORC_BLK;
  double *part = xcalloc(nb, sizeof(double));
#pragma omp parallel for
  for (long b = 0; b < nb; ++b) {
    long e = (b + 1) * ORC_BLK < n ? (b + 1) * ORC_BLK : n;
    double acc = 0.0;
    for (long i = b * ORC_BLK; i < e; ++i) acc += s->tr[i] * s->tr[i];
    part[b] = acc;
  }
  double tot = 0.0;
  for (long b = 0; b < nb; ++b) tot += part[b];
  free(part);
  return sqrt(tot / (double)n);
}

/* ------------------------------------------------------------------ */
/* modified Anderson: anderson.cu:13-112, anderson_cuda.cu:1-66        */
/* H slots in a ring; H = 2 is the reference's AN2 exactly.            */
/* ------------------------------------------------------------------ */
void orc_anderson_reset(orc_sys *s) {
  s->an_count = s->count;
  s->an_write = 0;
  s->an_filled = 0;
}

/* Gaussian elimination without pivoting (anderson.cu:87-112) */
static void leg_solve(double *a, double *x, double *b, int n) {
  for (int k = 0; k < n - 1; ++k) {
    double *ak = a + k * n;
    double p = ak[k];
    for (int j = k + 1; j < n; ++j) ak[j] /= p;
    double r = b[k] /= p;
    for (int i = k + 1; i < n; ++i) {
      double *ai = a + i * n;
      double q = ai[k];
      for (int j = k + 1; j < n; ++j) ai[j] -= q * ak[j];
      b[i] -= q * r;
    }
  }
  x[n - 1] = b[n - 1] / a[(n - 1) * n + (n - 1)];
  for (int k = n - 2; k >= 0; --k) {
    double sacc = b[k];
    for (int j = k + 1; j < n; ++j) sacc -= a[k * n + j] * x[j];
    x[k] = sacc;
  }
}

void orc_anderson(orc_sys *s) {
  ensure_fields(s);
  long N = s->ng;
  int V = s->natv, H = s->nhist;
  size_t VN = (size_t)V * N;
  int w = s->an_write;
  if (s->an_count > 0) {            /* warm-up: newdt0, anderson.cu:43-49 */
    --s->an_count;
#pragma omp parallel for
    for (size_t i = 0; i < VN; ++i) {
      s->tp[(size_t)w * VN + i] = s->t[i];
      s->rp[(size_t)w * VN + i] = s->tr[i];
      s->t[i] += s->tr[i];
    }
  } else {
    /* slots in use: the nf most recently written (all H once filled) */
    int nf = s->an_filled < H ? s->an_filled : H;
    int slot[64];
    for (int i = 0; i < nf; ++i) slot[i] = i;  /* slot order = index order */
    if (nf < H) {                                /* ring not yet wrapped */
      for (int i = 0; i < nf; ++i) slot[i] = i;
    }
    int nd = nf + nf * (nf + 1) / 2;  /* c_i and upper A_ij */
    double tot[64 + 64 * 65 / 2];
    memset(tot, 0, sizeof tot);
    long nb = (N + ORC_BLK - 1) / ORC_BLK;
    double *part = xcalloc((size_t)nb * nd, sizeof(double));
    for (int v = 0; v < V; ++v) {       /* theta30 per site, weight 1/rho */
      double irho = 1.0 / s->rhov[v];
#pragma omp parallel for
      for (long b = 0; b < nb; ++b) {
        long e = (b + 1) * ORC_BLK < N ? (b + 1) * ORC_BLK : N;
        double acc[64 + 64 * 65 / 2];
        memset(acc, 0, sizeof(double) * nd);
        double d[64];
        for (long ip = b * ORC_BLK; ip < e; ++ip) {
          size_t i = (size_t)v * N + ip;
          double R = s->tr[i];
          for (int a = 0; a < nf; ++a) d[a] = R - s->rp[(size_t)slot[a] * VN + i];
          int k = 0;
          for (int a = 0; a < nf; ++a) acc[k++] += R * d[a];
          for (int a = 0; a < nf; ++a)
            for (int c = a; c < nf; ++c) acc[k++] += d[a] * d[c];
        }
        for (int k = 0; k < nd; ++k) part[(size_t)b * nd + k] = acc[k];
      }
      for (int k = 0; k < nd; ++k) {
        double sv = 0.0;
        for (long b = 0; b < nb; ++b) sv += part[(size_t)b * nd + k];
        tot[k] += sv * irho;
      }
    }
    free(part);
    double A[64 * 64], c[64], th[64];
    int k = 0;
    for (int a = 0; a < nf; ++a) c[a] = tot[k++];
    for (int a = 0; a < nf; ++a)
      for (int b2 = a; b2 < nf; ++b2) { A[a * nf + b2] = tot[k]; A[b2 * nf + a] = tot[k]; ++k; }
    if (nf > 0) leg_solve(A, th, c, nf);
    /* mp goes to the most recently written slot (anderson.cu:54-55) */
    int recent = (w - 1 + H) % H;
    for (int a = 0; a < nf; ++a)
      if (slot[a] == recent) th[a] += s->mp;
    double mm = s->m;
#pragma omp parallel for
    for (size_t i = 0; i < VN; ++i) {     /* newdt, anderson_cuda.cu:11-23 */
      double t = s->t[i], R = s->tr[i];
      double u = t, vv = t + R;
      for (int a = 0; a < nf; ++a) {
        double tpa = s->tp[(size_t)slot[a] * VN + i], rpa = s->rp[(size_t)slot[a] * VN + i];
        u += th[a] * (tpa - t);
        vv += th[a] * (tpa + rpa - t - R);
      }
      s->tp[(size_t)w * VN + i] = t;
      s->rp[(size_t)w * VN + i] = R;
      s->t[i] = u + mm * (vv - u);
    }
  }
  s->an_write = (w + 1) % H;
  if (s->an_filled < H) ++s->an_filled;
}

/* ------------------------------------------------------------------ */
/* driver: iterate.cu:6-90, initialize_tuv.cu, add_tuv.cu              */
/* ------------------------------------------------------------------ */
int orc_iterate(orc_sys *s, int cu, int verbose, const char *tuv_in, int *steps,
                double *rms_out) {
  ensure_fields(s);
  long N = s->ng;
  int V = s->natv;
  double cf = 1.0, cuf = 0.0;
  orc_anderson_reset(s);
  int saved = 0;
  if (tuv_in && tuv_in[0]) {
    FILE *f = fopen(tuv_in, "r");
    if (f) { fclose(f); saved = 1; }
  }
  if (saved) {
    if (orc_read_tuv(s, tuv_in)) return ORC_ERR;
    cu = 0; cf = 1.0;
  } else if (cu == 0) {
    if (verbose) printf("synthesizing initial estimate for Tuv ...\n");
    for (int v = 0; v < V; ++v)
      for (long ip = 0; ip < N; ++ip) s->t[(size_t)v * N + ip] = s->qv[v] * s->fr[ip];
    cf = 1.0;
  } else {
    memset(s->t, 0, sizeof(double) * V * N);
    cuf = 1.0 / cu; cf = 0.0;
  }
  int status = ORC_CONVERGED, total = 0;
  double rms = 0.0;
  for (int c = 0; c <= cu; ++c) {
    if (c > 0) {
      cf += cuf;
      if (cf > 1.0) cf = 1.0;
      for (int v = 0; v < V; ++v)
        for (long ip = 0; ip < N; ++ip) s->t[(size_t)v * N + ip] += s->qv[v] * cuf * s->fr[ip];
    }
    if (verbose) printf("relaxing 3D UV RISM: Charge Up Factor = %g\n", cf);
    int conver = 0, diverge = 0;
    for (int istep = 1; istep <= s->maxstep; ++istep) {
      orc_calculate(s, cf);
      rms = orc_rms(s);
      ++total;
      diverge = !isfinite(rms);
      if (diverge) break;
      if (rms <= s->tol) conver = 1;
      else orc_anderson(s);
      if (verbose) printf(" Step = %d Reside = %g\n", istep, rms);
      if (conver) break;
    }
    if (diverge) {
      if (verbose) printf("Calculation diverged.\n");
      status = ORC_DIVERGED;
    } else if (!conver) {
      if (verbose) printf("3D UV RISM: reached limit # of relaxation steps: %d\n", s->maxstep);
      status = ORC_MAXSTEP;
      break;
    } else {
      status = ORC_CONVERGED;
    }
  }
  s->steps_total = total;
  s->last_rms = rms;
  if (steps) *steps = total;
  if (rms_out) *rms_out = rms;
  return status;
}

/* ------------------------------------------------------------------ */
/* thermodynamics: cal_exchem.cu:4-154, cal_pmv.cu:5-55,               */
/* cal_pressure.cu:6-20, cal_euv.cu:7-76, output_xmu.cu:7-111          */
/* ------------------------------------------------------------------ */
static double site_sum(const orc_sys *s, int v, int kind) {
  long N = s->ng;
  long nb = (N + ORC_BLK - 1) / ORC_BLK;
  double *part = xcalloc(nb, sizeof(double));
  const double complex *hu = s->huv + (size_t)v * N;
  const double *t = s->t + (size_t)v * N;
#pragma omp parallel for
  for (long b = 0; b < nb; ++b) {
    long e = (b + 1) * ORC_BLK < N ? (b + 1) * ORC_BLK : N;
    double acc = 0.0;
    for (long ip = b * ORC_BLK; ip < e; ++ip) {
      double h = creal(hu[ip]), tt = t[ip];
      double val;
      switch (kind) {
        case 0:  /* ekh */
          val = (h > 0.0) ? -(h * 0.5 + 1.0) * (h - tt) : h * 0.5 * tt - (h - tt);
          break;
        case 1:  /* ehnc */
          val = h * 0.5 * tt - (h - tt);
          break;
        case 2:  /* egf */
          val = -(h * 0.5 + 1.0) * (h - tt);
          break;
        default: /* pmv_cuv */
          val = h - tt;
      }
      acc += val;
    }
    part[b] = acc;
  }
  double tot = 0.0;
  for (long b = 0; b < nb; ++b) tot += part[b];
  free(part);
  return tot;
}

/* per-(u,v) solute-solvent energy (cal_euv.cu:35-76); returns LJ, ES sums */
static void euv_sum(const orc_sys *s, int v, int u, double sig, double eps,
                    double *elj, double *ees) {
  long N = s->ng;
  int nx = s->n[0], ny = s->n[1];
  const double cc = ORC_HARTREE * ORC_BOHR * ORC_NA;
  const long Np = (s->sample > 0 && s->sample < N) ? s->sample : N;
  long nb = (Np + ORC_BLK - 1) / ORC_BLK;
  double *pl = xcalloc(nb, sizeof(double)), *pe = xcalloc(nb, sizeof(double));
  const double complex *g = s->guv + (size_t)v * N;
  double qq = s->qu[u] * s->qv[v];
  for (long b = 0; b < nb; ++b) {
    long e = (b + 1) * ORC_BLK < Np ? (b + 1) * ORC_BLK : Np;
    double al = 0.0, ae = 0.0;
    for (long ip = b * ORC_BLK; ip < e; ++ip) {
      int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
      double dx = (x - nx / 2) * s->dr[0] - s->ru[3 * u];
      double dy = (y - ny / 2) * s->dr[1] - s->ru[3 * u + 1];
      double dz = (z - s->n[2] / 2) * s->dr[2] - s->ru[3 * u + 2];
      double r2 = dx * dx + dy * dy + dz * dz, r1 = sqrt(r2);
#if ORC_MUTANT == 1 /* planted mistake (test of the SE pins): exclusion at sigma */
      if (r1 < sig) continue;
#else
      if (r1 < sig * 0.5) continue;
#endif
      double rs2i = sig * sig / r2, rs6i = rs2i * rs2i * rs2i;
      double gg = creal(g[ip]);
      al += eps * 4.0 * rs6i * (rs6i - 1.0) * gg;
#if ORC_MUTANT == 4 /* planted mistake: cc without Avogadro's number */
      ae += qq / r1 * (ORC_HARTREE * ORC_BOHR) * gg;
#else
      ae += qq / r1 * cc * gg;
#endif
    }
    pl[b] = al; pe[b] = ae;
  }
  double tl = 0.0, te = 0.0;
  for (long b = 0; b < nb; ++b) { tl += pl[b]; te += pe[b]; }
  free(pl); free(pe);
  *elj = tl; *ees = te;
}

void orc_thermo(orc_sys *s, orc_xmu *x) {
  int V = s->natv;
  memset(x, 0, sizeof *x);
  double ibeta = ORC_NA * ORC_KB * s->temper;
  double dv = s->dv;
  double sc = 0, schnc = 0, gf = 0;
  for (int v = 0; v < V; ++v) {
    double a = site_sum(s, v, s->clos == 0 ? 0 : 1) * s->rhov[v] * dv;
    double g = site_sum(s, v, 2) * s->rhov[v] * dv;
    x->sfec_sc[v] = ibeta * a;
    x->sfec_gf[v] = ibeta * g;
    sc += a; gf += g;
    if (s->clos == 0) {
      double b = site_sum(s, v, 1) * s->rhov[v] * dv;
      x->sfec_sc_hnc[v] = ibeta * b;
      schnc += b;
    }
  }
  x->sfe_sc = ibeta * sc;
  x->sfe_gf = ibeta * gf;
  x->sfe_sc_hnc = ibeta * schnc;
  /* SE: sum over (v,u), rho_v factor, x dv */
  double selj = 0.0, sees = 0.0;
  for (int v = 0; v < V; ++v) {
    double lj_v = 0.0, es_v = 0.0;
#pragma omp parallel for reduction(+ : lj_v, es_v) schedule(dynamic)
    for (int u = 0; u < s->natu; ++u) {
#if ORC_MUTANT == 2
      double sg = sqrt(s->sigu[u] * s->sigv[v]);
#else
      double sg = (s->sigu[u] + s->sigv[v]) * 0.5;
#endif
      double ep = sqrt(s->epsu[u] * s->epsv[v]);
      double a, b;
      euv_sum(s, v, u, sg, ep, &a, &b);
#if ORC_MUTANT == 3
      lj_v += a;
      es_v += b;
#else
      lj_v += a * s->rhov[v];
      es_v += b * s->rhov[v];
#endif
    }
    selj += lj_v; sees += es_v;
  }
  x->se_lj = selj * dv;
  x->se_es = sees * dv;
  x->se = (selj + sees) * dv;
  /* PMV (cal_pmv.cu:5-22) */
  double cuv = 0.0;
  for (int v = 0; v < V; ++v) cuv += site_sum(s, v, 3) * s->rhov[v];
  cuv *= dv;
  double pmv = s->xt * (1.0 - cuv) * ibeta;
  /* pressure (cal_pressure.cu:6-20) */
  double pr = 0.0;
  for (int v = 0; v < V; ++v) pr += s->rhov[v];
  pr *= 1.0e30 / ORC_NA;
  pr = 0.5 * (pr + 1.0 / (s->xt * ibeta));
  x->pmv = pmv * 1000.0;
  x->pressure = ibeta * pr;
  x->correction = -pr * pmv * ibeta;
}

/* cal_grad.cu:8-145 with output_grad.cu:15-24 scaling (x dv) */
/* cal_euv per (atom, site) x dv, as output_euv.cu:14-27 writes it:
   out[(kind * V + v) * natu + u], kind 0 = LJ, 1 = Coulomb (cal_euv.cu:7-33) */
void orc_euv(orc_sys *s, double *out) {
  const int V = s->natv, nu = s->natu;
  const double dv = s->dr[0] * s->dr[1] * s->dr[2];
  for (int v = 0; v < V; ++v) {
#pragma omp parallel for schedule(dynamic)
    for (int u = 0; u < nu; ++u) {
#if ORC_MUTANT == 2 /* planted mistake: geometric sigma mixing */
      double sg = sqrt(s->sigu[u] * s->sigv[v]);
#else
      double sg = (s->sigu[u] + s->sigv[v]) * 0.5;
#endif
      double ep = sqrt(s->epsu[u] * s->epsv[v]);
      double a, b;
      euv_sum(s, v, u, sg, ep, &a, &b);
#if ORC_MUTANT == 3 /* planted mistake: the rho_v factor dropped */
      out[(size_t)v * nu + u] = a * dv;
      out[(size_t)(V + v) * nu + u] = b * dv;
#else
      out[(size_t)v * nu + u] = a * s->rhov[v] * dv;
      out[(size_t)(V + v) * nu + u] = b * s->rhov[v] * dv;
#endif
    }
  }
}

/* ------------------------------------------------------------------ */
/* RMDFT free energy: cal_rmdft.cu:5-114 with solvent_spline2.cu:7-92  */
/* and the .cvk/.xsv tables of solvent_read.cc:88-143                  */
/* ------------------------------------------------------------------ */
/* cvk: [ntab][V][V] in file order (the .cvk table, same k grid as the .xvk);
   khs/chs/whs: the .xsv columns (k, hard-sphere c(k), weight w(k)) of the
   reference fluid, given for site 0 only, as the reference reads them; the
   reference leaves the other sites' arrays unset (solvent_read.cc:90-91,
   128-130, alloc2D.cc:4-9) and this restatement takes them as zero, what
   the reference's fresh allocations hold.  Returns SFE_RMDFT in J/mol
   (output_xmu.cu:25-28).  Uses huv of the last calculate. */
static double excp_fd_o(double pf, double dens) { /* excp_fd.cc:1-6 */
  return (4.0 - 2.0 * pf) / ((1.0 - pf) * (1.0 - pf) * (1.0 - pf)) * (pf / dens);
}

int orc_rmdft(orc_sys *s, const double *cvk, int nhs, const double *khs, const double *chs,
              const double *whs, double pfhs0, double w00, double *out) {
  const int V = s->natv, nt = s->ntab, nga = s->nga;
  const long N = s->ng;
  if (!s->huv || !s->indga) { set_err("orc_rmdft: no converged fields"); return ORC_ERR; }
  if (s->ga[nga - 1] > s->ttab[nt - 1]) { set_err("insufficient maximal T tabulated"); return ORC_ERR; }
  /* splines at the unique |k| (solvent_spline.cu:36-60 for cvv, spline2 for chs, wfk) */
  double *cva = xcalloc((size_t)V * V * nga, sizeof(double));
  double *chsa = xcalloc((size_t)V * nga, sizeof(double)), *wfka = xcalloc((size_t)V * nga, sizeof(double));
  double *y = xcalloc(nt > nhs ? nt : nhs, sizeof(double));
  double *c0 = xcalloc(nt > nhs ? nt : nhs, sizeof(double)), *c1 = xcalloc(nt > nhs ? nt : nhs, sizeof(double));
  double *c2 = xcalloc(nt > nhs ? nt : nhs, sizeof(double));
  for (int v2 = 0; v2 < V; ++v2)
    for (int v1 = 0; v1 < V; ++v1) {
      for (int i = 0; i < nt; ++i) y[i] = cvk[((size_t)i * V + v2) * V + v1];
      orc_spline_coef(s->ttab, y, nt, c0, c1, c2);
      for (int i = 0; i < nga; ++i)
        cva[((size_t)v2 * V + v1) * nga + i] = orc_splint(s->ttab, y, c0, c1, c2, nt, s->ga[i]);
    }
  orc_spline_coef(khs, chs, nhs, c0, c1, c2);
  for (int i = 0; i < nga; ++i) chsa[i] = orc_splint(khs, chs, c0, c1, c2, nhs, s->ga[i]);
  orc_spline_coef(khs, whs, nhs, c0, c1, c2);
  for (int i = 0; i < nga; ++i) wfka[i] = orc_splint(khs, whs, c0, c1, c2, nhs, s->ga[i]);
  const double rho0 = s->rhov[0];
  for (int v = 0; v < V; ++v)                                  /* solvent_spline2.cu:64-69 */
    for (int i = 0; i < nga; ++i) cva[((size_t)v * V + v) * nga + i] -= chsa[(size_t)v * nga + i] * rho0 / s->rhov[v];
  double *pf = xcalloc(V, sizeof(double)), *w0 = xcalloc(V, sizeof(double));
  pf[0] = pfhs0;
  w0[0] = w00;
  double complex *H = xcalloc((size_t)V * N, sizeof(double complex));
  double complex *W = xcalloc((size_t)N, sizeof(double complex));
  double xmu1 = 0.0, xmu2 = 0.0, xmu3 = 0.0, eda = 0.0;
  for (int v = 0; v < V; ++v) {                                /* cal_rmdft.cu:22-28 */
    double acc = 0.0;
    for (long ip = 0; ip < N; ++ip) acc += creal(s->huv[(size_t)v * N + ip]);
    xmu1 -= acc * s->rhov[v];
  }
  for (int v = 0; v < V; ++v) {
    memcpy(H + (size_t)v * N, s->huv + (size_t)v * N, N * sizeof(double complex));
    orc_fft3d(H + (size_t)v * N, s->n, s->box, -1);
  }
  for (int v = 0; v < V; ++v) {
    /* cal_dwork (cal_rmdft.cu:167-188): sum_v' h_v'(k) c~_vv'(k) rho_v', back to r */
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) {
      const int ig = s->indga[ip];
      double complex w = 0.0;
      for (int v1 = 0; v1 < V; ++v1) w += H[(size_t)v1 * N + ip] * cva[((size_t)v * V + v1) * nga + ig] * s->rhov[v1];
      W[ip] = w;
    }
    orc_fft3d(W, s->n, s->box, +1);
    double a2 = 0.0, a3 = 0.0;                                 /* cal_rmdft.cu:45-60 */
    for (long ip = 0; ip < N; ++ip) {
      a2 += creal(W[ip]);
      a3 += creal(W[ip]) * creal(s->huv[(size_t)v * N + ip]);
    }
    xmu2 += a2 * s->rhov[v];
    xmu3 += a3 * s->rhov[v];
    /* rho0wfkhk, rho0, rho0guvfmbex, wfka, cal_eda (cal_rmdft.cu:66-112) */
    const double *wf = wfka + (size_t)v * nga;
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) W[ip] = rho0 * wf[s->indga[ip]] * H[(size_t)v * N + ip];
    orc_fft3d(W, s->n, s->box, +1);
    const double bex = excp_fd_o(pf[v], rho0), av = pf[v] / rho0;
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) {
      const double dens = creal(W[ip]) + rho0;
      const double p = av * dens;
      const double fd = (4.0 - 2.0 * p) / ((1.0 - p) * (1.0 - p) * (1.0 - p)) * (p / dens);
      W[ip] = rho0 * ((creal(s->huv[(size_t)v * N + ip]) + 1.0) * fd - bex);
    }
    orc_fft3d(W, s->n, s->box, -1);
#pragma omp parallel for
    for (long ip = 0; ip < N; ++ip) W[ip] *= wf[s->indga[ip]];
    orc_fft3d(W, s->n, s->box, +1);
    const double e0 = rho0 * bex * w0[v];
    double ae = 0.0;
    for (long ip = 0; ip < N; ++ip) ae += (creal(W[ip]) + e0) * (creal(s->huv[(size_t)v * N + ip]) + 1.0) - e0;
    eda += ae * s->rhov[v];
  }
  const double ibeta = ORC_NA * ORC_KB * s->temper;
  *out = ibeta * (xmu1 + xmu2 + 0.5 * xmu3 - eda) * s->dv;
  free(cva); free(chsa); free(wfka); free(y); free(c0); free(c1); free(c2); free(pf); free(w0);
  free(H); free(W);
  return 0;
}

void orc_grad(orc_sys *s, double *dulj, double *due) {
  long N = s->ng;
  int V = s->natv, nu = s->natu;
  int nx = s->n[0], ny = s->n[1], nz = s->n[2];
  const double cc = ORC_HARTREE * ORC_BOHR * ORC_NA;
  for (int i = 0; i < 3 * nu; ++i) dulj[i] = due[i] = 0.0;
#pragma omp parallel for schedule(dynamic)
  for (int u = 0; u < nu; ++u) {
    for (int v = 0; v < V; ++v) {
      double sig = (s->sigu[u] + s->sigv[v]) * 0.5;
      double eps = sqrt(s->epsu[u] * s->epsv[v]);
      double qq = s->qu[u] * s->qv[v];
      const double complex *g = s->guv + (size_t)v * N;
      double al[3] = {0, 0, 0}, ae[3] = {0, 0, 0};
      for (long ip = 0; ip < N; ++ip) {
        int x = (int)(ip % nx), y = (int)((ip / nx) % ny), z = (int)(ip / ((long)nx * ny));
        double d[3] = {(x - nx / 2) * s->dr[0] - s->ru[3 * u], (y - ny / 2) * s->dr[1] - s->ru[3 * u + 1],
                       (z - nz / 2) * s->dr[2] - s->ru[3 * u + 2]};
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r1 = sqrt(r2);
        if (r1 < sig * 0.5) continue;
        double rs2i = sig * sig / r2, rs6i = rs2i * rs2i * rs2i;
        double gg = creal(g[ip]);
        double flj = eps * 24.0 * rs6i / r2 * (2.0 * rs6i - 1.0) * gg;
        double fes = qq / (r2 * r1) * cc * gg;
        for (int k = 0; k < 3; ++k) { al[k] += flj * d[k]; ae[k] += fes * d[k]; }
      }
      for (int k = 0; k < 3; ++k) {
        dulj[3 * u + k] += al[k] * s->rhov[v];
        due[3 * u + k] += ae[k] * s->rhov[v];
      }
    }
  }
  for (int i = 0; i < 3 * nu; ++i) { dulj[i] *= s->dv; due[i] *= s->dv; }
}

/* output_xmu.cu:7-111 layout (rmdft block omitted: out of scope) */
int orc_write_xmu(const orc_sys *s, const orc_xmu *x, const char *path) {
  FILE *f = fopen(path, "w");
  if (!f) { set_err("cannot write %s", path); return ORC_ERR; }
  int V = s->natv;
  fprintf(f, " $RESULT\n");
  fprintf(f, "SFE_SC= %.5f !(J/mol)\n", x->sfe_sc);
  for (int v = 0; v < V; ++v) fprintf(f, "  SFEC_SC(%d)= %.5f\n", v, x->sfec_sc[v]);
  fprintf(f, "\n");
  if (s->clos == 0) {
    fprintf(f, "SFE_SC_HNC= %.5f !(J/mol)\n", x->sfe_sc_hnc);
    for (int v = 0; v < V; ++v) fprintf(f, "  SFEC_SC_HNC(%d)= %.5f\n", v, x->sfec_sc_hnc[v]);
    fprintf(f, "\n");
  }
  fprintf(f, "SFE_GF= %.5f !(J/mol)\n", x->sfe_gf);
  for (int v = 0; v < V; ++v) fprintf(f, "  SFEC_GF(%d)= %.5f\n", v, x->sfec_gf[v]);
  fprintf(f, "\n");
  fprintf(f, "SE= %.5f !(J/mol)\n", x->se);
  fprintf(f, "  SE_LJ= %.5f\n", x->se_lj);
  fprintf(f, "  SE_ES= %.5f\n", x->se_es);
  fprintf(f, "\n");
  fprintf(f, "PMV= %.5f !(L/mol)\n", x->pmv);
  fprintf(f, "Pressure= %.5f !(J/m^3)\n", x->pressure);
  fprintf(f, "Correction_Term= %.5f !(J/mol)\n", x->correction);
  fprintf(f, " $END\n");
  fclose(f);
  return 0;
}

/* write_tuv.cu:7-33 (default stream precision: 6 significant digits) */
int orc_write_tuv(const orc_sys *s, const char *path) {
  FILE *f = fopen(path, "w");
  if (!f) { set_err("cannot write %s", path); return ORC_ERR; }
  fprintf(f, "%d %d %d %d\n", s->n[0], s->n[1], s->n[2], s->natv);
  size_t VN = (size_t)s->natv * s->ng;
  for (size_t i = 0; i < VN; ++i) fprintf(f, "%g\n", s->t[i]);
  fclose(f);
  return 0;
}

/* read_tuv.cu:6-34 (a size mismatch only warns, as in the reference) */
int orc_read_tuv(orc_sys *s, const char *path) {
  ensure_fields(s);
  FILE *f = fopen(path, "r");
  if (!f) { set_err("cannot open %s", path); return ORC_ERR; }
  int a, b, c, d;
  if (fscanf(f, "%d %d %d %d", &a, &b, &c, &d) != 4) {
    fclose(f); set_err("bad tuv header"); return ORC_ERR;
  }
  if (a != s->n[0] || b != s->n[1] || c != s->n[2] || d != s->natv)
    fprintf(stderr, " RXRISM:  actual   NGr = %ld   NatV = %d\n           saved   NGr = %ld   NatV = %d\n",
            s->ng, s->natv, (long)a * b * c, d);
  size_t VN = (size_t)s->natv * s->ng;
  for (size_t i = 0; i < VN; ++i)
    if (fscanf(f, "%lf", &s->t[i]) != 1) { fclose(f); set_err("tuv truncated"); return ORC_ERR; }
  fclose(f);
  return 0;
}

/* ------------------------------------------------------------------ */
/* accessors                                                           */
/* ------------------------------------------------------------------ */
long orc_ng(const orc_sys *s) { return s->ng; }
int orc_natv(const orc_sys *s) { return s->natv; }
int orc_natu(const orc_sys *s) { return s->natu; }

void orc_set_t(orc_sys *s, const double *t) {
  ensure_fields(s);
  memcpy(s->t, t, sizeof(double) * s->natv * s->ng);
}

/* which: 0 t, 1 tr, 2 g (guv.re), 3 h (huv.re), 4 du, 5 de, 6 fr,
   7 fk (interleaved re,im), 8 chi on grid [V*V*ng], 9 solute xyz [natu*3] */
void orc_copy_field(const orc_sys *s, int which, double *out) {
  long N = s->ng;
  size_t VN = (size_t)s->natv * N;
  switch (which) {
    case 0: memcpy(out, s->t, VN * sizeof(double)); break;
    case 1: memcpy(out, s->tr, VN * sizeof(double)); break;
    case 2: for (size_t i = 0; i < VN; ++i) out[i] = creal(s->guv[i]); break;
    case 3: for (size_t i = 0; i < VN; ++i) out[i] = creal(s->huv[i]); break;
    case 4: memcpy(out, s->du, VN * sizeof(double)); break;
    case 5: memcpy(out, s->de, N * sizeof(double)); break;
    case 6: memcpy(out, s->fr, N * sizeof(double)); break;
    case 7: for (long i = 0; i < N; ++i) { out[2 * i] = creal(s->fk[i]); out[2 * i + 1] = cimag(s->fk[i]); } break;
    case 8: {
      int V = s->natv;
      for (int p = 0; p < V * V; ++p)
        for (long i = 0; i < N; ++i) out[(size_t)p * N + i] = s->chik[(size_t)p * s->nga + s->indga[i]];
      break;
    }
    case 9: memcpy(out, s->ru, 3 * s->natu * sizeof(double)); break;
    default: break;
  }
}

void orc_set_nhist(orc_sys *s, int nhist) {
  free_fields(s);
  s->nhist = nhist > 0 ? nhist : 2;
}

void orc_set_iter(orc_sys *s, double tol, int maxstep, int count, double m, double mp) {
  s->tol = tol; s->maxstep = maxstep; s->count = count; s->m = m; s->mp = mp;
}

int orc_alloc_test(orc_sys *s, int natv, long ng, int nhist, const double *rho) {
  free_fields(s);
  FREE0(s->rhov);
  s->natv = natv;
  s->ng = ng;
  s->nhist = nhist > 0 ? nhist : 2;
  s->rhov = xcalloc(natv, sizeof(double));
  memcpy(s->rhov, rho, natv * sizeof(double));
  ensure_fields(s);
  orc_anderson_reset(s);
  return 0;
}

/* test hook: g (guv, real) for the thermodynamic sums, e.g. g = 1 for the SE
   closed forms */
void orc_set_g(orc_sys *s, const double *g) {
  ensure_fields(s);
  const size_t VN = (size_t)s->natv * s->ng;
  for (size_t i = 0; i < VN; ++i) s->guv[i] = g[i];
}

void orc_set_tr(orc_sys *s, const double *tr) {
  ensure_fields(s);
  memcpy(s->tr, tr, sizeof(double) * s->natv * s->ng);
}

void orc_copy_hist(const orc_sys *s, int which, int slot, double *out) {
  size_t VN = (size_t)s->natv * s->ng;
  memcpy(out, (which ? s->rp : s->tp) + (size_t)slot * VN, VN * sizeof(double));
}

void orc_set_atom_xyz(orc_sys *s, int u, const double *xyz) {
  if (u < 0 || u >= s->natu) return;
  for (int k = 0; k < 3; ++k) s->ru[3 * u + k] = xyz[k];
}
