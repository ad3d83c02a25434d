/*
 * rism_oracle.h -- FP64 host oracle for the 3D-RISM calculate/iterate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C (C11 + OpenMP) restatement of
 * the reference's iteration scheme, used as the correctness checker by the
 * tests, by __graft_entry__.smoke() and as the CPU baseline leg of bench.py.
 * The product library (3d-rism-cuda_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference repository root).  Where the reference is defective the
 * oracle implements the intended mathematics and says so:
 *   - FFT half-shift phase: the reference omits a factor (-1)^(n/2) per axis,
 *     which is 1 for n % 4 == 0 (src-cuda/fft3d.cu:71-94).  The oracle
 *     includes it, so it equals the reference exactly for n % 4 == 0.
 *   - cufftPlan3d(nx,ny,nz) on x-fastest data (src-cuda/fft3d.cu:16) is only
 *     right for cubic grids; the oracle transforms x-fastest data correctly.
 *   - Anderson history slots never written (count < nhist) are excluded from
 *     the least-squares system instead of read uninitialised
 *     (src-cuda/anderson.cu:17-18 allocates them without clearing).
 *
 * Parity pins (tests/test_oracle_*.py): FFT vs numpy DFT golden vectors and
 * the Gaussian closed form; spline exact on cubics; potentials vs brute force
 * and single-atom closed forms; Anderson vs a hand-derived linear recurrence;
 * end-to-end vs example/test*.xmu and example_rmdft/benzene.xmu values.
 */
#ifndef RISM_ORACLE_H
#define RISM_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys {
  /* ---- input deck (read_input.cu:21-89) ---- */
  char outlist[64];
  int ksave;
  int clos;               /* 0 KH, 1 HNC */
  char fsolvent[4096];
  double tol;
  int maxstep;
  int count;              /* Anderson warm-up (Picard) steps */
  double m, mp;           /* Anderson mixing parameters */
  int nhist;              /* history slots; 2 = reference AN2 */
  double box[3];
  int n[3];
  /* cell (cell.cc:8-22) */
  double dr[3], dv, volume;
  long ng;                /* nx*ny*nz */
  long sample;            /* > 0: potentials and SE cover only the first `sample`
                             points (CPU-baseline timing sample; results partial) */
  /* solute (solute.cu) */
  int natu;
  double *sigu, *epsu, *qu, *ru;  /* ru: [natu][3] */
  double shift[3];
  /* ---- solvent (solvent_read.cc:10-86) ---- */
  int nsite, natv, ntab;
  double temper, xt;
  double *rhov, *qv, *sigv, *epsv;   /* [natv] */
  double *xvv;            /* [(v*natv+v')*ntab + i] */
  double *ttab;           /* [ntab] k_i = dk*(i+1) */
  /* ---- k grid (initialize_g.cu:7-50) ---- */
  int nga;
  double *ga;             /* unique |k| ascending */
  int *indga;             /* [ng] */
  double *chik;           /* [(v*natv+v')*nga + iga] splined chi */
  /* ---- potentials (cal_LJ.cu, cal_Coulomb.cu) ---- */
  double *du;             /* [natv*ng] beta*u_LJ */
  double *de;             /* [ng] beta*phi per unit solvent charge */
  double *fr;             /* [ng] long-range real part */
  double complex *fk;     /* [ng] long-range k part */
  /* ---- fields (iterate.cu:12-20) ---- */
  double *t, *tr;         /* [natv*ng] */
  double complex *guv, *huv;   /* [natv*ng]; guv.re = g after tr() */
  /* ---- Anderson (anderson.cu) ---- */
  double *tp, *rp;        /* [nhist*natv*ng] */
  int an_count, an_write, an_filled;
  /* ---- diagnostics ---- */
  int steps_total;
  double last_rms;
  int nthreads;
} orc_sys;

typedef struct orc_xmu {
  double sfe_sc, sfe_sc_hnc, sfe_gf;     /* J/mol */
  double sfec_sc[16], sfec_sc_hnc[16], sfec_gf[16];
  double se, se_lj, se_es;               /* J/mol */
  double pmv;                            /* L/mol */
  double pressure;                       /* J/m^3 */
  double correction;                     /* J/mol */
} orc_xmu;

/* status codes shared with the CLI */
enum { ORC_CONVERGED = 0, ORC_DIVERGED = 1, ORC_MAXSTEP = 2, ORC_ERR = -1 };

orc_sys *orc_new(void);
void orc_free(orc_sys *s);
const char *orc_last_error(void);

int orc_read_inp(orc_sys *s, const char *control, const char *structure,
                 int centering, int zero);
int orc_read_xvk(orc_sys *s, const char *path);
/* programmatic set-up (tests): solute xyz [natu][3] used as given */
int orc_set_deck(orc_sys *s, const int n[3], const double box[3], int clos,
                 double tol, int maxstep, int count, double m, double mp,
                 int nhist, int natu, const double *sig, const double *eps,
                 const double *q, const double *xyz);
int orc_setup(orc_sys *s);          /* k grid, chi spline, potentials */
int orc_kgrid(orc_sys *s);
int orc_spline_chi(orc_sys *s);
int orc_potentials(orc_sys *s);

/* one fixed-point map (calculate.cu:4-58): t -> tr, guv, huv */
void orc_calculate(orc_sys *s, double cf);
double orc_rms(const orc_sys *s);
void orc_anderson(orc_sys *s);
void orc_anderson_reset(orc_sys *s);
/* full iterate (iterate.cu:6-90); prints like the reference if verbose */
int orc_iterate(orc_sys *s, int cu, int verbose, const char *tuv_in,
                int *steps, double *rms);

void orc_thermo(orc_sys *s, orc_xmu *x);
void orc_grad(orc_sys *s, double *dulj, double *due);
void orc_euv(orc_sys *s, double *out);  /* [2][natv][natu], x dv */
/* RMDFT (cal_rmdft.cu): cvk [ntab][V][V], site-0 hard-sphere tables; J/mol */
int orc_rmdft(orc_sys *s, const double *cvk, int nhs, const double *khs, const double *chs,
              const double *whs, double pfhs0, double w00, double *out);  /* [natu*3] each */
int orc_write_xmu(const orc_sys *s, const orc_xmu *x, const char *path);
int orc_write_tuv(const orc_sys *s, const char *path);
int orc_read_tuv(orc_sys *s, const char *path);

/* reference-convention 3D FFT: dir=-1 forward (x dv), dir=+1 inverse
   (/volume).  a: x-fastest complex[nx*ny*nz]. */
void orc_fft3d(double complex *a, const int n[3], const double box[3],
               int dir);
/* plain unnormalised 1D DFT kernel used by orc_fft3d (sign -1 or +1) */
void orc_fft1d(double complex *a, int n, int sign);

/* natural cubic spline (spline.cc, splint.cc) */
void orc_spline_coef(const double *x, const double *y, int np, double *c0,
                     double *c1, double *c2);
double orc_splint(const double *x, const double *y, const double *c0,
                  const double *c1, const double *c2, int np, double xv);

/* accessors for ctypes */
long orc_ng(const orc_sys *s);
int orc_natv(const orc_sys *s);
int orc_natu(const orc_sys *s);
void orc_copy_field(const orc_sys *s, int which, double *out);
void orc_set_t(orc_sys *s, const double *t);
void orc_set_threads(int n);
void orc_set_nhist(orc_sys *s, int nhist);
/* test hooks: a bare field set for exercising the Anderson step alone */
int orc_alloc_test(orc_sys *s, int natv, long ng, int nhist, const double *rho);
void orc_set_tr(orc_sys *s, const double *tr);
/* test hook: set g (real) for the thermodynamic sums */
void orc_set_g(orc_sys *s, const double *g);
void orc_copy_hist(const orc_sys *s, int which, int slot, double *out);
void orc_set_iter(orc_sys *s, double tol, int maxstep, int count, double m, double mp);
/* test hook: move solute atom u (its grid potentials are not recomputed) */
void orc_set_atom_xyz(orc_sys *s, int u, const double *xyz);

#ifdef __cplusplus
}
#endif
#endif
