/*
 * rism3d_amd.h -- C ABI of the MI355X (gfx950) 3D-RISM calculate/iterate
 * library (librism3d_amd.so).
 *
 * The reference (rismical-dev/3D-RISM-CUDA) has no library API: its seam is
 * the class RISM3D { initialize, iterate, output } (src-cuda/rism3d.h:25-27)
 * driven by the CLI in src-cuda/main.cu.  Each entry point below names the
 * reference routine it replaces.  Conventions:
 *   - plain C types only; host inputs are copied in, the handle owns all
 *     device memory;
 *   - every call returns an int status (RISM_OK == 0 or a RISM_ERR_* code;
 *     rism_solve returns a RISM_SOLVE_* outcome) and never exits the process
 *     (the reference prints and calls exit(1), e.g. read_input.cu:31-32);
 *   - rism_last_error() gives a thread-local message for the last failure;
 *   - one handle is used from one host thread; calls on one handle are
 *     ordered on the handle's HIP stream.
 * Units follow the reference: lengths in Angstrom, LJ epsilon in J/mol,
 * charges in e, densities in 1/Angstrom^3, energies out in J/mol.
 */
#ifndef RISM3D_AMD_H
#define RISM3D_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define RISM_OK 0
#define RISM_SOLVE_CONVERGED 0 /* iterate.cu:58-59 */
#define RISM_SOLVE_DIVERGED 1  /* iterate.cu:54-57 "Calculation diverged." */
#define RISM_SOLVE_MAXSTEP 2   /* iterate.cu:76-79 "reached limit" */
#define RISM_ERR_ARG -1        /* bad argument / unsupported size */
#define RISM_ERR_IO -2         /* file missing or unreadable */
#define RISM_ERR_FORMAT -3     /* malformed .inp/.xvk/.tuv content */
#define RISM_ERR_VERSION -4    /* .inp version tag != 20220301 (read_input.cu:30-33) */
#define RISM_ERR_RANGE -5      /* |k| beyond the chi table (solvent_spline.cu:18-21) */
#define RISM_ERR_HIP -6        /* HIP runtime failure (message has details) */
#define RISM_ERR_NOMEM -7      /* device allocation failed */
#define RISM_ERR_STATE -8      /* call out of order (e.g. thermo before solve) */

/* closure switch (read_input.cu:35-44) */
#define RISM_CLOSURE_KH 0
#define RISM_CLOSURE_HNC 1

/* arithmetic of the iteration */
#define RISM_PREC_FP64 0  /* everything FP64 (reference arithmetic) */
#define RISM_PREC_MIXED 1 /* FP32 FFT/closure/convolution, FP64 state,
                             Anderson dot products, solve and free-energy sums */

/* fields for rism_get_field */
#define RISM_FIELD_T 0   /* t = h - c, [V][N] (iterate.cu dt) */
#define RISM_FIELD_G 1   /* g, [V][N] (dguv.x) */
#define RISM_FIELD_H 2   /* h, [V][N] (dhuv.x) */
#define RISM_FIELD_C 3   /* c = h - t, [V][N] */
#define RISM_FIELD_RES 4 /* last residual R = h + 1 - g, [V][N] (dtr) */
#define RISM_FIELD_ULJ 5 /* beta*u_LJ, [V][N] (du, cal_LJ.cu:47-72) */
#define RISM_FIELD_DE 6  /* beta*phi per unit charge, [N] (de) */
#define RISM_FIELD_FR 7  /* long-range real-space part, [N] (dfr) */
#define RISM_FIELD_FK 8  /* long-range k-space part, interleaved complex [N] (dfk) */
#define RISM_FIELD_CHI 9 /* chi_vv'(|k|) on the grid, [V][V][N] (sv->dx) */

typedef struct rism_handle rism_handle;

/* grid/cell: the .inp box and grid lines (read_input.cu:48-49, cell.cc) */
typedef struct rism_grid {
  int n[3];       /* points per axis; even, 2,3,5-smooth, <= 1024 */
  double box[3];  /* Angstrom */
} rism_grid;

/* solvent, as read from a .xvk file (solvent_read.cc:10-86) */
typedef struct rism_solvent {
  int nsite;          /* V: reduced (symmetry-merged) sites */
  int ntab;           /* k-table length */
  double dk;          /* table spacing; k_i = dk*(i+1), 1/Angstrom */
  double temper;      /* K */
  double xt;          /* isothermal compressibility, 1/Pa */
  const double *rho;  /* [V] number density, 1/Angstrom^3 */
  const double *q;    /* [V] charge, e */
  const double *sig;  /* [V] LJ sigma, Angstrom */
  const double *eps;  /* [V] LJ epsilon, J/mol */
  const double *chi;  /* [ntab][V][V]: h_v = sum_v' chi[v][v'] c_v' */
} rism_solvent;

/* solute (read_input.cu:72-76, solute.cu); coordinates used as given */
typedef struct rism_solute {
  int natom;
  const double *sig;  /* [n] Angstrom */
  const double *eps;  /* [n] J/mol */
  const double *q;    /* [n] e */
  const double *xyz;  /* [n][3] Angstrom */
} rism_solute;

/* solver options (.inp lines 2,4,5 and main.cu -c; plus extensions) */
typedef struct rism_opts {
  int closure;    /* RISM_CLOSURE_KH / _HNC */
  double tol;     /* RMS residual target (co->convergence) */
  int maxstep;    /* per charge-up stage */
  int count;      /* Anderson Picard warm-up steps (ma->count) */
  double m;       /* Anderson mixing (ma->m) */
  double mp;      /* extra weight of the most recent slot (ma->mp) */
  int nhist;      /* Anderson history slots, 1..10; 2 = the reference's AN2 */
  int precision;  /* RISM_PREC_FP64 / RISM_PREC_MIXED */
  int chargeup;   /* charge-up stages (main.cu -c); 0 = start from t = q_v*fr */
  int verbose;    /* print the reference's progress lines on stdout */
  int ksave;      /* .inp ksave: write the .tuv every ksave steps (iterate.cu:64-71) */
} rism_opts;

/* the .xmu quantities (output_xmu.cu:7-111), J/mol unless noted */
typedef struct rism_xmu {
  int nsite;
  double sfe_sc, sfe_sc_hnc, sfe_gf;
  double sfec_sc[16], sfec_sc_hnc[16], sfec_gf[16];
  double se, se_lj, se_es;
  double pmv;        /* L/mol */
  double pressure;   /* J/m^3 */
  double correction; /* J/mol */
  int has_rmdft;     /* 1: sfe_rmdft is set and the .xmu gets its line (output_xmu.cu:25-28) */
  double sfe_rmdft;  /* J/mol */
} rism_xmu;

/* RMDFT input (solvent_read.cc:88-143): c_vv'(k) of the solvent (.cvk) and
   the reference hard-sphere fluid (.xsv).  The .xsv holds one site's data;
   the reference reads it into site 0 and leaves the other sites' arrays
   unset, which hold zeros in practice: callers pass zeros for them. */
typedef struct rism_rmdft_input {
  const double *cvk;   /* [ntab][V][V], the k grid and order of the solvent's chi table */
  int nhs;             /* hard-sphere table length */
  const double *khs;   /* [nhs] k, ascending */
  const double *chs;   /* [V][nhs] hard-sphere c(k) per reduced site */
  const double *whs;   /* [V][nhs] weight function w(k) per reduced site */
  const double *pfhs;  /* [V] packing fraction */
  const double *w0;    /* [V] w(k = 0) */
} rism_rmdft_input;

/* a parsed deck: .inp (+ optional separate solute file) and its .xvk */
typedef struct rism_deck {
  char outlist[64];
  int ksave;
  char solvent_path[4096];
  char fname[4096];     /* output base name (set_fname.cu:5-12) */
  rism_grid grid;
  rism_opts opts;
  int natom;
  double *sig, *eps, *q, *xyz;   /* solute, centred if requested */
  double shift[3];
  int nsite, ntab;
  double dk, temper, xt;
  double *rho, *sq, *ssig, *seps, *chi;
} rism_deck;

/* ---- library ---------------------------------------------------------- */
const char *rism_last_error(void);
const char *rism_version(void);
/* number of HIP devices visible (0 without a GPU) */
int rism_device_count(void);

/* ---- input surface (host only, no GPU needed) ------------------------- */
/* read_input.cu:21-89 + solute.cu:13-49 + solvent_read.cc:10-86.
   control: .inp path; structure: separate solute file or NULL;
   centering: solute.cu:13-43 shift; zero: main.cu -z (zero solute charges).
   The solvent path in the deck is opened as written, relative to the
   current directory, as the reference does; if absent there, it is tried
   relative to the directory of the control file. */
int rism_deck_read(const char *control, const char *structure, int centering,
                   int zero, rism_deck **out);
void rism_deck_free(rism_deck *d);
/* views of a deck as create() arguments (pointers into the deck) */
void rism_deck_views(const rism_deck *d, rism_grid *g, rism_solvent *sv,
                     rism_solute *su, rism_opts *o);
/* ESP cube reader (cal_Coulomb.cu:40-96): checks the axis counts against
   n[3] and stores the values x-fastest in out[N] (Hartree/e) */
int rism_read_cube(const char *path, const int *n, double *out);
/* RMDFT tables (solvent_read.cc:88-143).  .cvk: the table part of a file
   with the .xvk layout, [ntab][V][V].  .xsv: call with k == NULL to get the
   length in *nhs, then with arrays of at least cap entries: k, c_hs(k),
   w(k), packing fraction and w(0) of the reference fluid. */
int rism_read_cvk(const char *path, int natv, int ntab, double *out);
int rism_read_xsv(const char *path, int cap, int *nhs, double *k, double *c, double *w, double *pf,
                  double *w0);
/* .xvk reader alone (solvent_read.cc:10-86); arrays owned by the deck */
int rism_xvk_read(const char *path, rism_deck *into);

/* ---- solve lifecycle (RISM3D::initialize/iterate/output) -------------- */
/* RISM3D::initialize (initialize.cu:4-13) minus parsing: k-grid + chi
   spline (initialize_g.cu, solvent_spline.cu), device allocation.  device
   selects the HIP device (main.cu -d / cudaSetDevice). */
int rism_create(const rism_grid *grid, const rism_solvent *solvent,
                const rism_solute *solute, const rism_opts *opts, int device,
                rism_handle **out);
/* RISM3D::cal_potential (cal_potential.cu:3-14): LJ, Coulomb, fr, fk */
int rism_build_potentials(rism_handle *h);
/* main.cu -e with cal_Coulomb.cu:40-96: replace the solute's point-charge
   potential de by an electrostatic potential on the grid, esp[N] in
   Hartree/e, x-fastest (rism_read_cube gives it from a cube file).  The
   long-range parts fr, fk stay those of the point charges.  Call after
   rism_build_potentials. */
int rism_set_esp(rism_handle *h, const double *esp);
/* RISM3D::iterate (iterate.cu:6-90).  If tuv_path names an existing file
   the solve restarts from it and charge-up is skipped (iterate.cu:25-33).
   Returns RISM_SOLVE_* or RISM_ERR_*. */
int rism_solve(rism_handle *h, const char *tuv_path, int *steps, double *rms);
/* cal_exchem/cal_pmv/cal_pressure/cal_euv (output.cu:10-27) */
int rism_thermo(rism_handle *h, rism_xmu *out);
/* cal_grad (cal_grad.cu:8-48) scaled as written by output_grad.cu:15-24:
   out[n][6] = (dLJ_x, dLJ_y, dLJ_z, dES_x, dES_y, dES_z), J/mol/Angstrom */
int rism_grad(rism_handle *h, double *out);
/* rism_grad that allocates no device memory, for callers that cannot
   allocate (code under HIP graph capture, tight serving or MD loops,
   harnesses that forbid hipMalloc).  No counterpart in the reference.  Its
   partial sums use a static per-GPU device array of
   RISM_NOALLOC_SCRATCH_BYTES; the atoms are taken in batches that fit it.
   Returns RISM_ERR_ARG when one atom's partials plus the output do not fit
   (48 * (natom + ceil(N / 4096)) bytes above the limit, i.e. natom plus
   N / 4096 above about 700 000).  The results equal rism_grad's.  Calls that overlap on
   one GPU share the array and race; the handle's stream orders calls on one
   handle. */
#define RISM_NOALLOC_SCRATCH_BYTES (32L << 20)
int rism_grad_noalloc(rism_handle *h, double *out);
/* Adjoint gradient of Delta-mu (SFE_SC, J/mol) with respect to the solute
   parameters, at the converged state of the last rism_solve.  New: the
   reference has no counterpart (its cal_grad, rism_grad above, is the
   explicit formula sum_v rho_v sum_r g dU/dR, which the adjoint reproduces
   when the functional is stationary).  Solves the transposed fixed-point
   system lambda = (g'-1)(K^T lambda - lambda) + b with the Anderson scheme
   of the forward solve to rms <= rtol * rms(b), then sums the field
   adjoints against the derivatives of the potentials (cal_LJ.cu,
   cal_Coulomb.cu).  Outputs (any may be NULL): dxyz[n][3] J/mol/Angstrom,
   dq[n] J/mol/e, dsig[n] J/mol/Angstrom, deps[n] (J/mol per J/mol).
   Precision follows the handle: RISM_PREC_MIXED runs the iterations' K^T
   applications with FP32 transforms (rtol floored at 1e-6), everything else
   in FP64.  Returns RISM_OK, RISM_SOLVE_MAXSTEP (outputs from the last
   iterate) or an error; steps receives the adjoint iteration count. */
int rism_adjoint_grad(rism_handle *h, double rtol, double *dxyz, double *dq, double *dsig,
                      double *deps, int *steps);
/* copy a field (RISM_FIELD_*) to host memory */
int rism_get_field(rism_handle *h, int which, double *out);
/* replace t (host [V][N]), e.g. a restart from memory */
int rism_set_t(rism_handle *h, const double *t);
/* write_tuv.cu:7-33 / read_tuv.cu:6-34 (.tuv text, 6 significant digits) */
int rism_save_tuv(rism_handle *h, const char *path);
int rism_load_tuv(rism_handle *h, const char *path);
/* .xmu writer (output_xmu.cu:7-111) and .gra writer (output_grad.cu:7-30) */
int rism_write_xmu(const rism_xmu *x, int closure, const char *path);
int rism_write_gra(const double *grad, int natom, const char *path);
/* cal_rmdft (cal_rmdft.cu:5-114, solvent_spline2.cu:7-92) at the converged
   state: the RMDFT solvation free energy, J/mol (the .xmu SFE_RMDFT line) */
int rism_rmdft(rism_handle *h, const rism_rmdft_input *in, double *sfe_rmdft);
/* cal_euv per (atom, site) (cal_euv.cu:7-76) at the converged state, x dv as
   output_euv.cu:14-27 writes it: out[(kind * V + v) * natom + u], kind 0 =
   LJ, 1 = Coulomb, J/mol */
int rism_euv(rism_handle *h, double *out);
/* cal_qv (cal_qv.cu:6-27): solvent charge density sum_v rho_v q_v g_v, [N] */
int rism_qv(rism_handle *h, double *out);
/* field writers.  .guv binary (output_guv.cu:11-21): int V, int n[3],
   double box[3], double shift[3] (NULL: zeros), g [V][N] */
int rism_write_guv(const char *path, int natv, const int *n, const double *box, const double *shift,
                   const double *g);
/* .huv / .cuv text (output_huv.cu:9-42, output_cuv.cu:9-42): box, grid, the
   solute rows q sig eps x y z, V, then one row of V values per grid point
   (f = h for .huv, c = h - t for .cuv) */
int rism_write_huv(const char *path, const int *n, const double *box, int natom, const double *q,
                   const double *sig, const double *eps, const double *xyz, int natv, const double *f);
/* .euv text (output_euv.cu:14-27) from rism_euv's array */
int rism_write_euv(const char *path, int natv, int natom, const double *euv);
/* .qv text (output_qv.cu:15-32) from rism_qv's array */
int rism_write_qv(const char *path, const int *n, const double *box, const double *qv);
/* per-solve statistics: steps, final rms, and timings in seconds */
int rism_stats(rism_handle *h, int *steps, double *rms, double *t_setup,
               double *t_solve, double *t_thermo);
/* rms after each step of the last solve (n <= maxstep); returns count */
int rism_rms_history(rism_handle *h, double *out, int n);
/* release the handle.  Its device buffers, stream and pinned control block
   go to a process-wide cache that later handles of the same sizes reuse
   (RISM_POOL_BYTES bytes of buffers at most, 8 GiB by default; 0 frees at
   once): releasing them synchronises the device, which stalls every other
   handle's stream when many handles come and go. */
void rism_destroy(rism_handle *h);
/* free everything that cache holds, on every device (no counterpart in the
   reference); returns the number of buffer bytes freed.  Handles in use are
   not affected. */
long rism_release_cache(void);

/* ---- many solutes (no counterpart in the reference: one solve per run) ---- */
/* n independent solves sharing one solvent, each solute on its own grid
   (grids[i], solutes[i]), on one device.  The solutes are grouped by grid
   class (grid and box) and each class runs in batched launches: one launch
   per kernel covers every solute of the class (up to RISM_BATCH_MAX = 128
   per launch), each solute's control block gating its own workgroups (its
   convergence mask), the iteration chunks replayed as one HIP graph;
   nworkers host threads take classes from a shared queue.  With
   RISM_BATCH_LAUNCH=0 (or chargeup / ksave set) nworkers threads instead
   each own a handle and its HIP stream and take solutes one at a time.
   Per solute: status[i] (RISM_SOLVE_* or RISM_ERR_*), steps[i], rms[i],
   xmu[i] (any of steps/rms/xmu may be NULL); seconds receives the wall
   time.  Returns RISM_OK once every solve has run. */
int rism_batch_solve(int n, const rism_grid *grids, const rism_solvent *solvent,
                     const rism_solute *solutes, const rism_opts *opts, int device,
                     int nworkers, int *status, int *steps, double *rms, rism_xmu *xmu,
                     double *seconds);

/* ---- solvent susceptibility chi_vv(k): the table an .xvk file carries ---- */
/* The reference reads this table (solvent_read.cc:10-86) and never computes
   it.  These entry points compute it with a 1D XRISM solve (KH or HNC) on
   the DST-I grid r_j = j dr, k_i = i pi/(ntab dr), the scheme that reproduces
   the reference's own example .xvk files, or with its dielectrically
   consistent form DRISM (rism_chi_opts.dielec), and write it in that format.
   They run on the host (a 4096-16384 point problem, once per solvent). */
typedef struct rism_chi_site {  /* one .xvk site row (solvent_read.cc:42-47) */
  char name[8];
  int species;    /* molecule index, 1-based: sites of one species form one rigid molecule */
  int sym;        /* reduced-site index, 1-based; negative = equivalent to |sym| */
  double sig;     /* LJ sigma, Angstrom */
  double eps;     /* LJ epsilon, J/mol */
  double q;       /* charge, e */
  double xyz[3];  /* position in the molecule, Angstrom */
  double dens;    /* concentration of the species, mol/L */
} rism_chi_site;

typedef struct rism_chi_opts {
  int ntab;       /* table length: a power of two (the examples use 4096 and 16384) */
  double dr;      /* r spacing, Angstrom */
  double temper;  /* K */
  int closure;    /* RISM_CLOSURE_KH / _HNC */
  double tol;     /* rms residual of t per charge-up stage */
  int maxstep;    /* per stage */
  int nhist;      /* Anderson (DIIS form) vectors */
  double mix;     /* mixing of the residual */
  int chargeup;   /* stages n: Coulomb scaled by 0, 1/n, .., 1 (0: one full-charge stage) */
  double dielec;  /* > 0: DRISM (Perkyns & Pettitt) to this dielectric constant for the one
                     polar species; 0: XRISM */
  double adbcor;  /* DRISM: smoothing length a of h_c(k) = h_c0 exp(-k^2 a^2 / 4), Angstrom */
} rism_chi_opts;

/* ntab 16384, dr 0.01, 298.15 K, KH, tol 1e-10, maxstep 10000, 10 vectors,
   mix 0.3, 4 charge-up stages, XRISM (dielec 0), adbcor 0.5 */
void rism_chi_default_opts(rism_chi_opts *o);
/* number of reduced sites V = max |sym| */
int rism_chi_nreduced(const rism_chi_site *sites, int nsite);
/* chi[ntab][V][V] in the reader's order (h_v = sum_v' chi[v][v'] c_v'),
   xt (1/Pa), the total step count and the last rms (any output may be NULL).
   Returns RISM_SOLVE_CONVERGED, RISM_SOLVE_MAXSTEP (tables of the last
   iterate), RISM_SOLVE_DIVERGED or RISM_ERR_*. */
int rism_chi_solve(const rism_chi_site *sites, int nsite, const rism_chi_opts *opts, double *chi,
                   double *xt, int *steps, double *rms);
/* .xvk writer: the header lines and table layout solvent_read.cc:31-58 reads */
int rism_write_xvk(const char *path, const rism_chi_site *sites, int nsite, int ntab, double dr,
                   double temper, double xt, const double *chi);
/* the site rows, ntab, dr and temperature of an .xvk header, as a model to
   solve again (solvent_read.cc:31-48) */
int rism_xvk_read_sites(const char *path, rism_chi_site *sites, int maxsite, int *nsite, int *ntab,
                        double *dr, double *temper);

/* ---- measurement (no counterpart in the reference, which has no timers) */
/* kernel ids for rism_kernel_times */
#define RISM_K_CLOSURE_FX 0   /* Anderson update + closure + forward x FFT */
#define RISM_K_FFT_Y_FWD 1    /* forward y FFT */
#define RISM_K_FFT_Z_OZ 2     /* z FFT + k-space OZ mixing + inverse z FFT */
#define RISM_K_FFT_Y_INV 3    /* inverse y FFT */
#define RISM_K_RESID 4        /* inverse x FFT + residual + Anderson dots */
#define RISM_K_SOLVE 5        /* single-workgroup reduction + Anderson solve */
#define RISM_K_POT 6          /* LJ + Coulomb grid potentials, with the SE atom sums */
#define RISM_K_FK 7           /* long-range k-space term */
#define RISM_K_EXCHEM 8       /* free-energy / PMV sums */
#define RISM_K_EUV 9          /* solute-solvent energy: weighted sum over the grid */
#define RISM_K_REFRESH 10     /* mixed precision: FP64 passes of a refinement step */
#define RISM_NKERNEL 11
/* enable (1) or disable (0) HIP-event timing of every launch; resets the
   accumulators.  While enabled the solver reads its control block after
   every iteration. */
int rism_profile(rism_handle *h, int enable);
/* accumulated milliseconds and launch counts per kernel id */
int rism_kernel_times(rism_handle *h, double *ms, long *count);
/* the HBM ceiling the iteration kernels are compared with: a plain copy of
   `bytes` (a multiple of 16) between two fresh device buffers, one 16-byte
   element per thread, timed with HIP events over `reps` launches.  gbps
   receives (read + written bytes) / time, us the time per copy. */
int rism_stream_copy(int device, size_t bytes, int reps, double *gbps, double *us);
/* Tile autotuner (no counterpart in the reference): times one iteration of
   the field passes for candidate tile shapes around the defaults (x-line
   group and workgroup count, y tile, z tile; `reps` timed launches each)
   and keeps the fastest.  Call after rism_build_potentials and before
   rism_solve; the state is reset to the initial estimate.  Results are
   cached per grid, site count, history class and precision for the life of
   the process.  Setting RISM_AUTOTUNE=<reps> makes rism_solve call it.
   geom_out (may be NULL) receives rism_geometry's ten values. */
int rism_autotune(rism_handle *h, int reps, int *geom_out);
/* launch geometry: Lx, Ty, Tz, ept_x, ept_y, ept_z, xgroups, xblocks, packs, history class */
int rism_geometry(rism_handle *h, int *out /* [10] */);

/* ---- hot-path seams on caller device memory --------------------------- */
/* FFT3D::execute (fft3d.cu:22-38): in-place transform of one x-fastest
   complex field with the reference's convention (dir -1: x dv with the
   half-shifted k grid; +1: / volume).  data: device pointer to
   double2[N] (precision FP64) or float2[N] (mixed).  Runs on the handle's
   stream and synchronises before returning. */
int rism_fft3d(rism_handle *h, void *data, int dir);
/* RISM3D::calculate (calculate.cu:4-58) for charge factor 1 on device
   memory: t [V][N] in, residual [V][N] out (both double). */
int rism_calculate(rism_handle *h, const double *t_dev, double *res_dev);

#ifdef __cplusplus
}
#endif
#endif /* RISM3D_AMD_H */
