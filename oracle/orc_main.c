/*
 * orc_main.c -- command-line driver for the FP64 oracle (test
 * infrastructure / CPU baseline).  Flags follow src-cuda/main.cu:20-54:
 *   -c N  charge-up stages   -i inp   -s solute   -f no centring   -z zero q
 * plus oracle-only flags:
 *   -H n  Anderson history slots (2 = reference)   -T n  OpenMP threads
 *   -n K  cap the iteration count (bounded CPU-baseline sample)
 *   -j    print a JSON timing record on stdout after the run
 *   -q    quiet (no per-step lines)
 *   -G f  write g of the final state to f as raw float64 [V][N]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "rism_oracle.h"

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv) {
  int cu = 0, centering = 1, zero = 0, nhist = 2, threads = 0, cap = 0, json = 0, quiet = 0;
  const char *input = NULL, *structure = NULL, *gout = NULL;
  int ch;
  long sdiv = 1;
  while ((ch = getopt(argc, argv, "c:d:i:s:fzH:T:n:x:jqG:")) != -1) {
    switch (ch) {
      case 'c': cu = atoi(optarg); break;
      case 'd': break;  /* device index: meaningless for the host oracle */
      case 'i': input = optarg; break;
      case 's': structure = optarg; break;
      case 'f': centering = 0; break;
      case 'z': zero = 1; break;
      case 'H': nhist = atoi(optarg); break;
      case 'T': threads = atoi(optarg); break;
      case 'n': cap = atoi(optarg); break;
      case 'x': sdiv = atol(optarg); break;  /* timing sample: 1/x of the points for potentials and SE */
      case 'j': json = 1; break;
      case 'q': quiet = 1; break;
      case 'G': gout = optarg; break;
      default: return 1;
    }
  }
  if (!input || !structure) {
    if (optind >= argc) { printf("No input file!\n"); return 1; }
    input = argv[optind];
    structure = NULL;
  }
  /* per-step lines reach a pipe as they happen */
  setvbuf(stdout, NULL, _IOLBF, 0);
  if (threads > 0) orc_set_threads(threads);
  int nthr = 1;
#ifdef _OPENMP
  nthr = omp_get_max_threads();
#endif
  orc_sys *s = orc_new();
  s->nhist = nhist;
  if (!quiet) printf("reading input data file  : %s\n", input);
  if (orc_read_inp(s, input, structure, centering, zero)) {
    printf("%s\n", orc_last_error());
    return 1;
  }
  s->nhist = nhist;
  if (cap > 0) s->maxstep = cap;
  /* fname: structure (if given) else control, minus extension (set_fname.cu) */
  char fname[4096];
  snprintf(fname, sizeof fname, "%s", structure ? structure : input);
  char *dot = strrchr(fname, '.');
  if (dot) *dot = '\0';
  if (orc_read_xvk(s, s->fsolvent)) { printf("%s\n", orc_last_error()); return 1; }
  for (int d = 0; d < 3; ++d)
    if (s->n[d] < 2 || (s->n[d] & 1)) { printf("grid sizes must be even\n"); return 1; }
  if (sdiv > 1) s->sample = s->ng / sdiv;
  double t0 = now();
  if (orc_kgrid(s) || orc_spline_chi(s)) { printf("%s\n", orc_last_error()); return 1; }
  double t0p = now();
  if (orc_potentials(s)) { printf("%s\n", orc_last_error()); return 1; }
  double t1 = now();
  char tuv[4200];
  snprintf(tuv, sizeof tuv, "%s.tuv", fname);
  int steps = 0;
  double rms = 0.0;
  int st = orc_iterate(s, cu, !quiet, tuv, &steps, &rms);
  double t2 = now();
  orc_xmu x;
  orc_thermo(s, &x);
  double t3 = now();
  char out[4200];
  snprintf(out, sizeof out, "%s%s", fname, zero ? ".xmu0" : ".xmu");
  if (strchr(s->outlist, 'm') || strchr(s->outlist, 'M')) orc_write_xmu(s, &x, out);
  if (gout) {
    const size_t VN = (size_t)s->natv * s->ng;
    double *g = malloc(VN * sizeof(double));
    FILE *f = fopen(gout, "wb");
    if (!g || !f || (orc_copy_field(s, 2, g), fwrite(g, sizeof(double), VN, f)) != VN) {
      printf("cannot write %s\n", gout);
      st = -1;
    }
    if (f) fclose(f);
    free(g);
  }
  if (json)
    printf("{\"status\": %d, \"steps\": %d, \"rms\": %.6e, \"t_setup\": %.6f, \"t_pot\": %.6f, "
           "\"t_iter\": %.6f, \"t_thermo\": %.6f, \"threads\": %d, \"ng\": %ld, \"sample_div\": %ld, "
           "\"sfe_sc\": %.10f, \"sfe_gf\": %.10f, \"pmv\": %.10f, \"se\": %.10f}\n",
           st, steps, rms, t1 - t0, t1 - t0p, t2 - t1, t3 - t2, nthr, s->ng, sdiv, x.sfe_sc, x.sfe_gf, x.pmv, x.se);
  orc_free(s);
  return st < 0 ? 1 : 0;
}
