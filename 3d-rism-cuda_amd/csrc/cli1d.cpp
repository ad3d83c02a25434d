// cli1d.cpp -- rism1d-amd: computes the chi_vv(k) table of a solvent model
// and writes it as an .xvk file, the input the 3D solve reads
// (src-cuda/solvent_read.cc:10-86).  The model is the header of an .xvk
// file (its site rows, ntab, dr and temperature); options override the grid.
//
//   rism1d-amd [-c KH|HNC] [-n ntab] [-r dr] [-T K] [-t tol] [-u stages]
//              [-m vectors] [-x mix] [-e eps [-a adbcor]] model.xvk out.xvk
// -e eps selects DRISM to dielectric constant eps (XRISM without it).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rism3d_amd.h"

static void usage() {
  std::fprintf(stderr,
               "usage: rism1d-amd [-c KH|HNC] [-n ntab] [-r dr] [-T K] [-t tol] [-u chargeup] [-m nhist] "
               "[-x mix] [-e eps (DRISM) [-a adbcor]] model.xvk out.xvk\n");
  std::exit(1);
}

int main(int argc, char **argv) {
  rism_chi_opts o;
  rism_chi_default_opts(&o);
  int ntab = 0, closure = -1, i = 1;
  double dr = 0.0, temp = 0.0;
  for (; i < argc && argv[i][0] == '-' && argv[i][1]; ++i) {
    const char f = argv[i][1];
    if (i + 1 >= argc) usage();
    const char *v = argv[++i];
    switch (f) {
      case 'c':
        if (!std::strcmp(v, "KH")) closure = RISM_CLOSURE_KH;
        else if (!std::strcmp(v, "HNC")) closure = RISM_CLOSURE_HNC;
        else usage();
        break;
      case 'n': ntab = std::atoi(v); break;
      case 'r': dr = std::atof(v); break;
      case 'T': temp = std::atof(v); break;
      case 't': o.tol = std::atof(v); break;
      case 'u': o.chargeup = std::atoi(v); break;
      case 'm': o.nhist = std::atoi(v); break;
      case 'x': o.mix = std::atof(v); break;
      case 'e': o.dielec = std::atof(v); break;
      case 'a': o.adbcor = std::atof(v); break;
      default: usage();
    }
  }
  if (argc - i != 2) usage();
  std::vector<rism_chi_site> sites(64);
  int ns = 0;
  if (rism_xvk_read_sites(argv[i], sites.data(), (int)sites.size(), &ns, &o.ntab, &o.dr, &o.temper) != RISM_OK) {
    std::fprintf(stderr, "%s\n", rism_last_error());
    return 1;
  }
  if (ntab) o.ntab = ntab;
  if (dr > 0.0) o.dr = dr;
  if (temp > 0.0) o.temper = temp;
  if (closure >= 0) o.closure = closure;
  const int V = rism_chi_nreduced(sites.data(), ns);
  std::vector<double> chi((size_t)o.ntab * V * V);
  double xt = 0.0, rms = 0.0;
  int steps = 0;
  const auto t0 = std::chrono::steady_clock::now();
  const int rc = rism_chi_solve(sites.data(), ns, &o, chi.data(), &xt, &steps, &rms);
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (rc < 0 || rc == RISM_SOLVE_DIVERGED) {
    std::fprintf(stderr, "%s\n", rism_last_error());
    return 1;
  }
  if (rc == RISM_SOLVE_MAXSTEP) std::printf("warning: step limit reached (rms %.3e)\n", rms);
  if (rism_write_xvk(argv[i + 1], sites.data(), ns, o.ntab, o.dr, o.temper, xt, chi.data()) != RISM_OK) {
    std::fprintf(stderr, "%s\n", rism_last_error());
    return 1;
  }
  std::printf("%s, %d sites (%d reduced), ntab %d, dr %.4f, %.2f K: %d steps, rms %.3e, xt %.8e 1/Pa, %.2f s\n",
              o.dielec > 0.0 ? "DRISM" : "XRISM", ns, V, o.ntab, o.dr, o.temper, steps, rms, xt, sec);
  return 0;
}
