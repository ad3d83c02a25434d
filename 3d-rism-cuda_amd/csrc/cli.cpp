// cli.cpp -- rism3d-amd: drop-in for the reference executable 3drism-cuda
// (src-cuda/main.cu:6-64 with RISM3D::output, output.cu:6-62).
//
//   rism3d-amd [-c N] [-d dev] [-i inp] [-s solute] [-f] [-z] [-p prec] [-H n] [inp]
//
// Reference flags: -c charge-up stages, -d device, -i control deck, -s solute
// file, -f no centring, -z zero solute charges (writes .xmu0).  Added flags:
// -p 0|1 FP64 (default) or mixed precision, -H Anderson history slots
// (default 2, the reference's).  -r (RMDFT) and -e (ESP cube) are not part
// of this build and are rejected with a message.
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "rism3d_amd.h"

int main(int argc, char **argv) {
  int cu = 0, dn = 0, prec = RISM_PREC_FP64, nhist = 2;
  std::string input, structure;
  bool centering = true, q0 = false;
  int ch;
  while ((ch = getopt(argc, argv, "c:d:i:s:r:e:fzp:H:")) != -1) {
    switch (ch) {
      case 'c': cu = std::atoi(optarg); break;
      case 'd': dn = std::atoi(optarg); break;
      case 'i': input = optarg; break;
      case 's': structure = optarg; break;
      case 'r': std::printf("RMDFT (-r) is not supported by this build.\n"); return 1;
      case 'e': std::printf("ESP cube input (-e) is not supported by this build.\n"); return 1;
      case 'f': centering = false; break;
      case 'z': q0 = true; break;
      case 'p': prec = std::atoi(optarg); break;
      case 'H': nhist = std::atoi(optarg); break;
      default: return 1;
    }
  }
  if (input.empty() || structure.empty()) {  // main.cu:48-54
    if (optind >= argc) {
      std::printf("No input file!\n");
      return 1;
    }
    input = argv[optind];
  }
  std::printf("Set device %d\n", dn);
  if (cu > 0) std::printf("Charge up %d\n", cu);
  std::printf("reading input data file  : %s\n", input.c_str());
  if (!structure.empty()) std::printf("reading solute data file : %s\n", structure.c_str());
  rism_deck *deck = nullptr;
  int rc = rism_deck_read(input.c_str(), structure.empty() ? nullptr : structure.c_str(), centering, q0, &deck);
  if (rc != RISM_OK) {
    std::printf("%s\n", rism_last_error());
    return 1;
  }
  rism_grid g;
  rism_solvent sv;
  rism_solute su;
  rism_opts o;
  rism_deck_views(deck, &g, &sv, &su, &o);
  o.chargeup = cu;
  o.precision = prec;
  o.nhist = nhist;
  o.verbose = 1;
  o.ksave = deck->ksave;
  if (g.box[0] / g.n[0] > 0.5 || g.box[1] / g.n[1] > 0.5 || g.box[2] / g.n[2] > 0.5) {  // cell.cc:16-21
    std::printf("##########################################\n");
    std::printf("WARRING: Grid spacing is greater than 0.5.\n");
    std::printf("##########################################\n");
  }
  rism_handle *h = nullptr;
  rc = rism_create(&g, &sv, &su, &o, dn, &h);
  if (rc != RISM_OK) {
    std::printf("%s\n", rism_last_error());
    rism_deck_free(deck);
    return 1;
  }
  std::printf("tabulating solute Lennard-Jones potential ...\n");
  std::printf("synthesizing solute Coulomb potential ...\n");
  if (rism_build_potentials(h) != RISM_OK) {
    std::printf("%s\n", rism_last_error());
    return 1;
  }
  const std::string fname = deck->fname;
  const std::string tuv = fname + ".tuv";
  int steps = 0;
  double rms = 0.0;
  rc = rism_solve(h, tuv.c_str(), &steps, &rms);
  if (rc < 0) {
    std::printf("%s\n", rism_last_error());
    return 1;
  }
  std::string outlist = deck->outlist;
  std::transform(outlist.begin(), outlist.end(), outlist.begin(), ::tolower);
  if (outlist.find('m') != std::string::npos) {
    rism_xmu x;
    if (rism_thermo(h, &x) != RISM_OK) {
      std::printf("%s\n", rism_last_error());
      return 1;
    }
    if (rism_write_xmu(&x, o.closure, (fname + (q0 ? ".xmu0" : ".xmu")).c_str()) != RISM_OK) {
      std::printf("%s\n", rism_last_error());
      return 1;
    }
  }
  if (outlist.find('d') != std::string::npos) {
    std::vector<double> gr((size_t)su.natom * 6);
    std::printf("outputting grad to file:  %s  ...\n", (fname + ".gra").c_str());
    if (rism_grad(h, gr.data()) != RISM_OK) {
      std::printf("%s\n", rism_last_error());
      return 1;
    }
    if (rism_write_gra(gr.data(), su.natom, (fname + ".gra").c_str()) != RISM_OK) {
      std::printf("%s\n", rism_last_error());
      return 1;
    }
    std::printf("done.\n");
  }
  for (char c : std::string("cgheq"))
    if (outlist.find(c) != std::string::npos)
      std::printf("note: field writer '%c' is not part of this build\n", c);
  rism_destroy(h);
  rism_deck_free(deck);
  return 0;
}
