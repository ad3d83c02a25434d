// cli.cpp -- rism3d-amd: drop-in for the reference executable 3drism-cuda
// (src-cuda/main.cu:6-64 with RISM3D::output, output.cu:6-62).
//
//   rism3d-amd [-c N] [-d dev] [-i inp] [-s solute] [-e cube] [-f x] [-z] [-p prec] [-H n] [inp]
//
// Reference flags: -c charge-up stages, -d device, -i control deck, -s solute
// file, -e ESP cube replacing the point-charge potential (it also turns
// centring off, as main.cu:37-41 falls through to -f), -f <any> no centring
// (the reference's getopt string "f:" makes -f take an argument), -z zero
// solute charges (writes .xmu0), -r <file.xsv> RMDFT free energy with the
// reference fluid's hard-sphere tables (the solvent's c(k) is read from the
// .cvk beside its .xvk, solvent_read.cc:91-95).  Added flags: -p 0|1 FP64
// (default) or mixed precision, -H Anderson history slots (default 2, the
// reference's).  The outlist letters of the
// deck select the outputs as in output.cu:6-62: m .xmu, d .gra, c .cuv,
// g .guv, h .huv, e .euv, q .qv.
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
  std::string input, structure, esp, hs;
  bool centering = true, q0 = false;
  int ch;
  while ((ch = getopt(argc, argv, "c:d:i:s:r:e:f:zp:H:")) != -1) {
    switch (ch) {
      case 'c': cu = std::atoi(optarg); break;
      case 'd': dn = std::atoi(optarg); break;
      case 'i': input = optarg; break;
      case 's': structure = optarg; break;
      case 'r': hs = optarg; break;
      case 'e': esp = optarg; centering = false; break;
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
  if (!esp.empty()) {
    std::vector<double> e((size_t)g.n[0] * g.n[1] * g.n[2]);
    if (rism_read_cube(esp.c_str(), g.n, e.data()) != RISM_OK || rism_set_esp(h, e.data()) != RISM_OK) {
      std::printf("%s\n", rism_last_error());
      return 1;
    }
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
    if (!hs.empty()) {  // cal_rmdft (output.cu:24): c(k) from <solvent>.cvk, hard spheres from -r
      const std::string sp = deck->solvent_path;
      const std::string cvk = sp.substr(0, sp.rfind('.')) + ".cvk";
      const int V = sv.nsite;
      std::vector<double> c((size_t)sv.ntab * V * V);
      int nh = 0;
      if (rism_read_cvk(cvk.c_str(), V, sv.ntab, c.data()) != RISM_OK ||
          rism_read_xsv(hs.c_str(), 0, &nh, nullptr, nullptr, nullptr, nullptr, nullptr) != RISM_OK) {
        std::printf("%s\n", rism_last_error());
        return 1;
      }
      std::vector<double> kh(nh), ch((size_t)V * nh, 0.0), wh((size_t)V * nh, 0.0), pf(V, 0.0), w0(V, 0.0);
      if (rism_read_xsv(hs.c_str(), nh, &nh, kh.data(), ch.data(), wh.data(), &pf[0], &w0[0]) != RISM_OK) {
        std::printf("%s\n", rism_last_error());
        return 1;
      }
      rism_rmdft_input ri = {c.data(), nh, kh.data(), ch.data(), wh.data(), pf.data(), w0.data()};
      if (rism_rmdft(h, &ri, &x.sfe_rmdft) != RISM_OK) {
        std::printf("%s\n", rism_last_error());
        return 1;
      }
      x.has_rmdft = 1;
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
  // field writers in output.cu:38-61 order: c, g, h, e, q
  const size_t N = (size_t)g.n[0] * g.n[1] * g.n[2];
  auto fail = [&]() {
    std::printf("%s\n", rism_last_error());
    return 1;
  };
  struct Writer {
    char key;
    const char *label, *ext;
    int field;
  };
  const Writer tw[] = {{'c', "Cuv", ".cuv", RISM_FIELD_C}, {'g', "Guv", ".guv", RISM_FIELD_G},
                       {'h', "Huv", ".huv", RISM_FIELD_H}};
  for (const Writer &w : tw) {
    if (outlist.find(w.key) == std::string::npos) continue;
    const std::string path = fname + w.ext;
    std::printf("outputting %s to file:  %s  ...\n", w.label, path.c_str());
    std::vector<double> f(N * sv.nsite);
    if (rism_get_field(h, w.field, f.data()) != RISM_OK) return fail();
    const int rc2 = w.key == 'g'
                        ? rism_write_guv(path.c_str(), sv.nsite, g.n, g.box, centering ? deck->shift : nullptr, f.data())
                        : rism_write_huv(path.c_str(), g.n, g.box, su.natom, su.q, su.sig, su.eps, su.xyz, sv.nsite,
                                         f.data());
    if (rc2 != RISM_OK) return fail();
    std::printf("done.\n");
  }
  if (outlist.find('e') != std::string::npos) {
    const std::string path = fname + ".euv";
    std::printf("outputting euv to file:  %s  ...\n", path.c_str());
    std::vector<double> e((size_t)2 * sv.nsite * su.natom);
    if (rism_euv(h, e.data()) != RISM_OK || rism_write_euv(path.c_str(), sv.nsite, su.natom, e.data()) != RISM_OK)
      return fail();
    std::printf("done.\n");
  }
  if (outlist.find('q') != std::string::npos) {
    const std::string path = fname + ".qv";
    std::printf("outputting qv to file:  %s  ...\n", path.c_str());
    std::vector<double> qv(N);
    if (rism_qv(h, qv.data()) != RISM_OK || rism_write_qv(path.c_str(), g.n, g.box, qv.data()) != RISM_OK)
      return fail();
    std::printf("done.\n");
  }
  rism_destroy(h);
  rism_deck_free(deck);
  return 0;
}
