// batch_stress.cpp -- host-side stress of the library's multi-threaded paths,
// built with AddressSanitizer on the host code (`make asan`; the device code
// is the same as in the regular build).
//
// With the solvent of a deck (argv[1], e.g. the reference's test1.inp and
// its h2o_vv.xvk) it solves 24 seeded solutes (6-30 LJ atoms with charges) on
// 48^3, 60^3 and 64^3 grids at 0.5 A spacing:
//   1. rism_batch_solve on the per-solute path (RISM_BATCH_LAUNCH=0: worker
//      threads that each own a handle and stream, reuse destroyed handles'
//      buffers, streams and pinned control blocks, and swap solutes on a
//      handle) with 1 and 6 workers, and on the batched-launch path with 1
//      and 3 workers, in FP64: every solve converges, and every run gives
//      the same step counts and SFE_SC as the first (same arithmetic);
//   2. four threads that each create, solve, read the thermodynamics of and
//      destroy handles on the same device at once, against the same results.
// Exit status 0 when all of that holds; AddressSanitizer aborts the process
// on an invalid host memory access.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "rism3d_amd.h"

namespace {

struct Mol {
  rism_grid grid;
  std::vector<double> sig, eps, q, xyz;
  rism_solute view() const {
    rism_solute s;
    s.natom = (int)q.size();
    s.sig = sig.data();
    s.eps = eps.data();
    s.q = q.data();
    s.xyz = xyz.data();
    return s;
  }
};

// deterministic 64-bit LCG (Knuth's MMIX constants), uniform in [0, 1)
struct Lcg {
  unsigned long long s;
  double next() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
};

std::vector<Mol> make_solutes(int count) {
  std::vector<Mol> out;
  Lcg r{20220301ULL};
  const int sizes[3] = {48, 60, 64};
  for (int i = 0; i < count; ++i) {
    Mol m;
    const int n = sizes[i % 3];
    m.grid.n[0] = m.grid.n[1] = m.grid.n[2] = n;
    m.grid.box[0] = m.grid.box[1] = m.grid.box[2] = 0.5 * n;
    const int na = 6 + (int)(r.next() * 25.0);
    double qsum = 0.0;
    for (int a = 0; a < na; ++a) {
      m.sig.push_back(2.5 + 1.3 * r.next());
      m.eps.push_back(200.0 + 600.0 * r.next());
      m.q.push_back(r.next() - 0.5);
      qsum += m.q.back();
      for (int k = 0; k < 3; ++k) m.xyz.push_back(-3.0 + 6.0 * r.next());
    }
    for (double &c : m.q) c -= qsum / na;  // neutral solutes
    out.push_back(m);
  }
  return out;
}

bool same(double a, double b) { return std::fabs(a - b) <= 1e-12 * std::fmax(1.0, std::fabs(b)); }

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s deck.inp\n", argv[0]);
    return 2;
  }
  rism_deck *d = nullptr;
  if (rism_deck_read(argv[1], nullptr, 0, 0, &d) != RISM_OK) {
    std::fprintf(stderr, "deck: %s\n", rism_last_error());
    return 2;
  }
  rism_grid g0;
  rism_solvent sv;
  rism_solute su0;
  rism_opts o;
  rism_deck_views(d, &g0, &sv, &su0, &o);
  o.precision = RISM_PREC_FP64;
  o.verbose = 0;
  o.ksave = 0;
  o.chargeup = 0;
  const std::vector<Mol> mols = make_solutes(24);
  const int n = (int)mols.size();
  std::vector<rism_grid> grids(n);
  std::vector<rism_solute> sus(n);
  for (int i = 0; i < n; ++i) {
    grids[i] = mols[i].grid;
    sus[i] = mols[i].view();
  }
  int bad = 0;
  std::vector<int> ref_steps;
  std::vector<double> ref_sfe;
  const struct { const char *launch; int workers; } runs[] = {{"0", 1}, {"0", 6}, {"1", 1}, {"1", 3}};
  for (const auto &run : runs) {
    setenv("RISM_BATCH_LAUNCH", run.launch, 1);
    std::vector<int> status(n), steps(n);
    std::vector<double> rms(n);
    std::vector<rism_xmu> xmu(n);
    double sec = 0.0;
    std::printf("batch: batched launches %s, %d workers ... ", run.launch, run.workers);
    std::fflush(stdout);
    const int rc = rism_batch_solve(n, grids.data(), &sv, sus.data(), &o, 0, run.workers, status.data(),
                                    steps.data(), rms.data(), xmu.data(), &sec);
    if (rc != RISM_OK) {
      std::printf("failed: %s\n", rism_last_error());
      return 1;
    }
    std::printf("%.2f s\n", sec);
    for (int i = 0; i < n; ++i) {
      if (status[i] != RISM_SOLVE_CONVERGED) {
        std::printf("  solute %d: status %d\n", i, status[i]);
        ++bad;
      }
      if (ref_steps.empty()) continue;
      if (steps[i] != ref_steps[i] || !same(xmu[i].sfe_sc, ref_sfe[i])) {
        std::printf("  solute %d: %d steps, SFE_SC %.12g against %d, %.12g\n", i, steps[i], xmu[i].sfe_sc,
                    ref_steps[i], ref_sfe[i]);
        ++bad;
      }
    }
    if (ref_steps.empty()) {
      ref_steps = steps;
      for (int i = 0; i < n; ++i) ref_sfe.push_back(xmu[i].sfe_sc);
    }
  }
  // handles created, solved and destroyed by four threads at once
  std::printf("threads: 4 x create/solve/thermo/destroy ... ");
  std::fflush(stdout);
  std::vector<int> tbad(4, 0);
  std::vector<std::thread> th;
  for (int t = 0; t < 4; ++t)
    th.emplace_back([&, t]() {
      for (int i = t; i < n; i += 4) {
        rism_handle *h = nullptr;
        int st = 0;
        double rr = 0.0;
        rism_xmu x;
        if (rism_create(&grids[i], &sv, &sus[i], &o, 0, &h) != RISM_OK || rism_build_potentials(h) != RISM_OK ||
            rism_solve(h, nullptr, &st, &rr) != RISM_SOLVE_CONVERGED || rism_thermo(h, &x) != RISM_OK ||
            st != ref_steps[i] || !same(x.sfe_sc, ref_sfe[i]))
          ++tbad[t];
        if (h) rism_destroy(h);
      }
    });
  for (auto &t : th) t.join();
  for (int t = 0; t < 4; ++t) bad += tbad[t];
  std::printf("%s\n", bad ? "mismatch" : "ok");
  rism_deck_free(d);
  std::printf("%s: %d problems\n", bad ? "FAIL" : "PASS", bad);
  return bad ? 1 : 0;
}
