// batch.cpp -- many independent 3D-RISM solves (one solvent, many solutes
// on their own grids) run concurrently on one GPU.
//
// Small grids (48^3 - 96^3) leave most of the device idle when solved one
// at a time: a 64^3 iteration is a few dozen microseconds of launches.
// Here nworkers host threads each own a handle (and so a HIP stream) and
// take solutes from a shared counter: potentials, solve, thermo.  A worker
// keeps its handle while the grid stays the same and only swaps the solute
// (rism::set_solute), so the field buffers and the captured iteration
// graphs carry over; a new grid, or an error, gets a new handle.  The
// streams' kernels overlap on the device; every solve keeps its own
// convergence test (per-solute masks are the handles' control blocks).
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "rism3d_amd.h"
#include "rism_host.h"

namespace {
bool same_grid(const rism_grid &a, const rism_grid &b) {
  for (int k = 0; k < 3; ++k)
    if (a.n[k] != b.n[k] || a.box[k] != b.box[k]) return false;
  return true;
}
}  // namespace

extern "C" int rism_batch_solve(int n, const rism_grid *grids, const rism_solvent *solvent,
                                const rism_solute *solutes, const rism_opts *opts, int device, int nworkers,
                                int *status, int *steps, double *rms, rism_xmu *xmu, double *seconds) {
  if (n < 0 || (n > 0 && (!grids || !solvent || !solutes || !opts || !status))) return RISM_ERR_ARG;
  if (nworkers < 1) nworkers = 1;
  if (nworkers > n) nworkers = n > 0 ? n : 1;
  std::atomic<int> next(0);
  const auto t0 = std::chrono::steady_clock::now();
  auto work = [&]() {
    rism_handle *h = nullptr;
    rism_grid hg{};
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= n) break;
      int st = RISM_OK;
      if (h && same_grid(hg, grids[i])) {
        st = rism::set_solute(h, &solutes[i]);
      } else {
        if (h) rism_destroy(h);
        h = nullptr;
        st = rism_create(&grids[i], solvent, &solutes[i], opts, device, &h);
        hg = grids[i];
      }
      int nstep = 0;
      double r = 0.0;
      if (st == RISM_OK) st = rism_build_potentials(h);
      if (st == RISM_OK) st = rism_solve(h, nullptr, &nstep, &r);
      if (st >= 0 && xmu) {
        const int t = rism_thermo(h, &xmu[i]);
        if (t < 0) st = t;
      }
      if (st < 0) std::fprintf(stderr, "rism_batch_solve: solute %d: %s\n", i, rism_last_error());
      status[i] = st;
      if (steps) steps[i] = nstep;
      if (rms) rms[i] = r;
      if (st < 0) {  // a handle that failed is not used again
        rism_destroy(h);
        h = nullptr;
      }
    }
    if (h) rism_destroy(h);
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < nworkers; ++w) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return RISM_OK;
}
