// batch.cpp -- many independent 3D-RISM solves (one solvent, many solutes
// on their own grids) on one GPU.
//
// Small grids (48^3 - 96^3) leave most of the device idle when solved one
// at a time: a 64^3 iteration is a few dozen microseconds of launches.
// The solutes are grouped by grid class (grid and box); each class is
// solved in batched launches (rism::batch_solve_class): one launch per
// kernel covers every solute of the class, each solute's control block
// gating its own workgroups (its convergence mask), with the iteration
// chunks captured as one HIP graph.  nworkers host threads take classes
// (in waves of at most RISM_BATCH_MAX solutes, 128 by default) from a
// shared queue, so the classes' launches overlap on the device.
// RISM_BATCH_LAUNCH=0 selects the per-solute path instead: each worker owns
// a handle and its stream and takes solutes one at a time, keeping its
// handle while the grid stays the same (rism::set_solute).
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>

#include "rism3d_amd.h"
#include "rism_host.h"

namespace {
bool same_grid(const rism_grid &a, const rism_grid &b) {
  for (int k = 0; k < 3; ++k)
    if (a.n[k] != b.n[k] || a.box[k] != b.box[k]) return false;
  return true;
}
}  // namespace

namespace {
// batched launches per grid class
void solve_by_class(int n, const rism_grid *grids, const rism_solvent *solvent, const rism_solute *solutes,
                    const rism_opts *opts, int device, int nworkers, int *status, int *steps, double *rms,
                    rism_xmu *xmu) {
  const char *env = std::getenv("RISM_BATCH_MAX");
  const int smax = (env && std::atoi(env) > 0) ? std::atoi(env) : 128;
  std::vector<std::vector<int>> waves;  // solute indices of one class, at most smax
  std::vector<std::vector<int>> classes;
  for (int i = 0; i < n; ++i) {
    int c = -1;
    for (size_t k = 0; k < classes.size(); ++k)
      if (same_grid(grids[classes[k][0]], grids[i])) { c = (int)k; break; }
    if (c < 0) {
      classes.push_back({});
      c = (int)classes.size() - 1;
    }
    classes[c].push_back(i);
  }
  for (auto &c : classes)
    for (size_t a = 0; a < c.size(); a += smax)
      waves.emplace_back(c.begin() + a, c.begin() + std::min(c.size(), a + smax));
  // largest waves first, so the long ones start early
  std::sort(waves.begin(), waves.end(), [](const std::vector<int> &a, const std::vector<int> &b) { return a.size() > b.size(); });
  std::atomic<int> next(0);
  auto work = [&]() {
    for (;;) {
      const int w = next.fetch_add(1);
      if (w >= (int)waves.size()) break;
      const std::vector<int> &ix = waves[w];
      const int S = (int)ix.size();
      std::vector<const rism_solute *> sus(S);
      for (int s = 0; s < S; ++s) sus[s] = &solutes[ix[s]];
      std::vector<int> st(S), sp(S);
      std::vector<double> r(S);
      std::vector<rism_xmu> x(S);
      const int rc = rism::batch_solve_class(&grids[ix[0]], solvent, sus.data(), S, opts, device, st.data(), sp.data(),
                                             r.data(), xmu ? x.data() : nullptr);
      if (rc < 0) std::fprintf(stderr, "rism_batch_solve: grid class of solute %d: %s\n", ix[0], rism_last_error());
      for (int s = 0; s < S; ++s) {
        status[ix[s]] = st[s];
        if (steps) steps[ix[s]] = sp[s];
        if (rms) rms[ix[s]] = r[s];
        if (xmu) xmu[ix[s]] = x[s];
      }
    }
  };
  const int nw = std::max(1, std::min(nworkers, (int)waves.size()));
  std::vector<std::thread> pool;
  for (int w = 1; w < nw; ++w) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
}
}  // namespace

extern "C" int rism_batch_solve(int n, const rism_grid *grids, const rism_solvent *solvent,
                                const rism_solute *solutes, const rism_opts *opts, int device, int nworkers,
                                int *status, int *steps, double *rms, rism_xmu *xmu, double *seconds) {
  if (n < 0 || (n > 0 && (!grids || !solvent || !solutes || !opts || !status))) return RISM_ERR_ARG;
  if (nworkers < 1) nworkers = 1;
  if (nworkers > n) nworkers = n > 0 ? n : 1;
  const auto t0 = std::chrono::steady_clock::now();
  const char *mode = std::getenv("RISM_BATCH_LAUNCH");
  if (!(mode && mode[0] == '0') && opts->chargeup == 0 && opts->ksave == 0) {
    // grids that a solve cannot use fail in their own class only
    for (int i = 0; i < n; ++i) status[i] = RISM_ERR_STATE;
    solve_by_class(n, grids, solvent, solutes, opts, device, nworkers, status, steps, rms, xmu);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return RISM_OK;
  }
  std::atomic<int> next(0);
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
