// rism_host.h -- host-side helpers shared by the C ABI implementation and
// the command-line driver (no device code here).
#pragma once

#include <string>
#include <vector>

#include "rism3d_amd.h"

namespace rism {

// physical constants; the values are those of src-cuda/physical.h:5-17 so
// that results agree digit for digit
constexpr double kAvogadro = 6.02214076e23;
constexpr double kBohr = 0.5291771500001;
constexpr double kBoltzmann = 1.380649e-23;
constexpr double kHartree = 4.359744722207185e-18;
constexpr double kKcal2J = 4184.0;

void set_error(const std::string &msg);
const char *last_error();

// natural cubic spline through (x_i, y_i): per interval i the value at
// x_i + t is ((c2 t + c1) t + c0) t + y_i  (spline.cc / splint.cc form)
void natural_spline(const double *x, const double *y, int n, double *c0, double *c1, double *c2);

// interval index on the uniform table x_i = dk (i + 1), bisection-equivalent
int table_interval(double k, double dk, int ntab);

bool is_supported_length(int n);

// A new solute on an existing handle of the same grid, solvent and options
// (the batched solves reuse a worker's handle per grid shape).
int set_solute(rism_handle *h, const rism_solute *su);

// S solutes on one grid (same grid, box, solvent and options) in batched
// launches: one launch per kernel covers all of them, each gated by its own
// control block; then each solute's thermodynamics.  Per solute status,
// steps, rms, xmu as rism_batch_solve reports them.
int batch_solve_class(const rism_grid *grid, const rism_solvent *sv, const rism_solute *const *sus, int S,
                      const rism_opts *o, int device, int *status, int *steps, double *rms, rism_xmu *xmu);

}  // namespace rism
