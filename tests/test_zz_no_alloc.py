"""rism_grad_noalloc, the cal_grad entry point that allocates no device
memory (its partial sums sit in the library's static per-GPU scratch),
against rism_grad on the same handles: two calls in a row, the second on a
smaller system, must give rism_grad's forces exactly (the same kernels and
summation order); a system whose single-atom batch cannot fit the scratch
is rejected with RISM_ERR_ARG and a message naming the limit."""
import numpy as np
import pytest

import rism3d
from rism3d import configs

pytestmark = pytest.mark.gpu


def _solved(w, sv, **opts):
    s = rism3d.Solver(w.n, w.box, sv, w.solute, rism3d.Options(nhist=w.nhist, **opts))
    s.build_potentials()
    st, _, _ = s.solve()
    return s, st


def test_noalloc_grad_equals_rism_grad(gpu):
    sv = configs.solvent("water")
    for idx in (2, 0):  # 128^3 / 200 atoms, then 64^3 / 16 atoms
        w = configs.config(idx)
        print(f"noalloc grad config {idx}: {w.n} {len(w.solute.q)} atoms", flush=True)
        s, st = _solved(w, sv)
        assert st == rism3d.SOLVE_CONVERGED
        a = s.grad_noalloc()
        b = s.grad()
        c = s.grad_noalloc()
        s.close()
        assert np.array_equal(a, b), np.abs(a - b).max()
        assert np.array_equal(c, b)
        assert np.abs(b).max() > 0


def test_noalloc_grad_rejects_what_does_not_fit(gpu):
    # natom + N / 4096 above RISM_NOALLOC_SCRATCH_BYTES / 48 (about 700 000)
    n, box = (8, 8, 8), (4.0, 4.0, 4.0)
    na = 720000
    rng = np.random.default_rng(5)
    su = rism3d.Solute(np.full(na, 1.0), np.full(na, 10.0), np.zeros(na), rng.uniform(-1.5, 1.5, (na, 3)))
    sv = configs.solvent("water")
    print(f"noalloc grad rejection: {na} atoms on {n}", flush=True)
    s = rism3d.Solver(n, box, sv, su, rism3d.Options(maxstep=1))
    s.build_potentials()
    s.solve()
    with pytest.raises(RuntimeError, match="rism_grad_noalloc.*the limit is"):
        s.grad_noalloc()
    s.close()
