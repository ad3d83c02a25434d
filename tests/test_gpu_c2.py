"""The adjoint-gradient workload at full size: 128^3 / 64 A with the seeded
200-atom KH solute (rism3d.configs.config(2)), water from the library's 1D
DRISM-KH table.

FP64: rism_adjoint_grad (the transposed fixed-point solve) against central
finite differences of complete FP64 re-solves to RMS 1e-11 on 20 seeded
random atoms, with the prescribed steps: 1e-3 A along one random axis,
1e-4 e, 1e-4 of sigma and of epsilon.  Bar: 1e-4 relative, with a floor of
1e-5 of the largest derivative of the same kind (derivatives that nearly
vanish).  The reference's cal_grad force formula (rism_grad, LJ + ES
columns; cal_grad.cu:8-48) is the stationary-functional approximation of
the same derivative and is held to 5% of the largest component.

FP32: the mixed path (solve to RMS 1e-8, adjoint iterations with FP32
transforms to 1e-6 of rms(b)) against the FP64 adjoint; its error is
printed and held to 1e-3 of the largest component of each kind.
"""
import time

import numpy as np
import pytest

import rism3d
from rism3d import configs

pytestmark = pytest.mark.gpu

NCHECK = 20


def _solve(w, sv, sig, eps, q, xyz, prec=rism3d.FP64, tol=1e-11):
    s = rism3d.Solver(w.n, w.box, sv, rism3d.Solute(sig, eps, q, xyz),
                      rism3d.Options(precision=prec, tol=tol, maxstep=20000, nhist=w.nhist))
    s.build_potentials()
    st, steps, _ = s.solve()
    assert st == rism3d.SOLVE_CONVERGED, (st, steps)
    return s, s.thermo()["SFE_SC"]


@pytest.fixture(scope="module")
def c2():
    w = configs.config(2)
    sv = configs.solvent("water")
    su = w.solute
    sig, eps, q, xyz = (np.array(a, dtype=np.float64) for a in (su.sig, su.eps, su.q, su.xyz))
    t0 = time.perf_counter()
    s, f0 = _solve(w, sv, sig, eps, q, xyz)
    g = s.adjoint_grad(1e-10)
    cg = s.grad()
    s.close()
    print(f"c2 fp64 adjoint: {g['steps']} steps, {time.perf_counter() - t0:.2f} s with the solve", flush=True)
    return w, sv, sig, eps, q, xyz, g, cg


def test_c2_adjoint_matches_central_differences(gpu, c2):
    w, sv, sig, eps, q, xyz, g, _ = c2
    rng = np.random.default_rng(20220301 + 2)
    atoms = rng.choice(len(q), NCHECK, replace=False)
    axes = rng.integers(0, 3, NCHECK)
    scale = {"x": np.abs(g["xyz"]).max(), "q": np.abs(g["q"]).max(), "sig": np.abs(g["sig"]).max(),
             "eps": np.abs(g["eps"]).max()}
    worst = {k: 0.0 for k in scale}
    for i, ax in zip(atoms, axes):
        for name in ("x", "q", "sig", "eps"):
            vals = []
            for sgn in (1, -1):
                X, Q, S, E = xyz.copy(), q.copy(), sig.copy(), eps.copy()
                if name == "x":
                    h = 1e-3
                    X[i, ax] += sgn * h
                elif name == "q":
                    h = 1e-4
                    Q[i] += sgn * h
                elif name == "sig":
                    h = 1e-4 * sig[i]
                    S[i] += sgn * h
                else:
                    h = 1e-4 * eps[i]
                    E[i] += sgn * h
                s2, f = _solve(w, sv, S, E, Q, X)
                s2.close()
                vals.append(f)
            fd = (vals[0] - vals[1]) / (2 * h)
            ad = g["xyz"][i, ax] if name == "x" else g[name][i]
            err = abs(ad - fd) / max(abs(fd), 1e-300)
            worst[name] = max(worst[name], abs(ad - fd) / scale[name])
            print(f"atom {i} {name}{'xyz'[ax] if name == 'x' else ''}: fd {fd:.8e} adjoint {ad:.8e} rel {err:.2e}",
                  flush=True)
            assert abs(ad - fd) <= 1e-4 * abs(fd) + 1e-5 * scale[name], (i, name, fd, ad)
    print("worst |adjoint - fd| / max |d| per kind", worst, flush=True)


def test_c2_adjoint_close_to_cal_grad_formula(gpu, c2):
    *_, g, cg = c2
    ref = cg[:, :3] + cg[:, 3:]
    err = np.abs(g["xyz"] - ref).max() / np.abs(ref).max()
    print(f"c2 adjoint vs cal_grad: max |diff| / max |force| = {err:.3e}", flush=True)
    assert err <= 0.05


def test_c2_fp32_adjoint_against_fp64(gpu, c2):
    w, sv, sig, eps, q, xyz, g64, _ = c2
    t0 = time.perf_counter()
    s, _ = _solve(w, sv, sig, eps, q, xyz, prec=rism3d.MIXED, tol=1e-8)
    t1 = time.perf_counter()
    g32 = s.adjoint_grad(1e-6)
    t2 = time.perf_counter()
    s.close()
    errs = {}
    for k, a in (("xyz", "xyz"), ("q", "q"), ("sig", "sig"), ("eps", "eps")):
        errs[k] = float(np.abs(g32[a] - g64[a]).max() / np.abs(g64[a]).max())
    print(f"c2 fp32 adjoint: {g32['steps']} steps, solve {t1 - t0:.2f} s, adjoint {t2 - t1:.2f} s; "
          f"max |fp32 - fp64| / max |fp64|: {errs}", flush=True)
    for k, e in errs.items():
        assert e <= 1e-3, (k, e)
