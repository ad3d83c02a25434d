"""The headline workload, 256^3 / 128 A with the 846-atom KH solute
(rism3d.configs.config(1)), in FP64 and in mixed precision against the FP64
oracle's solution of the same deck (tests/golden/c1_oracle.npz, written by
tests/golden/make_c1_oracle.py):

  tol6  the oracle's solve to RMS 1e-6 (337 steps, DRISM-KH water);
  conv  the oracle's solve to RMS 1e-10 (the fixed point).

The FP64 path follows the oracle's Anderson trajectory: the same step
count, the free energies within 0.05 J/mol and g within 1e-6 of tol6.

The mixed path converges to RMS 1e-6 along another trajectory, so it is
held to the fixed point: Delta-mu (SFE_SC) within 0.01 kcal/mol
(41.84 J/mol), and g within twice the tolerance ball.  The ball is the
oracle's own distance between its RMS-1e-6 state and the fixed point:
0.62 J/mol in SFE_SC, and 9.4e-3 in g (max-abs over the grid), which is
above the 1e-4 a g(r) bar at RMS 1e-6 would ask for; no solver that stops
at RMS 1e-6 meets that bar here, the oracle included.  g is compared on the
committed sub-lattice (every second point of the central 96^3, +-24 A
around the solute), relative to max(1, |g|) for the FP64 path (the
fixture's float32 values carry 6e-8 of relative rounding).
"""
import json
import os

import numpy as np
import pytest

import rism3d
from rism3d import configs
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

KCAL = 4184.0
SUB = slice(80, 176, 2)


@pytest.fixture(scope="module")
def ref():
    z = np.load(os.path.join(GOLDEN, "c1_oracle.npz"))
    return json.loads(str(z["meta"])), z["g_tol6"].astype(np.float64), z["g_conv"].astype(np.float64)


def _solve(prec):
    w = configs.config(1)
    sv = configs.solvent("water")
    s = rism3d.Solver(w.n, w.box, sv, w.solute, rism3d.Options(precision=prec, nhist=w.nhist))
    s.build_potentials()
    st, steps, rms = s.solve()
    x = s.thermo()
    g = s.field(rism3d.FIELD_G).reshape(sv.nsite, *w.n[::-1])
    s.close()
    return st, steps, rms, x, g


def test_c1_fp64_follows_the_oracle(gpu, ref):
    meta, g6, _ = ref
    print("c1 fp64", flush=True)
    st, steps, rms, x, g = _solve(rism3d.FP64)
    o = meta["tol6"]
    # relative to max(1, |g|): the fixture holds g in float32 (6e-8 relative)
    err_g = float((np.abs(g[:, SUB, SUB, SUB] - g6) / np.maximum(1.0, np.abs(g6))).max())
    print(f"c1 fp64: {steps} steps (oracle {o['steps']}), SFE_SC - oracle {x['SFE_SC'] - o['SFE_SC']:.3e} J/mol, "
          f"g max-abs vs oracle {err_g:.3e}", flush=True)
    assert st == rism3d.SOLVE_CONVERGED and rms <= 1e-6
    assert abs(steps - o["steps"]) <= 1
    for k in ("SFE_SC", "SFE_SC_HNC", "SFE_GF"):
        assert x[k] == pytest.approx(o[k], abs=0.05), k
    for k in ("SE_LJ", "SE_ES"):
        assert x[k] == pytest.approx(o[k], abs=0.05), k
    assert x["PMV"] == pytest.approx(o["PMV"], abs=1e-8)
    assert x["Pressure"] == pytest.approx(o["Pressure"], rel=1e-12)
    assert err_g < 2e-7


def test_c1_mixed_within_bar_of_the_fixed_point(gpu, ref):
    meta, g6, gc = ref
    print("c1 mixed", flush=True)
    st, steps, rms, x, g = _solve(rism3d.MIXED)
    c, o = meta["conv"], meta["tol6"]
    ball_sfe = abs(o["SFE_SC"] - c["SFE_SC"])
    err_g = float(np.abs(g[:, SUB, SUB, SUB] - gc).max())
    err_g6 = float(np.abs(g[:, SUB, SUB, SUB] - g6).max())
    print(f"c1 mixed: {steps} steps, SFE_SC - fixed point {x['SFE_SC'] - c['SFE_SC']:.4f} J/mol "
          f"(oracle at 1e-6: {o['SFE_SC'] - c['SFE_SC']:.4f}), PMV - fixed point {x['PMV'] - c['PMV']:.3e}, "
          f"g max-abs vs fixed point {err_g:.3e} (ball {meta['ball_g']:.3e}), vs oracle at 1e-6 {err_g6:.3e}",
          flush=True)
    assert st == rism3d.SOLVE_CONVERGED and rms <= 1e-6
    assert abs(x["SFE_SC"] - c["SFE_SC"]) <= 0.01 * KCAL
    assert abs(x["SFE_SC"] - o["SFE_SC"]) <= 0.01 * KCAL
    assert abs(x["SFE_GF"] - c["SFE_GF"]) <= 0.01 * KCAL
    assert abs(x["PMV"] - c["PMV"]) <= max(1e-4, 2 * abs(o["PMV"] - c["PMV"]))
    assert err_g <= 2 * meta["ball_g"]
    assert ball_sfe < 1.0  # the fixture's own ball, as documented above
