"""The large electrolyte workload at full size: 512^3 / 256 A with the
seeded 5000-atom solute of net charge -12 (rism3d.configs.config(4)),
water plus 0.15 M NaCl (5 sites, V = 4 reduced; chi from the library's 1D
DRISM-KH solve), KH, 10 Anderson histories.

Size-independent properties (an oracle run at this size takes hours on the
host): both paths converge to RMS 1e-6; Delta-mu of the mixed path is within
0.01 kcal/mol (41.84 J/mol) of the FP64 path's, and PMV within 1e-3 L/mol;
the adjoint forces on all 5000 atoms (from the mixed solve) agree with the
cal_grad formula (cal_grad.cu:8-48) to 1% of the largest component; each
path's device memory is printed.
"""
import time

import numpy as np
import pytest

import rism3d
from rism3d import configs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_c4_mixed_fp64_and_adjoint(gpu):
    w = configs.config(4)
    sv = configs.solvent("nacl")
    res = {}
    for name, prec in (("mixed", rism3d.MIXED), ("fp64", rism3d.FP64)):
        rism3d.release_cache()
        free0, _ = torch.cuda.mem_get_info()
        print(f"c4 {name}: create", flush=True)
        s = rism3d.Solver(w.n, w.box, sv, w.solute, rism3d.Options(precision=prec, nhist=w.nhist))
        free1, _ = torch.cuda.mem_get_info()
        t0 = time.perf_counter()
        s.build_potentials()
        st, steps, rms = s.solve()
        x = s.thermo()
        dt = time.perf_counter() - t0
        print(f"c4 {name}: {st} after {steps} steps (rms {rms:.2e}), {dt:.2f} s, "
              f"{(free0 - free1) / 1e9:.1f} GB, SFE_SC {x['SFE_SC']:.4f}, PMV {x['PMV']:.5f}", flush=True)
        assert st == rism3d.SOLVE_CONVERGED and rms <= 1e-6
        if name == "mixed":
            t1 = time.perf_counter()
            g = s.adjoint_grad(1e-6)
            cg = s.grad()
            ref = cg[:, :3] + cg[:, 3:]
            err = float(np.abs(g["xyz"] - ref).max() / np.abs(ref).max())
            print(f"c4 adjoint on {len(w.solute.q)} atoms: {g['steps']} steps, {time.perf_counter() - t1:.2f} s, "
                  f"vs cal_grad max |diff| / max |force| {err:.3e}", flush=True)
            assert err <= 0.01
        res[name] = x
        s.close()
    rism3d.release_cache()
    d = res["mixed"]["SFE_SC"] - res["fp64"]["SFE_SC"]
    print(f"c4 mixed - fp64: SFE_SC {d:.4f} J/mol, PMV {res['mixed']['PMV'] - res['fp64']['PMV']:.2e} L/mol", flush=True)
    assert abs(d) <= 41.84
    assert abs(res["mixed"]["PMV"] - res["fp64"]["PMV"]) <= 1e-3
