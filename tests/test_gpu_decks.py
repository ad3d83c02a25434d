"""The reference's own example decks through the rism3d-amd command line on
the GPU, and the outputs no other test reads back.

test3.inp (754 atoms, 64 A / 128^3) and test4.inp (protein G, 846 atoms,
128 A / 256^3: the deck of every published timing, README.md:9-23) run in
FP64 and in mixed precision (-p 1) and are compared with
  - the FP64 oracle on the same deck (tests/golden/reference_decks_oracle.json,
    written by tests/golden/make_deck_oracle.py), at RMS 1e-6 and at the
    fixed point (RMS 1e-9): FP64 follows its Anderson trajectory (the same
    step count) to the same fixed point (every .xmu quantity within
    0.05 J/mol, or 1% of the tolerance balls named next).  Rounding
    differences grow over some 300 steps, so the two RMS-1e-6 states differ
    by up to the sum of their distances to the fixed point (the tolerance
    balls, printed; test4: 0.5 and 1.0 J/mol in SFE_SC, 0.16 and 10 J/mol
    in SE_LJ);
  - the reference's own test3.xmu / test4.xmu.  Those came from an older
    build (they lack the SE lines the current output_xmu.cu:75-97 writes).
    SFE_SC is the small difference of two per-site sums of ~7e9 J/mol
    (test4: +7.19e9 and -7.19e9 give 1.1e5), and the oracle itself sits
    8e-7 and 1e-5 (relative) from test3.xmu's two per-site values, so these
    files pin SFE_SC and SFE_GF only to 2e-5 of the per-site magnitude, and
    PMV to 2e-4 L/mol.
The mixed path converges to RMS 1e-6 along another trajectory; its bar is
0.01 kcal/mol or twice the tolerance ball, whichever is larger, around the
FP64 fixed point (the FP64 path continued to RMS 1e-9), with the ball
measured in the test as |SFE(RMS 1e-6) - SFE(fixed point)| of the FP64 path.

The .gra file (output_grad.cu:16-25) is read back against the oracle's
cal_grad sums, and the periodic .tuv writes of ksave (iterate.cu:64-71)
against the oracle's t at the same step.
"""
import json
import os
import subprocess
import time

import numpy as np
import pytest

import rism3d
from conftest import GOLDEN, PKG, example
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

CLI = os.path.join(PKG, "_build", "rism3d-amd")
KCAL = 4184.0


def _xmu(path):
    out = {}
    for ln in open(path):
        if "=" in ln:
            k, v = ln.split("=", 1)
            out[k.strip()] = float(v.split()[0])
    return out


def _oracle(deck):
    with open(os.path.join(GOLDEN, "reference_decks_oracle.json")) as f:
        return json.load(f)[deck]


def _run_cli(args, timeout=600):
    t0 = time.perf_counter()
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)
    dt = time.perf_counter() - t0
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout, dt


def _steps(stdout):
    steps = [int(ln.split()[2]) for ln in stdout.splitlines() if ln.startswith(" Step = ")]
    return steps[-1] if steps else None


def _with_tol(deck, tol, out):
    lines = open(f"{deck}.inp").read().split("\n")
    lines[3] = f" {tol:g}   10000"
    open(out, "w").write("\n".join(lines))


@pytest.mark.parametrize("deck", ["test3", "test4"])
def test_reference_deck_fp64_and_mixed(gpu, deckdir, deck):
    print(f"deck {deck}", flush=True)
    orc_rec = _oracle(deck)
    ref = _xmu(example(f"{deck}.xmu"))
    per_site = sum(abs(v) for k, v in ref.items() if k.startswith("SFEC_SC("))
    res = {}
    for prec in (0, 1):
        for f in (f"{deck}.xmu", f"{deck}.tuv"):
            if os.path.exists(f):
                os.remove(f)
        out, dt = _run_cli(["-p", str(prec), f"{deck}.inp"])
        assert "reached limit" not in out and "diverged" not in out
        res[prec] = (_xmu(f"{deck}.xmu"), _steps(out), dt)
        print(f"{deck} -p {prec}: {dt:.2f} s whole program, {res[prec][1]} steps, "
              f"SFE_SC {res[prec][0]['SFE_SC']:.5f}, PMV {res[prec][0]['PMV']:.5f}", flush=True)
    x64, st64, _ = res[0]
    xmx, _, _ = res[1]
    # the fixed point: the FP64 path continued to RMS 1e-9
    _with_tol(deck, 1e-9, "fp.inp")
    _run_cli(["fp.inp"])
    xf = _xmu("fp.xmu")
    keys = ("SFE_SC", "SFE_GF", "SE_LJ", "SE_ES", "PMV", "Correction_Term")
    ox, oxf = orc_rec["xmu"], orc_rec["fixed_point_xmu"]
    ball = {k: abs(x64[k] - xf[k]) for k in keys}        # this path's RMS-1e-6 state to the fixed point
    oball = {k: abs(ox[k] - oxf[k]) for k in keys}       # the oracle's
    print(f"{deck} FP64 - oracle: " + ", ".join(f"{k} {x64[k] - ox[k]:.4g}" for k in keys), flush=True)
    print(f"{deck} fixed points, FP64 - oracle: " + ", ".join(f"{k} {xf[k] - oxf[k]:.4g}" for k in keys), flush=True)
    print(f"{deck} balls (RMS 1e-6 vs fixed point), FP64 / oracle: "
          + ", ".join(f"{k} {ball[k]:.4g} / {oball[k]:.4g}" for k in keys), flush=True)
    print(f"{deck} mixed - fixed point: " + ", ".join(f"{k} {xmx[k] - xf[k]:.4g}" for k in keys), flush=True)
    # FP64: the oracle's trajectory (the step count) and the oracle's fixed
    # point (the same equations); the RMS-1e-6 states, reached along
    # trajectories whose rounding differs, within the two balls.  test4's
    # trajectory turns on last-bit rounding: the same z/OZ and y kernels
    # compiled with and without FMA contraction take 299, 308 and 325 steps
    # (one source, two compilations with contraction, one without; all at
    # the same fixed point), so its step count is held to 10%; test3's
    # trajectory is stable to a step.
    slack = 1 if deck == "test3" else max(1, round(0.1 * orc_rec["steps"]))
    assert abs(st64 - orc_rec["steps"]) <= slack, (st64, orc_rec["steps"])
    # at RMS 1e-9 both states sit about 1e-3 of their RMS-1e-6 balls from the
    # fixed point: 0.05 J/mol (1e-5 L/mol) or 1% of the two balls
    floor = {"SFE_SC": 0.05, "SFE_GF": 0.05, "SE_LJ": 0.05, "SE_ES": 0.05, "PMV": 1e-5, "Correction_Term": 0.05}
    for k in keys:
        assert abs(xf[k] - oxf[k]) <= max(floor[k], 0.01 * (ball[k] + oball[k])), (k, xf[k], oxf[k])
    assert x64["Pressure"] == pytest.approx(ox["Pressure"], rel=1e-12)
    # the reference's older-build .xmu
    for x in (x64, xmx):
        assert x["SFE_SC"] == pytest.approx(ref["SFE_SC"], abs=2e-5 * per_site)
        assert x["SFE_GF"] == pytest.approx(ref["SFE_GF"], abs=2e-5 * per_site)
        assert x["PMV"] == pytest.approx(ref["PMV"], abs=2e-4)
    # mixed: within the tolerance ball of the FP64 fixed point
    bar = max(0.01 * KCAL, 2 * ball["SFE_SC"])
    assert abs(xmx["SFE_SC"] - xf["SFE_SC"]) <= bar
    assert abs(xmx["SFE_GF"] - xf["SFE_GF"]) <= max(0.01 * KCAL, 2 * ball["SFE_GF"])
    assert abs(xmx["PMV"] - xf["PMV"]) <= max(1e-4, 2 * ball["PMV"])


def test_gra_file_matches_oracle_grad(gpu, deckdir):
    """outlist 'd': the .gra file, 6 columns of LJ and Coulomb forces x dv
    (output_grad.cu:16-25), read back against the oracle's cal_grad sums on
    the same FP64 trajectory."""
    print("gra file", flush=True)
    lines = open("test1.inp").read().split("\n")
    lines[0] = " md    0\t 20220301"
    open("g1.inp", "w").write("\n".join(lines))
    out, _ = _run_cli(["g1.inp"])
    assert "outputting grad to file:  g1.gra  ..." in out
    gra = np.loadtxt("g1.gra", ndmin=2)
    # the file holds the library's own rism_grad values to its %.8e format
    s = rism3d.Solver.from_deck(rism3d.read_deck("g1.inp")).build_potentials()
    s.solve()
    mine = s.grad()
    s.close()
    assert gra.shape == mine.shape
    np.testing.assert_allclose(gra, mine, rtol=1e-8, atol=0)
    # and those are the oracle's cal_grad sums on the same trajectory (the
    # per-point terms cancel around the atom, so the bar is relative to the
    # result at 2e-6, as in the parity tests' 1e-6 on larger forces)
    o = orc.Oracle.from_deck("g1.inp").read_xvk("h2o_vv.xvk").setup()
    o.iterate()
    want = o.grad()
    np.testing.assert_allclose(gra, want, rtol=2e-6, atol=1e-9 * np.abs(want).max())
    txt = open("g1.gra").read().split("\n")[0]
    assert len(txt.split()) == 6 and all("e" in f for f in txt.split())


def test_ksave_writes_tuv_every_k_steps(gpu, deckdir):
    """ksave = 10 with maxstep 25: the .tuv is written after the Anderson
    update of steps 10 and 20 (iterate.cu:64-66) and not at the step-limit
    stop, so the file holds t after step 20, which is the oracle's t after a
    20-step run."""
    print("ksave", flush=True)
    lines = open("test1.inp").read().split("\n")
    lines[0] = " m    10\t 20220301"
    lines[3] = " 1.0e-12   25"
    open("k1.inp", "w").write("\n".join(lines))
    out, _ = _run_cli(["k1.inp"])
    assert out.count("saving Tuv to file:  k1.tuv") == 2, out[-3000:]
    assert "reached limit # of relaxation steps: 25" in out
    hdr = open("k1.tuv").readline().split()
    t_gpu = np.loadtxt("k1.tuv", skiprows=1)
    o = orc.Oracle.from_deck("k1.inp").read_xvk("h2o_vv.xvk").setup()
    o.set_iter(1e-12, 20, 2, 0.6, 0.2)
    st, steps, _ = o.iterate()
    assert steps == 20
    o.write_tuv("orc.tuv")
    assert open("orc.tuv").readline().split() == hdr
    t_orc = np.loadtxt("orc.tuv", skiprows=1)
    # both written with 6 significant digits
    np.testing.assert_allclose(t_gpu, t_orc, rtol=2e-5, atol=1e-5 * np.abs(t_orc).max())
    # and the restart from it continues at that state
    s = rism3d.Solver.from_deck(rism3d.read_deck("k1.inp"), maxstep=1, tol=1e-12).build_potentials()
    s.load_tuv("k1.tuv")
    t_loaded = s.field(rism3d.FIELD_T).reshape(-1)
    s.close()
    assert np.abs(t_loaded - o.field(orc.F_T)).max() <= 1e-5 * np.abs(t_loaded).max()


@pytest.mark.parametrize("maxstep", [10, 23])
def test_step_limit_stop_applies_the_pending_update_once(gpu, deckdir, maxstep):
    """A step-limit stop inside a launch chunk (test1 runs chunks of 16
    steps): the stop leaves the Anderson update of its last step pending
    (iterate.cu:61, 76-79), to be applied exactly once, whatever else of the
    chunk was queued behind it.  t after the stop equals the oracle's t after
    a run of the same number of steps."""
    print(f"step limit {maxstep}", flush=True)
    s = rism3d.Solver.from_deck(rism3d.read_deck("test1.inp"), maxstep=maxstep, tol=1e-12).build_potentials()
    st, steps, _ = s.solve()
    t = s.field(rism3d.FIELD_T).reshape(-1)
    s.close()
    assert st == rism3d.SOLVE_MAXSTEP and steps == maxstep
    o = orc.Oracle.from_deck("test1.inp").read_xvk("h2o_vv.xvk").setup()
    o.set_iter(1e-12, maxstep, 2, 0.6, 0.2)
    _, osteps, _ = o.iterate()
    assert osteps == maxstep
    ot = o.field(orc.F_T)
    err = np.abs(t - ot).max() / max(1.0, np.abs(ot).max())
    print(f"max |t - t_oracle| / max|t| = {err:.3e}", flush=True)
    assert err < 1e-9
