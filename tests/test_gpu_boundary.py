"""The drop-in boundary on the GPU: the rism3d-amd command line against the
oracle's command line (same deck, same flags, the .xmu files compared), and
the .tuv restart files exchanged between the HIP solver and the oracle."""
import os
import subprocess

import numpy as np
import pytest

import rism3d
from conftest import PKG
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CLI = os.path.join(PKG, "_build", "rism3d-amd")


def _xmu(path):
    out = {}
    for ln in open(path):
        if "=" in ln:
            k, v = ln.split("=", 1)
            out[k.strip()] = float(v.split()[0])
    return out


@pytest.mark.parametrize("flags,ext", [([], ".xmu"), (["-c", "3"], ".xmu"), (["-z"], ".xmu0")])
def test_cli_matches_oracle_cli(gpu, deckdir, flags, ext):
    print(f"cli {flags}", flush=True)
    for f in ("test1.xmu", "test1.xmu0", "test1.tuv"):  # the example directory holds old outputs
        if os.path.exists(f):
            os.remove(f)
    r = subprocess.run([CLI, *flags, "test1.inp"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "reading input data file" in r.stdout
    gpu_x = _xmu("test1" + ext)
    os.rename("test1" + ext, "gpu" + ext)
    r2 = subprocess.run([orc.EXE, "-q", *flags, "test1.inp"], capture_output=True, text=True, timeout=600)
    assert r2.returncode == 0, r2.stdout[-2000:]
    ref_x = _xmu("test1" + ext)
    assert set(gpu_x) == set(ref_x), (sorted(gpu_x), sorted(ref_x))
    bad = {}
    for k in gpu_x:
        tol = 1e-5 if k == "PMV" else (1.0 if abs(ref_x[k]) > 1e6 else 0.05)
        if abs(gpu_x[k] - ref_x[k]) > tol:
            bad[k] = (gpu_x[k], ref_x[k])
    print("cli diff", bad, flush=True)
    assert not bad, bad


def test_cli_rejects_rmdft_and_missing_input(gpu, deckdir):
    r = subprocess.run([CLI, "-r", "x", "test1.inp"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "not supported" in r.stdout
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "No input file" in r.stdout


def test_tuv_restart_both_ways(gpu, deckdir):
    """A .tuv written by the HIP solver restarts the oracle at its fixed
    point, and one written by the oracle restarts the HIP solver; both then
    converge in a few steps to the same Delta-mu."""
    deck = rism3d.read_deck("test1.inp")
    s = rism3d.Solver.from_deck(deck).build_potentials()
    st, steps, _ = s.solve()
    assert st == rism3d.SOLVE_CONVERGED
    x0 = s.thermo()
    s.save_tuv("gpu.tuv")
    s.close()
    o = orc.Oracle.from_deck("test1.inp").read_xvk("h2o_vv.xvk").setup()
    ost, osteps, _ = o.iterate(tuv="gpu.tuv")
    assert ost == 0 and osteps < steps / 2
    ox = o.thermo()
    assert abs(ox["SFE_SC"] - x0["SFE_SC"]) < 1.0
    o2 = orc.Oracle.from_deck("test1.inp").read_xvk("h2o_vv.xvk").setup()
    o2.iterate()
    t_ref = o2.field(orc.F_T)
    o2.write_tuv("orc.tuv")
    s2 = rism3d.Solver.from_deck(deck).build_potentials()
    s2.load_tuv("orc.tuv")
    t_loaded = s2.field(rism3d.FIELD_T).reshape(-1)
    # 6 significant digits in the file
    assert np.abs(t_loaded - t_ref).max() <= 1e-5 * np.abs(t_ref).max()
    st2, steps2, _ = s2.solve("orc.tuv")
    assert st2 == rism3d.SOLVE_CONVERGED and steps2 < steps / 2
    assert abs(s2.thermo()["SFE_SC"] - x0["SFE_SC"]) < 1.0
    s2.close()
