"""The drop-in boundary on the GPU: the rism3d-amd command line against the
oracle's command line (same deck, same flags, the .xmu files compared), and
the .tuv restart files exchanged between the HIP solver and the oracle."""
import os
import shutil
import subprocess
import sys

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


def test_cli_rmdft_without_cvk_and_missing_input(gpu, deckdir):
    # -r needs the solvent's c(k) beside its .xvk (solvent_read.cc:91-100)
    r = subprocess.run([CLI, "-r", "x.xsv", "test1.inp"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "cvk file could not be opened" in r.stdout
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


def test_cli_field_writers_match_oracle(gpu, deckdir):
    """outlist "mdcgheq": .guv/.huv/.cuv/.euv/.qv (output_guv.cu, output_huv.cu,
    output_cuv.cu, output_euv.cu, output_qv.cu) against the oracle's fields
    at the same convergence."""
    print("cli field writers", flush=True)
    lines = open("test1.inp").read().split("\n")
    lines[0] = " mdcgheq    0\t 20220301"
    open("w1.inp", "w").write("\n".join(lines))
    r = subprocess.run([CLI, "w1.inp"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for ext, label in ((".cuv", "Cuv"), (".guv", "Guv"), (".huv", "Huv"), (".euv", "euv"), (".qv", "qv")):
        assert f"outputting {label} to file:  w1{ext}  ..." in r.stdout
    o = orc.Oracle.from_deck("w1.inp").read_xvk("h2o_vv.xvk").setup()
    o.set_iter(1e-6, 10000, 2, 0.6, 0.2)
    o.iterate()
    g_ref = o.field(orc.F_G)
    h_ref = o.field(orc.F_H)
    c_ref = h_ref - o.field(orc.F_T)   # c = h - t, t the full iterated t
    deck = rism3d.read_deck("w1.inp")
    n, box = deck.n, deck.box
    N = int(np.prod(n))
    V = deck.solvent.nsite
    # .guv: int V, int n[3], double box[3], double shift[3], g [V][N]
    raw = open("w1.guv", "rb").read()
    hdr = np.frombuffer(raw[:16], dtype=np.int32)
    assert list(hdr) == [V, *n]
    dbl = np.frombuffer(raw[16:], dtype=np.float64)
    np.testing.assert_allclose(dbl[:3], box)
    np.testing.assert_allclose(dbl[3:6], deck.shift)
    g = dbl[6:].reshape(V, N)
    assert np.abs(g - g_ref.reshape(V, N)).max() < 1e-6   # same FP64 trajectory as the oracle

    def text_field(path):
        rows = open(path).read().split("\n")
        assert [float(x) for x in rows[0].split()] == list(box)
        assert [int(x) for x in rows[1].split()] == list(n)
        na = int(rows[2])
        atoms = np.array([[float(x) for x in rows[3 + u].split()] for u in range(na)])
        assert int(rows[3 + na]) == V
        vals = np.array(" ".join(rows[4 + na:]).split(), float).reshape(N, V).T
        return atoms, vals

    atoms, h = text_field("w1.huv")
    np.testing.assert_allclose(atoms[:, 0], deck.solute.q, atol=1e-6)
    np.testing.assert_allclose(atoms[:, 3:6], deck.solute.xyz, atol=1e-5)
    hr = h_ref.reshape(V, N)
    assert np.abs(h - hr).max() < 1e-5 * max(1.0, np.abs(hr).max())   # %g: 6 significant digits
    _, c = text_field("w1.cuv")
    cr = c_ref.reshape(V, N)
    assert np.abs(c - cr).max() < 1e-5 * max(1.0, np.abs(cr).max())
    # .euv: per atom, LJ with each site then Coulomb with each site, x dv
    e = np.loadtxt("w1.euv", ndmin=2)
    eref = o.euv()
    want = np.concatenate([eref[0].T, eref[1].T], axis=1)
    np.testing.assert_allclose(e, want, rtol=1e-6, atol=1e-9)
    x = _xmu("w1.xmu")   # SE_LJ / SE_ES of the .xmu are the sums of the .euv columns
    assert e[:, :V].sum() == pytest.approx(x["SE_LJ"], rel=1e-5, abs=1e-4)
    assert e[:, V:].sum() == pytest.approx(x["SE_ES"], rel=1e-5, abs=1e-4)
    # .qv: grid coordinates and sum_v rho_v q_v g_v dv
    q = np.loadtxt("w1.qv")
    assert q.shape == (N, 4)
    dv = np.prod(np.array(box) / np.array(n))
    qref = (deck.solvent.rho[:, None] * deck.solvent.q[:, None] * g_ref.reshape(V, N)).sum(0) * dv
    assert np.abs(q[:, 3] - qref).max() < 1e-6 * max(1e-12, np.abs(qref).max())
    ix = np.arange(N)
    np.testing.assert_allclose(q[:, 0], (ix % n[0] - n[0] // 2) * box[0] / n[0], atol=1e-4)
    np.testing.assert_allclose(q[:, 2], (ix // (n[0] * n[1]) - n[2] // 2) * box[2] / n[2], atol=1e-4)


def test_cli_esp_cube_of_the_point_charges_reproduces_the_solve(gpu, deckdir):
    """-e with a cube that holds the solute's own point-charge potential
    (Hartree/e) gives the point-charge solve (-e also turns centring off,
    main.cu:37-41, so it is compared with -f)."""
    print("cli esp cube", flush=True)
    deck = rism3d.read_deck("test1.inp", centering=False)
    n, box = deck.n, deck.box
    dr = np.array(box) / np.array(n)
    bohr = 0.5291771500001
    ix = np.arange(n[0]) - n[0] // 2
    X, Y, Z = np.meshgrid(ix * dr[0], (np.arange(n[1]) - n[1] // 2) * dr[1], (np.arange(n[2]) - n[2] // 2) * dr[2],
                          indexing="ij")
    esp = np.zeros(X.shape)
    for q, (ax, ay, az) in zip(deck.solute.q, deck.solute.xyz):
        r = np.sqrt((X - ax) ** 2 + (Y - ay) ** 2 + (Z - az) ** 2)
        esp += np.where(r >= 1e-5, q * bohr / np.maximum(r, 1e-300), 0.0)
    with open("esp.cube", "w") as f:  # values with z fastest, as the reader expects
        f.write("esp of test1\npoint charges\n")
        f.write(f"{len(deck.solute.q):5d} 0.0 0.0 0.0 1\n")
        for a in range(3):
            e = [0.0, 0.0, 0.0]
            e[a] = dr[a] / bohr
            f.write(f"{n[a]:5d} {e[0]:.6f} {e[1]:.6f} {e[2]:.6f}\n")
        for q, (ax, ay, az) in zip(deck.solute.q, deck.solute.xyz):
            f.write(f"    1 {q:.6f} {ax / bohr:.6f} {ay / bohr:.6f} {az / bohr:.6f}\n")
        np.savetxt(f, esp.reshape(-1), fmt="%.15e")
    for f in ("test1.xmu", "test1.tuv"):
        if os.path.exists(f):
            os.remove(f)
    r = subprocess.run([CLI, "-e", "esp.cube", "test1.inp"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    x_esp = _xmu("test1.xmu")
    os.remove("test1.xmu")
    os.remove("test1.tuv") if os.path.exists("test1.tuv") else None
    r = subprocess.run([CLI, "-f", "x", "test1.inp"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    x_pc = _xmu("test1.xmu")
    for k in x_pc:
        assert x_esp[k] == pytest.approx(x_pc[k], rel=1e-8, abs=1e-4), k
    # a cube on another grid is an error, with the reference's message
    lines = open("esp.cube").read().split("\n")
    lines[3] = f"{n[0] + 2:5d} " + " ".join(lines[3].split()[1:])
    open("bad.cube", "w").write("\n".join(lines))
    r = subprocess.run([CLI, "-e", "bad.cube", "test1.inp"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "doesn't match" in r.stdout


def test_milc_workflow_on_the_reference_deck(gpu, deckdir):
    """example_milc/README: a charged run, a -z run, then the MILC tool, on
    the reference's mobley_3053621 deck (256^3 / 64 A, 12-atom solute given
    with -s).  Its ctip3p.xvk is not shipped, so the in-project water table
    stands in and the values are not compared with the example's .xmu; the
    checks are the workflow and that the report is the MILC formula of the
    two files written."""
    print("milc workflow", flush=True)
    from rism3d import configs, milc
    shutil.copy(configs.solvent_file("water"), "ctip3p.xvk")
    shutil.copy("milc_3drism.dat", "3drism.dat")
    for extra in ([], ["-z"]):
        r = subprocess.run([CLI, *extra, "-i", "3drism.dat", "-s", "mobley_3053621"], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert "reading solute data file : mobley_3053621" in r.stdout
    x, x0 = _xmu("mobley_3053621.xmu"), _xmu("mobley_3053621.xmu0")
    assert x0["SE_ES"] == 0.0 and x["SE_ES"] != 0.0  # -z zeroes the solute charges
    assert x["Pressure"] == x0["Pressure"]
    r = subprocess.run([sys.executable, "-m", "rism3d.milc", "mobley_3053621"], capture_output=True, text=True,
                       timeout=60, env={**os.environ, "PYTHONPATH": PKG})
    assert r.returncode == 0, r.stderr
    want = milc.milc(milc.read_xmu("mobley_3053621.xmu"), milc.read_xmu("mobley_3053621.xmu0"))
    assert f"SFE(MILC)= {want['Sub']!r} (J/mol)" in r.stdout
    assert f"SFE(MILC)= {want['HNC']!r} (J/mol)" in r.stdout


def test_rms_history_of_chunked_solve_matches_oracle_progress_lines(gpu, deckdir):
    """The residual history of a solve replayed in chunks of 16 steps as HIP
    graphs (the host reads each chunk's rms from the control block's window
    of the last 16 steps) equals the oracle's progress lines step for step
    (iterate.cu:63, printed with 6 significant digits)."""
    print("rms history test1", flush=True)
    r = subprocess.run([orc.EXE, "test1.inp"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    orms = [float(ln.split()[-1]) for ln in r.stdout.splitlines() if ln.startswith(" Step = ")]
    s = rism3d.Solver.from_deck(rism3d.read_deck("test1.inp")).build_potentials()
    st, steps, rms = s.solve()
    hist = s.rms_history()
    s.close()
    assert st == rism3d.SOLVE_CONVERGED and steps > 16
    assert len(hist) == steps and hist[-1] == rms
    assert abs(len(orms) - steps) <= 1
    n = min(len(orms), steps)
    np.testing.assert_allclose(hist[:n], orms[:n], rtol=1e-5)
