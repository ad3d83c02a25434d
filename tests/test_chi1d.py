"""The chi_vv(k) generator: the library's 1D XRISM solve (rism_chi_solve,
csrc/chi1d.cpp) and its numpy oracle (oracle/rism1d.py).

Pins:
  * the reference's own tables (example/h2o_vv.xvk, example/naclaq_vv.xvk,
    example_rmdft/h2o_vv.xvk) are reproduced from their own header parameters,
    table and compressibility, by the oracle and by the library;
  * the oracle converges at first order in dr to the Percus-Yevick
    hard-sphere closed form (Wertheim-Thiele c(r) and S(0));
  * the library follows the oracle step for step on small grids (KH, HNC,
    a four-site electrolyte);
  * exact limits: the last table row is the bare omega, the XRISM dielectric
    constant is the ideal-dipole 1 + 3y;
  * DRISM: the table's longitudinal dielectric constant at k -> 0 is the
    requested one (two targets), DRISM at eps = 1 + 3y is XRISM, the library
    equals the oracle's restatement, and two planted mistakes in the oracle
    (no density on zeta, eps in place of eps - 1) fail the dielectric pin;
  * .xvk writer -> the library's reader round trip, and error paths.
No GPU is needed (the generator runs on the host).
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import example

import rism3d
from oracle import rism1d as r1

TABLES = ["h2o_vv_16384.xvk", "h2o_vv.xvk", "naclaq_vv.xvk"]


def _oracle_from_file(name, closure="KH"):
    sites, T, ntab, dr, natv, xt = r1.read_header(example(name))
    return r1.Rism1D(sites, T, ntab, dr, closure=closure), xt


@pytest.mark.parametrize("name", TABLES)
def test_oracle_reproduces_reference_table(name):
    o, xt_file = _oracle_from_file(name)
    o.solve_chargeup(4, tol=1e-10)
    got = o.reduced()
    ref = r1.read_table(example(name))
    # the files' own precision / the producer's convergence: about 1e-5
    assert np.abs(got - ref).max() < 2e-5
    assert o.compressibility() == pytest.approx(xt_file, rel=1e-5)


@pytest.mark.parametrize("name", TABLES)
def test_library_reproduces_reference_table(name):
    sites, nt, dr, T = rism3d.solvent_model(example(name))
    sv, info = rism3d.chi_solve(sites, nt, dr, T)
    assert info["status"] == rism3d.SOLVE_CONVERGED and info["rms"] <= 1e-10
    ref = rism3d.read_xvk(example(name))
    assert sv.chi.shape == ref.chi.shape
    assert np.abs(sv.chi - ref.chi).max() < 2e-5
    assert sv.xt == pytest.approx(ref.xt, rel=1e-5)
    np.testing.assert_allclose(sv.rho, ref.rho, rtol=1e-15)
    assert sv.dk == pytest.approx(ref.dk, rel=1e-15)


def test_oracle_percus_yevick_hard_spheres_first_order():
    """Single hard-sphere site, PY closure: c(r) = -l1 - 6 eta l2 x - eta l1 x^3 / 2
    inside, 0 outside; S(0) = (1-eta)^4/(1+2 eta)^2 (compressibility route)."""
    eta = 0.3
    rho = 6 * eta / np.pi
    l1 = (1 + 2 * eta) ** 2 / (1 - eta) ** 4
    l2 = -(1 + eta / 2) ** 2 / (1 - eta) ** 4
    errs, s0errs, cs = [], [], []
    for N, dr in ((2048, 0.005), (4096, 0.0025)):
        site = r1.Site("A", 1, 1, 1.0, 0.0, 0.0, [0, 0, 0], rho / (r1.NA * 1e-27))
        o = r1.Rism1D([site], 300.0, N, dr, closure="PY", hard_sphere=1.0)
        o.solve(tol=1e-13)
        _, h, c, hk = o.map(o.t)
        x = o.r
        exact = np.where(x < 1, -l1 - 6 * eta * l2 * x - 0.5 * eta * l1 * x ** 3, 0.0)
        inside = x < 0.95
        errs.append(np.abs(c[0, 0] - exact)[inside].max() / l1)
        assert np.abs(c[0, 0][x > 1.0]).max() < 1e-12  # PY: c = 0 outside the core
        s0 = 1.0 / (1.0 - rho * 4 * np.pi * dr * (x ** 2 * c[0, 0]).sum())
        s0errs.append(s0 / ((1 - eta) ** 4 / (1 + 2 * eta) ** 2) - 1.0)
        cs.append(np.interp(np.linspace(0.05, 0.9, 18), x, c[0, 0]))
    assert errs[1] < 0.01 and abs(s0errs[1]) < 0.01
    assert 1.9 < errs[0] / errs[1] < 2.1          # first order in dr
    assert 1.9 < s0errs[0] / s0errs[1] < 2.1
    xs = np.linspace(0.05, 0.9, 18)
    rich = 2 * cs[1] - cs[0]                       # Richardson: the O(dr) term removed
    exact = -l1 - 6 * eta * l2 * xs - 0.5 * eta * l1 * xs ** 3
    assert np.abs(rich - exact).max() / l1 < 2e-3


def _small_model(kind):
    if kind == "water":
        return rism3d.make_sites([
            ("O", 1, 1, 3.1507, 636.39, -0.834, (0.0, 0.0, 0.0), 55.36),
            ("H", 1, 2, 0.4, 192.46, 0.417, (0.75695, 0.58588, 0.0), 55.36),
            ("H", 1, -2, 0.4, 192.46, 0.417, (-0.75695, 0.58588, 0.0), 55.36)])
    return rism3d.make_sites([
        ("O", 1, 1, 3.1507, 636.39, -0.834, (0.0, 0.0, 0.0), 55.36),
        ("H", 1, 2, 0.4, 192.46, 0.417, (0.75695, 0.58588, 0.0), 55.36),
        ("H", 1, -2, 0.4, 192.46, 0.417, (-0.75695, 0.58588, 0.0), 55.36),
        ("Na", 2, 3, 2.35, 543.92, 1.0, (0.0, 0.0, 0.0), 0.15),
        ("Cl", 3, 4, 4.40, 418.40, -1.0, (0.0, 0.0, 0.0), 0.15)])


def _oracle_sites(sites):
    return [r1.Site(s.name.decode(), s.species, s.sym, s.sig, s.eps, s.q, list(s.xyz), s.dens) for s in sites]


@pytest.mark.parametrize("kind,closure", [("water", rism3d.KH), ("water", rism3d.HNC), ("nacl", rism3d.KH)])
def test_library_matches_oracle_step_for_step(kind, closure):
    sites = _small_model(kind)
    ntab, dr, T = 1024, 0.04, 298.15
    sv, info = rism3d.chi_solve(sites, ntab, dr, T, closure=closure, tol=1e-11, chargeup=2)
    o = r1.Rism1D(_oracle_sites(sites), T, ntab, dr, closure="KH" if closure == rism3d.KH else "HNC")
    steps, _ = o.solve_chargeup(2, tol=1e-11)
    assert info["status"] == rism3d.SOLVE_CONVERGED
    # same DIIS recurrence: equal step counts up to rounding on short runs; the
    # 500-step electrolyte run drifts apart by a few percent
    assert abs(info["steps"] - steps) <= (2 if kind == "water" else 0.05 * steps)
    ref = o.reduced()
    assert np.abs(sv.chi - ref).max() < 1e-8 * max(1.0, np.abs(ref).max())
    assert sv.xt == pytest.approx(o.compressibility(), rel=1e-8)


def test_last_row_is_omega_and_dielectric_is_ideal_dipole():
    sites = _small_model("water")
    ntab, dr, T = 2048, 0.02, 298.15
    sv, info = rism3d.chi_solve(sites, ntab, dr, T, tol=1e-11)
    kmax = np.pi / dr
    l_oh = np.hypot(0.75695, 0.58588)
    l_hh = 2 * 0.75695
    j0 = lambda x: np.sin(x) / x  # noqa: E731
    np.testing.assert_allclose(sv.chi[-1], [[1.0, 2 * j0(kmax * l_oh)], [j0(kmax * l_oh), 1 + j0(kmax * l_hh)]],
                               atol=1e-14)
    # XRISM: lim (4 pi beta rho / k^2) sum_ag q_a q_g chi_ag = 3y / (1 + 3y), i.e. eps = 1 + 3y
    beta_c = r1.HARTREE * r1.BOHR / (r1.KB * T)
    rho = sv.rho[0]
    q = sv.q
    n = np.array([1, 2])
    k = sv.dk * np.arange(1, 4)
    X = np.array([4 * np.pi * beta_c * rho * sum(n[v] * q[v] * q[w] * sv.chi[i, v, w] for v in range(2)
                                                  for w in range(2)) / k[i] ** 2 for i in range(3)])
    X0 = X[0] - (X[1] - X[0]) / 3.0  # X = X0 + b k^2
    mu = 2 * 0.417 * 0.58588
    y = 4 * np.pi / 9 * beta_c * rho * mu * mu
    assert 1.0 / (1.0 - X0) == pytest.approx(1 + 3 * y, rel=2e-3)


def test_xvk_write_read_round_trip(tmp_path):
    sites = _small_model("nacl")
    sv, _ = rism3d.chi_solve(sites, 512, 0.05, 298.15, tol=1e-9)
    p = str(tmp_path / "gen.xvk")
    rism3d.write_xvk(p, sites, sv, 0.05)
    back = rism3d.read_xvk(p)
    np.testing.assert_allclose(back.chi, sv.chi, rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(back.rho, sv.rho, rtol=1e-9)  # densities are written in M with 8 decimals
    assert back.xt == pytest.approx(sv.xt, rel=1e-8) and back.temper == pytest.approx(298.15)
    np.testing.assert_allclose(back.q, [-0.834, 0.417, 1.0, -1.0])
    # and the oracle's reader agrees with the header the writer produced
    s2, T, nt, dr, natv, xt = r1.read_header(p)
    assert (nt, natv, len(s2)) == (512, 4, 5) and dr == pytest.approx(0.05)


def test_errors():
    L = rism3d.lib()
    sites = _small_model("water")
    o = rism3d.ChiOpts()
    L.rism_chi_default_opts(C.byref(o))
    o.ntab = 1000  # not a power of two
    chi = np.zeros(4 * 1000)
    assert L.rism_chi_solve(sites, 3, C.byref(o), rism3d._dptr(chi), None, None, None) == rism3d.ERR_ARG
    bad = _small_model("water")
    bad[0].q = -0.9  # not electroneutral
    L.rism_chi_default_opts(C.byref(o))
    o.ntab = 64
    chi = np.zeros(4 * 64)
    assert L.rism_chi_solve(bad, 3, C.byref(o), rism3d._dptr(chi), None, None, None) == rism3d.ERR_ARG
    assert "electroneutral" in rism3d.last_error()
    assert L.rism_xvk_read_sites(b"/nonexistent.xvk", sites, 3, None, None, None, None) == rism3d.ERR_IO


def test_cli_writes_the_reference_table(tmp_path):
    import subprocess
    exe = os.path.join(os.path.dirname(rism3d.LIB_PATH), "rism1d-amd")
    out = str(tmp_path / "w.xvk")
    r = subprocess.run([exe, example("h2o_vv.xvk"), out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = rism3d.read_xvk(out)
    ref = rism3d.read_xvk(example("h2o_vv.xvk"))
    assert np.abs(got.chi - ref.chi).max() < 2e-5
    assert got.xt == pytest.approx(ref.xt, rel=1e-5)


# ---------------------------------------------------------------- DRISM
def _water_charge_eps(sv, T, mult=(1, 2)):
    """Longitudinal dielectric constant of a water table at k -> 0:
    X(k) = (4 pi beta rho / k^2) sum_ag q_a q_g chi_ag, extrapolated as
    X0 + b k^2 from the first two rows; eps = 1 / (1 - X0)."""
    beta_c = r1.HARTREE * r1.BOHR / (r1.KB * T)
    rho, q, n = sv.rho[0], sv.q, np.array(mult)
    k = sv.dk * np.arange(1, 3)
    X = [4 * np.pi * beta_c * rho * sum(n[v] * q[v] * q[w] * sv.chi[i, v, w] for v in range(2) for w in range(2))
         / k[i] ** 2 for i in range(2)]
    X0 = (X[0] * k[1] ** 2 - X[1] * k[0] ** 2) / (k[1] ** 2 - k[0] ** 2)
    return 1.0 / (1.0 - X0)


def _water_y(T):
    beta_c = r1.HARTREE * r1.BOHR / (r1.KB * T)
    sites = _small_model("water")
    mu = sum(s.q * np.array(s.xyz) for s in sites)
    return 4 * np.pi / 9 * beta_c * sites[0].dens * 6.02214076e23 * 1e-27 * float(mu @ mu)


@pytest.mark.parametrize("eps", [78.4, 40.0])
def test_drism_recovers_the_dielectric_constant(eps):
    sites = _small_model("water")
    ntab, dr, T = 4096, 0.04, 298.15
    sv, info = rism3d.chi_solve(sites, ntab, dr, T, tol=1e-10, dielec=eps)
    assert info["status"] == rism3d.SOLVE_CONVERGED
    assert _water_charge_eps(sv, T) == pytest.approx(eps, rel=1e-4)
    # the last row is still the bare omega (zeta vanishes at k_max), and the
    # compressibility stays within a few percent of XRISM's (zeta(0) = 0)
    x, _ = rism3d.chi_solve(sites, ntab, dr, T, tol=1e-10)
    np.testing.assert_allclose(sv.chi[-1], x.chi[-1], atol=1e-14)
    assert sv.xt == pytest.approx(x.xt, rel=0.05)
    assert abs(_water_charge_eps(x, T) - eps) > 10  # XRISM alone is far from it


def test_drism_at_the_ideal_dipole_value_is_xrism():
    sites = _small_model("water")
    ntab, dr, T = 1024, 0.04, 298.15
    x, _ = rism3d.chi_solve(sites, ntab, dr, T, tol=1e-11)
    d, _ = rism3d.chi_solve(sites, ntab, dr, T, tol=1e-11, dielec=1 + 3 * _water_y(T))
    assert np.abs(d.chi - x.chi).max() < 1e-9 * np.abs(x.chi).max()


def test_drism_library_matches_oracle():
    sites = _small_model("water")
    ntab, dr, T = 1024, 0.04, 298.15
    sv, info = rism3d.chi_solve(sites, ntab, dr, T, tol=1e-11, chargeup=2, dielec=78.4)
    o = r1.Rism1D(_oracle_sites(sites), T, ntab, dr, closure="KH", dielectric=78.4)
    steps, _ = o.solve_chargeup(2, tol=1e-11)
    assert info["status"] == rism3d.SOLVE_CONVERGED
    ref = o.reduced()
    assert np.abs(sv.chi - ref).max() < 1e-8 * max(1.0, np.abs(ref).max())
    assert sv.xt == pytest.approx(o.compressibility(), rel=1e-8)


@pytest.mark.parametrize("mutant", [None, "no_rho", "eps_over_y"])
def test_drism_planted_mistakes_fail_the_dielectric_pin(mutant, monkeypatch):
    monkeypatch.setattr(r1, "_DRISM_MUTANT", mutant)
    sites = _small_model("water")
    ntab, dr, T = 2048, 0.08, 298.15
    o = r1.Rism1D(_oracle_sites(sites), T, ntab, dr, closure="KH", dielectric=78.4)
    try:
        o.solve_chargeup(4, tol=1e-9, maxstep=1500)
    except (RuntimeError, FloatingPointError):
        assert mutant is not None  # the mistaken equations do not even converge
        return
    red = o.reduced()

    class Tab:
        chi, rho, q, dk = red, o.rho[[0, 1]], o.q[[0, 1]], o.dk
    eps = _water_charge_eps(Tab, T)
    if mutant is None:
        assert eps == pytest.approx(78.4, rel=1e-4)
    else:
        assert abs(eps - 78.4) > 0.5, (mutant, eps)


def test_drism_errors():
    L = rism3d.lib()
    o = rism3d.ChiOpts()
    L.rism_chi_default_opts(C.byref(o))
    o.ntab, o.dielec = 64, 78.4
    chi = np.zeros(16 * 64)
    ions = rism3d.make_sites([("Na", 1, 1, 2.35, 543.92, 1.0, (0.0, 0.0, 0.0), 0.15),
                              ("Cl", 2, 2, 4.40, 418.40, -1.0, (0.0, 0.0, 0.0), 0.15)])
    assert L.rism_chi_solve(ions, 2, C.byref(o), rism3d._dptr(chi), None, None, None) == rism3d.ERR_ARG
    assert "polar" in rism3d.last_error()
    o.dielec = -1.0
    assert L.rism_chi_solve(_small_model("water"), 3, C.byref(o), rism3d._dptr(chi), None, None, None) == rism3d.ERR_ARG
