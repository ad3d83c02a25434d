"""Pins for the FP64 host oracle (CPU only).

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a sign, a transposed operand, a wrong index) fails it: closed forms,
library routines (numpy.fft, scipy's natural spline), brute-force sums on
tiny grids, the least-squares characterisation of the Anderson step, and
the reference's own example outputs.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import example
from oracle import oracle as orc

NA, BOHR, KB, HARTREE = 6.02214076e23, 0.5291771500001, 1.380649e-23, 4.359744722207185e-18


# ---------------------------------------------------------------- FFT ----
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 15, 16, 20, 27, 30, 32, 45, 48, 49, 60,
                               64, 77, 96, 128, 250, 256])
def test_fft1d_matches_numpy(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(orc.fft1d(a, -1), np.fft.fft(a), rtol=0, atol=1e-12 * max(1, n))
    assert np.allclose(orc.fft1d(a, +1), np.fft.ifft(a) * n, rtol=0, atol=1e-12 * max(1, n))


def _direct_transform(f, n, box, direction):
    """F(k) = dv sum_r f(r) exp(-i k.r) (direction -1) or
    f(r) = (1/vol) sum_k F(k) exp(+i k.r) (direction +1) by brute force,
    with r = (i - n/2) dr and k = 2 pi (j - n/2 + 1/2) / L."""
    axes_r = [(np.arange(n[d]) - n[d] // 2) * (box[d] / n[d]) for d in range(3)]
    axes_k = [2 * np.pi * (np.arange(n[d]) - n[d] / 2 + 0.5) / box[d] for d in range(3)]
    out = f.astype(np.complex128)
    for d in range(3):  # separable; array axis order is (z, y, x)
        ax = 2 - d
        phase = np.exp(direction * 1j * np.outer(axes_k[d], axes_r[d]))
        if direction > 0:
            phase = phase.T
        out = np.moveaxis(np.tensordot(phase, out, axes=([1], [ax])), 0, ax)
    vol = box[0] * box[1] * box[2]
    return out * (vol / np.prod(n) if direction < 0 else 1.0 / vol)


@pytest.mark.parametrize("n", [(4, 6, 8), (6, 6, 6), (8, 10, 12), (12, 8, 4), (10, 14, 6)])
def test_fft3d_half_shift_convention_brute_force(n):
    """Pins the reference convention (fft3d.cu:22-94) independently of the
    oracle's twiddle tables, including n % 4 == 2 axes."""
    rng = np.random.default_rng(sum(n))
    box = (n[0] * 0.37, n[1] * 0.41, n[2] * 0.53)
    f = rng.standard_normal(n[::-1]) + 1j * rng.standard_normal(n[::-1])
    for direction in (-1, 1):
        ref = _direct_transform(f, n, box, direction)
        got = orc.fft3d(f, n, box, direction)
        assert np.abs(got - ref).max() < 1e-11 * np.abs(ref).max()


def test_fft3d_gaussian_closed_form():
    """exp(-a r^2) -> (pi/a)^(3/2) exp(-k^2 / 4a) on the half-shifted k grid.
    The grid sum is the continuous transform plus its images at k + 2 pi m / dr
    (Poisson summation; x_j = (j - n/2) dr makes every image phase one), so
    the reference includes the m = -1, 0, 1 images on each axis."""
    n, box, a = (48, 48, 48), (24.0, 24.0, 24.0), 0.6
    r = [(np.arange(n[d]) - n[d] // 2) * box[d] / n[d] for d in range(3)]
    z, y, x = np.meshgrid(r[2], r[1], r[0], indexing="ij")
    f = np.exp(-a * (x * x + y * y + z * z))
    F = orc.fft3d(f, n, box, -1)
    g1 = []
    for d in range(3):
        k = 2 * np.pi * (np.arange(n[d]) - n[d] / 2 + 0.5) / box[d]
        kp = 2 * np.pi * n[d] / box[d]
        g1.append(sum(np.exp(-(k + m * kp) ** 2 / (4 * a)) for m in (-1, 0, 1)))
    ref = (np.pi / a) ** 1.5 * np.einsum("k,j,i->kji", g1[2], g1[1], g1[0])
    assert np.abs(F - ref).max() < 1e-12
    # without the images the edge of the grid is off by the aliasing term
    assert np.abs(F - (np.pi / a) ** 1.5 * np.einsum("k,j,i->kji", *[
        np.exp(-(2 * np.pi * (np.arange(n[d]) - n[d] / 2 + 0.5) / box[d]) ** 2 / (4 * a)) for d in (2, 1, 0)
    ])).max() > 1e-8
    back = orc.fft3d(F, n, box, +1)
    assert np.abs(back - f).max() < 1e-12


# ------------------------------------------------------------- spline ----
def _spline(x, y):
    c0, c1, c2 = (np.zeros(len(x)) for _ in range(3))
    x = np.ascontiguousarray(x, float)
    y = np.ascontiguousarray(y, float)
    orc.lib().orc_spline_coef(orc._d(x), orc._d(y), len(x), orc._d(c0), orc._d(c1), orc._d(c2))
    return lambda v: orc.lib().orc_splint(orc._d(x), orc._d(y), orc._d(c0), orc._d(c1), orc._d(c2),
                                          len(x), float(v))


def test_spline_matches_scipy_natural():
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(5)
    x = np.cumsum(rng.uniform(0.05, 0.3, 40))
    y = np.sin(3 * x) + 0.1 * x * x
    f = _spline(x, y)
    cs = CubicSpline(x, y, bc_type="natural")
    for v in np.linspace(x[0], x[-1], 301):
        assert f(v) == pytest.approx(cs(v), abs=1e-12)


def test_spline_exact_on_lines_and_knots():
    x = np.linspace(0.1, 5.0, 25)
    f = _spline(x, 2.5 * x - 1.0)
    for v in np.linspace(0.1, 5.0, 97):
        assert f(v) == pytest.approx(2.5 * v - 1.0, abs=1e-13)
    y = np.cos(x)
    g = _spline(x, y)
    for xi, yi in zip(x, y):
        assert g(xi) == pytest.approx(yi, abs=1e-14)


# --------------------------------------------------------- potentials ----
def _tiny_system(n=(12, 10, 8), box=(6.0, 5.0, 4.4)):
    rng = np.random.default_rng(9)
    k = 3
    xyz = rng.uniform(-1.5, 1.5, (k, 3))
    sig = rng.uniform(2.8, 3.6, k)
    eps = rng.uniform(300, 700, k)
    q = np.array([0.6, -0.35, -0.25])
    o = orc.Oracle.from_arrays(n, box, example("h2o_vv.xvk"), sig, eps, q, xyz).setup()
    return o, n, box, sig, eps, q, xyz


def test_potentials_brute_force():
    o, n, box, sig, eps, q, xyz = _tiny_system()
    sv = _xvk_sites(example("h2o_vv.xvk"))
    T = sv["T"]
    r = [(np.arange(n[d]) - n[d] // 2) * box[d] / n[d] for d in range(3)]
    Z, Y, X = np.meshgrid(r[2], r[1], r[0], indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    d = np.linalg.norm(P[:, None, :] - xyz[None, :, :], axis=2)
    ubeta = HARTREE * BOHR / (KB * T)
    de = ubeta * (q / d).sum(1)
    fr = ubeta * (q * (1 - np.exp(-d)) / d).sum(1)
    # sums of signed terms: absolute tolerance at the scale of the field
    assert np.allclose(o.field(orc.F_DE), de, rtol=1e-13, atol=1e-14 * np.abs(de).max())
    assert np.allclose(o.field(orc.F_FR), fr, rtol=1e-13, atol=1e-14 * np.abs(fr).max())
    du = []
    for v in range(2):
        s = 0.5 * (sig + sv["sig"][v])
        e = np.sqrt(eps * sv["eps"][v])
        dd = np.maximum(d, 0.01)
        du.append((4 * e * ((s / dd) ** 12 - (s / dd) ** 6)).sum(1) / (NA * KB * T))
    assert np.allclose(o.field(orc.F_DU), np.concatenate(du), rtol=1e-12, atol=0)
    k = [2 * np.pi * (np.arange(n[dd]) - n[dd] / 2 + 0.5) / box[dd] for dd in range(3)]
    KZ, KY, KX = np.meshgrid(k[2], k[1], k[0], indexing="ij")
    K = np.stack([KX.ravel(), KY.ravel(), KZ.ravel()], 1)
    k2 = (K * K).sum(1)
    fk = ubeta * 4 * np.pi / (k2 * (k2 + 1)) * (q[None, :] * np.exp(-1j * K @ xyz.T)).sum(1)
    assert np.allclose(o.field(orc.F_FK), fk, rtol=1e-12, atol=1e-14 * np.abs(fk).max())


def _xvk_sites(path):
    lines = open(path).read().splitlines()
    num, natv = int(lines[1].split()[1]), int(lines[1].split()[2])
    T = float(lines[2].split()[2])
    sig, eps = [0.0] * natv, [0.0] * natv
    for ln in lines[4:4 + num]:
        t = ln.split()
        k = abs(int(t[3])) - 1
        sig[k], eps[k] = float(t[4]), float(t[5])
    return {"T": T, "sig": sig, "eps": eps}


def test_fk_is_transform_of_long_range_kernel():
    """FT identity behind the long-range split: (1 - e^-r)/r <-> 4 pi / (k^2 (k^2+1)).
    A single unit charge at the origin: fk must equal the closed form."""
    n, box = (16, 16, 16), (8.0, 8.0, 8.0)
    o = orc.Oracle.from_arrays(n, box, example("h2o_vv.xvk"), [3.0], [500.0], [1.0],
                               [[0.0, 0.0, 0.0]]).setup()
    T = _xvk_sites(example("h2o_vv.xvk"))["T"]
    ubeta = HARTREE * BOHR / (KB * T)
    k = [2 * np.pi * (np.arange(n[d]) - n[d] / 2 + 0.5) / box[d] for d in range(3)]
    KZ, KY, KX = np.meshgrid(k[2], k[1], k[0], indexing="ij")
    k2 = (KX ** 2 + KY ** 2 + KZ ** 2).ravel()
    assert np.allclose(o.field(orc.F_FK), ubeta * 4 * np.pi / (k2 * (k2 + 1)), rtol=1e-13)
    fr = o.field(orc.F_FR).reshape(n[::-1])
    assert fr[n[2] // 2, n[1] // 2, n[0] // 2] == pytest.approx(ubeta)   # r -> 0 limit: q
    assert fr[n[2] // 2, n[1] // 2, n[0] // 2 + 2] == pytest.approx(ubeta * (1 - math.exp(-1.0)) / 1.0)


# ----------------------------------------------------------- Anderson ----
def _weighted_theta(R, hist_r, slots, w):
    """theta minimising ||R + sum_i theta_i (R_i - R)||_w (normal equations)."""
    D = np.stack([hist_r[s] - R for s in slots], 1)
    A = D.T @ (w[:, None] * D)
    b = -D.T @ (w * R)
    return np.linalg.solve(A, b)


@pytest.mark.parametrize("nhist", [2, 3, 5])
def test_anderson_step_is_weighted_least_squares(nhist):
    """After `count` Picard steps (t += R, slot filled), each Anderson step
    gives t + sum theta_i (t_i - t) + m (R + sum theta_i (R_i - R)) where
    theta minimises the 1/rho-weighted residual norm, with mp added to the
    most recently written slot; the oldest slot is then overwritten."""
    V, N, m, mp, count = 2, 50, 0.6, 0.2, 2
    rho = np.array([0.0333, 0.0667])
    w = np.repeat(1.0 / rho, N)
    o = orc.Oracle.bare(V, N, nhist, rho, count, m, mp)
    rng = np.random.default_rng(nhist)
    t = rng.standard_normal(V * N)
    o.set_t(t)
    slots_t, slots_r, order = {}, {}, []
    for step in range(8):
        R = rng.standard_normal(V * N) * 0.5 ** step
        o.set_tr(R)
        o.anderson()
        got = o.field(orc.F_T)
        w_slot = step % nhist
        if step < count:
            expect = t + R
        else:
            slots = sorted(slots_t)
            th = _weighted_theta(R, slots_r, slots, w)
            recent = order[-1]
            th = th + np.array([mp if s == recent else 0.0 for s in slots])
            tb = t + sum(th[i] * (slots_t[s] - t) for i, s in enumerate(slots))
            rb = R + sum(th[i] * (slots_r[s] - R) for i, s in enumerate(slots))
            expect = tb + m * rb
        assert np.allclose(got, expect, rtol=1e-11, atol=1e-12), step
        slots_t[w_slot], slots_r[w_slot] = t.copy(), R.copy()
        order.append(w_slot)
        assert np.array_equal(o.hist(0, w_slot), t) and np.array_equal(o.hist(1, w_slot), R)
        t = got


# ------------------------------------------------- end to end (decks) ----
def _read_xmu(path):
    out = {}
    for ln in open(path):
        if "=" in ln:
            k, v = ln.split("=", 1)
            out[k.strip()] = float(v.split()[0])
    return out


def test_pressure_matches_current_format_output(tmp_path):
    """cal_pressure.cu on example_rmdft/h2o_vv.xvk reproduces benzene.xmu's
    Pressure and, with its PMV, the Correction_Term."""
    o = orc.Oracle.from_arrays((8, 8, 8), (4.0, 4.0, 4.0), example("h2o_vv_16384.xvk"), [3.0],
                               [100.0], [0.0], [[0, 0, 0]]).setup()
    o.calculate(1.0)
    x = o.thermo()
    ref = _read_xmu(example("benzene.xmu"))
    assert x["Pressure"] == pytest.approx(ref["Pressure"], abs=1e-4)
    # Correction_Term = -P * PMV with the unrounded PMV: the PMV it implies
    # with the oracle's pressure rounds to the printed 5-decimal PMV
    implied_pmv = -ref["Correction_Term"] / x["Pressure"] * 1000.0  # L/mol
    assert round(implied_pmv, 5) == pytest.approx(ref["PMV"], abs=1e-9)


def test_test1_deck_matches_reference_xmu(deckdir):
    """example/test1 (older-format output: SFE_SC, SFE_GF, PMV comparable).
    Per-site values ~9e7 J/mol at RMS 1e-6 allow ~1e2 J/mol of drift."""
    o = orc.Oracle.from_deck("test1.inp").read_xvk("h2o_vv.xvk").setup()
    st, steps, rms = o.iterate()
    assert st == 0 and rms <= 1e-6
    x = o.thermo()
    ref = _read_xmu("test1.xmu")
    assert x["SFE_SC"] == pytest.approx(ref["SFE_SC"], abs=100.0)
    assert x["SFE_GF"] == pytest.approx(ref["SFE_GF"], abs=100.0)
    assert round(x["PMV"], 5) == pytest.approx(ref["PMV"], abs=1e-5)
    assert x["SFEC_SC"][0] == pytest.approx(ref["SFEC_SC(0)"], rel=1e-6)


def test_test2_deck_matches_reference_xmu(deckdir):
    o = orc.Oracle.from_deck("test2.inp").read_xvk("h2o_vv.xvk").setup()
    st, steps, rms = o.iterate()
    assert st == 0
    x = o.thermo()
    ref = _read_xmu("test2.xmu")
    assert x["SFE_SC"] == pytest.approx(ref["SFE_SC"], abs=1.0)
    assert x["SFE_GF"] == pytest.approx(ref["SFE_GF"], abs=1.0)
    assert round(x["PMV"], 5) == pytest.approx(ref["PMV"], abs=1e-5)


def test_oracle_cli_writes_xmu(deckdir):
    exe = orc.EXE
    if not os.path.exists(exe):
        orc.build()
    r = subprocess.run([exe, "-q", "-j", "test1.inp"], capture_output=True, text=True, check=True)
    assert '"status": 0' in r.stdout
    x = _read_xmu("test1.xmu")
    assert "SFE_SC_HNC" in x and "SE" in x and "Correction_Term" in x


def test_bad_version_is_rejected(tmp_path):
    p = tmp_path / "bad.inp"
    p.write_text(" m 0 20200101\n KH\n h2o_vv.xvk\n")
    with pytest.raises(RuntimeError, match="old version"):
        orc.Oracle.from_deck(str(p))


# ------------------------------------------------- gradient vs energy ----
def test_grad_is_derivative_of_solute_solvent_energy():
    """cal_grad.cu's per-atom sums are the position derivatives of the
    SE_LJ / SE_ES sums (cal_euv.cu) at fixed g: central differences of the
    oracle's SE with one atom moved by +-h must equal its grad rows.  A sign,
    index or factor mistake in either formula fails this."""
    o, n, box, sig, eps, q, xyz = _tiny_system(n=(16, 16, 16), box=(8.0, 8.0, 8.0))
    o.calculate(1.0)  # a g field to weigh the sums with
    g = o.grad()
    h = 1e-4
    for u in range(len(q)):
        for k in range(3):
            e = []
            for sgn in (+1, -1):
                p = xyz[u].copy()
                p[k] += sgn * h
                o.set_atom_xyz(u, p)
                x = o.thermo()
                e.append((x["SE_LJ"], x["SE_ES"]))
            o.set_atom_xyz(u, xyz[u])
            d_lj = (e[0][0] - e[1][0]) / (2 * h)
            d_es = (e[0][1] - e[1][1]) / (2 * h)
            assert d_lj == pytest.approx(g[u, k], rel=1e-5, abs=1e-6 * np.abs(g[:, :3]).max())
            assert d_es == pytest.approx(g[u, 3 + k], rel=1e-5, abs=1e-6 * np.abs(g[:, 3:]).max())


# ------------------------------------------------------ SE closed forms ----
def _box_inv_r(lo, hi):
    """Integral of dV / |r| over the box [lo, hi] (corners relative to the
    point): the Newtonian potential of a uniform rectangular prism, in the
    closed form of its corner function
    F = yz ln(x+r) + xz ln(y+r) + xy ln(z+r)
        - x^2/2 atan(yz/(xr)) - y^2/2 atan(xz/(yr)) - z^2/2 atan(xy/(zr))."""
    def F(x, y, z):
        r = math.sqrt(x * x + y * y + z * z)
        return (y * z * math.log(x + r) + x * z * math.log(y + r) + x * y * math.log(z + r)
                - 0.5 * x * x * math.atan(y * z / (x * r)) - 0.5 * y * y * math.atan(x * z / (y * r))
                - 0.5 * z * z * math.atan(x * y / (z * r)))
    s = 0.0
    for i, x in enumerate((lo[0], hi[0])):
        for j, y in enumerate((lo[1], hi[1])):
            for k, z in enumerate((lo[2], hi[2])):
                s += (-1) ** (i + j + k + 1) * F(x, y, z)
    return s


def test_box_inv_r_closed_form():
    # the potential at the centre of a unit cube, 2.3800774... (a tabulated constant)
    assert _box_inv_r((-0.5,) * 3, (0.5,) * 3) == pytest.approx(2.380077363979553, rel=1e-12)
    # additivity: two halves make the whole
    a = _box_inv_r((-1.0, -2.0, -0.5), (3.0, 1.0, 2.0))
    b = _box_inv_r((-1.0, -2.0, -0.5), (1.0, 1.0, 2.0)) + _box_inv_r((1.0, -2.0, -0.5), (3.0, 1.0, 2.0))
    assert a == pytest.approx(b, rel=1e-12)


def _xvk_reduced_sites(path):
    """(sigma, eps, q, density 1/A^3) per reduced site of an .xvk header,
    equivalent sites (negative symmetry index) merged by summing densities."""
    rows = [ln.split() for ln in open(path).read().splitlines()[4:] if ln.startswith("##") and len(ln.split()) > 9]
    sites = {}
    for r in rows:
        sym = abs(int(r[3]))
        sig, eps, q, conc = float(r[4]), float(r[5]), float(r[6]), float(r[10])
        if sym in sites:
            sites[sym][3] += conc * NA * 1e-27
        else:
            sites[sym] = [sig, eps, q, conc * NA * 1e-27]
    return [sites[k] for k in sorted(sites)]


def se_closed_form_errors():
    """Relative deviations of the oracle's SE sums (cal_euv.cu:35-76, per
    atom and site through orc_euv, totals through orc_thermo) from closed
    forms, for g = 1 on a 128^3 / 12.8 A grid around a handful of LJ atoms
    placed near the centre, off the lattice.  With g = 1 each (atom u, site
    v) term is an integral of its pair potential alone:
      LJ  rho_v * int_{r >= s/2} 4 eps_uv [(s/r)^12 - (s/r)^6] dV
          = 16 pi rho_v eps_uv s^3 (2^9/9 - 2^3/3)       (s = sigma_uv)
      ES  rho_v q_u q_v E_h a0 N_A [int_cell-box dV/r - 2 pi (s/2)^2]
    with sigma_uv = (sigma_u + sigma_v)/2 and eps_uv = sqrt(eps_u eps_v)
    (Lorentz-Berthelot, cal_LJ.cu:24-31).  The lattice sum meets the LJ
    integral to about 1% (the r^-12 wall at s/2 cuts through the lattice)
    and the ES integral to about 1e-5."""
    n, box = (128, 128, 128), (12.8, 12.8, 12.8)
    dr = box[0] / n[0]
    xvk = example("h2o_vv.xvk")
    sites = _xvk_reduced_sites(xvk)
    rng = np.random.default_rng(11)
    sig = np.array([2.0, 2.4, 2.8, 3.1, 3.4])
    eps = rng.uniform(200.0, 900.0, sig.size)
    q = rng.uniform(-1.0, 1.0, sig.size)
    xyz = rng.uniform(-0.05, 0.05, (sig.size, 3))
    o = orc.Oracle.from_arrays(n, box, xvk, sig, eps, q, xyz).setup()
    V = o.V
    o.set_g(np.ones(V * o.N))
    e = o.euv()
    x = o.thermo()
    # constants from CODATA (the reference's bohr differs from CODATA 2018 in the 7th digit)
    cc = 4.3597447222071e-18 * 0.529177210903 * 6.02214076e23
    lo = np.full(3, (-(n[0] // 2) - 0.5) * dr)
    hi = np.full(3, (n[0] - 1 - n[0] // 2 + 0.5) * dr)
    err_lj, err_es = [], []
    for v, (sv, ev, qv, rho) in enumerate(sites):
        for u in range(sig.size):
            s = 0.5 * (sig[u] + sv)
            euv = math.sqrt(eps[u] * ev)
            lj = 16 * math.pi * rho * euv * s ** 3 * (2 ** 9 / 9 - 2 ** 3 / 3)
            es = rho * q[u] * qv * cc * (_box_inv_r(lo - xyz[u], hi - xyz[u]) - 2 * math.pi * (s / 2) ** 2)
            err_lj.append(e[0, v, u] / lj - 1)
            err_es.append(e[1, v, u] / es - 1)
    tot = {"lj": x["SE_LJ"] / e[0].sum() - 1, "es": x["SE_ES"] / e[1].sum() - 1}
    return {"lj": float(np.abs(err_lj).max()), "es": float(np.abs(err_es).max()),
            "lj_mean": float(np.mean(err_lj)), "tot_lj": tot["lj"], "tot_es": tot["es"]}


SE_TOL = {"lj": 0.03, "es": 2e-4, "tot": 1e-12}


def _se_within(r):
    return r["lj"] < SE_TOL["lj"] and r["es"] < SE_TOL["es"] and \
        abs(r["tot_lj"]) < SE_TOL["tot"] and abs(r["tot_es"]) < SE_TOL["tot"]


def test_se_matches_closed_forms():
    r = se_closed_form_errors()
    assert _se_within(r), r


@pytest.mark.parametrize("mutant", [1, 2, 3, 4],
                         ids=["exclusion_at_sigma", "geometric_sigma", "rho_dropped", "cc_without_NA"])
def test_se_closed_forms_reject_planted_mistakes(mutant, tmp_path):
    """The same checks on oracle builds with one planted mistake each
    (ORC_MUTANT in rism_oracle.c) must fail."""
    here = os.path.dirname(orc.HERE)
    so = tmp_path / f"liboracle_m{mutant}.so"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-fPIC", "-std=gnu11", "-shared", f"-DORC_MUTANT={mutant}",
                    "-o", str(so), os.path.join(orc.HERE, "rism_oracle.c"), "-lm"], check=True)
    code = ("import json, sys; sys.path[:0] = [%r, %r]; import test_oracle as t; "
            "print(json.dumps(t.se_closed_form_errors()))" % (os.path.join(here, "tests"), here))
    env = dict(os.environ, RISM_ORACLE_LIB=str(so))
    out = subprocess.run([os.sys.executable, "-c", code], env=env, capture_output=True, text=True, check=True)
    r = __import__("json").loads(out.stdout.strip().splitlines()[-1])
    assert not _se_within(r), r
