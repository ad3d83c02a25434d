"""Seeded synthetic workloads (solutes, grids) for the benchmark configurations.

Every generator uses numpy.random.default_rng(20220301 + index), so the same
index always yields the same molecule.  LJ epsilon is drawn in kcal/mol and
converted to J/mol (x 4184, src-cuda/physical.h:17); coordinates are centred
with the reference's integer-rounded shift (src-cuda/solute.cu:13-43).

  c0  64^3 / 32 A box, 16 atoms: random walk of 1.5 A steps kept within 6 A
      of the centre; sigma U[3.0,3.8] A, eps U[0.05,0.2] kcal/mol,
      q ~ N(0, 0.4) shifted to sum 0; 5 Anderson histories.
  c1  256^3 / 128 A box, 846 atoms: compact self-avoiding chain of 3.8 A
      steps confined to a sphere (radius of gyration ~14 A), same parameter
      distributions, net charge -4.
  c2  128^3 / 64 A box, 200 atoms (same chain model, net charge 0).
  c3  many small solutes (batch_config): molecules of 8-40 atoms (uniform
      count), random walks of 1.5 A steps within 6 A of the centre, on
      0.5 A grids sized to the solute extent plus a 12 A margin on each
      side, rounded up to a 2,3,5-smooth multiple of 4 within 48..96.
  c4  512^3 / 256 A box, 5000 atoms, net charge -12.

Solvents: chi_vv(k) comes from the library's own 1D DRISM-KH solve
(rism_chi_solve, 16384 points at dr 0.01 A, 298.15 K, dielectric constant
78.4; dielec=0 gives XRISM, the scheme of the reference's own tables):
  water  3-site rigid water: O sigma 3.1507 A, eps 0.1521 kcal/mol,
         q -0.834; H sigma 0.4 A, eps 0.046 kcal/mol, q +0.417; OH 0.9572 A,
         HOH 104.52 deg; rho 0.03334 1/A^3
  nacl   the same water plus Na+ (sigma 2.35 A, eps 0.13 kcal/mol, q +1) and
         Cl- (sigma 4.40 A, eps 0.10 kcal/mol, q -1) at 0.15 M
Generated tables are cached as .xvk files next to the built library.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import LIB_PATH, Solute, chi_solve, make_sites, read_xvk, write_xvk

SEED0 = 20220301
KCAL = 4184.0


@dataclass
class Workload:
    name: str
    n: tuple
    box: tuple
    solute: Solute
    nhist: int = 2


WATER_RHO = 0.03334  # 1/A^3
AVOGADRO_PER_A3 = 6.02214076e23 * 1e-27  # mol/L -> 1/A^3 (solvent_read.cc:78-80)


def water_rows():
    """Site rows of the 3-site water model (species 1)."""
    half = np.radians(104.52) / 2
    x, y = 0.9572 * np.sin(half), 0.9572 * np.cos(half)
    conc = WATER_RHO / AVOGADRO_PER_A3
    return [("O", 1, 1, 3.1507, 0.1521 * KCAL, -0.834, (0.0, 0.0, 0.0), conc),
            ("H", 1, 2, 0.4, 0.046 * KCAL, 0.417, (x, y, 0.0), conc),
            ("H", 1, -2, 0.4, 0.046 * KCAL, 0.417, (-x, y, 0.0), conc)]


def solvent_rows(kind: str):
    if kind == "water":
        return water_rows()
    if kind == "nacl":
        return water_rows() + [("Na", 2, 3, 2.35, 0.13 * KCAL, 1.0, (0.0, 0.0, 0.0), 0.15),
                               ("Cl", 3, 4, 4.40, 0.10 * KCAL, -1.0, (0.0, 0.0, 0.0), 0.15)]
    raise ValueError(f"unknown solvent {kind!r}")


WATER_DIELEC = 78.4  # water at 298.15 K


def solvent_file(kind: str = "water", ntab: int = 16384, dr: float = 0.01, temper: float = 298.15,
                 dielec: float = WATER_DIELEC) -> str:
    """Path of the .xvk of `kind`, computed by the library's 1D DRISM-KH solve
    to dielectric constant `dielec` (0: XRISM) on first use and cached beside
    the library."""
    d = os.path.join(os.path.dirname(LIB_PATH), "solvents")
    scheme = f"drism{dielec:g}" if dielec > 0 else "xrism"
    path = os.path.join(d, f"{kind}_{ntab}_{dr:g}_{temper:g}_{scheme}.xvk")
    if not os.path.exists(path):
        os.makedirs(d, exist_ok=True)
        sites = make_sites(solvent_rows(kind))
        sv, info = chi_solve(sites, ntab, dr, temper, dielec=dielec)
        if info["status"] != 0:
            raise RuntimeError(f"1D RISM for {kind} did not converge: {info}")
        tmp = f"{path}.{os.getpid()}.tmp"
        write_xvk(tmp, sites, sv, dr)
        os.replace(tmp, path)
    return path


def solvent(kind: str = "water", **kw):
    """The Solvent of `kind` (see solvent_file)."""
    return read_xvk(solvent_file(kind, **kw))


def centre(xyz: np.ndarray) -> np.ndarray:
    lo, hi = xyz.min(0), xyz.max(0)
    shift = np.round(-(hi - lo) / 2 - lo)
    return xyz + shift


def _params(rng, n, net):
    sig = rng.uniform(3.0, 3.8, n)
    eps = rng.uniform(0.05, 0.2, n) * KCAL
    q = rng.normal(0.0, 0.4, n)
    q = q - (q.sum() - net) / n
    return sig, eps, q


def random_walk(rng, n, step, radius):
    """Random walk of fixed step length, each step redrawn until it stays
    within `radius` of the origin."""
    xyz = np.zeros((n, 3))
    for i in range(1, n):
        while True:
            d = rng.normal(size=3)
            p = xyz[i - 1] + step * d / np.linalg.norm(d)
            if np.linalg.norm(p) <= radius:
                xyz[i] = p
                break
    return xyz


def compact_chain(rng, n, step=3.8, radius=18.0, min_dist=2.2, tries=200, max_rounds=200000):
    """Self-avoiding chain of fixed bond length confined to a sphere: each new
    bead is drawn on the sphere of radius `step` around the previous one and
    accepted if inside the confining radius and at least `min_dist` from all
    earlier beads; when no draw succeeds the chain backs up a few beads."""
    xyz = np.zeros((n, 3))
    i = 1
    fails = 0
    rounds = 0
    while i < n:
        rounds += 1
        if rounds > max_rounds:
            raise RuntimeError("compact_chain: packing too dense for these parameters")
        ok = False
        for _ in range(tries):
            d = rng.normal(size=3)
            p = xyz[i - 1] + step * d / np.linalg.norm(d)
            if np.linalg.norm(p) > radius:
                continue
            if i > 1 and np.min(np.linalg.norm(xyz[: i - 1] - p, axis=1)) < min_dist:
                continue
            xyz[i] = p
            ok = True
            break
        if ok:
            i += 1
            fails = 0
        else:
            fails += 1
            i = max(1, i - min(4 * fails, 20))
    return xyz


def radius_of_gyration(xyz):
    c = xyz - xyz.mean(0)
    return float(np.sqrt((c * c).sum(1).mean()))


def config(index: int) -> Workload:
    rng = np.random.default_rng(SEED0 + index)
    if index == 0:
        xyz = random_walk(rng, 16, 1.5, 6.0)
        sig, eps, q = _params(rng, 16, 0.0)
        return Workload("c0: 64^3 / 32 A, 16-atom random walk, KH", (64, 64, 64), (32.0, 32.0, 32.0),
                        Solute(sig, eps, q, centre(xyz)), nhist=5)
    if index == 1:
        xyz = compact_chain(rng, 846, radius=18.0)
        sig, eps, q = _params(rng, 846, -4.0)
        return Workload("c1: 256^3 / 128 A, 846-atom compact chain (net -4), KH", (256, 256, 256),
                        (128.0, 128.0, 128.0), Solute(sig, eps, q, centre(xyz)))
    if index == 2:
        xyz = compact_chain(rng, 200, radius=11.0)
        sig, eps, q = _params(rng, 200, 0.0)
        return Workload("c2: 128^3 / 64 A, 200-atom chain, KH", (128, 128, 128), (64.0, 64.0, 64.0),
                        Solute(sig, eps, q, centre(xyz)))
    if index == 4:
        xyz = compact_chain(rng, 5000, radius=33.0)
        sig, eps, q = _params(rng, 5000, -12.0)
        return Workload("c4: 512^3 / 256 A, 5000-atom chain (net -12), KH", (512, 512, 512),
                        (256.0, 256.0, 256.0), Solute(sig, eps, q, centre(xyz)), nhist=10)
    raise ValueError(f"no single-solute workload {index}")


def smooth_grid(n: int, lo: int = 48, hi: int = 96) -> int:
    """Smallest multiple of 4 >= n whose only prime factors are 2, 3, 5,
    clamped to [lo, hi]."""
    m = max(lo, n + (-n) % 4)
    while True:
        r = m
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1 or m >= hi:
            return min(m, hi)
        m += 4


def batch_config(count: int = 256, spacing: float = 0.5, margin: float = 12.0) -> list:
    """The batched-solute workload: `count` seeded small molecules, each on
    its own grid (index 3 of the seeded configurations)."""
    rng = np.random.default_rng(SEED0 + 3)
    out = []
    for i in range(count):
        na = int(rng.integers(8, 41))
        xyz = random_walk(rng, na, 1.5, 6.0)
        sig, eps, q = _params(rng, na, 0.0)
        xyz = centre(xyz)
        ext = float((xyz.max(0) - xyz.min(0)).max())
        n = smooth_grid(int(np.ceil((ext + 2 * margin) / spacing)))
        b = n * spacing
        out.append(Workload(f"c3[{i}]: {n}^3 / {b:g} A, {na} atoms, KH", (n, n, n), (b, b, b),
                            Solute(sig, eps, q, xyz)))
    return out


def write_inp(path: str, w: Workload, solvent_file: str, tol=1e-6, maxstep=10000, count=2, m=0.6,
              mp=0.2, closure="KH"):
    """Write a workload as a reference-format .inp deck (read_input.cu:21-89)."""
    with open(path, "w") as f:
        f.write(" m    0\t 20220301\n")
        f.write(f" {closure}\n {solvent_file}\n {tol:g}   {maxstep}\n {count}   {m}   {mp}\n")
        f.write("  {:.4f} {:.4f} {:.4f}\n".format(*w.box))
        f.write("  {} {} {}\n".format(*w.n))
        s = w.solute
        f.write(f" {len(s.q)}\n")
        for i in range(len(s.q)):
            f.write("{:.10f} {:.10f} {:.10f} {:.10f} {:.10f} {:.10f}\n".format(
                s.sig[i], s.eps[i], s.q[i], *s.xyz[i]))
