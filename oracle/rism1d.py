"""FP64 numpy oracle of the 1D site-site RISM solver that produces chi_vv(k)
(.xvk files).  TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
the library's generator (`rism_chi_solve`); the product never uses it.

The reference consumes chi_vv(k) and does not produce it: its reader
(`src-cuda/solvent_read.cc:10-86`) fixes the file layout, the k grid
k_i = i*pi/(ntab*dr) for i = 1..ntab (`:82-86`), the density unit (M, `:78-80`)
and the reduction of equivalent sites (`:71-77`).  The reference's own tables
(`example_rmdft/h2o_vv.xvk`, `example/h2o_vv.xvk`) equal the intramolecular
omega exactly in their last row and imply the ideal-dipole dielectric constant
1 + 3y, which is what an XRISM solve on a DST-I grid (r_j = j*dr, k_i = i*dk,
dk = pi/(ntab*dr)) gives; that is the scheme restated here:

  omega_ag(k) = delta_ag + [same molecule] j0(k l_ag)
  chi = omega + rho h,    h = omega c omega + omega c rho h   (XRISM)
  closures: KH (Kovalenko & Hirata 1999), HNC, PY; hard spheres for PY tests
  long-range Coulomb: c = c_s - beta q q erf(r/eta)/r; all transformed
  functions are short-ranged (t_s = h - c_s in k space)
  charge-up: the Coulomb terms are scaled by lam = 1/n .. 1 (or 0 .. 1),
  each stage starting from the previous stage's t_s.

DRISM (Perkyns & Pettitt, Chem. Phys. Lett. 190, 626 (1992)), `dielectric`
given: for the one polar species (molecular density rho_p, dipole mu), with
its sites in a frame centred on the centre of mass, z along mu,
  zeta_ag(k) = f_a(k) h_c(k) f_g(k),  f_a = j0(k x_a) j0(k y_a) j1(k z_a),
  h_c(k) = h_c0 exp(-k^2 a^2 / 4),
  omega' = omega + rho zeta,  h' = h - zeta,  h' = omega' c omega' + omega' c rho h'.
As k -> 0, chi = (1 - rho omega' c)^-1 omega' and the charge is a left
eigenvector of 1 - rho omega c_sr (omega(0) sums the charge of a molecule,
zero), so (4 pi beta rho / k^2) q.chi.q -> 3y'/(1 + 3y') with
3y' = 3y (1 + rho h_c0 / 3), y = 4 pi beta rho mu^2 / 9: the longitudinal
relation 1 - 1/eps then gives eps = 1 + 3y', so
  h_c0 = ((eps - 1) / y - 3) / rho_p,
which is zero (XRISM) at eps = 1 + 3y.  The solve reaches it from the XRISM
solution by scaling zeta up in stages.

Physical constants are those of src-cuda/physical.h:5-17.
"""
from __future__ import annotations

import numpy as np
from scipy import fft as sfft
from scipy.special import erfc

# atomic masses (g/mol) for the centre of mass of the polar species in DRISM,
# by the element leading the site name
MASS = {"H": 1.008, "He": 4.0026, "Li": 6.94, "C": 12.011, "N": 14.007, "O": 15.999, "F": 18.998,
        "Na": 22.990, "Mg": 24.305, "P": 30.974, "S": 32.06, "Cl": 35.45, "K": 39.098, "Ca": 40.078,
        "Br": 79.904, "I": 126.90}

# a planted mistake for the tests that show the DRISM checks catch one
# ("no_rho": omega' = omega + zeta; "eps_over_y": h_c0 = (eps / y - 3) / rho)
_DRISM_MUTANT = None

NA = 6.02214076e23
BOHR = 0.5291771500001
KB = 1.380649e-23
HARTREE = 4.359744722207185e-18


class Site:
    def __init__(self, name, species, sym, sig, eps, q, xyz, dens_M):
        self.name, self.species, self.sym = name, int(species), int(sym)
        self.sig, self.eps, self.q = float(sig), float(eps), float(q)
        self.xyz = np.asarray(xyz, float)
        self.dens_M = float(dens_M)


def read_header(path):
    """Sites, temperature, ntab and dr of an .xvk file (solvent_read.cc:31-48)."""
    with open(path) as f:
        lines = f.read().split("\n")
    h1 = lines[1].split()
    num, natv, ntab, dr = int(h1[1]), int(h1[2]), int(h1[3]), float(h1[4])
    h2 = lines[2].split()
    temp, xt = float(h2[2]), float(h2[3])
    sites = []
    for l in lines[4:4 + num]:
        w = l.split()[1:]
        sites.append(Site(w[0], w[1], w[2], w[3], w[4], w[5], [float(w[6]), float(w[7]), float(w[8])], w[9]))
    return sites, temp, ntab, dr, natv, xt


def read_table(path):
    """chi table [ntab][V][V] of an .xvk file (solvent_read.cc:51-58)."""
    sites, temp, ntab, dr, natv, xt = read_header(path)
    with open(path) as f:
        body = [l for l in f.read().split("\n")[4 + len(sites):] if not l.startswith("##")]
    return np.array(" ".join(body).split(), float).reshape(ntab, natv, natv)


def dst1(f):
    """sum_{j=1}^{N-1} f_j sin(pi i j / N), i = 1..N-1 (last axis)."""
    return 0.5 * sfft.dst(f, type=1, axis=-1)


def _element(name):
    """The element a site name starts with (two letters when known: Na, Cl)."""
    return name[:2] if name[:2] in MASS else name[:1]


class Rism1D:
    """XRISM/DRISM solver on the DST-I grid; `sites` as in an .xvk header."""

    def __init__(self, sites, temp, ntab, dr, closure="KH", eta=1.0,
                 hard_sphere=None, dielectric=None, adbcor=0.5):
        self.sites, self.T, self.N, self.dr = sites, float(temp), int(ntab), float(dr)
        self.closure, self.eta = closure.upper(), eta
        S = len(sites)
        self.S = S
        N = self.N
        self.r = dr * np.arange(1, N)
        self.dk = np.pi / (N * dr)
        self.k = self.dk * np.arange(1, N)
        self.rho = np.array([s.dens_M for s in sites]) * NA * 1e-27
        self.beta_lj = 1.0 / (NA * KB * self.T)
        self.beta_c = HARTREE * BOHR / (KB * self.T)
        q = np.array([s.q for s in sites])
        self.q = q
        sig = np.array([s.sig for s in sites])
        eps = np.array([s.eps for s in sites])
        # pair potentials (S, S, N-1)
        r = self.r[None, None, :]
        sg = 0.5 * (sig[:, None] + sig[None, :])[..., None]
        ep = np.sqrt(eps[:, None] * eps[None, :])[..., None]
        qq = (q[:, None] * q[None, :])[..., None]
        s6 = (sg / r) ** 6
        self.bu_lj = self.beta_lj * 4.0 * ep * (s6 * s6 - s6)
        if hard_sphere is not None:  # hard spheres of diameter sigma (PY tests)
            self.bu_lj = np.where(r < hard_sphere, np.inf, 0.0) * np.ones_like(self.bu_lj)
        self.bu_sr_coul = self.beta_c * qq * erfc(r / eta) / r
        self.bu_l_k = (4.0 * np.pi * self.beta_c * qq * np.exp(-(self.k[None, None, :] * eta) ** 2 / 4.0)
                       / self.k[None, None, :] ** 2)
        self.omega = self._omega(self.k)
        self.lam = 1.0
        self.t = np.zeros((S, S, N - 1))
        self.zeta = np.zeros((S, S, N - 1))
        self.zfull = self._zeta(dielectric, adbcor) if dielectric else None

    def _zeta(self, eps, a):
        """zeta_ag(k) of the one polar species (module docstring)."""
        q = self.q
        spc = sorted({s.species for s in self.sites})
        polar = []
        for sp in spc:
            idx = [i for i, s in enumerate(self.sites) if s.species == sp]
            mu = sum(q[i] * self.sites[i].xyz for i in idx)
            if np.linalg.norm(mu) > 1e-8:
                polar.append((idx, mu))
        if len(polar) != 1:
            raise ValueError("DRISM needs exactly one polar species")
        idx, mu = polar[0]
        xyz = np.array([self.sites[i].xyz for i in idx])
        m = np.array([MASS[_element(self.sites[i].name)] for i in idx])
        r = xyz - (m[:, None] * xyz).sum(0) / m.sum()
        mun = np.linalg.norm(mu)
        ez = mu / mun
        ex = None
        for v in r:  # x axis: the first site off the dipole axis
            t = v - (v @ ez) * ez
            if np.linalg.norm(t) > 1e-8:
                ex = t / np.linalg.norm(t)
                break
        if ex is None:
            ex = np.cross(ez, [1.0, 0.0, 0.0] if abs(ez[0]) < 0.9 else [0.0, 1.0, 0.0])
            ex /= np.linalg.norm(ex)
        ey = np.cross(ez, ex)
        rho = self.rho[idx[0]]
        self.y = 4.0 * np.pi * self.beta_c * rho * mun ** 2 / 9.0
        self.hc0 = ((eps - (0.0 if _DRISM_MUTANT == "eps_over_y" else 1.0)) / self.y - 3.0) / rho
        k = self.k
        j1 = lambda x: np.where(np.abs(x) < 1e-3, x / 3.0 - x ** 3 / 30.0,
                                (np.sin(x) / np.where(x == 0, 1, x) - np.cos(x)) / np.where(x == 0, 1, x))
        f = np.zeros((self.S, len(k)))
        for n, i in enumerate(idx):
            f[i] = np.sinc(k * (r[n] @ ex) / np.pi) * np.sinc(k * (r[n] @ ey) / np.pi) * j1(k * (r[n] @ ez))
        hc = self.hc0 * np.exp(-(k * a) ** 2 / 4.0)
        return f[:, None, :] * f[None, :, :] * hc[None, None, :]

    # ------------------------------------------------------------------
    def _omega(self, k):
        S = self.S
        w = np.zeros((S, S, len(k)))
        for a in range(S):
            for g in range(S):
                if a == g:
                    w[a, g] = 1.0
                elif self.sites[a].species == self.sites[g].species:
                    l = np.linalg.norm(self.sites[a].xyz - self.sites[g].xyz)
                    w[a, g] = np.sinc(k * l / np.pi)
        return w

    # ------------------------------------------------------------------
    def fwd(self, f):
        return 4.0 * np.pi * self.dr / self.k * dst1(self.r * f)

    def inv(self, F):
        return self.dk / (2.0 * np.pi ** 2 * self.r) * dst1(self.k * F)

    def closure_h(self, ts):
        d = -self.bu_lj - self.lam * self.bu_sr_coul + ts
        if self.closure == "KH":
            return np.where(d > 0.0, d, np.expm1(np.minimum(d, 0.0)))
        if self.closure == "HNC":
            return np.expm1(d)
        if self.closure == "PY":  # g = exp(-beta u)(1 + t); no Coulomb for PY
            return np.exp(-self.bu_lj) * (1.0 + ts) - 1.0
        raise ValueError(self.closure)

    def rism_k(self, cs_k):
        """h(k) from c_s(k): (I - omega c rho) h = omega c omega per k."""
        c = cs_k - self.lam * self.bu_l_k
        rz = 1.0 if _DRISM_MUTANT == "no_rho" else self.rho[:, None, None]
        W = np.moveaxis(self.omega + rz * self.zeta, -1, 0)
        Ck = np.moveaxis(c, -1, 0)
        rho = self.rho
        A = np.eye(self.S)[None] - W @ Ck * rho[None, None, :]
        B = W @ Ck @ W
        return np.moveaxis(np.linalg.solve(A, B), 0, -1) + self.zeta

    def map(self, ts):
        h = self.closure_h(ts)
        cs = h - ts
        cs_k = self.fwd(cs)
        h_k = self.rism_k(cs_k)
        return self.inv(h_k - cs_k), h, cs, h_k

    def solve(self, tol=1e-12, maxstep=20000, nhist=10, mix=0.3, verbose=False):
        """Anderson (DIIS form) on t_s; stops at rms residual <= tol."""
        ts = self.t
        T, Rs = [], []
        rms = np.inf
        for it in range(1, maxstep + 1):
            new = self.map(ts)[0]
            R = new - ts
            rms = np.sqrt(np.mean(R * R))
            if verbose and it % 50 == 0:
                print(it, rms)
            if not np.isfinite(rms):
                raise FloatingPointError("1D RISM diverged")
            if rms <= tol:
                self.t = ts
                self.steps, self.rms = it, rms
                return it, rms
            T.append(ts.copy())
            Rs.append(R.copy())
            if len(T) > nhist:
                T.pop(0)
                Rs.pop(0)
            n = len(Rs)
            if n == 1:
                ts = ts + mix * R
                continue
            M = np.array([[np.vdot(Rs[i], Rs[j]) for j in range(n)] for i in range(n)])
            Bm = np.zeros((n + 1, n + 1))
            Bm[:n, :n] = M / M.max()
            Bm[n, :n] = Bm[:n, n] = -1.0
            rhs = np.zeros(n + 1)
            rhs[n] = -1.0
            try:
                cf = np.linalg.solve(Bm, rhs)[:n]
            except np.linalg.LinAlgError:
                T, Rs = T[-1:], Rs[-1:]
                ts = ts + mix * R
                continue
            ts = sum(cf[i] * (T[i] + mix * Rs[i]) for i in range(n))
        raise RuntimeError(f"1D RISM did not converge: rms {rms:.3e} after {maxstep} steps")

    def solve_chargeup(self, nstage=4, from_zero=True, zstages=10, **kw):
        """Stages lam = 0 (if from_zero), 1/nstage, ..., 1, then (DRISM) zeta
        scaled by 1/zstages, ..., 1; returns total steps."""
        lams = ([0.0] if from_zero else []) + [i / nstage for i in range(1, nstage + 1)]
        total = 0
        self.zeta = 0.0 * self.zeta
        for lam in lams:
            self.lam = lam
            total += self.solve(**kw)[0]
        if self.zfull is not None:
            for i in range(1, zstages + 1):
                self.zeta = self.zfull * (i / zstages)
                total += self.solve(**kw)[0]
        return total, self.rms

    # ------------------------------------------------------------------
    def chi_full(self):
        """chi_ag(k) = omega_ag + rho_a h_ag on k_i = i dk, i = 1..N (row N: h = 0)."""
        _, h, cs, h_k = self.map(self.t)
        chi = self.omega + self.rho[:, None, None] * h_k
        kN = np.array([self.N * self.dk])
        wN = self._omega(kN)
        return np.concatenate([chi, wN], axis=-1)

    def reduced(self, chi=None):
        """xvv[v][v'] = sum_{s' in v'} chi_{s' s} with s the first site of v
        (the reduction solvent_read.cc:71-77 assumes), as [ntab][V][V]."""
        chi = self.chi_full() if chi is None else chi
        syms = [abs(s.sym) for s in self.sites]
        V = max(syms)
        first = [syms.index(v + 1) for v in range(V)]
        out = np.zeros((chi.shape[-1], V, V))
        for v in range(V):
            s = first[v]
            for vp in range(V):
                for sp in range(self.S):
                    if syms[sp] == vp + 1:
                        out[:, v, vp] += chi[sp, s]
        return out

    def compressibility(self):
        """Isothermal compressibility [1/Pa]: rho kT chi_T = 1/(1 - sum_ag rho_a
        rho_g c_ag(k=0) / rho_tot), c_ag(0) = 4 pi sum_j r_j^2 c_s(r_j) dr
        (the Coulomb parts cancel by electroneutrality).  Reproduces the xt
        of the reference's .xvk files to a few ppm."""
        _, h, cs, _ = self.map(self.t)
        c0 = 4.0 * np.pi * self.dr * (self.r ** 2 * cs).sum(-1)
        rm = {}
        for a, s in enumerate(self.sites):
            rm[s.species] = self.rho[a]
        rtot = sum(rm.values())
        acc = float(np.einsum("a,g,ag->", self.rho, self.rho, c0))
        return 1.0 / (1.0 - acc / rtot) / (rtot * 1e30 * KB * self.T)

    def dielectric(self, chi=None, i=0):
        """1/(1 - X), X = (4 pi beta rho / k_i^2) sum_ag q_a q_g chi_ag(k_i) for a
        one-species solvent: the longitudinal dielectric relation; XRISM gives
        the ideal-dipole value 1 + 3y as k -> 0."""
        chi = self.chi_full() if chi is None else chi
        Q = np.einsum("a,g,ag->", self.q, self.q, chi[:, :, i])
        X = 4.0 * np.pi * self.beta_c * self.rho[0] * Q / self.k[i] ** 2
        return 1.0 / (1.0 - X)
