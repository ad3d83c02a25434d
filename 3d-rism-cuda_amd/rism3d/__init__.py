"""rism3d -- Python view of librism3d_amd.so (MI355X / gfx950 3D-RISM).

A thin ctypes layer over the C ABI declared in include/rism3d_amd.h.  It
mirrors the reference's RISM3D life cycle (src-cuda/rism3d.h:25-27:
initialize -> iterate -> output) as :class:`Solver` (create + potentials ->
solve -> thermo / grad / fields), plus the .inp/.xvk deck reader.  There is
no fallback: if the shared library is missing, or no GPU is present when a
GPU call is made, a RuntimeError is raised.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# RISM3D_LIBRARY selects another build of the library (e.g. one compiled
# with different tuning macros); the default is the in-tree build.
LIB_PATH = os.environ.get("RISM3D_LIBRARY") or os.path.join(os.path.dirname(_HERE), "_build", "librism3d_amd.so")

# status codes (include/rism3d_amd.h)
OK = 0
SOLVE_CONVERGED, SOLVE_DIVERGED, SOLVE_MAXSTEP = 0, 1, 2
ERR_ARG, ERR_IO, ERR_FORMAT, ERR_VERSION, ERR_RANGE, ERR_HIP, ERR_NOMEM, ERR_STATE = (
    -1, -2, -3, -4, -5, -6, -7, -8)
KH, HNC = 0, 1
FP64, MIXED = 0, 1
FIELD_T, FIELD_G, FIELD_H, FIELD_C, FIELD_RES, FIELD_ULJ, FIELD_DE, FIELD_FR, FIELD_FK, FIELD_CHI = range(10)

EXPORTED = [
    "rism_last_error", "rism_version", "rism_device_count", "rism_deck_read", "rism_deck_free",
    "rism_deck_views", "rism_xvk_read", "rism_create", "rism_build_potentials", "rism_solve",
    "rism_thermo", "rism_grad", "rism_grad_noalloc", "rism_get_field", "rism_set_t", "rism_save_tuv", "rism_load_tuv",
    "rism_write_xmu", "rism_write_gra", "rism_stats", "rism_rms_history", "rism_destroy", "rism_release_cache",
    "rism_fft3d", "rism_calculate", "rism_profile", "rism_kernel_times", "rism_geometry",
    "rism_adjoint_grad", "rism_batch_solve", "rism_chi_default_opts", "rism_chi_nreduced",
    "rism_chi_solve", "rism_write_xvk", "rism_xvk_read_sites", "rism_stream_copy",
    "rism_euv", "rism_qv", "rism_write_guv", "rism_write_huv", "rism_write_euv", "rism_write_qv",
    "rism_read_cube", "rism_set_esp", "rism_rmdft", "rism_read_cvk", "rism_read_xsv", "rism_autotune",
]

KERNELS = ["closure_fx", "fft_y_fwd", "fft_z_oz", "fft_y_inv", "resid", "solve", "potentials", "fk",
           "exchem", "euv", "refresh_fp64"]


class Grid(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("box", C.c_double * 3)]


class SolventC(C.Structure):
    _fields_ = [("nsite", C.c_int), ("ntab", C.c_int), ("dk", C.c_double), ("temper", C.c_double),
                ("xt", C.c_double), ("rho", C.POINTER(C.c_double)), ("q", C.POINTER(C.c_double)),
                ("sig", C.POINTER(C.c_double)), ("eps", C.POINTER(C.c_double)),
                ("chi", C.POINTER(C.c_double))]


class SoluteC(C.Structure):
    _fields_ = [("natom", C.c_int), ("sig", C.POINTER(C.c_double)), ("eps", C.POINTER(C.c_double)),
                ("q", C.POINTER(C.c_double)), ("xyz", C.POINTER(C.c_double))]


class Opts(C.Structure):
    _fields_ = [("closure", C.c_int), ("tol", C.c_double), ("maxstep", C.c_int), ("count", C.c_int),
                ("m", C.c_double), ("mp", C.c_double), ("nhist", C.c_int), ("precision", C.c_int),
                ("chargeup", C.c_int), ("verbose", C.c_int), ("ksave", C.c_int)]


class Xmu(C.Structure):
    _fields_ = [("nsite", C.c_int), ("sfe_sc", C.c_double), ("sfe_sc_hnc", C.c_double),
                ("sfe_gf", C.c_double), ("sfec_sc", C.c_double * 16), ("sfec_sc_hnc", C.c_double * 16),
                ("sfec_gf", C.c_double * 16), ("se", C.c_double), ("se_lj", C.c_double),
                ("se_es", C.c_double), ("pmv", C.c_double), ("pressure", C.c_double),
                ("correction", C.c_double), ("has_rmdft", C.c_int), ("sfe_rmdft", C.c_double)]

    def as_dict(self):
        v = self.nsite
        d = {
            "SFE_SC": self.sfe_sc, "SFE_SC_HNC": self.sfe_sc_hnc, "SFE_GF": self.sfe_gf,
            "SFEC_SC": list(self.sfec_sc[:v]), "SFEC_SC_HNC": list(self.sfec_sc_hnc[:v]),
            "SFEC_GF": list(self.sfec_gf[:v]), "SE": self.se, "SE_LJ": self.se_lj, "SE_ES": self.se_es,
            "PMV": self.pmv, "Pressure": self.pressure, "Correction_Term": self.correction,
        }
        if self.has_rmdft:
            d["SFE_RMDFT"] = self.sfe_rmdft
        return d


class RmdftInput(C.Structure):
    """include/rism3d_amd.h rism_rmdft_input (solvent_read.cc:88-143)."""
    _fields_ = [("cvk", C.POINTER(C.c_double)), ("nhs", C.c_int), ("khs", C.POINTER(C.c_double)),
                ("chs", C.POINTER(C.c_double)), ("whs", C.POINTER(C.c_double)),
                ("pfhs", C.POINTER(C.c_double)), ("w0", C.POINTER(C.c_double))]


class ChiSite(C.Structure):
    """One .xvk site row (solvent_read.cc:42-47), include/rism3d_amd.h rism_chi_site."""
    _fields_ = [("name", C.c_char * 8), ("species", C.c_int), ("sym", C.c_int), ("sig", C.c_double),
                ("eps", C.c_double), ("q", C.c_double), ("xyz", C.c_double * 3), ("dens", C.c_double)]


class ChiOpts(C.Structure):
    _fields_ = [("ntab", C.c_int), ("dr", C.c_double), ("temper", C.c_double), ("closure", C.c_int),
                ("tol", C.c_double), ("maxstep", C.c_int), ("nhist", C.c_int), ("mix", C.c_double),
                ("chargeup", C.c_int), ("dielec", C.c_double), ("adbcor", C.c_double)]


class DeckC(C.Structure):
    _fields_ = [("outlist", C.c_char * 64), ("ksave", C.c_int), ("solvent_path", C.c_char * 4096),
                ("fname", C.c_char * 4096), ("grid", Grid), ("opts", Opts), ("natom", C.c_int),
                ("sig", C.POINTER(C.c_double)), ("eps", C.POINTER(C.c_double)),
                ("q", C.POINTER(C.c_double)), ("xyz", C.POINTER(C.c_double)), ("shift", C.c_double * 3),
                ("nsite", C.c_int), ("ntab", C.c_int), ("dk", C.c_double), ("temper", C.c_double),
                ("xt", C.c_double), ("rho", C.POINTER(C.c_double)), ("sq", C.POINTER(C.c_double)),
                ("ssig", C.POINTER(C.c_double)), ("seps", C.POINTER(C.c_double)),
                ("chi", C.POINTER(C.c_double))]


_lib = None


def lib():
    """Load librism3d_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # torch bundles its own libamdhip64.so.7 with the same SONAME as the
        # system one this library links; loading torch first makes both share
        # one HIP runtime (loading ours first leaves torch without devices)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run __graft_entry__.build() or make -C 3d-rism-cuda_amd")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp = P(C.c_double)
        L.rism_last_error.restype = C.c_char_p
        L.rism_version.restype = C.c_char_p
        L.rism_device_count.restype = C.c_int
        L.rism_deck_read.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, P(P(DeckC))]
        L.rism_deck_free.argtypes = [P(DeckC)]
        L.rism_deck_views.argtypes = [P(DeckC), P(Grid), P(SolventC), P(SoluteC), P(Opts)]
        L.rism_xvk_read.argtypes = [C.c_char_p, P(DeckC)]
        L.rism_create.argtypes = [P(Grid), P(SolventC), P(SoluteC), P(Opts), C.c_int, P(C.c_void_p)]
        L.rism_build_potentials.argtypes = [C.c_void_p]
        L.rism_solve.argtypes = [C.c_void_p, C.c_char_p, P(C.c_int), dp]
        L.rism_thermo.argtypes = [C.c_void_p, P(Xmu)]
        L.rism_grad.argtypes = [C.c_void_p, dp]
        L.rism_grad_noalloc.argtypes = [C.c_void_p, dp]
        L.rism_get_field.argtypes = [C.c_void_p, C.c_int, dp]
        L.rism_set_t.argtypes = [C.c_void_p, dp]
        L.rism_save_tuv.argtypes = [C.c_void_p, C.c_char_p]
        L.rism_load_tuv.argtypes = [C.c_void_p, C.c_char_p]
        L.rism_write_xmu.argtypes = [P(Xmu), C.c_int, C.c_char_p]
        L.rism_write_gra.argtypes = [dp, C.c_int, C.c_char_p]
        L.rism_stats.argtypes = [C.c_void_p, P(C.c_int), dp, dp, dp, dp]
        L.rism_rms_history.argtypes = [C.c_void_p, dp, C.c_int]
        L.rism_destroy.argtypes = [C.c_void_p]
        L.rism_destroy.restype = None
        L.rism_release_cache.argtypes = []
        L.rism_release_cache.restype = C.c_long
        L.rism_fft3d.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.rism_calculate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rism_adjoint_grad.argtypes = [C.c_void_p, C.c_double, dp, dp, dp, dp, C.POINTER(C.c_int)]
        ip = P(C.c_int)
        L.rism_batch_solve.argtypes = [C.c_int, P(Grid), P(SolventC), P(SoluteC), P(Opts), C.c_int, C.c_int,
                                       ip, ip, dp, P(Xmu), dp]
        L.rism_profile.argtypes = [C.c_void_p, C.c_int]
        L.rism_kernel_times.argtypes = [C.c_void_p, dp, P(C.c_long)]
        L.rism_geometry.argtypes = [C.c_void_p, P(C.c_int)]
        L.rism_stream_copy.argtypes = [C.c_int, C.c_size_t, C.c_int, dp, dp]
        ip = P(C.c_int)
        L.rism_read_cube.argtypes = [C.c_char_p, ip, dp]
        L.rism_autotune.argtypes = [C.c_void_p, C.c_int, ip]
        L.rism_rmdft.argtypes = [C.c_void_p, P(RmdftInput), dp]
        L.rism_read_cvk.argtypes = [C.c_char_p, C.c_int, C.c_int, dp]
        L.rism_read_xsv.argtypes = [C.c_char_p, C.c_int, P(C.c_int), dp, dp, dp, dp, dp]
        L.rism_set_esp.argtypes = [C.c_void_p, dp]
        L.rism_euv.argtypes = [C.c_void_p, dp]
        L.rism_qv.argtypes = [C.c_void_p, dp]
        L.rism_write_guv.argtypes = [C.c_char_p, C.c_int, ip, dp, dp, dp]
        L.rism_write_huv.argtypes = [C.c_char_p, ip, dp, C.c_int, dp, dp, dp, dp, C.c_int, dp]
        L.rism_write_euv.argtypes = [C.c_char_p, C.c_int, C.c_int, dp]
        L.rism_write_qv.argtypes = [C.c_char_p, ip, dp, dp]
        L.rism_chi_default_opts.argtypes = [P(ChiOpts)]
        L.rism_chi_default_opts.restype = None
        L.rism_chi_nreduced.argtypes = [P(ChiSite), C.c_int]
        L.rism_chi_solve.argtypes = [P(ChiSite), C.c_int, P(ChiOpts), dp, dp, P(C.c_int), dp]
        L.rism_write_xvk.argtypes = [C.c_char_p, P(ChiSite), C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, dp]
        L.rism_xvk_read_sites.argtypes = [C.c_char_p, P(ChiSite), C.c_int, P(C.c_int), P(C.c_int), dp, dp]
        _lib = L
    return _lib


def last_error() -> str:
    return lib().rism_last_error().decode()


def _check(rc: int, what: str):
    if rc < 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")
    return rc


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _arr(p, n):
    return np.ctypeslib.as_array(p, shape=(n,)).copy() if n > 0 else np.zeros(0)


@dataclass
class Solvent:
    """Reduced solvent model as read from a .xvk file (solvent_read.cc:10-86)."""
    rho: np.ndarray
    q: np.ndarray
    sig: np.ndarray
    eps: np.ndarray
    chi: np.ndarray  # [ntab][V][V]
    dk: float
    temper: float
    xt: float

    @property
    def nsite(self):
        return len(self.rho)


@dataclass
class Solute:
    sig: np.ndarray
    eps: np.ndarray
    q: np.ndarray
    xyz: np.ndarray  # [n][3]


@dataclass
class Options:
    closure: int = KH
    tol: float = 1e-6
    maxstep: int = 10000
    count: int = 2
    m: float = 0.6
    mp: float = 0.2
    nhist: int = 2
    precision: int = FP64
    chargeup: int = 0
    verbose: int = 0
    ksave: int = 0


@dataclass
class Deck:
    """A parsed .inp deck with its solvent (read_input.cu:21-89)."""
    outlist: str
    ksave: int
    fname: str
    n: tuple
    box: tuple
    opts: Options
    solute: Solute
    solvent: Solvent
    shift: tuple = field(default=(0.0, 0.0, 0.0))


def read_deck(control: str, structure: Optional[str] = None, centering: bool = True,
              zero: bool = False) -> Deck:
    L = lib()
    p = C.POINTER(DeckC)()
    _check(L.rism_deck_read(control.encode(), structure.encode() if structure else None,
                            int(centering), int(zero), C.byref(p)), "rism_deck_read")
    d = p.contents
    try:
        V, nt, n = d.nsite, d.ntab, d.natom
        o = d.opts
        opts = Options(o.closure, o.tol, o.maxstep, o.count, o.m, o.mp, 2, FP64, 0, 0, d.ksave)
        su = Solute(_arr(d.sig, n), _arr(d.eps, n), _arr(d.q, n), _arr(d.xyz, 3 * n).reshape(n, 3))
        sv = Solvent(_arr(d.rho, V), _arr(d.sq, V), _arr(d.ssig, V), _arr(d.seps, V),
                     _arr(d.chi, nt * V * V).reshape(nt, V, V), d.dk, d.temper, d.xt)
        return Deck(d.outlist.decode(), d.ksave, d.fname.decode(), tuple(d.grid.n), tuple(d.grid.box),
                    opts, su, sv, tuple(d.shift))
    finally:
        L.rism_deck_free(p)


def read_xvk(path: str) -> Solvent:
    """Read a .xvk solvent file (solvent_read.cc:10-86)."""
    L = lib()
    d = DeckC()
    try:
        _check(L.rism_xvk_read(path.encode(), C.byref(d)), "rism_xvk_read")
        V, nt = d.nsite, d.ntab
        return Solvent(_arr(d.rho, V), _arr(d.sq, V), _arr(d.ssig, V), _arr(d.seps, V),
                       _arr(d.chi, nt * V * V).reshape(nt, V, V), d.dk, d.temper, d.xt)
    finally:
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        for name in ("rho", "sq", "ssig", "seps", "chi"):
            p = getattr(d, name)
            if p:
                libc.free(C.cast(p, C.c_void_p))


@dataclass
class RmdftTables:
    """The RMDFT inputs: c_vv'(k) [ntab][V][V] and the reference fluid's
    hard-sphere tables, per reduced site (sites without data hold zeros)."""
    cvk: np.ndarray
    khs: np.ndarray
    chs: np.ndarray  # [V][nhs]
    whs: np.ndarray  # [V][nhs]
    pfhs: np.ndarray  # [V]
    w0: np.ndarray  # [V]


def read_rmdft(cvk_path: str, xsv_path: str, solvent: Solvent) -> RmdftTables:
    """.cvk and .xsv files as the reference reads them (solvent_read.cc:88-143):
    the hard-sphere data go to site 0, zeros to the other sites."""
    L = lib()
    V, nt = solvent.nsite, solvent.chi.shape[0]
    cvk = np.zeros((nt, V, V))
    _check(L.rism_read_cvk(cvk_path.encode(), V, nt, _dptr(cvk)), "rism_read_cvk")
    n = C.c_int()
    _check(L.rism_read_xsv(xsv_path.encode(), 0, C.byref(n), None, None, None, None, None), "rism_read_xsv")
    k, c, w = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
    pf, w0 = C.c_double(), C.c_double()
    _check(L.rism_read_xsv(xsv_path.encode(), n.value, C.byref(n), _dptr(k), _dptr(c), _dptr(w), C.byref(pf),
                           C.byref(w0)), "rism_read_xsv")
    chs, whs = np.zeros((V, n.value)), np.zeros((V, n.value))
    chs[0], whs[0] = c, w
    pfv, w0v = np.zeros(V), np.zeros(V)
    pfv[0], w0v[0] = pf.value, w0.value
    return RmdftTables(cvk, k, chs, whs, pfv, w0v)


def read_cube(path: str, n) -> np.ndarray:
    """ESP cube values (Hartree/e) on the grid n, x-fastest [N] (rism_read_cube)."""
    nn = (C.c_int * 3)(*n)
    out = np.zeros(int(np.prod(n)))
    _check(lib().rism_read_cube(path.encode(), nn, _dptr(out)), "rism_read_cube")
    return out


def stream_copy(bytes_: int = 1 << 30, reps: int = 20, device: int = 0):
    """(GB/s, us per copy) of the library's plain device copy (rism_stream_copy)."""
    g, u = C.c_double(), C.c_double()
    _check(lib().rism_stream_copy(device, bytes_, reps, C.byref(g), C.byref(u)), "rism_stream_copy")
    return g.value, u.value


def release_cache() -> int:
    """Free the device buffers, streams and pinned blocks that destroyed
    handles left for reuse (rism_release_cache); returns the bytes freed."""
    return int(lib().rism_release_cache())


def solvent_model(path: str):
    """The site rows, ntab, dr and temperature of an .xvk header as a model
    for chi_solve: (sites ChiSite array, ntab, dr, temperature)."""
    sites = (ChiSite * 64)()
    ns, nt, dr, T = C.c_int(), C.c_int(), C.c_double(), C.c_double()
    _check(lib().rism_xvk_read_sites(path.encode(), sites, 64, C.byref(ns), C.byref(nt), C.byref(dr), C.byref(T)),
           "rism_xvk_read_sites")
    out = (ChiSite * ns.value)()
    C.memmove(out, sites, C.sizeof(ChiSite) * ns.value)
    return out, nt.value, dr.value, T.value


def make_sites(rows) -> C.Array:
    """ChiSite array from rows (name, species, sym, sig A, eps J/mol, q e, (x, y, z) A, conc M)."""
    out = (ChiSite * len(rows))()
    for s, (name, spc, sym, sig, eps, q, xyz, dens) in zip(out, rows):
        s.name, s.species, s.sym = name.encode()[:7], spc, sym
        s.sig, s.eps, s.q, s.dens = sig, eps, q, dens
        s.xyz[:] = list(xyz)
    return out


def chi_solve(sites, ntab: int = 16384, dr: float = 0.01, temper: float = 298.15, closure: int = KH,
              tol: float = 1e-10, maxstep: int = 10000, nhist: int = 10, mix: float = 0.3, chargeup: int = 4,
              dielec: float = 0.0, adbcor: float = 0.5):
    """chi_vv(k) of a solvent model by the library's 1D XRISM solve, or DRISM
    to the dielectric constant `dielec` (> 0) (rism_chi_solve): returns
    (Solvent, info dict with steps, rms, xt, status)."""
    L = lib()
    o = ChiOpts()
    L.rism_chi_default_opts(C.byref(o))
    o.ntab, o.dr, o.temper, o.closure, o.tol = ntab, dr, temper, closure, tol
    o.maxstep, o.nhist, o.mix, o.chargeup = maxstep, nhist, mix, chargeup
    o.dielec, o.adbcor = dielec, adbcor
    V = L.rism_chi_nreduced(sites, len(sites))
    chi = np.zeros((ntab, max(V, 1), max(V, 1)))
    xt, rms, steps = C.c_double(), C.c_double(), C.c_int()
    st = _check(L.rism_chi_solve(sites, len(sites), C.byref(o), _dptr(chi), C.byref(xt), C.byref(steps),
                                 C.byref(rms)), "rism_chi_solve")
    if st == SOLVE_DIVERGED:
        raise RuntimeError(f"rism_chi_solve diverged: {last_error()}")
    rho, q, sig, eps = (np.zeros(V) for _ in range(4))
    for s in sites:  # solvent_read.cc:71-80
        v = abs(s.sym) - 1
        sig[v], eps[v], q[v] = s.sig, s.eps, s.q
        rho[v] += s.dens * 6.02214076e23 * 1e-27
    sv = Solvent(rho, q, sig, eps, chi, np.pi / (dr * ntab), temper, xt.value)
    return sv, {"status": st, "steps": steps.value, "rms": rms.value, "xt": xt.value}


def write_xvk(path: str, sites, solvent: Solvent, dr: float):
    """Write `solvent` (from chi_solve on `sites`) as an .xvk file."""
    chi = np.ascontiguousarray(solvent.chi, dtype=np.float64)
    _check(lib().rism_write_xvk(path.encode(), sites, len(sites), chi.shape[0], dr, solvent.temper, solvent.xt,
                                _dptr(chi)), "rism_write_xvk")


class Solver:
    """One 3D-RISM solve on one GPU (handle over the C ABI)."""

    def __init__(self, n, box, solvent: Solvent, solute: Solute, opts: Options = Options(),
                 device: int = 0):
        L = lib()
        self._keep = []
        g = Grid((C.c_int * 3)(*n), (C.c_double * 3)(*box))

        def keep(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return _dptr(a)

        sv = SolventC(solvent.nsite, solvent.chi.shape[0], solvent.dk, solvent.temper, solvent.xt,
                      keep(solvent.rho), keep(solvent.q), keep(solvent.sig), keep(solvent.eps),
                      keep(solvent.chi))
        su = SoluteC(len(solute.q), keep(solute.sig), keep(solute.eps), keep(solute.q),
                     keep(np.asarray(solute.xyz).reshape(-1)))
        o = Opts(opts.closure, opts.tol, opts.maxstep, opts.count, opts.m, opts.mp, opts.nhist,
                 opts.precision, opts.chargeup, opts.verbose, opts.ksave)
        h = C.c_void_p()
        _check(L.rism_create(C.byref(g), C.byref(sv), C.byref(su), C.byref(o), device, C.byref(h)),
               "rism_create")
        self.h = h
        self.n = tuple(n)
        self.N = int(np.prod(n))
        self.V = solvent.nsite
        self.natom = len(solute.q)
        self.opts = opts

    @classmethod
    def from_deck(cls, deck: Deck, **over):
        o = Options(**{**deck.opts.__dict__, **over})
        return cls(deck.n, deck.box, deck.solvent, deck.solute, o)

    def build_potentials(self):
        _check(lib().rism_build_potentials(self.h), "rism_build_potentials")
        return self

    def set_esp(self, esp: np.ndarray):
        """Replace the point-charge potential by esp[N] (Hartree/e), as the
        reference's -e cube input (rism_set_esp); after build_potentials."""
        e = np.ascontiguousarray(esp, dtype=np.float64).reshape(-1)
        if e.size != self.N:
            raise ValueError(f"esp has {e.size} values, the grid {self.N}")
        _check(lib().rism_set_esp(self.h, _dptr(e)), "rism_set_esp")
        return self

    def solve(self, tuv_path: Optional[str] = None):
        steps, rms = C.c_int(), C.c_double()
        st = _check(lib().rism_solve(self.h, tuv_path.encode() if tuv_path else None, C.byref(steps),
                                     C.byref(rms)), "rism_solve")
        return st, steps.value, rms.value

    def thermo(self) -> dict:
        x = Xmu()
        _check(lib().rism_thermo(self.h, C.byref(x)), "rism_thermo")
        self._xmu = x
        return x.as_dict()

    def write_xmu(self, path: str):
        _check(lib().rism_write_xmu(C.byref(self._xmu), self.opts.closure, path.encode()), "rism_write_xmu")

    def grad(self) -> np.ndarray:
        out = np.zeros(self.natom * 6)
        _check(lib().rism_grad(self.h, _dptr(out)), "rism_grad")
        return out.reshape(self.natom, 6)

    def grad_noalloc(self) -> np.ndarray:
        """rism_grad through the entry point that allocates no device memory
        (partial sums in the library's static per-GPU scratch)."""
        out = np.zeros(self.natom * 6)
        _check(lib().rism_grad_noalloc(self.h, _dptr(out)), "rism_grad_noalloc")
        return out.reshape(self.natom, 6)

    def adjoint_grad(self, rtol: float = 1e-10) -> dict:
        """dDelta-mu/d(solute parameters) by the adjoint of the converged
        fixed point (rism_adjoint_grad): {"xyz": [n][3], "q": [n], "sig": [n],
        "eps": [n], "steps": adjoint iterations}, J/mol per unit."""
        n = self.natom
        dxyz, dq, dsig, deps = np.zeros(3 * n), np.zeros(n), np.zeros(n), np.zeros(n)
        steps = C.c_int()
        rc = _check(lib().rism_adjoint_grad(self.h, float(rtol), _dptr(dxyz), _dptr(dq), _dptr(dsig),
                                            _dptr(deps), C.byref(steps)), "rism_adjoint_grad")
        if rc == SOLVE_DIVERGED:
            raise RuntimeError(f"rism_adjoint_grad: adjoint solve diverged: {last_error()}")
        return {"xyz": dxyz.reshape(n, 3), "q": dq, "sig": dsig, "eps": deps, "steps": steps.value}

    def rmdft(self, tabs: RmdftTables) -> float:
        """RMDFT solvation free energy (J/mol) of the converged state (rism_rmdft)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (tabs.cvk, tabs.khs, tabs.chs, tabs.whs, tabs.pfhs, tabs.w0)]
        inp = RmdftInput(_dptr(arrs[0]), len(arrs[1]), _dptr(arrs[1]), _dptr(arrs[2]), _dptr(arrs[3]),
                         _dptr(arrs[4]), _dptr(arrs[5]))
        out = C.c_double()
        _check(lib().rism_rmdft(self.h, C.byref(inp), C.byref(out)), "rism_rmdft")
        return out.value

    def euv(self) -> np.ndarray:
        """per-(atom, site) solute-solvent energy x dv, [2 (LJ, ES)][V][natom] J/mol (rism_euv)"""
        V, n = self.V, self.natom
        out = np.zeros(2 * V * n)
        _check(lib().rism_euv(self.h, _dptr(out)), "rism_euv")
        return out.reshape(2, V, n)

    def qv(self) -> np.ndarray:
        """solvent charge density sum_v rho_v q_v g_v, [N] (rism_qv)"""
        out = np.zeros(self.N)
        _check(lib().rism_qv(self.h, _dptr(out)), "rism_qv")
        return out

    def field(self, which: int) -> np.ndarray:
        if which == FIELD_FK:
            out = np.zeros(2 * self.N)
            _check(lib().rism_get_field(self.h, which, _dptr(out)), "rism_get_field")
            return (out[0::2] + 1j * out[1::2]).reshape(self.n[::-1])
        size = {FIELD_DE: self.N, FIELD_FR: self.N, FIELD_CHI: self.V * self.V * self.N}.get(
            which, self.V * self.N)
        out = np.zeros(size)
        _check(lib().rism_get_field(self.h, which, _dptr(out)), "rism_get_field")
        shape = self.n[::-1]
        if which in (FIELD_DE, FIELD_FR):
            return out.reshape(shape)
        if which == FIELD_CHI:
            return out.reshape((self.V, self.V) + shape)
        return out.reshape((self.V,) + shape)

    def set_t(self, t: np.ndarray):
        t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1)
        _check(lib().rism_set_t(self.h, _dptr(t)), "rism_set_t")

    def save_tuv(self, path):
        _check(lib().rism_save_tuv(self.h, path.encode()), "rism_save_tuv")

    def load_tuv(self, path):
        _check(lib().rism_load_tuv(self.h, path.encode()), "rism_load_tuv")

    def stats(self) -> dict:
        s, r, a, b, c = C.c_int(), C.c_double(), C.c_double(), C.c_double(), C.c_double()
        _check(lib().rism_stats(self.h, C.byref(s), C.byref(r), C.byref(a), C.byref(b), C.byref(c)),
               "rism_stats")
        return {"steps": s.value, "rms": r.value, "t_setup": a.value, "t_solve": b.value,
                "t_thermo": c.value}

    def rms_history(self) -> np.ndarray:
        n = lib().rism_rms_history(self.h, None, 0)
        out = np.zeros(max(n, 1))
        lib().rism_rms_history(self.h, _dptr(out), n)
        return out[:n]

    def profile(self, enable: bool = True):
        """Start (and reset) or stop HIP-event timing of every kernel launch."""
        _check(lib().rism_profile(self.h, int(enable)), "rism_profile")

    def kernel_times(self) -> dict:
        """{kernel: (total_ms, launches)} since profile(True)."""
        ms = (C.c_double * len(KERNELS))()
        n = (C.c_long * len(KERNELS))()
        _check(lib().rism_kernel_times(self.h, ms, n), "rism_kernel_times")
        return {k: (ms[i], n[i]) for i, k in enumerate(KERNELS)}

    def autotune(self, reps: int = 3) -> dict:
        """Pick the fastest tile shapes by timing (rism_autotune); returns the geometry."""
        g = (C.c_int * 10)()
        _check(lib().rism_autotune(self.h, reps, g), "rism_autotune")
        keys = ["Lx", "Ty", "Tz", "ept_x", "ept_y", "ept_z", "xgroups", "xblocks", "packs", "hclass"]
        return dict(zip(keys, list(g)))

    def geometry(self) -> dict:
        g = (C.c_int * 10)()
        _check(lib().rism_geometry(self.h, g), "rism_geometry")
        keys = ["Lx", "Ty", "Tz", "ept_x", "ept_y", "ept_z", "xgroups", "xblocks", "packs", "hist_class"]
        return dict(zip(keys, list(g)))

    def fft3d(self, data_ptr: int, direction: int):
        """In-place transform of caller device memory (FFT3D::execute seam)."""
        _check(lib().rism_fft3d(self.h, C.c_void_p(data_ptr), int(direction)), "rism_fft3d")

    def calculate(self, t_ptr: int, res_ptr: int):
        """One fixed-point map t -> R on caller device memory (calculate.cu)."""
        _check(lib().rism_calculate(self.h, C.c_void_p(t_ptr), C.c_void_p(res_ptr)), "rism_calculate")

    def close(self):
        if getattr(self, "h", None):
            lib().rism_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_solve(systems, solvent: Solvent, opts: Options = Options(), workers: int = 8, device: int = 0):
    """Solve many solutes concurrently (rism_batch_solve).  systems: list of
    (n, box, Solute).  Returns (results, seconds) with one dict per system:
    status, steps, rms and the rism_thermo quantities."""
    L = lib()
    keep = []

    def dptr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _dptr(a)

    n = len(systems)
    grids = (Grid * n)()
    sols = (SoluteC * n)()
    for i, (g, box, su) in enumerate(systems):
        grids[i] = Grid((C.c_int * 3)(*g), (C.c_double * 3)(*box))
        sols[i] = SoluteC(len(su.q), dptr(su.sig), dptr(su.eps), dptr(su.q), dptr(np.asarray(su.xyz).reshape(-1)))
    sv = SolventC(solvent.nsite, solvent.chi.shape[0], solvent.dk, solvent.temper, solvent.xt,
                  dptr(solvent.rho), dptr(solvent.q), dptr(solvent.sig), dptr(solvent.eps), dptr(solvent.chi))
    o = Opts(opts.closure, opts.tol, opts.maxstep, opts.count, opts.m, opts.mp, opts.nhist,
             opts.precision, opts.chargeup, opts.verbose, opts.ksave)
    status, steps = (C.c_int * n)(), (C.c_int * n)()
    rms = (C.c_double * n)()
    xmu = (Xmu * n)()
    sec = C.c_double()
    _check(L.rism_batch_solve(n, grids, C.byref(sv), sols, C.byref(o), device, workers, status, steps, rms, xmu,
                              C.byref(sec)), "rism_batch_solve")
    out = []
    for i in range(n):
        d = {"status": status[i], "steps": steps[i], "rms": rms[i]}
        if status[i] >= 0:
            d.update(xmu[i].as_dict())
        out.append(d)
    return out, sec.value
