"""ctypes view of the FP64 host oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker.  The product never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RISM_ORACLE_LIB names another build of the same sources (the SE pin tests
# load builds with a planted mistake this way, in a subprocess of their own)
LIB = os.environ.get("RISM_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")
EXE = os.path.join(HERE, "_build", "rism_oracle")


def build(force: bool = False):
    args = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


class XmuC(C.Structure):
    _fields_ = [("sfe_sc", C.c_double), ("sfe_sc_hnc", C.c_double), ("sfe_gf", C.c_double),
                ("sfec_sc", C.c_double * 16), ("sfec_sc_hnc", C.c_double * 16),
                ("sfec_gf", C.c_double * 16), ("se", C.c_double), ("se_lj", C.c_double),
                ("se_es", C.c_double), ("pmv", C.c_double), ("pressure", C.c_double),
                ("correction", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.orc_new.restype = vp
        L.orc_free.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        L.orc_read_inp.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_int, C.c_int]
        L.orc_read_xvk.argtypes = [vp, C.c_char_p]
        L.orc_set_deck.argtypes = [vp, ip, dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_int, C.c_int, dp, dp, dp, dp]
        L.orc_setup.argtypes = [vp]
        L.orc_calculate.argtypes = [vp, C.c_double]
        L.orc_rms.argtypes = [vp]
        L.orc_rms.restype = C.c_double
        L.orc_anderson.argtypes = [vp]
        L.orc_anderson_reset.argtypes = [vp]
        L.orc_iterate.argtypes = [vp, C.c_int, C.c_int, C.c_char_p, ip, dp]
        L.orc_thermo.argtypes = [vp, C.POINTER(XmuC)]
        L.orc_grad.argtypes = [vp, dp, dp]
        L.orc_euv.argtypes = [vp, dp]
        L.orc_rmdft.argtypes = [vp, dp, C.c_int, dp, dp, dp, C.c_double, C.c_double, dp]
        L.orc_write_xmu.argtypes = [vp, C.POINTER(XmuC), C.c_char_p]
        L.orc_fft3d.argtypes = [C.c_void_p, ip, dp, C.c_int]
        L.orc_fft1d.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_spline_coef.argtypes = [dp, dp, C.c_int, dp, dp, dp]
        L.orc_splint.argtypes = [dp, dp, dp, dp, dp, C.c_int, C.c_double]
        L.orc_splint.restype = C.c_double
        L.orc_ng.argtypes = [vp]
        L.orc_ng.restype = C.c_long
        L.orc_natv.argtypes = [vp]
        L.orc_natu.argtypes = [vp]
        L.orc_copy_field.argtypes = [vp, C.c_int, dp]
        L.orc_set_t.argtypes = [vp, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_nhist.argtypes = [vp, C.c_int]
        L.orc_set_iter.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_alloc_test.argtypes = [vp, C.c_int, C.c_long, C.c_int, dp]
        L.orc_set_tr.argtypes = [vp, dp]
        L.orc_set_g.argtypes = [vp, dp]
        L.orc_copy_hist.argtypes = [vp, C.c_int, C.c_int, dp]
        L.orc_set_atom_xyz.argtypes = [vp, C.c_int, dp]
        L.orc_write_tuv.argtypes = [vp, C.c_char_p]
        L.orc_read_tuv.argtypes = [vp, C.c_char_p]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fft3d(a: np.ndarray, n, box, direction: int) -> np.ndarray:
    """Reference-convention 3D transform of an x-fastest complex field.
    a: complex128 array of shape (nz, ny, nx)."""
    out = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().orc_fft3d(out.ctypes.data, (C.c_int * 3)(*n), (C.c_double * 3)(*box), int(direction))
    return out


def fft1d(a: np.ndarray, sign: int) -> np.ndarray:
    out = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().orc_fft1d(out.ctypes.data, len(out), int(sign))
    return out


F_T, F_TR, F_G, F_H, F_DU, F_DE, F_FR, F_FK, F_CHI, F_XYZ = range(10)


class Oracle:
    """One FP64 host solve (mirrors RISM3D on the CPU)."""

    def __init__(self):
        self.s = lib().orc_new()

    @classmethod
    def from_deck(cls, control, structure=None, centering=True, zero=False, nhist=2, setup=True):
        o = cls()
        L = lib()
        if L.orc_read_inp(o.s, control.encode(), structure.encode() if structure else None,
                          int(centering), int(zero)):
            raise RuntimeError(L.orc_last_error().decode())
        L.orc_set_nhist(o.s, nhist)
        return o

    def read_xvk(self, path):
        if lib().orc_read_xvk(self.s, path.encode()):
            raise RuntimeError(lib().orc_last_error().decode())
        return self

    @classmethod
    def from_arrays(cls, n, box, xvk_path, sig, eps, q, xyz, closure=0, tol=1e-6, maxstep=10000,
                    count=2, m=0.6, mp=0.2, nhist=2):
        o = cls()
        L = lib()
        sig, eps, q = (np.ascontiguousarray(x, dtype=np.float64) for x in (sig, eps, q))
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        L.orc_set_deck(o.s, (C.c_int * 3)(*n), (C.c_double * 3)(*box), closure, tol, maxstep, count, m,
                       mp, nhist, len(q), _d(sig), _d(eps), _d(q), _d(xyz))
        o.read_xvk(xvk_path)
        return o

    def setup(self):
        if lib().orc_setup(self.s):
            raise RuntimeError(lib().orc_last_error().decode())
        return self

    def set_iter(self, tol, maxstep, count, m, mp):
        lib().orc_set_iter(self.s, tol, maxstep, count, m, mp)

    @property
    def N(self):
        return lib().orc_ng(self.s)

    @property
    def V(self):
        return lib().orc_natv(self.s)

    def iterate(self, cu=0, verbose=False, tuv=None):
        steps, rms = C.c_int(), C.c_double()
        st = lib().orc_iterate(self.s, cu, int(verbose), tuv.encode() if tuv else None, C.byref(steps),
                               C.byref(rms))
        return st, steps.value, rms.value

    def calculate(self, cf=1.0):
        lib().orc_calculate(self.s, cf)

    def rms(self):
        return lib().orc_rms(self.s)

    def anderson_reset(self):
        lib().orc_anderson_reset(self.s)

    def anderson(self):
        lib().orc_anderson(self.s)

    def set_t(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1)
        lib().orc_set_t(self.s, _d(t))

    @classmethod
    def bare(cls, natv, ng, nhist, rho, count, m, mp):
        """A field set with no grid, for exercising the Anderson step alone."""
        o = cls()
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        lib().orc_alloc_test(o.s, natv, ng, nhist, _d(rho))
        lib().orc_set_iter(o.s, 1e-6, 1000, count, m, mp)
        lib().orc_anderson_reset(o.s)
        return o

    def set_tr(self, tr):
        tr = np.ascontiguousarray(tr, dtype=np.float64).reshape(-1)
        lib().orc_set_tr(self.s, _d(tr))

    def set_g(self, g):
        """Overwrite g (test hook) for the thermodynamic sums."""
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
        lib().orc_set_g(self.s, _d(g))

    def write_tuv(self, path):
        """write_tuv.cu format: header n[3] V, then t, 6 significant digits."""
        if lib().orc_write_tuv(self.s, path.encode()) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def set_atom_xyz(self, u, xyz):
        """Move atom u without recomputing the potentials (test hook)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(3)
        lib().orc_set_atom_xyz(self.s, int(u), _d(xyz))

    def hist(self, which, slot):
        out = np.zeros(self.V * self.N)
        lib().orc_copy_hist(self.s, which, slot, _d(out))
        return out

    def thermo(self) -> dict:
        x = XmuC()
        lib().orc_thermo(self.s, C.byref(x))
        self._x = x
        v = self.V
        return {"SFE_SC": x.sfe_sc, "SFE_SC_HNC": x.sfe_sc_hnc, "SFE_GF": x.sfe_gf,
                "SFEC_SC": list(x.sfec_sc[:v]), "SFEC_SC_HNC": list(x.sfec_sc_hnc[:v]),
                "SFEC_GF": list(x.sfec_gf[:v]), "SE": x.se, "SE_LJ": x.se_lj, "SE_ES": x.se_es,
                "PMV": x.pmv, "Pressure": x.pressure, "Correction_Term": x.correction}

    def grad(self) -> np.ndarray:
        nu = lib().orc_natu(self.s)
        a, b = np.zeros(3 * nu), np.zeros(3 * nu)
        lib().orc_grad(self.s, _d(a), _d(b))
        return np.concatenate([a.reshape(nu, 3), b.reshape(nu, 3)], axis=1)

    def rmdft(self, cvk: np.ndarray, khs, chs, whs, pfhs: float, w0: float) -> float:
        """SFE_RMDFT (J/mol) of the last calculate (cal_rmdft.cu:5-114)"""
        cvk = np.ascontiguousarray(cvk, dtype=np.float64)
        khs, chs, whs = (np.ascontiguousarray(a, dtype=np.float64) for a in (khs, chs, whs))
        out = C.c_double()
        if lib().orc_rmdft(self.s, _d(cvk), len(khs), _d(khs), _d(chs), _d(whs), pfhs, w0, C.byref(out)):
            raise RuntimeError(lib().orc_last_error().decode())
        return out.value

    def euv(self) -> np.ndarray:
        """per-(atom, site) SE x dv, [2 (LJ, ES)][V][natu] (cal_euv.cu:7-76)"""
        nu = lib().orc_natu(self.s)
        out = np.zeros(2 * self.V * nu)
        lib().orc_euv(self.s, _d(out))
        return out.reshape(2, self.V, nu)

    def field(self, which: int, shape=None) -> np.ndarray:
        N, V = self.N, self.V
        nu = lib().orc_natu(self.s)
        size = {F_DE: N, F_FR: N, F_FK: 2 * N, F_CHI: V * V * N, F_XYZ: 3 * nu}.get(which, V * N)
        out = np.zeros(size)
        lib().orc_copy_field(self.s, which, _d(out))
        if which == F_FK:
            out = out[0::2] + 1j * out[1::2]
        return out

    def __del__(self):
        try:
            lib().orc_free(self.s)
        except Exception:
            pass
