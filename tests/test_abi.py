"""The C-ABI library: it loads, and it exports every entry point that
include/rism3d_amd.h declares (no compute calls, so this runs without a GPU).
Also the host-side readers behind the ABI, which need no device."""
import ctypes
import os
import re

import numpy as np
import pytest

import rism3d
from conftest import REPO, example

HEADER = os.path.join(REPO, "include", "rism3d_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rism_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("rism_create", "rism_build_potentials", "rism_solve", "rism_thermo", "rism_grad",
                 "rism_destroy", "rism_fft3d", "rism_calculate", "rism_write_xmu"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = rism3d.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_match_header():
    assert sorted(rism3d.EXPORTED) == declared_functions()


def test_version_and_error_string():
    lib = rism3d.lib()
    assert lib.rism_version().decode().startswith("rism3d_amd")
    assert isinstance(lib.rism_last_error(), bytes)


def test_xvk_reader_matches_file():
    """rism_xvk_read: site table and chi(k) of the reference's water file."""
    sv = rism3d.read_xvk(example("h2o_vv.xvk"))
    assert sv.nsite == 2
    nt = sv.chi.shape[0]
    assert nt > 1000
    # the table rows are chi(k)[v][v'] in file order, one k per line
    rows = [ln for ln in open(example("h2o_vv.xvk")) if not ln.startswith("#") and ln.strip()]
    assert len(rows) == nt
    for i in (0, 1, nt // 2, nt - 1):
        want = np.array([float(x) for x in rows[i].replace("E", "e").split()]).reshape(2, 2)
        assert np.allclose(sv.chi[i], want, rtol=1e-14, atol=0)
    # chi -> omega at large k: the diagonal tends to one
    assert np.allclose(np.diagonal(sv.chi[-1]), 1.0, atol=0.01)
    # neutral solvent: sum_v rho_v q_v = 0 (H has twice the O density)
    assert abs(float(np.sum(sv.rho * sv.q))) < 1e-6 * float(np.abs(sv.rho * sv.q).max())
    assert sv.temper > 0 and sv.dk > 0


def test_deck_reader_test1():
    d = rism3d.read_deck(example("test1.inp"))
    assert tuple(d.n) == (64, 64, 64)
    assert tuple(d.box) == (32.0, 32.0, 32.0)
    assert d.opts.closure == rism3d.KH
    assert d.opts.tol == pytest.approx(1e-6)
    assert d.opts.maxstep == 10000
    assert len(d.solute.q) == 1


def test_bad_deck_is_an_error(tmp_path):
    p = tmp_path / "bad.inp"
    p.write_text("garbage\n")
    with pytest.raises(RuntimeError):
        rism3d.read_deck(str(p))
