"""Writes tests/golden/rmdft_benzene_256_oracle.json: the FP64 oracle's
SFE_RMDFT (and SFE_SC) for the reference's example_rmdft deck (benzene,
64 A / 256^3, KH, h2o_vv.xvk + h2o_vv.cvk + 287.xsv), with the reference's
convergence settings, and the same functional on four copies of the .xsv
table moved within its printed precision.  The reference's own
benzene.xmu holds -4393.43772.

    python tests/golden/make_rmdft_golden.py      (about 2 minutes on 8 cores)
"""
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [REPO]
from oracle import oracle as orc, rism1d as r1  # noqa: E402

EX = os.path.join(HERE, "reference_examples")


def read_xsv(path):
    lines = open(path).read().split("\n")
    n = int(lines[0])
    pf, w0 = map(float, lines[1].split())
    tab = np.array([[float(x) for x in ln.split()] for ln in lines[2:2 + n]])
    return tab[:, 0], tab[:, 1], tab[:, 2], pf, w0


def main():
    d = tempfile.mkdtemp()
    old = os.getcwd()
    try:
        shutil.copy(os.path.join(EX, "benzene_3drism.dat"), os.path.join(d, "3drism.dat"))
        shutil.copy(os.path.join(EX, "benzene"), os.path.join(d, "benzene"))
        shutil.copy(os.path.join(EX, "h2o_vv_16384.xvk"), os.path.join(d, "h2o_vv.xvk"))
        os.chdir(d)
        o = orc.Oracle.from_deck("3drism.dat", "benzene").read_xvk("h2o_vv.xvk").setup()
        st, steps, rms = o.iterate()
        k, c, w, pf, w0 = read_xsv(os.path.join(EX, "287.xsv"))
        cvk = r1.read_table(os.path.join(EX, "h2o_vv_16384.cvk"))
        val = o.rmdft(cvk, k, c, w, pf, w0)
        # the same functional with the .xsv columns (c, w, pf, w0) moved by up
        # to half a unit of their printed 6th significant digit: the spread a
        # differently rounded copy of the table would give
        jit = []
        for seed in (1, 2, 3, 4):
            rng = np.random.default_rng(seed)

            def rnd(a):
                a = np.asarray(a, dtype=float)
                e = np.floor(np.log10(np.abs(a) + 1e-300)) - 5
                return a + (rng.random(a.shape) - 0.5) * 10.0 ** e
            jit.append(o.rmdft(cvk, k, rnd(c), rnd(w), float(rnd(pf)), float(rnd(w0))))
        x = o.thermo()
    finally:
        os.chdir(old)
        shutil.rmtree(d, ignore_errors=True)
    out = {"deck": "example_rmdft benzene, 64 A / 256^3, KH, tol 1e-6, Anderson 2 0.6 0.2",
           "steps": steps, "rms": rms, "SFE_RMDFT": val, "SFE_SC": x["SFE_SC"], "SFE_GF": x["SFE_GF"],
           "SFE_RMDFT_xsv_rounding_jitter": jit,
           "reference_benzene_xmu": {"SFE_RMDFT": -4393.43772, "SFE_SC": 66573.52045, "SFE_GF": 39653.45450}}
    with open(os.path.join(HERE, "rmdft_benzene_256_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
