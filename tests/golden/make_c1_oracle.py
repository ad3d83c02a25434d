"""Writes tests/golden/c1_oracle.npz: the FP64 oracle's solution of the
256^3 / 128 A, 846-atom KH workload (rism3d.configs.config(1), water chi from
the library's 1D DRISM-KH solve (eps 78.4), Anderson 2 0.6 0.2).

Two states are recorded:
  tol6   the oracle's own solve from the initial estimate to RMS <= 1e-6
         (steps, rms, every .xmu quantity, and g on a sub-lattice);
  conv   a solve of the same deck to RMS <= 1e-10 (the fixed point to far
         below the tolerance), with the same quantities.
g is stored (float32) on the sub-lattice x, y, z in {80, 82, ..., 174}
(every second point of the central 96^3, i.e. +-24 A around the solute's
centre), per site.  `ball_g` is max |g_tol6 - g_conv| over the whole grid:
how far a solver that stops at RMS 1e-6 may sit from the fixed point.

    python tests/golden/make_c1_oracle.py       (about 15 minutes on 8 cores)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "3d-rism-cuda_amd"), REPO]
from oracle import oracle as orc  # noqa: E402
from rism3d import configs  # noqa: E402

SUB = slice(80, 176, 2)


def sub(g, V, n):
    return g.reshape(V, n[2], n[1], n[0])[:, SUB, SUB, SUB].astype(np.float32)


def main():
    w = configs.config(1)
    xvk = configs.solvent_file("water")
    t0 = time.perf_counter()
    o = orc.Oracle.from_arrays(w.n, w.box, xvk, w.solute.sig, w.solute.eps, w.solute.q, w.solute.xyz,
                               tol=1e-6, nhist=w.nhist).setup()
    t_setup = time.perf_counter() - t0
    V = o.V
    out = {}
    meta = {"workload": w.name, "threads": os.cpu_count(), "t_setup_s": round(t_setup, 1)}
    for tag, tol in (("tol6", 1e-6), ("conv", 1e-10)):
        o.set_iter(tol, 10000, 2, 0.6, 0.2)
        t1 = time.perf_counter()
        st, steps, rms = o.iterate(verbose=False)
        t_it = time.perf_counter() - t1
        x = o.thermo()
        g = o.field(orc.F_G)
        out[f"g_{tag}"] = sub(g, V, w.n)
        if tag == "tol6":
            g6 = g
        else:
            meta["ball_g"] = float(np.abs(g6 - g).max())
        meta[tag] = {"status": st, "steps": steps, "rms": rms, "t_iter_s": round(t_it, 1),
                     **{k: x[k] for k in ("SFE_SC", "SFE_SC_HNC", "SFE_GF", "SE", "SE_LJ", "SE_ES", "PMV",
                                          "Pressure", "Correction_Term")}}
        print(tag, json.dumps(meta[tag]), flush=True)
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, "c1_oracle.npz"), **out)
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
