"""Writes tests/golden/reference_decks_oracle.json: the FP64 oracle's results
for the reference's own example decks test3.inp (754 atoms, 64 A / 128^3)
and test4.inp (protein G, 846 atoms, 128 A / 256^3; the deck behind the
reference's published timings), with their own convergence settings
(KH, tol 1e-6, Anderson 2 0.6 0.2) and solvent table h2o_vv.xvk, and for
the same decks converged to RMS 1e-9 (the fixed point).

    python tests/golden/make_deck_oracle.py [deck ...]   (about 35 minutes on 8 cores)
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [REPO]
from oracle import oracle as orc  # noqa: E402

EX = os.path.join(HERE, "reference_examples")


def xmu(path):
    out = {}
    for ln in open(path):
        if "=" in ln:
            k, v = ln.split("=", 1)
            out[k.strip()] = float(v.split()[0])
    return out


def main():
    if not os.path.exists(orc.EXE):
        orc.build()
    path = os.path.join(HERE, "reference_decks_oracle.json")
    res = json.load(open(path)) if os.path.exists(path) else {}
    for deck in sys.argv[1:] or ("test3", "test4"):
        rec = {}
        for tag, tol in (("", None), ("fixed_point_", 1e-9)):
            d = tempfile.mkdtemp()
            try:
                for f in (f"{deck}.inp", "h2o_vv.xvk"):
                    shutil.copy(os.path.join(EX, f), os.path.join(d, f))
                if tol is not None:  # the same deck to RMS 1e-9: the fixed point
                    path_inp = os.path.join(d, f"{deck}.inp")
                    lines = open(path_inp).read().split("\n")
                    lines[3] = f" {tol:g}   10000"
                    open(path_inp, "w").write("\n".join(lines))
                r = subprocess.run([orc.EXE, "-q", "-j", f"{deck}.inp"], cwd=d, capture_output=True, text=True,
                                   check=True)
                j = json.loads(r.stdout.strip().splitlines()[-1])
                x = xmu(os.path.join(d, f"{deck}.xmu"))
            finally:
                shutil.rmtree(d, ignore_errors=True)
            rec.update({f"{tag}steps": j["steps"], f"{tag}rms": j["rms"], f"{tag}xmu": x})
            if tol is None:
                rec.update({"threads": j["threads"], "t_total_s": round(j["t_setup"] + j["t_iter"] + j["t_thermo"], 1)})
        res[deck] = rec
        print(deck, json.dumps(res[deck]), flush=True)
    with open(path, "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
