"""Writes tests/golden/reference_decks_oracle.json: the FP64 oracle's results
for the reference's own example decks test3.inp (754 atoms, 64 A / 128^3)
and test4.inp (protein G, 846 atoms, 128 A / 256^3; the deck behind the
reference's published timings), with their own convergence settings
(KH, tol 1e-6, Anderson 2 0.6 0.2) and solvent table h2o_vv.xvk.

    python tests/golden/make_deck_oracle.py [deck ...]   (about 8 minutes on 8 cores)
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
        d = tempfile.mkdtemp()
        try:
            for f in (f"{deck}.inp", "h2o_vv.xvk"):
                shutil.copy(os.path.join(EX, f), os.path.join(d, f))
            r = subprocess.run([orc.EXE, "-q", "-j", f"{deck}.inp"], cwd=d, capture_output=True, text=True,
                               check=True)
            rec = json.loads(r.stdout.strip().splitlines()[-1])
            x = xmu(os.path.join(d, f"{deck}.xmu"))
        finally:
            shutil.rmtree(d, ignore_errors=True)
        res[deck] = {"steps": rec["steps"], "rms": rec["rms"], "threads": rec["threads"],
                     "t_total_s": round(rec["t_setup"] + rec["t_iter"] + rec["t_thermo"], 1), "xmu": x}
        print(deck, json.dumps(res[deck]), flush=True)
    with open(path, "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
