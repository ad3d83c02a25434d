"""MILC solvation free energy from a charged (.xmu) and an uncharged (.xmu0,
rism3d-amd -z) run: the reference's post-processing tool (tools/milc.py:1-66)
with the same coefficient sets, PMV scaling and printed lines.

    python -m rism3d.milc basename      # reads basename.xmu and basename.xmu0
"""
from __future__ import annotations

import re
import sys

KEYS = ["SFE_SC", "SFE_SC_HNC", "SFE_GF", "PMV"]
# tools/milc.py:7-16: weights of (SFE_SC, SFE_SC_HNC, SFE_GF, PMV x 4184)
COEFFS = {
    "Sub": {"xmu": [-0.718451957, 0.0, 1.62878139, 0.0],
            "xmu0": [5.19413792, 0.0, -3.58716288, -750.513679]},
    "HNC": {"xmu": [0.0, 0.99554824, 0.0, 0.0],
            "xmu0": [0.0, 1.50999895, 0.0, -682.54621964]},
}
PMV_SCALE = 4184.0
_LINE = re.compile(r"^\s*(\w+)\s*=\s*([-\d\.Ee+]+)")


def read_xmu(path: str) -> dict:
    """The MILC inputs of an .xmu file; PMV is scaled by 4184."""
    out = {}
    with open(path) as f:
        for line in f:
            m = _LINE.match(line)
            if m and m.group(1) in KEYS:
                v = float(m.group(2))
                out[m.group(1)] = v * PMV_SCALE if m.group(1) == "PMV" else v
    return out


def _weighted(values: dict, w) -> float:
    s = 0.0
    for k, c in zip(KEYS, w):
        if k not in values:
            raise RuntimeError(f"{k} not found")
        s += c * values[k]
    return s


def milc(xmu: dict, xmu0: dict) -> dict:
    """SFE(MILC) in J/mol for each coefficient set."""
    return {name: _weighted(xmu, c["xmu"]) + _weighted(xmu0, c["xmu0"]) for name, c in COEFFS.items()}


def report(base: str) -> str:
    """The text tools/milc.py prints for basename `base`."""
    res = milc(read_xmu(f"{base}.xmu"), read_xmu(f"{base}.xmu0"))
    lines = []
    for name, total in res.items():
        lines += [f"== {name} ==", f"SFE(MILC)= {total!r} (J/mol)", f"         = {total / 4184!r} (Kcal/mol)", ""]
    return "\n".join(lines) + "\n"


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("Usage: python -m rism3d.milc [basename]")
        return 1
    sys.stdout.write(report(argv[0]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
