"""Many small solutes solved concurrently (rism_batch_solve), the batched
workload of seeded 8-40-atom molecules on 60^3-72^3 grids.

Checks: every solve converges; every batched FP64 result equals the same
solute solved alone (same arithmetic, so to 1e-12 relative), with one
worker (each same-grid neighbour through the solute swap) and with six; one solute
against the FP64 oracle (SFE terms within 0.05 J/mol, as for single
solves); mixed precision against FP64 within 0.01 kcal/mol (41.84 J/mol)
for every solute.
"""
import numpy as np
import pytest

import rism3d
from rism3d import configs
from conftest import example
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

COUNT = 12


@pytest.fixture(scope="module")
def batch():
    sv = rism3d.read_xvk(example("h2o_vv.xvk"))
    ws = configs.batch_config(COUNT)
    systems = [(w.n, w.box, w.solute) for w in ws]
    return sv, ws, systems


def test_batch_fp64_matches_single_solves_and_oracle(gpu, batch):
    """Every batched FP64 result equals its standalone solve.  With one
    worker the solutes run in order on one handle per grid, so every
    consecutive pair on the same grid goes through the solute swap
    (rism::set_solute, which keeps the field buffers and captured graphs),
    including swaps to a larger solute, which grow the per-atom arrays."""
    sv, ws, systems = batch
    swaps = [(ws[i - 1].n == ws[i].n, len(ws[i].solute.q) > len(ws[i - 1].solute.q)) for i in range(1, len(ws))]
    assert any(same and grow for same, grow in swaps), "workload has no same-grid swap to a larger solute"
    single = []
    for w in ws:
        s = rism3d.Solver(w.n, w.box, sv, w.solute, rism3d.Options())
        s.build_potentials()
        st, steps, _ = s.solve()
        single.append((st, steps, s.thermo()))
        s.close()
    for workers in (1, 6):
        print(f"batch fp64 n={len(systems)} workers={workers}", flush=True)
        res, sec = rism3d.batch_solve(systems, sv, rism3d.Options(), workers=workers)
        print("seconds", sec, "solutes/s", len(res) / sec, flush=True)
        assert all(r["status"] == rism3d.SOLVE_CONVERGED for r in res), [r["status"] for r in res]
        for i, (st, steps, x) in enumerate(single):
            assert st == rism3d.SOLVE_CONVERGED
            assert steps == res[i]["steps"], (workers, i)
            for k in ("SFE_SC", "SFE_GF", "PMV"):
                assert x[k] == pytest.approx(res[i][k], rel=1e-12, abs=1e-9), (workers, i, k)
    w = ws[0]
    su = w.solute
    o = orc.Oracle.from_arrays(w.n, w.box, example("h2o_vv.xvk"), su.sig, su.eps, su.q, su.xyz).setup()
    ost, _, _ = o.iterate()
    ox = o.thermo()
    assert ost == 0
    for k in ("SFE_SC", "SFE_GF"):
        assert abs(res[0][k] - ox[k]) < 0.05, k


def test_batch_mixed_within_bar_of_fp64(gpu, batch):
    sv, ws, systems = batch
    print(f"batch mixed n={len(systems)}", flush=True)
    r64, _ = rism3d.batch_solve(systems, sv, rism3d.Options(), workers=6)
    rmx, sec = rism3d.batch_solve(systems, sv, rism3d.Options(precision=rism3d.MIXED), workers=6)
    assert all(r["status"] == rism3d.SOLVE_CONVERGED for r in rmx)
    err = max(abs(a["SFE_SC"] - b["SFE_SC"]) for a, b in zip(rmx, r64))
    print("max |dmu mixed - fp64| J/mol", err, "seconds", sec, flush=True)
    assert err < 41.84


def test_batch_reports_errors_per_solute(gpu, batch):
    sv, ws, systems = batch
    bad = [((62, 62, 62), ws[0].box, ws[0].solute)] + systems[:1]  # 62 = 2 x 31: not 2,3,5-smooth
    res, _ = rism3d.batch_solve(bad, sv, rism3d.Options(), workers=2)
    assert res[0]["status"] < 0
    assert res[1]["status"] == rism3d.SOLVE_CONVERGED
