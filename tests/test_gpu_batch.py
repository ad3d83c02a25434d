"""Many small solutes (rism_batch_solve): the batched workload of seeded
8-40-atom molecules on 60^3-72^3 grids, plus two of them re-gridded to
48^3 / 24 A and 96^3 / 48 A so that every grid class from 48^3 to 96^3
occurs.  By default every solute of a grid class runs in batched launches
(one launch per kernel for the whole class, each solute gated by its own
control block); RISM_BATCH_LAUNCH=0 selects the per-solute path.

Checks: every solve converges; every batched FP64 result equals the same
solute solved alone (same kernels and arithmetic, so to 1e-12 relative) in
both paths; the 48^3, 72^3 and 96^3 members against the FP64 oracle (SFE
terms within 0.05 J/mol, as for single solves); mixed precision against
FP64 within 0.01 kcal/mol (41.84 J/mol) for every solute.
"""
import numpy as np
import pytest

import rism3d
from rism3d import configs
from conftest import example
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

COUNT = 64


@pytest.fixture(scope="module")
def batch():
    sv = rism3d.read_xvk(example("h2o_vv.xvk"))
    ws = configs.batch_config(COUNT)
    extra = []
    for w, n in ((ws[1], 48), (ws[2], 96)):  # the smallest and largest grid classes
        extra.append(configs.Workload(f"{w.name} on {n}^3", (n, n, n), (n * 0.5,) * 3, w.solute))
    ws = ws + extra
    systems = [(w.n, w.box, w.solute) for w in ws]
    return sv, ws, systems


def test_batch_fp64_matches_single_solves_and_oracle(gpu, batch, monkeypatch):
    """Every batched FP64 result equals its standalone solve, in both
    batch paths: batched launches (one launch per kernel for every solute of
    a grid class) and the per-solute path (RISM_BATCH_LAUNCH=0).  In the
    per-solute path with one worker the solutes run in order on one handle
    per grid, so every consecutive pair on the same grid goes through the
    solute swap (rism::set_solute, which keeps the field buffers and
    captured graphs), including swaps to a larger solute, which grow the
    per-atom arrays."""
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
    for mode, workers in (("1", 1), ("1", 3), ("0", 1), ("0", 6)):
        monkeypatch.setenv("RISM_BATCH_LAUNCH", mode)
        print(f"batch fp64 n={len(systems)} batched launches={mode} workers={workers}", flush=True)
        res, sec = rism3d.batch_solve(systems, sv, rism3d.Options(), workers=workers)
        print("seconds", sec, "solutes/s", len(res) / sec, flush=True)
        assert all(r["status"] == rism3d.SOLVE_CONVERGED for r in res), [r["status"] for r in res]
        for i, (st, steps, x) in enumerate(single):
            assert st == rism3d.SOLVE_CONVERGED
            assert steps == res[i]["steps"], (mode, workers, i)
            for k in ("SFE_SC", "SFE_GF", "PMV"):
                assert x[k] == pytest.approx(res[i][k], rel=1e-12, abs=1e-9), (mode, workers, i, k)
    for n in (48, 72, 96):  # one member of each of three grid classes against the oracle
        i = next(k for k, w in enumerate(ws) if w.n[0] == n)
        w = ws[i]
        su = w.solute
        o = orc.Oracle.from_arrays(w.n, w.box, example("h2o_vv.xvk"), su.sig, su.eps, su.q, su.xyz).setup()
        ost, osteps, _ = o.iterate()
        ox = o.thermo()
        assert ost == 0
        print(f"oracle {n}^3 member {i}: steps {osteps} vs {res[i]['steps']}, "
              f"SFE_SC diff {res[i]['SFE_SC'] - ox['SFE_SC']:.3e}", flush=True)
        for k in ("SFE_SC", "SFE_GF"):
            assert abs(res[i][k] - ox[k]) < 0.05, (n, k)


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


def test_batch_step_limit_stop_matches_single_solves(gpu, batch, monkeypatch):
    """maxstep 7 inside the batched solves' 16-step chunks: every solute
    stops with the update of step 7 applied once (iterate.cu:61, 76-79), in
    both batch paths, as its standalone solve with the same limit does."""
    sv, ws, systems = batch
    sub = systems[:6]
    opts = rism3d.Options(maxstep=7, tol=1e-12)
    single = []
    for n, box, su in sub:
        s = rism3d.Solver(n, box, sv, su, opts)
        s.build_potentials()
        st, steps, _ = s.solve()
        single.append((st, steps, s.thermo()))
        s.close()
    for mode in ("1", "0"):
        monkeypatch.setenv("RISM_BATCH_LAUNCH", mode)
        print(f"batch step limit, batched launches={mode}", flush=True)
        res, _ = rism3d.batch_solve(sub, sv, opts, workers=2)
        for i, (st, steps, x) in enumerate(single):
            assert st == rism3d.SOLVE_MAXSTEP and steps == 7
            assert res[i]["status"] == rism3d.SOLVE_MAXSTEP and res[i]["steps"] == 7, (mode, i)
            for k in ("SFE_SC", "SFE_GF", "PMV"):
                assert x[k] == pytest.approx(res[i][k], rel=1e-12, abs=1e-9), (mode, i, k)
