"""The library's host code under AddressSanitizer: the `make asan` variant
(3d-rism-cuda_amd/_build_asan/, built by __graft_entry__.build(); device code
as in the regular build) runs check/batch_stress.cpp on the test1 deck's
water.  24 seeded solutes on 48^3-64^3 grids go through rism_batch_solve on
the per-solute path (worker threads owning handles and streams, reusing
destroyed handles' buffers, streams and pinned control blocks, swapping
solutes on a handle) with 1 and 6 workers and on the batched-launch path
with 1 and 3 workers, then four threads create, solve and destroy handles at
once.  Every solve converges with the same steps and SFE_SC in every run,
and AddressSanitizer finds no invalid host memory access (it aborts the
process on one)."""
import os
import subprocess

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu

BIN = os.path.join(PKG, "_build_asan", "rism-batch-stress")


def test_asan_batch_paths_and_handle_threads(gpu, deckdir):
    if not os.path.exists(BIN):
        pytest.skip("the AddressSanitizer variant is not built (make -C 3d-rism-cuda_amd asan)")
    print("asan batch stress", flush=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    env.pop("RISM_BATCH_LAUNCH", None)
    r = subprocess.run([BIN, "test1.inp"], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "PASS: 0 problems" in r.stdout
