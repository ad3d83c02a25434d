"""Shared fixtures: import paths, the gpu marker, reference example files."""
import os
import shutil
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "3d-rism-cuda_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
EXAMPLES = os.path.join(GOLDEN, "reference_examples")
for p in (PKG, REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: long-running CPU oracle case")


def example(name: str) -> str:
    return os.path.join(EXAMPLES, name)


@pytest.fixture
def deckdir(tmp_path):
    """A scratch directory holding copies of the reference example decks
    (their .inp files name the solvent file relative to the run directory)."""
    for f in os.listdir(EXAMPLES):
        shutil.copy(os.path.join(EXAMPLES, f), tmp_path / f)
    old = os.getcwd()
    os.chdir(tmp_path)
    yield tmp_path
    os.chdir(old)


@pytest.fixture(scope="session")
def gpu():
    """Fail loudly (not skip) when a gpu-marked test runs without the HIP
    library or without a device."""
    import rism3d
    n = rism3d.lib().rism_device_count()
    if n < 1:
        pytest.fail("gpu test selected but no HIP device is visible")
    return n
