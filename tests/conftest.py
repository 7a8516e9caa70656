"""Shared test set-up: markers, import paths, cached inputs.

`-m gpu` tests run the HIP path on cuda:0 through the C ABI and compare with the
FP64 CPU oracle (oracle/); `-m "not gpu"` tests pin the oracle itself and check
the host-side logic and the library's exports without touching a GPU.
"""
import functools
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long-running")


@functools.lru_cache(maxsize=None)
def cu_tables(n=5000):
    from exastamp_amd import potentials
    return potentials.johnson_tables(nr=n, nrho=n)


@functools.lru_cache(maxsize=None)
def cuni_tables(n=2000):
    from exastamp_amd import potentials
    return potentials.johnson_tables((potentials.CU, potentials.NI), nr=n, nrho=n)


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    pyoracle.load()
    return pyoracle
