"""Shared pytest set-up: markers and import paths.

`-m gpu` tests drive the HIP library through its C-ABI on a real MI355X;
everything else runs on the CPU (oracle pinning, host logic, ABI exports).
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "nabla-ai_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU reference run")


@pytest.fixture(scope="session")
def orc():
    import oracle_py
    if not oracle_py.available("orc"):
        pytest.skip("oracle library not built")
    return oracle_py.OracleLib("orc")


@pytest.fixture(scope="session")
def ref():
    import oracle_py
    if not oracle_py.available("ref"):
        pytest.skip("reference library (oracle/_ref) not built here")
    return oracle_py.OracleLib("ref")


@pytest.fixture(scope="session")
def ctx():
    import nbl
    c = nbl.Ctx(0)
    yield c
    c.close()
