import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "lbmengine.jl_amd"
for p in (str(PKG), str(ROOT / "oracle"), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblbm_hip.so on the device)")


@pytest.fixture(scope="session")
def oracle():
    """The oracle's ctypes harness, building liblbm_oracle.so if needed."""
    import harness
    if not harness.ORACLE_LIB.exists():
        harness.build_oracle()
    return harness


@pytest.fixture(scope="session")
def hiplib():
    """liblbm_hip.so path; built here (CPU cross-compile) when absent."""
    from lbmengine_hip import build as b
    if not b.LIB_PATH.exists():
        b.build()
    return b.LIB_PATH


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is visible; the HIP library must already be built."""
    from lbmengine_hip import LIB_PATH, _capi
    if not LIB_PATH.exists():
        pytest.fail(f"{LIB_PATH} is missing: build with `python __graft_entry__.py build` before running GPU tests")
    if _capi.device_count() < 1:
        pytest.skip("no GPU visible")
    import lbmengine_hip
    return lbmengine_hip
