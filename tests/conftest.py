import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (HIP) and the built libcubam_amd.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o

    if os.environ.get("CUBAM_AMD_NO_ORACLE_BUILD") is None:
        o.build()
    return o


@pytest.fixture(scope="session")
def synth():
    from cubam_amd import synth as s

    if not s._LIB_PATH.exists():
        s.build()
    return s
