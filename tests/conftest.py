import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "xpcs-eigen_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def xmt():
    import xpcs_mt
    xpcs_mt.lib()
    return xpcs_mt


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
