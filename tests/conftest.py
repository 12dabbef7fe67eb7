import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "xpcs-eigen_amd", ROOT / "oracle", ROOT / "tests", ROOT / "tests" / "golden"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def xmt():
    # torch ships its own HIP runtime; tests that hand torch device memory to
    # the library need torch's device initialised before the library's first
    # HIP call (the order bench.py uses)
    try:
        import torch
        if torch.cuda.is_available():
            torch.zeros(1, device="cuda")
    except ImportError:
        pass
    import xpcs_mt
    xpcs_mt.lib()
    return xpcs_mt


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def pytest_terminal_summary(terminalreporter):
    # name the benchmark-workload tests in the run's last lines
    done = [r.nodeid.split("::")[-1] for r in terminalreporter.stats.get("passed", []) if "::test_config" in r.nodeid]
    if done:
        terminalreporter.write_line("benchmark-workload tests passed: " + ", ".join(done))
