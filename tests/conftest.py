import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sdpa-dd_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)
if os.path.join(ROOT, "tests") not in sys.path:
    sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def _oracle_built():
    return os.path.exists(os.path.join(ROOT, "oracle", "_build", "liboracle.so")) and os.path.exists(
        os.path.join(ROOT, "oracle", "_build", "sdpa_oracle"))


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (built on demand if missing; test infrastructure only)."""
    if not _oracle_built():
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-j4"], check=True, capture_output=True)
    import oracle_lib
    return oracle_lib.Oracle()


@pytest.fixture(scope="session")
def gpu():
    import sdpadd
    dev = sdpadd.Device(0)
    yield dev
    dev.close()
