"""Shared pytest setup: the `gpu` marker, repo root on sys.path, and helpers
to tell whether the oracle / reference build / GPU library are usable."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcpet_hip.so)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O

    if not O.ORACLE_SO.exists():
        import subprocess

        subprocess.run(["make", "-C", str(ROOT / "oracle"), "oracle"], check=True,
                       capture_output=True)
    return O


@pytest.fixture(scope="session")
def gpu_ctx():
    """A libcpet_hip context on device 0; fails (does not skip) when the
    library is missing, so a GPU run can never pass without the HIP path."""
    from pycpet_amd import _native as nat

    nat.load()
    if nat.device_count() < 1:
        pytest.fail("no GPU visible to libcpet_hip")
    return nat.default_context()
