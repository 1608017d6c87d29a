"""pytest configuration: the `gpu` marker and import paths.

CPU tests (-m "not gpu") cover the oracle, host-side logic and the C ABI
exports; GPU tests (-m gpu) are the parity tests through libqob_hip.so.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "quantumopticsbase.jl_amd", ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and the built libqob_hip.so")


def gpu_available() -> bool:
    return Path("/dev/kfd").exists()


@pytest.fixture(scope="session", autouse=True)
def _announce(request):
    yield


@pytest.fixture(autouse=True)
def _gpu_context(request):
    """GPU tests run with the library context already created.  Creating it
    applies the packaged launch-shape file, so a test that sets tunables
    before its first library call would otherwise have them reset by that
    call when it happens to be the first GPU test in the process."""
    if request.node.get_closest_marker("gpu") is not None:
        import qobhip
        qobhip.context()
    yield
