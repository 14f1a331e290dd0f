import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# torch ships its own HIP runtime: imported before libsmax_hip.so is loaded,
# the library binds to that runtime and both see the GPU; loaded after it, a
# second runtime finds no device.  Import it first so the tests that share
# streams with torch pass whichever files run.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an AMD GPU (gfx950) and the built libsmax_hip.so")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the whole GPU test session (one process)."""
    from smaxkernels_amd import capi
    c = capi.Context(int(os.environ.get("SMAX_HIP_DEVICE", "0")))
    yield c
    c.close()


@pytest.fixture(scope="session")
def sort_pin(tmp_path_factory):
    """libstdc++'s real std::sort, run the way the reference conversion runs
    it (tests/golden/sort_pin.cpp), compiled with the system g++."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("pin") / "sort_pin"
    subprocess.run([gxx, "-O2", "-std=c++17",
                    str(ROOT / "tests" / "golden" / "sort_pin.cpp"), "-o",
                    str(exe)], check=True)
    return exe
