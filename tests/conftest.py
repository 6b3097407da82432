import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


_capture = None


def pytest_configure(config):
    global _capture
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: larger CPU cases")
    _capture = config.pluginmanager.getplugin("capturemanager")


def progress(*a):
    """A progress line that pytest's output capture does not hold back: a long case (the
    R-MAT 26 one prepares its graph for minutes) shows as it goes that it is running, so
    a runner that takes minutes of silence for a hang sees it alive."""
    if _capture is not None:
        with _capture.global_and_fixture_disabled():
            print(*a, flush=True)
    else:
        print(*a, flush=True)


def _ensure_built():
    lib = ROOT / "blest_amd" / "libblest.so"
    if not lib.exists():
        subprocess.run([sys.executable, "-m", "blest_amd.build"], cwd=ROOT, check=True)
    orc = ROOT / "oracle" / "liboracle.so"
    if not orc.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def blest():
    import blest_amd
    return blest_amd


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.load()
    return o
