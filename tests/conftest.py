import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nse2d-channel-gpu_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running (minutes)")


@pytest.fixture
def progress(request):
    """Writes a line straight to the terminal, past pytest's output capture, so that a long test
    shows it is alive while it runs (and not only in its report)."""
    cm = request.config.pluginmanager.getplugin("capturemanager")

    def write(msg):
        if cm is not None:
            with cm.global_and_fixture_disabled():
                print(msg, flush=True)
        else:
            print(msg, flush=True)
    return write
