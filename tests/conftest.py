import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def ctx():
    """One library context on cuda:0 for the whole GPU session."""
    from sloam_amd.gpu import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def long_stream(tmp_path_factory):
    """The first 200 keyframes of the long keyframe stream (128 x 2048 scans
    0.2 m apart in a 260 m forest, odometry noise 0.0316 m / 0.316 deg per
    step), cast by eight processes in a child (the test process holds the
    GPU)."""
    import subprocess
    path = str(tmp_path_factory.mktemp("long_stream") / "long.bin")
    print("[long-stream] writing 200 keyframes", flush=True)
    subprocess.run([sys.executable, "-m", "sloam_amd.synth", path, "200", "0.2", "260",
                    "0.0316227766", "0.316227766", "8", "128"], check=True, timeout=300, cwd=ROOT)
    return path
