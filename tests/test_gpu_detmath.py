"""detmath.h's sine, cosine, arc tangent and arc cosine give the same bits on
the GPU as on the host (tests/cpp/detmath_gpu: about a million arguments per
function, over [-4, 4], [-1000, 1000] and magnitudes 1e-8 to 1e8, and the
exact points).  The pose composition, the angle-axis conversions and the
model gates of both estimators call these, so the host estimator and the
device estimator agree bit for bit; the program also reports how often the
two libms differ on the same arguments."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "detmath_gpu")


def test_detmath_host_device_bit_equal():
    assert os.path.exists(BIN), "build first (make -C tests/cpp)"
    print("[detmath] detmath_gpu 1048576", flush=True)
    r = subprocess.run([BIN, "1048576"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.count("sloam_dm host/device differ 0 of") == 4
