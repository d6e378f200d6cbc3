"""The whole keyframe on the GPU (KeyframeRunner's device mode: upload,
front end, tree fits on counted plans, sloam_kf_estimator, results, one
captured HIP graph per keyframe after the first) against the host path
(front end, host-planned tree fits, the host drop-in sloam::sloam), through
sloam_amd/keyframe_bench.

Bar: bit for bit.  Per keyframe (keyframe_bench --compare-device: a second
runner with the device estimator, from its own state, on the same scans,
estimates and submaps), the same success, the same XY-yaw / Z-roll-pitch
decisions and feature-match counts, the same pose bits, and the same
landmarks (models, ids, map matches); over a whole run
(--device-estimator), the same trajectory bits.  Both estimators evaluate
the same model functions (host/geometry.h) with the same elementary
functions (host/detmath.h: the two libms differ in last bits) and run the
same LM kernel."""
import json
import os
import subprocess

import numpy as np
import pytest

from sloam_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sloam_amd", "keyframe_bench")
WEIGHTS = os.path.join(ROOT, "sloam_amd", "weights", "cylinder_proposal.npy")
# the stream's estimator parameters: trunks up to 0.4 m radius pass the
# radius cap, and the tree association tolerates a few degrees of heading
PARAMS = ("--max-tree-radius", "0.45", "--tree-match-thresh", "1.0")


@pytest.fixture(scope="module")
def sequences(tmp_path_factory):
    d = tmp_path_factory.mktemp("kfdev")
    short = str(d / "short.bin")
    synth.write_keyframe_sequence(short, n_keyframes=12, H=128)
    stream = str(d / "stream.bin")
    synth.write_keyframe_sequence(stream, n_keyframes=40, H=128, spacing=0.2)
    return {"short": short, "stream": stream}


def _run(seq, tmp, model, extra=(), tag=""):
    traj = os.path.join(str(tmp), f"traj_{model}_{tag}.bin")
    cmd = [EXE, seq, "1", "1", "--tree-model", model, "--weights", WEIGHTS, "--traj-out", traj,
           *PARAMS, *extra]
    print("[kf-device]", " ".join(cmd[3:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return res, np.fromfile(traj, dtype=np.float64).reshape(-1, 7)


@pytest.mark.parametrize("which", ["short", "stream", "long"])
@pytest.mark.parametrize("model", ["plain", "learned"])
def test_device_estimator_matches_host_per_keyframe(sequences, long_stream, tmp_path, model,
                                                    which):
    assert os.path.exists(EXE), "build first"
    seq = long_stream if which == "long" else sequences[which]
    res, _ = _run(seq, tmp_path, model, extra=("--compare-device",), tag=which)
    dc = res["device_compare"]
    print(f"[kf-device] {model} {which}: {dc}; steps {res['steps']}", flush=True)
    assert dc["calls"] >= 10
    assert dc["success_differs"] == 0
    assert dc["pose_bit_equal"] == dc["calls"], dc
    assert dc["landmarks_equal"] == dc["calls"], dc
    assert dc["steps_equal"] == dc["calls"], dc


@pytest.mark.parametrize("which", ["stream", "long"])
def test_device_estimator_trajectory(sequences, long_stream, tmp_path, which):
    """A whole run with the device estimator gives the host estimator's
    trajectory bit for bit (the map built from its own outputs)."""
    seq = long_stream if which == "long" else sequences[which]
    host, t_host = _run(seq, tmp_path, "plain", tag="host")
    dev, t_dev = _run(seq, tmp_path, "plain", extra=("--device-estimator",), tag="dev")
    print(f"[kf-device] host success {host['success_rate']}, device {dev['success_rate']}; "
          f"latency p50 host {host['latency_ms']['p50']} ms, device {dev['latency_ms']['p50']} ms",
          flush=True)
    assert dev["device_estimator"] and not host["device_estimator"]
    assert np.array_equal(t_host, t_dev)
    assert host["success_rate"] == dev["success_rate"]


def test_device_keyframe_graph_matches_eager(sequences, tmp_path):
    """The captured device keyframe replays the eager sequence bit for bit."""
    import subprocess as sp
    trajs = []
    for graph in (1, 0):
        traj = os.path.join(str(tmp_path), f"traj_dev_{graph}.bin")
        cmd = [EXE, sequences["short"], "1", str(graph), "--tree-model", "learned", "--weights",
               WEIGHTS, "--traj-out", traj, "--device-estimator", *PARAMS]
        print("[kf-device]", " ".join(cmd[3:]), flush=True)
        r = sp.run(cmd, capture_output=True, text=True, timeout=150)
        assert r.returncode == 0, r.stderr[-2000:]
        trajs.append(np.fromfile(traj, dtype=np.float64))
    assert np.array_equal(trajs[0], trajs[1])
