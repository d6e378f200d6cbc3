"""The small-scan plain-method keyframe (sloam_amd/plain_keyframe.py) on the
GPU against the oracle's restatement of the same steps (oracle.plain_keyframe):
one 64 x 2048 scan of a 40-trunk forest (radius U(0.1, 0.4) m, tilt <= 5 deg,
height 2-8 m, sigma 2 cm; ground slope <= 5 deg, sigma 3 cm; 30 % clutter),
per-tree cylinder RANSAC with 200 hypotheses at 5 cm, ground-plane RANSAC
with 200 hypotheses, and the pose against the 40-trunk map from an estimate
0.3 m / 3 deg off the true pose.

Bar: ground list, tree points and their offsets bit-exact; per-tree and
ground hypotheses (samples), inlier counts, best hypotheses and inlier masks
bit-exact; tree and ground models within 1e-6 (m) of the FP64 oracle; the
same matches and residual points; the pose within 1e-4 (m, quaternion
components) of the FP64 oracle's LM.  Every error is printed."""
import math

import numpy as np
import pytest
import torch

import oracle
from sloam_amd import synth
from sloam_amd.plain_keyframe import PlainKeyframe

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [0, 1])
def test_c0_keyframe_against_oracle(ctx, seed):
    s, submap, plane, est = synth.make_small_keyframe(seed)
    print(f"[c0] seed {seed}: 64x2048 scan, {len(submap)} map trunks, estimate "
          f"{np.linalg.norm(est[:3]):.2f} m / "
          f"{math.degrees(2 * math.acos(min(1.0, est[6]))):.2f} deg off", flush=True)
    dev = torch.device("cuda", 0)
    x = torch.as_tensor(s["xyzi"]).to(dev)
    lab = torch.as_tensor(s["label"]).to(dev)
    torch.cuda.synchronize()
    kf = PlainKeyframe(ctx)
    g = kf.run(x, lab, est, submap, plane, keep=True)
    kf.close()
    o = oracle.plain_keyframe(s["xyzi"], s["label"], est, submap, plane, threads=8)
    # front end
    np.testing.assert_array_equal(g["ground_xyz"], o["ground_xyz"])
    np.testing.assert_array_equal(g["tree_offsets"], o["tree_offsets"])
    np.testing.assert_array_equal(g["tree_xyz"], o["tree_xyz"])
    nt = len(g["tree_offsets"]) - 1
    print(f"[c0] {nt} trees, {int(g['tree_offsets'][-1])} tree points, "
          f"{len(g['ground_xyz'])} ground points", flush=True)
    assert nt >= 10
    # per-tree cylinder RANSAC
    for k in ("samples", "counts", "best"):
        np.testing.assert_array_equal(g["cyl"][k], o["cyl"][k], err_msg=k)
    np.testing.assert_array_equal(g["cyl"]["mask"].view(np.uint32), o["cyl"]["mask"])
    ok = o["cyl"]["model"][:, 6] > 0
    assert np.array_equal(g["cyl"]["model"][:, 6] > 0, ok)
    dm = np.abs(g["cyl"]["model"][ok, :7] - o["cyl"]["model"][ok, :7]).max()
    print(f"[c0] tree models: largest difference from the FP64 oracle {dm:.2e}", flush=True)
    assert dm <= 1e-6
    # ground RANSAC
    for k in ("samples", "counts", "best"):
        np.testing.assert_array_equal(g["plane"][k], o["plane"][k], err_msg=k)
    np.testing.assert_array_equal(g["plane"]["mask"].view(np.uint32), o["plane"]["mask"])
    dg = np.abs(g["plane"]["model"][0, :7] - o["plane"]["model"][0, :7]).max()
    print(f"[c0] ground model: difference from the FP64 oracle {dg:.2e}", flush=True)
    assert dg <= 1e-6
    # association and pose
    np.testing.assert_array_equal(g["matches"], o["matches"])
    np.testing.assert_array_equal(g["problem"]["points"], o["problem"]["points"])
    assert int(g["status"][0]) == int(o["status"][0]) == 0
    dp = np.abs(g["pose"][:3] - o["pose"][:3]).max()
    qg, qo = g["pose"][3:], o["pose"][3:]
    dq = min(np.abs(qg - qo).max(), np.abs(qg + qo).max())
    et, er = synth.pose_error(g["pose"], np.array([0, 0, 0, 0, 0, 0, 1.0]))
    print(f"[c0] matched {int(np.sum(g['matches'] >= 0))} trees; pose vs FP64 oracle "
          f"{dp:.2e} m, {dq:.2e} (quaternion); vs truth {et * 100:.2f} cm, "
          f"{math.degrees(er):.3f} deg", flush=True)
    assert dp <= 1e-4 and dq <= 1e-4
    assert et < 0.05 and math.degrees(er) < 0.5
