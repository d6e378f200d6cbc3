"""CPU tests pinning the projection / partition oracle
(sloam/src/segmentation/inference.cpp:80-300): closed-form pixels, a
brute-force z-buffer, the partition invariants, the synthetic scan's beam
geometry and the reference's own organised tree cloud fixture."""
import math
import os

import numpy as np

import oracle
from sloam_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_closed_form_pixels():
    # (x, y, z) -> (col, row) for H = 64, W = 2048, fov +-22.5 deg
    pts = np.array([
        [10.0, 0.0, 0.0, 0],     # yaw 0 -> col 1024; pitch 0 -> row 32
        [0.0, 10.0, 0.0, 0],     # yaw -pi/2 -> col 512
        [0.0, -10.0, 0.0, 0],    # yaw +pi/2 -> col 1536
        [-10.0, 1e-3, 0.0, 0],   # yaw ~ -pi -> col 0
        [10.0, 0.0, 10.0 * math.tan(math.radians(30)), 0],   # above fov -> row 0
        [10.0, 0.0, -10.0 * math.tan(math.radians(30)), 0],  # below fov -> row 63
    ], np.float32)
    lab = np.array([2, 2, 1, 1, 0, 0], np.uint8)
    o = oracle.project_partition(pts, lab)
    assert o["proj"].tolist() == [[1024, 32], [512, 32], [1536, 32], [0, 32], [1024, 0],
                                  [1024, 63]]


def test_zbuffer_brute_force_and_invariants():
    rng = np.random.default_rng(0)
    n = 20000
    # many points per pixel: a small image
    H, W = 8, 32
    d = rng.normal(size=(n, 3))
    d[:, 2] *= 0.2
    r = rng.uniform(0.5, 30, n)
    r[:50] = r[50]  # exact range ties in some pixels
    xyz = d / np.linalg.norm(d, axis=1, keepdims=True) * r[:, None]
    xyzi = np.concatenate([xyz, rng.uniform(0, 1, (n, 1))], 1).astype(np.float32)
    xyzi[::97, 1] = np.nan
    lab = rng.integers(0, 3, n).astype(np.uint8)
    o = oracle.project_partition(xyzi, lab, H=H, W=W)
    proj = o["proj"]
    valid = np.isfinite(xyzi[:, :3]).all(1)
    assert np.all(proj[~valid] == -1)
    rng32 = np.sqrt((xyzi[:, :3].astype(np.float32) ** 2).sum(1, dtype=np.float32))
    # brute force: per pixel the minimum range, lowest index on ties
    best = {}
    for i in np.nonzero(valid)[0]:
        key = (proj[i, 1], proj[i, 0])
        if key not in best or rng32[i] < rng32[best[key]]:
            best[key] = i
    win = np.full((H, W), -1)
    for (rr, cc), i in best.items():
        win[rr, cc] = i
    assert np.array_equal(win, o["winner"])
    # mask: invalid below 1 m, else the winner's class
    for rr in range(H):
        for cc in range(W):
            w = win[rr, cc]
            exp = 0
            if w >= 0 and rng32[w] >= 1.0:
                exp = {2: 255, 1: 1, 0: 0}[int(lab[w])]
            assert o["mask"][rr, cc] == exp
    # ground list: input order, exactly the points whose pixel is ground
    pm = np.zeros(n, np.uint8)
    pm[valid] = o["mask"][proj[valid, 1], proj[valid, 0]]
    assert np.array_equal(o["ground"], xyzi[pm == 1])
    # dense tree cloud: the point where its pixel is tree, NaN elsewhere
    t = o["tree"]
    assert np.array_equal(t[pm == 255], xyzi[pm == 255])
    assert np.all(np.isnan(t[pm != 255, :3]))


def test_scan_beams_land_on_their_pixels():
    s = synth.make_scan(seed=3, H=64, W=2048)
    o = oracle.project_partition(s["xyzi"], s["label"])
    fin = np.isfinite(s["xyzi"][:, 0])
    idx = np.nonzero(fin)[0]
    assert np.array_equal(o["proj"][idx, 0], idx % 2048)
    assert np.array_equal(o["proj"][idx, 1], idx // 2048)


def test_destagger_shift():
    s = synth.make_scan(seed=4, H=8, W=64, n_trees=30)
    lab = np.full(len(s["label"]), 2, np.uint8)
    a = oracle.project_partition(s["xyzi"], lab, H=8, W=64, destagger=False)
    b = oracle.project_partition(s["xyzi"], lab, H=8, W=64, destagger=True)
    A = a["tree"].reshape(8, 64, 4)
    B = b["tree"].reshape(8, 64, 4)
    for r in range(8):
        if r % 2 == 0:
            assert np.array_equal(B[r, 32:, :3], A[r, :32, :3], equal_nan=True)
            assert np.array_equal(B[r, :32], A[r, :32], equal_nan=True)
            assert np.array_equal(B[r, :, 3], A[r, :, 3], equal_nan=True)
        else:
            assert np.array_equal(B[r], A[r], equal_nan=True)


def test_reference_tree_fixture_dense_layout():
    """The reference's organised 2048 x 64 tree cloud (still_tree_t0.pcd): all
    its finite points are trees, so the dense tree output of the partition is
    the input itself (NaN fill kept, input order)."""
    f = np.load(os.path.join(GOLDEN, "still_tree_t0.npz"))
    W, H = int(f["width"]), int(f["height"])
    assert (W, H) == (2048, 64)
    xyzi = np.full((W * H, 4), np.nan, np.float32)
    xyzi[f["index"], :3] = f["xyz"]
    xyzi[f["index"], 3] = 0.0
    lab = np.full(W * H, 2, np.uint8)
    o = oracle.project_partition(xyzi, lab, H=H, W=W)
    rng = np.linalg.norm(f["xyz"], axis=1)
    assert rng.min() >= 1.0
    assert np.array_equal(o["tree"][f["index"]], xyzi[f["index"]])
    assert len(o["ground"]) == 0
    # NaN fill points carry intensity 0, as the reference's default PointT
    assert np.isnan(o["tree"][:, :3]).sum() == np.isnan(xyzi[:, :3]).sum()
    assert np.all(o["tree"][np.isnan(xyzi[:, 0]), 3] == 0)
