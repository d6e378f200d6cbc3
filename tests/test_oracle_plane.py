"""CPU tests pinning the ground-plane oracle: the reference's plane_test
assertions, closed-form planes, a brute-force re-binning of the reference's
own ground cloud fixture (sloam/src/tests/aux/still_ground_t0.pcd) and the
plane RANSAC predicate against the FP64 point-to-plane distance."""
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _xyzi(xyz):
    x = np.zeros((len(xyz), 4), np.float32)
    x[:, :3] = xyz
    return x


def test_reference_plane_test_three_points():
    """plane_test.cpp:57-96: three points (0,0,0), (0,1,0), (1,0,0) with
    numGroundFeatures = 1 give a valid plane through them."""
    pts = _xyzi(np.array([[0, 0, 0], [0, 1, 0], [1, 0, 0]], np.float32))
    # a single cell that holds everything (1 x 1 bins, no reduction)
    o = oracle.ground_planes(pts, radii_bins=1, theta_bins=1, min_dist=-1.0, max_dist=30.0,
                             retain_thresh=0.1, num_features=1)
    assert o["info"][0, 0] == 1 and o["info"][0, 1] == 3
    n = o["model"][0, :3]
    assert np.allclose(n, [0, 0, 1], atol=1e-12)
    assert abs(o["model"][0, 3]) < 1e-12
    assert np.allclose(o["model"][0, 4:7], [1 / 3, 1 / 3, 0], atol=1e-7)
    # distance of pta to the plane ~ 0 (EXPECT_NEAR(0, d, 0.1))
    d = abs(n @ [0, 0, 0] + o["model"][0, 3]) / np.linalg.norm(n)
    assert d < 0.1


def test_closed_form_tilted_planes_per_cell():
    rng = np.random.default_rng(1)
    n = np.array([math.sin(0.05), 0.02, math.cos(0.05)])
    n /= np.linalg.norm(n)
    xy = rng.uniform(-25, 25, (20000, 2))
    z = -(n[0] * xy[:, 0] + n[1] * xy[:, 1] + 1.7) / n[2]
    pts = _xyzi(np.c_[xy, z])
    o = oracle.ground_planes(pts, radii_bins=2, theta_bins=18, min_dist=1.0, max_dist=25.0,
                             retain_thresh=0.05, num_features=5)
    valid = o["info"][:, 0] == 1
    assert valid.sum() >= 30
    for c in np.nonzero(valid)[0]:
        m = o["model"][c]
        assert np.allclose(m[:3], n, atol=2e-6)
        assert abs(m[3] - 1.7) < 2e-5
        # kept = floor(binned * retain) and kept points are the lowest by z
        assert o["info"][c, 1] == int(o["info"][c, 2] / 20.0)


def test_reference_ground_fixture_binning_brute_force():
    """Bins of the reference's ground fixture, recomputed independently with
    numpy (float32 radius, float32 atan2, double bin arithmetic)."""
    f = np.load(os.path.join(GOLDEN, "still_ground_t0.npz"))
    xyz = f["xyz"]
    pts = _xyzi(xyz)
    R, T, lo, hi = 1, 18, 0.0, 30.0
    o = oracle.ground_planes(pts, radii_bins=R, theta_bins=T, min_dist=lo, max_dist=hi,
                             retain_thresh=0.05, num_features=60)
    x32, y32 = xyz[:, 0], xyz[:, 1]
    a = (np.float32(0) - x32).astype(np.float64) ** 2
    b = (np.float32(0) - y32).astype(np.float64) ** 2
    r = np.sqrt(a.astype(np.float32) + b.astype(np.float32)).astype(np.float64)
    ok = (r < hi) & (r > lo)
    th = np.arctan2(y32, x32).astype(np.float64)
    rb = np.clip(np.floor(r / (hi / R)).astype(int), 0, R - 1)
    tb = np.clip(np.floor((3.14159265 + th) / (2 * 3.14159265 / T)).astype(int), 0, T - 1)
    cell = rb * T + tb
    counts = np.bincount(cell[ok], minlength=R * T)
    assert np.array_equal(o["info"][:, 2], counts)
    # features of each valid cell: its 60 lowest points (ties: input order)
    for c in range(R * T):
        if o["info"][c, 0] != 1:
            continue
        idx = np.nonzero(ok & (cell == c))[0]
        order = idx[np.argsort(xyz[idx, 2], kind="stable")]
        assert np.array_equal(o["features"][c, :, :3], xyz[order[:60]])


def test_plane_ransac_noise_free_and_predicate():
    rng = np.random.default_rng(2)
    n = np.array([0.03, -0.04, 1.0])
    n /= np.linalg.norm(n)
    xy = rng.uniform(-10, 10, (3000, 2))
    z = -(n[0] * xy[:, 0] + n[1] * xy[:, 1] - 0.5) / n[2]
    P = np.c_[xy, z]
    # outliers at least 1 m off the plane
    out_xy = rng.uniform(-10, 10, (600, 2))
    off_z = rng.uniform(1, 5, 600) * rng.choice([-1, 1], 600)
    out_pts = np.c_[out_xy, -(n[0] * out_xy[:, 0] + n[1] * out_xy[:, 1] - 0.5) / n[2] + off_z]
    xyz = np.concatenate([P, out_pts]).astype(np.float32)
    off = np.array([0, len(xyz)], np.int32)
    o = oracle.plane_fit_batch(xyz, off, 200, thr=0.05)
    m = o["model"][0]
    assert np.allclose(m[:3], n, atol=1e-6)
    assert abs(m[3] + 0.5) < 1e-5
    assert m[7] == 3000
    # counts of the FP32 predicate vs the FP64 distance |n.p + d| < thr of the
    # same hypothesis planes: equal except for points within 1e-5 of thr
    o2 = oracle.plane_fit_batch(xyz, off, 50, thr=0.05, samples=True)
    X = xyz.astype(np.float64)
    for h in range(50):
        i0, i1, i2 = o2["samples"][0, h]
        nn = np.cross(X[i1] - X[i0], X[i2] - X[i0])
        if np.linalg.norm(nn) == 0:
            continue
        nn /= np.linalg.norm(nn)
        dist = np.abs((X - X[i0]) @ nn)
        near = np.abs(dist - 0.05) < 1e-5
        lo = int(((dist < 0.05) & ~near).sum())
        hi = int(((dist < 0.05) | near).sum())
        assert lo <= o2["counts"][0, h] <= hi
