"""The CPU restatement of RunSloam (oracle/oracle_estimator.cpp) pinned by
what the mathematics fixes, independently of the drop-in:

* its Euler-angle reading of the rotation taking +z onto a plane normal
  against scipy's intrinsic x-y-z decomposition, mapped onto Eigen's
  documented range (first angle in [0, pi]);
* noise-free scenes built from known poses, trees and ground: the joint solve
  returns the true pose from a 0.2 m / 2 deg perturbation, the two-step solve
  from an x / y / yaw perturbation; every tree matches its map tree;
* the first scan's outputs (pose = estimate, no matches, models projected);
* each acceptance rule (tilted or overhead plane, radius cap, axis tilt, too
  short, out of range) on a scene where only that rule fails; the early
  return on an empty map."""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation as Rot

import oracle


def test_from_two_vectors_euler_matches_scipy():
    rng = np.random.default_rng(0)
    z = np.array([0.0, 0.0, 1.0])
    for i in range(3000):
        b = z + (0.3 if i % 2 else 2.0) * rng.normal(size=3)
        R, e = oracle.from_two_vectors_euler(z, b)
        bn = b / np.linalg.norm(b)
        assert np.allclose(R @ z, bn, atol=1e-12)
        # the smallest rotation: its angle is the angle between the vectors
        ang = np.linalg.norm(Rot.from_matrix(R).as_rotvec())
        assert abs(ang - math.acos(np.clip(z @ bn, -1, 1))) < 1e-9
        # scipy's intrinsic XYZ angles of R, mapped onto Eigen's range
        a, bb, c = Rot.from_matrix(R).as_euler("XYZ")
        if a < 0:
            a, bb, c = a + math.pi, math.copysign(math.pi, bb) - bb, c - math.copysign(math.pi, c)
        assert np.allclose(e, [a, bb, c], atol=1e-9), (e, (a, bb, c))
        assert 0.0 <= e[0] <= math.pi


def _pose(t, rotvec):
    q = Rot.from_rotvec(rotvec).as_quat()  # x, y, z, w
    return np.concatenate([t, q])


def _apply(pose, p):
    return Rot.from_quat(pose[3:]).apply(p) + pose[:3]


def _inv_apply(pose, p):
    return Rot.from_quat(pose[3:]).inv().apply(p - pose[:3])


def _compose(a, b):
    r = Rot.from_quat(a[3:]) * Rot.from_quat(b[3:])
    return np.concatenate([_apply(a, b[:3]), r.as_quat()])


def scene(truth, n_trees=10, radius=0.25, tree_tilt_deg=0.0, short=(), wide=(), far=(),
          plane_tilt=0.0, overhead=False, seed=0):
    """A keyframe record of a flat world (ground z = 0) seen from `truth`:
    36 ground cells (2 x 18) with 5 features each, n_trees vertical trunks
    with their vertices, points and exact cylinder fits, the map of the
    trunks (world frame) and the previous ground planes (world frame)."""
    rng = np.random.default_rng(seed)
    R = Rot.from_quat(truth[3:])
    # ground cells around the sensor, 6-9 m away
    cells, feats = [], []
    for k in range(36):
        th = 2 * math.pi * (k % 18) / 18 + 0.1
        rr = 6.5 if k < 18 else 8.5
        cw = np.array([truth[0] + rr * math.cos(th), truth[1] + rr * math.sin(th), 0.0])
        nw = np.array([0.0, 0.0, 1.0])
        if plane_tilt and k == 0:
            nw = np.array([math.sin(plane_tilt), 0.0, math.cos(plane_tilt)])
        if overhead and k == 1:
            cw = cw + np.array([0, 0, truth[2] + 1.0])
        pts_w = []
        for j in range(5):
            u = rng.uniform(-0.8, 0.8, 2)
            d = np.array([u[0], u[1], 0.0])
            d = d - (d @ nw) * nw
            pts_w.append(cw + d)
        pts_s = _inv_apply(truth, np.array(pts_w))
        ns = R.inv().apply(nw)
        cs = _inv_apply(truth, cw)
        cells.append(np.concatenate([ns, [-(ns @ cs)], cs]))
        feats.append(np.concatenate([pts_s, np.zeros((5, 1))], 1))
    cell_model = np.array(cells)
    prev = np.array([[0, 0, 1, 0, *(_apply(truth, c[4:7]))] for c in cells])
    # trunks on a ring 4-9 m around the sensor, at least 3 m apart
    trees_w = []
    while len(trees_w) < n_trees:
        th = rng.uniform(0, 2 * math.pi)
        rr = rng.uniform(4.0, 9.0)
        p = np.array([truth[0] + rr * math.cos(th), truth[1] + rr * math.sin(th), 0.0])
        if all(np.linalg.norm(p[:2] - q[:2]) > 3.0 for q in trees_w):
            trees_w.append(p)
    tree_off, vtx_int, vtx_f64, vtx_pt_off, pts = [0], [], [], [0], []
    cyl_ok, cyl_fit, map_models = [], [], []
    for t, root_w in enumerate(trees_w):
        # a trunk exactly parallel to the ground normal makes the reference's
        # tilt test acos(1 + ulp) = NaN (the tree is dropped): 0.05 deg off
        tilt = math.radians(tree_tilt_deg if t == 0 and tree_tilt_deg else 0.05)
        axis_w = np.array([math.sin(tilt), 0.0, math.cos(tilt)])
        rad = radius + (0.1 if t in wide else 0.0)
        if t in far:
            root_w = root_w + 30.0 * (root_w - truth[:3]) / np.linalg.norm(root_w[:2] - truth[:2])
            root_w[2] = 0.0
        height = 1.0 if t in short else 4.0
        n_v = 17
        e1 = np.cross(axis_w, [0, 1.0, 0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(axis_w, e1)
        for v in range(n_v):
            h = 0.3 + height * v / (n_v - 1)
            cw = root_w + h * axis_w
            angs = rng.uniform(0, 2 * math.pi, 6)
            pw = cw + rad * (np.cos(angs)[:, None] * e1 + np.sin(angs)[:, None] * e2)
            ps = _inv_apply(truth, pw)
            cs = _inv_apply(truth, cw)
            vtx_int.append((100 + t, 1))
            vtx_f64.append((rad, *cs.astype(np.float32)))
            pts.append(np.concatenate([ps, np.zeros((6, 1))], 1).astype(np.float32))
            vtx_pt_off.append(vtx_pt_off[-1] + 6)
        tree_off.append(tree_off[-1] + n_v)
        root_s = _inv_apply(truth, root_w)
        axis_s = R.inv().apply(axis_w)
        cyl_ok.append(1)
        cyl_fit.append((*root_s, *axis_s, rad))
        map_models.append((*root_w, *axis_w, rad))
    return {
        "pose": truth.copy(), "nf": 5, "cell_model": cell_model,
        "cell_valid": np.ones(36, np.int32), "cell_feat": np.array(feats, np.float32),
        "tree_off": np.array(tree_off, np.int32), "vtx_int": np.array(vtx_int, np.int32),
        "vtx_f64": np.array(vtx_f64), "vtx_pt_off": np.array(vtx_pt_off, np.int32),
        "pts": np.concatenate(pts), "line_int": np.zeros((0, 2), np.int32),
        "line_coeff": np.zeros((0, 6), np.float32), "cyl_ok": np.array(cyl_ok, np.int32),
        "cyl_fit": np.array(cyl_fit), "map": np.array(map_models), "prev_planes": prev,
    }


TRUTH = _pose(np.array([3.0, -2.0, 1.5]), np.radians([1.0, -0.5, 20.0]))


def _pose_err(a, b):
    dt = np.linalg.norm(a[:3] - b[:3])
    dq = min(np.linalg.norm(a[3:] - b[3:]), np.linalg.norm(a[3:] + b[3:]))
    return dt, dq


@pytest.mark.parametrize("tree_mode", [0, 2])
def test_joint_solve_recovers_true_pose(tree_mode):
    rec = scene(TRUTH)
    pert = _pose(np.array([0.15, -0.12, 0.05]), np.radians([1.0, -1.2, 1.5]))
    rec["pose"] = _compose(TRUTH, pert)
    est = oracle.SloamEstimator(twoStepOptim=0, treeMatchThresh=1.0)
    est.set_state(False, rec["prev_planes"])
    r = est.run(rec, tree_mode=tree_mode)
    assert r["ok"] and r["tree_step"] and r["ground_step"] and r["converged"]
    assert len(r["tm"]) == 10 and np.all(r["tree_valid"] == 1)
    assert r["tree_matches"] == 10 * 20 and r["plane_matches"] == 36 * 5
    dt, dq = _pose_err(r["pose"], TRUTH)
    assert dt < 1e-6 and dq < 1e-6, (dt, dq)
    np.testing.assert_array_equal(r["matches"], np.arange(10))
    # the kept planes, projected at the solved pose, lie on the world ground
    assert np.allclose(r["planes"][:, :3], [0, 0, 1], atol=1e-6)
    assert np.allclose(r["planes"][:, 3], 0, atol=1e-6)
    # trees projected at the solved pose lie on the map trees' axes: a root
    # stays at the bottom vertex's height (0.55 m) because the ground root
    # moves it only along a ray that points down (utils.h:41-52)
    assert np.allclose(r["tm"][:, 2], 0.55, atol=1e-4)
    foot = r["tm"][:, :3] - (r["tm"][:, 2:3] / r["tm"][:, 5:6]) * r["tm"][:, 3:6]
    assert np.allclose(foot[:, :2], rec["map"][:, :2], atol=1e-5)


def test_two_step_solve_recovers_xy_yaw():
    truth = _pose(np.array([3.0, -2.0, 1.5]), np.radians([0.0, 0.0, 20.0]))
    rec = scene(truth)
    rec["pose"] = _compose(truth, _pose(np.array([0.2, -0.15, 0.0]), np.radians([0, 0, 2.0])))
    est = oracle.SloamEstimator(twoStepOptim=1, treeMatchThresh=1.0)
    est.set_state(False, rec["prev_planes"])
    r = est.run(rec)
    assert r["ok"] and r["tree_step"] and r["ground_step"]
    assert np.all(r["delta"] == [0, 0, 0, 0, 0, 0, 1])  # T_Delta stays identity in two-step mode
    dt, dq = _pose_err(r["pose"], truth)
    assert dt < 1e-6 and dq < 1e-6, (dt, dq)


def test_first_scan_outputs():
    rec = scene(TRUTH)
    est = oracle.SloamEstimator()
    r = est.run(rec)
    assert r["first"] and r["ok"] and not r["early"]
    np.testing.assert_array_equal(r["pose"], TRUTH)
    assert np.all(r["matches"] == -1) and len(r["tm"]) == 10
    foot = r["tm"][:, :3] - (r["tm"][:, 2:3] / r["tm"][:, 5:6]) * r["tm"][:, 3:6]
    assert np.allclose(foot[:, :2], rec["map"][:, :2], atol=1e-5)
    # a second call is not a first scan; with an empty map it returns early
    rec2 = dict(rec)
    rec2["map"] = np.zeros((0, 7))
    r2 = est.run(rec2)
    assert not r2["first"] and not r2["ok"] and r2["early"]


def test_acceptance_rules():
    est = oracle.SloamEstimator()
    # one cell tilted by 0.2 rad, one above the sensor: both rejected
    r = est.run(scene(TRUTH, plane_tilt=0.2, overhead=True))
    assert len(r["planes"]) == 34
    # trunk 0 tilted 15 deg, trunk 1 too short, trunk 2 wider than the 0.3 m
    # cap, trunk 3 beyond maxLidarDist: each alone invalid
    for kw, bad in ((dict(tree_tilt_deg=15.0), 0), (dict(short=(1,)), 1), (dict(wide=(2,)), 2),
                    (dict(far=(3,)), 3)):
        for mode in (0, 2):
            e = oracle.SloamEstimator()
            r = e.run(scene(TRUTH, **kw), tree_mode=mode)
            want = np.ones(10, np.int32)
            want[bad] = 0
            np.testing.assert_array_equal(r["tree_valid"], want, err_msg=f"{kw} mode {mode}")
    # the radius cap is a parameter: 0.45 keeps the wide trunk
    e = oracle.SloamEstimator(maxTreeRadius=0.45)
    assert np.all(e.run(scene(TRUTH, wide=(2,)))["tree_valid"] == 1)
    # radius floor: a 0.12 m trunk gets defaultTreeRadius 0.2
    e = oracle.SloamEstimator()
    r = e.run(scene(TRUTH, radius=0.12), tree_mode=0)
    assert np.allclose(r["tm"][:, 6], 0.2)
