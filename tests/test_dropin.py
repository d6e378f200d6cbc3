"""The C++ drop-in classes (Plane, Cylinder, sloam::sloam in
sloam_amd/host/, linked into libsloam_gpu.so) through the test program
tests/cpp/dropin_test: the cases of the reference's gtests
(sloam/src/tests/plane_test.cpp, cylinder_test.cpp, core_test.cpp) on the
reference's own test data (tests/golden/still_*), and the models the drop-in
computes for those scans against the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "dropin_test")

# core_test.cpp:93-117 (the parameters the reference's SLOAMTest sets)
CORE = dict(radii_bins=1, theta_bins=18, min_dist=0.0, max_dist=30.0, retain_thresh=0.05,
            num_features=60)


def _binary():
    subprocess.run(["make", "-C", CPP, "-s"], check=True)
    assert os.path.exists(BIN)
    return BIN


def _landmarks(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    to, vo = d["tree_offsets"], d["vertex_offsets"]
    vi, vf, pts = d["vertex_int"], d["vertex_radius_coords"], d["points"]
    trees = []
    for t in range(len(to) - 1):
        verts = []
        for v in range(to[t], to[t + 1]):
            verts.append({"int": vi[v], "radius": float(vf[v, 0]),
                          "coords": vf[v, 1:4].astype(np.float32),
                          "points": pts[vo[v]:vo[v + 1]]})
        trees.append(verts)
    return trees


def _write_fixtures(d):
    for stamp in ("t0", "t1"):
        trees = _landmarks(f"still_landmarks_{stamp}")
        with open(os.path.join(d, f"still_landmarks_{stamp}.lm"), "wb") as f:
            f.write(struct.pack("<i", len(trees)))
            for verts in trees:
                f.write(struct.pack("<i", len(verts)))
                for v in verts:
                    f.write(np.asarray(v["int"], "<i4").tobytes())
                    f.write(struct.pack("<d", v["radius"]))
                    f.write(np.asarray(v["coords"], "<f4").tobytes())
                    f.write(struct.pack("<i", len(v["points"])))
                    f.write(np.ascontiguousarray(v["points"], "<f4").tobytes())
        tc = np.load(os.path.join(GOLDEN, f"still_tree_{stamp}.npz"))
        W, H = int(tc["width"]), int(tc["height"])
        org = np.full((H * W, 3), np.nan, np.float32)
        org[tc["index"]] = tc["xyz"]
        with open(os.path.join(d, f"still_tree_{stamp}.org"), "wb") as f:
            f.write(struct.pack("<ii", H, W))
            f.write(org.astype("<f4").tobytes())
        xyz = np.load(os.path.join(GOLDEN, f"still_ground_{stamp}.npz"))["xyz"]
        with open(os.path.join(d, f"still_ground_{stamp}.xyz"), "wb") as f:
            f.write(struct.pack("<i", len(xyz)))
            f.write(np.ascontiguousarray(xyz, "<f4").tobytes())


def _run(args, timeout=120):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    print(r.stderr)
    return r


def test_host_helpers(tmp_path):
    """Rotation helpers of the drop-in types (Euler angles, FromTwoVectors,
    angle-axis) against their defining identities: no device calls."""
    r = _run([_binary(), str(tmp_path), "host"])
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_reference_gtests(tmp_path):
    _write_fixtures(str(tmp_path))
    r = _run([_binary(), str(tmp_path), "gpu"])
    assert r.returncode == 0, r.stderr
    assert "0 failures" in r.stdout


def _parse_dump(path):
    runs, cur = [], None
    solve = {"tmatch": [], "gmatch": []}
    for line in open(path):
        w = line.split()
        if w[0] in ("tmatch", "gmatch"):
            solve[w[0]].append([float(x) for x in w[1:]])
        elif w[0] == "joint":
            solve["joint_ok"] = int(w[1])
            solve["joint"] = np.array([float(x) for x in w[2:]])
        elif w[0] in ("xyyaw", "zrollpitch"):
            solve[w[0]] = np.array([float(x) for x in w[1:]])
        elif w[0] == "scan":
            cur = {"twostep": int(w[3]), "planes": [], "trees": []}
            runs.append(cur)
        elif w[0] == "plane":
            cur["planes"].append([float(x) for x in w[1:]])
        elif w[0] == "tree":
            cur["trees"].append((int(w[1]), [float(x) for x in w[2:]]))
        elif w[0] == "pose":
            cur["ok"] = int(w[2])
            cur["pose"] = [float(x) for x in w[3:]]
        elif w[0] == "matches":
            cur["matches"] = [int(x) for x in w[1:]]
    return runs, solve


@pytest.mark.gpu
def test_models_against_oracle(tmp_path):
    """computeModels of scan t0 with the SLOAMTest parameters: every accepted
    ground plane is the oracle's fit of its cell (centroid bit-exact, plane
    within 1e-9), and every tree
    axis is the oracle's PCL-semantics line fit of the tree's vertex medians
    (bit-exact, float32 coefficients)."""
    _write_fixtures(str(tmp_path))
    out = os.path.join(str(tmp_path), "dump.txt")
    r = _run([_binary(), str(tmp_path), "dump", out])
    assert r.returncode == 0, r.stderr
    runs, _ = _parse_dump(out)
    assert len(runs) == 2
    xyz = np.load(os.path.join(GOLDEN, "still_ground_t0.npz"))["xyz"]
    pts = np.zeros((len(xyz), 4), np.float32)
    pts[:, :3] = xyz
    o = oracle.ground_planes(pts, **CORE)
    cells = [(o["model"][c, :4], o["model"][c, 4:7]) for c in range(len(o["info"]))
             if o["info"][c, 0]]
    trees = _landmarks("still_landmarks_t0")
    for run in runs:
        assert len(run["planes"]) > 0 and len(run["trees"]) > 0
        for p in run["planes"]:
            # float centroid summed in the same order: identical; normal and
            # offset: FP64 sums in another order
            hit = [c for c in cells if np.array_equal(c[1], p[4:7])]
            assert len(hit) == 1
            assert np.allclose(hit[0][0], p[:4], atol=1e-9, rtol=0)
        by_id = {int(t[2]["int"][0]): t for t in trees}
        for tid, m in run["trees"]:
            t = by_id[tid]
            med = np.array([v["coords"] for v in t], np.float32)
            ref = oracle.line_ransac(med)
            assert np.array_equal(np.asarray(m[3:6]), ref["coeff"][3:6].astype(np.float64))
        assert run["ok"] == 1
        assert any(m != -1 for m in run["matches"])
        if run["twostep"]:
            # the still sequence: the pose stays near the origin (the joint
            # solve is compared with the oracle's on its own matches below)
            assert np.linalg.norm(run["pose"][:3]) < 0.1


def _pose_problem(solve):
    """The matches as one oracle pose problem: consecutive matches against the
    same landmark share one table entry, as in the drop-in."""
    pts, cyl, planes = [], [], []
    for m in solve["tmatch"]:
        if not cyl or cyl[-1][:7] != m[3:10]:
            cyl.append(m[3:10] + [0.0])
        pts.append((m[:3], len(cyl) - 1))
    for m in solve["gmatch"]:
        if not planes or planes[-1] != m[3:7]:
            planes.append(m[3:7])
        pts.append((m[:3], -len(planes)))
    P = np.zeros((len(pts), 4), np.float32)
    for i, (f, lid) in enumerate(pts):
        P[i, :3] = f
        P[i, 3] = np.array([lid], np.int32).view(np.float32)[0]
    return P, np.array(cyl, np.float64).reshape(-1, 8), np.array(planes, np.float64).reshape(-1, 4)


@pytest.mark.gpu
def test_pose_solves_against_oracle(tmp_path):
    """OptimizePose, OptimizeXYYaw and OptimizeZRollPitch of the drop-in on the
    matches of scan t1 against the map of scan t0 (the SLOAMTest data) against
    the FP64 oracle's LM on the same residuals: within 1e-4 (m, and quaternion
    / angle-axis components)."""
    _write_fixtures(str(tmp_path))
    out = os.path.join(str(tmp_path), "dump.txt")
    r = _run([_binary(), str(tmp_path), "dump", out])
    assert r.returncode == 0, r.stderr
    _, solve = _parse_dump(out)
    assert len(solve["tmatch"]) > 0 and len(solve["gmatch"]) > 0
    P, cyl, planes = _pose_problem(solve)
    nt = len(solve["tmatch"])
    ident = np.array([0, 0, 0, 0, 0, 0, 1.0])
    # joint: all residuals, quaternion parametrisation
    o = oracle.pose_lm_batch({"points": P, "res_offsets": np.array([0, len(P)]), "cyl": cyl,
                              "planes": planes, "pose_init": ident[None]},
                             max_iters=50, stop_on_convergence=True)
    assert solve["joint_ok"] == int(o["status"][0, 0] == 0)
    if solve["joint_ok"]:
        q = o["pose"][0, 3:]
        g = solve["joint"]
        sign = 1.0 if np.dot(q, g[3:]) >= 0 else -1.0
        assert np.abs(g[:3] - o["pose"][0, :3]).max() < 1e-4
        assert np.abs(sign * g[3:] - q).max() < 1e-4
    # two-step: trees with z, roll, pitch fixed; ground with x, y, yaw fixed
    two = {"points": np.concatenate([P[:nt], P[nt:]]),
           "res_offsets": np.array([0, nt, len(P)]), "cyl": cyl, "planes": planes,
           "pose_init": np.stack([ident, ident]),
           "free_mask": np.array([0b100011, 0b011100], np.int32)}
    o = oracle.pose_lm_batch(two, max_iters=50, stop_on_convergence=True, parametrization=1)
    conv = o["status"][:, 0] == 0
    xy = o["params"][0, [0, 1, 5]] if conv[0] else np.zeros(3)
    zrp = o["params"][1, [2, 3, 4]] if conv[1] else np.zeros(3)
    assert np.abs(solve["xyyaw"] - xy).max() < 1e-4
    assert np.abs(solve["zrollpitch"] - zrp).max() < 1e-4


def _core_record(stamp):
    """SLOAMTest's input of scan `stamp` (core_test.cpp:78-117) as a record of
    the oracle's RunSloam: the ground cells fitted by the oracle, the
    fixture's trees with 5 points per vertex (zero-padded, as core_test
    pads them)."""
    xyz = np.load(os.path.join(GOLDEN, f"still_ground_{stamp}.npz"))["xyz"]
    pts = np.zeros((len(xyz), 4), np.float32)
    pts[:, :3] = xyz
    g = oracle.ground_planes(pts, **CORE)
    trees = _landmarks(f"still_landmarks_{stamp}")
    tree_off, vtx_int, vtx_f64, vtx_pt_off, vp = [0], [], [], [0], []
    for verts in trees:
        for v in verts:
            vtx_int.append((int(v["int"][0]), int(v["int"][3])))
            vtx_f64.append((v["radius"], *v["coords"].astype(np.float64)))
            p = np.zeros((5, 4), np.float32)
            k = min(5, len(v["points"]))
            p[:k, :3] = v["points"][:k]
            vp.append(p)
            vtx_pt_off.append(vtx_pt_off[-1] + 5)
        tree_off.append(tree_off[-1] + len(verts))
    return {"pose": np.array([0, 0, 0, 0, 0, 0, 1.0]), "nf": CORE["num_features"],
            "cell_model": g["model"][:, :7], "cell_valid": g["info"][:, 0].copy(),
            "cell_feat": g["features"], "tree_off": np.array(tree_off, np.int32),
            "vtx_int": np.array(vtx_int, np.int32), "vtx_f64": np.array(vtx_f64),
            "vtx_pt_off": np.array(vtx_pt_off, np.int32), "pts": np.concatenate(vp),
            "line_int": np.zeros((0, 2), np.int32), "line_coeff": np.zeros((0, 6), np.float32),
            "cyl_ok": np.zeros(0, np.int32), "cyl_fit": np.zeros((0, 7)),
            "map": np.zeros((0, 7))}


# core_params() of tests/cpp/dropin_test.cpp (core_test.cpp:93-117), in the
# oracle estimator's names
CORE_EST = dict(maxLidarDist=15.0, maxTreeRadius=0.3, maxAxisTheta=10.0, defaultTreeRadius=0.1,
                featuresPerTree=2, numGroundFeatures=60, groundRadiiBins=1, groundThetaBins=18,
                treeMatchThresh=1.0, AddNewTreeThreshDist=2.0)


@pytest.mark.gpu
def test_run_sloam_against_oracle_estimator(tmp_path):
    """SLOAMTest's two scans through the drop-in (computeModels of t0, then
    RunSloam t0 -> t1) against the CPU restatement of RunSloam on the same
    data, in both pose modes: the same accepted planes (within 1e-9), the same
    trees (ids; axis and radius within 1e-9), the same map matches and the
    pose within 1e-4."""
    _write_fixtures(str(tmp_path))
    out = os.path.join(str(tmp_path), "dump.txt")
    r = _run([_binary(), str(tmp_path), "dump", out])
    assert r.returncode == 0, r.stderr
    runs, _ = _parse_dump(out)
    t0, t1 = _core_record("t0"), _core_record("t1")
    for run in runs:
        est = oracle.SloamEstimator(twoStepOptim=run["twostep"], **CORE_EST)
        o0 = est.run(t0, tree_mode=0)
        assert o0["first"] and o0["ok"]
        gp = np.array(run["planes"])
        assert gp.shape == o0["planes"].shape
        np.testing.assert_allclose(o0["planes"], gp, rtol=0, atol=1e-9)
        assert [t[0] for t in run["trees"]] == list(o0["tm_id"])
        np.testing.assert_allclose(o0["tm"], np.array([t[1] for t in run["trees"]]), rtol=0,
                                   atol=1e-9)
        t1m = dict(t1)
        t1m["map"] = o0["tm"]
        o1 = est.run(t1m, tree_mode=0)
        print(f"[dropin] two-step {run['twostep']}: oracle pose {o1['pose']}, "
              f"drop-in {run['pose']}", flush=True)
        assert int(o1["ok"]) == run["ok"]
        assert list(o1["matches"]) == run["matches"]
        g = np.array(run["pose"])
        q = o1["pose"][3:]
        sign = 1.0 if np.dot(q, g[3:]) >= 0 else -1.0
        assert np.abs(g[:3] - o1["pose"][:3]).max() < 1e-4
        assert np.abs(sign * g[3:] - q).max() < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("destagger", [False, True])
def test_segmentation_dropin_against_oracle(tmp_path, destagger):
    """seg::Segmentation (run with given labels, then maskCloud as SLOAMNode
    calls it: ground with value 1, the dense tree cloud with 255) on a 64 x
    2048 scan against the oracle's projection and partition, bit for bit
    (the scan's beams point at pixel centres, far from the 1e-3 px band in
    which the device's atan2f / asinf may pick the neighbouring pixel); a
    third selection (value 0) against the
    oracle's projection and mask; and the same two clouds again from a mask
    edited in an unused value, which maskCloud selects from its kept
    projection instead of taking the GPU partition."""
    from sloam_amd import synth
    s = synth.make_scan(seed=9, H=64, W=2048)
    xyzi = np.ascontiguousarray(s["xyzi"], np.float32)
    lab = np.ascontiguousarray(s["label"], np.uint8)
    n = len(xyzi)
    with open(os.path.join(str(tmp_path), "scan.bin"), "wb") as f:
        f.write(struct.pack("<iiii", 64, 2048, n, int(destagger)))
        f.write(xyzi.tobytes())
        f.write(lab.tobytes())
    out = os.path.join(str(tmp_path), "seg.bin")
    r = _run([_binary(), str(tmp_path), "seg", out])
    assert r.returncode == 0, r.stderr
    buf = open(out, "rb").read()
    mask = np.frombuffer(buf, np.uint8, 64 * 2048).reshape(64, 2048)
    off = 64 * 2048
    clouds = []
    for _ in range(5):
        (cnt,) = struct.unpack_from("<i", buf, off)
        off += 4
        clouds.append(np.frombuffer(buf, np.float32, 4 * cnt, off).reshape(cnt, 4))
        off += 16 * cnt
    ground, tree, other, ground2, tree2 = clouds
    o = oracle.project_partition(xyzi, lab, H=64, W=2048, destagger=destagger)
    np.testing.assert_array_equal(mask, o["mask"].reshape(64, 2048))
    np.testing.assert_array_equal(ground, o["ground"])
    np.testing.assert_array_equal(tree, o["tree"].reshape(-1, 4))
    # value 0: points whose pixel is other / invalid, input order; points
    # without a pixel (non-finite) count as 0
    px = o["proj"]
    m = np.where(px[:, 0] >= 0, o["mask"].reshape(-1)[np.maximum(px[:, 1], 0) * 2048 +
                                                       np.maximum(px[:, 0], 0)], 0)
    np.testing.assert_array_equal(other, xyzi[m == 0])
    np.testing.assert_array_equal(ground2, ground)
    np.testing.assert_array_equal(tree2, tree)
