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
