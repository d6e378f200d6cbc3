"""CPU tests that pin the tree-instance oracle (organized connected components
and per-beam vertices, sloam/src/segmentation/trellis.cpp:15-132): the
reference's own tree clouds and the landmark archives made from them
(sloam/src/tests/aux/{still,moving}_tree_t{0,1}.pcd ->
{still,moving}_landmarks_t{0,1}), a breadth-first flood fill for the
components, and the vertex definitions checked point by point."""
import numpy as np
import pytest

import oracle
from tests_common import organized_cloud, load_landmarks

PAIRS = [("still", "t0"), ("still", "t1"), ("moving", "t0"), ("moving", "t1")]


@pytest.mark.parametrize("pre,t", PAIRS)
def test_reference_fixture_pairs(pre, t):
    """The archives store the older radius rule (half the first-last
    distance), so radius_scale = 0.5; both text formats print floats with
    about 8 significant digits, so coordinates agree to one float ulp."""
    cloud, H, W = organized_cloud(f"{pre}_tree_{t}")
    o = oracle.trellis(cloud, H, W, radius_scale=0.5)
    f = load_landmarks(f"{pre}_landmarks_{t}")
    assert np.array_equal(o["tree_offsets"], f["tree_offsets"])
    assert np.array_equal(o["vertex_int"][:, 0], f["vertex_int"][:, 0])  # treeId
    assert np.array_equal(o["point_offsets"], f["vertex_offsets"])
    ulp = np.spacing(np.abs(f["points"]).astype(np.float32))
    assert np.all(np.abs(o["points"][:, :3] - f["points"]) <= ulp)
    c = f["vertex_radius_coords"][:, 1:].astype(np.float32)
    assert np.all(np.abs(o["vertex_f64"][:, 1:] - c) <= np.spacing(np.abs(c)))
    r = f["vertex_radius_coords"][:, 0]
    assert np.allclose(o["vertex_f64"][:, 0], r, rtol=0, atol=1e-6)


def _dist(a, b):
    d = (a[:3] - b[:3]).astype(np.float32)
    return np.sqrt(np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))


def _flood_fill(cloud, H, W, thr):
    valid = np.isfinite(cloud[:, 0])
    lab = -np.ones(H * W, np.int64)
    n = 0
    for p in range(H * W):
        if not valid[p] or lab[p] >= 0:
            continue
        lab[p] = n
        stack = [p]
        while stack:
            q = stack.pop()
            r, c = divmod(q, W)
            for rr, cc in ((r, c - 1), (r, c + 1), (r - 1, c), (r + 1, c)):
                if 0 <= rr < H and 0 <= cc < W:
                    t = rr * W + cc
                    if valid[t] and lab[t] < 0 and _dist(cloud[q], cloud[t]) < thr:
                        lab[t] = n
                        stack.append(t)
        n += 1
    return lab, n


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_components_flood_fill(seed):
    rng = np.random.default_rng(seed)
    H, W = 6, 40
    cloud = np.zeros((H * W, 4), np.float32)
    cloud[:, :3] = rng.uniform(0, 2.5, (H * W, 3))
    cloud[rng.random(H * W) < 0.25, :3] = np.nan
    labels, n = oracle.trellis_clusters(cloud, H, W, 1.0)
    ref, nref = _flood_fill(cloud, H, W, np.float32(1.0))
    assert n == nref
    assert np.array_equal(labels, ref)


def test_vertex_definition_point_by_point():
    """Every vertex of a reference cloud: median = element floor(n/2) of each
    sorted axis over the beam's points of the tree, kept points = those within
    0.2 m in (z, y, x, column) order, radius = first-to-last distance."""
    cloud, H, W = organized_cloud("moving_tree_t0")
    o = oracle.trellis(cloud, H, W)
    lab = o["labels"]
    for v in range(len(o["vertex_int"])):
        tid, row, m, _ = o["vertex_int"][v]
        cols = [c for c in range(W) if lab[row * W + c] == tid]
        P = cloud[[row * W + c for c in cols]]
        med = np.array([np.sort(P[:, k])[len(P) // 2] for k in range(3)], np.float32)
        assert np.array_equal(o["vertex_f64"][v, 1:].astype(np.float32), med)
        keep = [i for i in range(len(P)) if _dist(P[i], med) < np.float32(0.2)]
        keep.sort(key=lambda i: (P[i, 2], P[i, 1], P[i, 0], cols[i]))
        pts = o["points"][o["point_offsets"][v]:o["point_offsets"][v + 1]]
        assert np.array_equal(pts, P[keep])
        assert o["vertex_f64"][v, 0] == float(_dist(P[keep[0]], P[keep[-1]]))
    sizes = np.diff(o["tree_offsets"])
    assert np.all((sizes > 16) & (sizes <= 56))
