"""GPU tree instances (sloam_trellis) against the CPU oracle: the reference's
four organized tree clouds, synthetic forest scans (64 and 128 beams), beams
longer than the LDS stage, empty input and too-small capacities.  Integer
outputs and point lists are bit-exact; medians and radii are exact too (the
same float operations on both sides)."""
import numpy as np
import pytest

import oracle
from sloam_amd import synth
from sloam_amd.gpu import trellis
from tests_common import organized_cloud

pytestmark = pytest.mark.gpu

KEYS = ("tree_offsets", "vertex_int", "vertex_f64", "point_offsets", "points")


def _same(g, o):
    assert np.array_equal(g["labels"], o["labels"])
    assert g["n_labels"] == o["n_labels"]
    assert g["overflow"] == 0
    for k in KEYS:
        assert g[k].shape == o[k].shape, k
        assert np.array_equal(g[k], o[k]), k


@pytest.mark.parametrize("name", ["still_tree_t0", "still_tree_t1", "moving_tree_t0",
                                  "moving_tree_t1"])
@pytest.mark.parametrize("scale", [1.0, 0.5])
def test_reference_clouds(ctx, name, scale):
    cloud, H, W = organized_cloud(name)
    print(f"[trellis] {name} {H}x{W} radius_scale {scale}", flush=True)
    g = trellis(ctx, cloud, H, W, radius_scale=scale)
    o = oracle.trellis(cloud, H, W, radius_scale=scale)
    _same(g, o)
    assert len(g["tree_offsets"]) > 1


@pytest.mark.parametrize("H,seed", [(64, 0), (128, 4)])
def test_synthetic_scans(ctx, H, seed):
    scan = synth.make_scan(seed=seed, H=H, W=2048)
    tree = oracle.project_partition(scan["xyzi"], scan["label"], H=H, W=2048)["tree"]
    print(f"[trellis] synthetic scan H={H} seed={seed}", flush=True)
    g = trellis(ctx, tree, H, 2048)
    o = oracle.trellis(tree, H, 2048)
    _same(g, o)
    assert len(g["tree_offsets"]) - 1 >= 5


def test_long_beams_and_ties(ctx):
    """Rows of 2000 connected points (beams longer than the 1024-point LDS
    stage), repeated z values (tie order), and isolated pixels."""
    H, W = 24, 2048
    rng = np.random.default_rng(7)
    cloud = np.full((H * W, 4), np.nan, np.float32)
    for r in range(H):
        c = np.arange(20, 2020)
        x = 5.0 + 0.001 * c
        y = 0.05 * np.sin(c / 50.0) + 0.0002 * r
        z = np.round(rng.normal(0.1 * r, 0.05, len(c)), 2)  # many equal z
        idx = r * W + c
        cloud[idx, 0], cloud[idx, 1], cloud[idx, 2] = x, y, z
        cloud[idx, 3] = r
    iso = rng.choice(H * W, 300, replace=False)
    cloud[iso, :3] = rng.uniform(20, 30, (300, 3))
    g = trellis(ctx, cloud, H, W, max_dist_to_centroid=0.5)
    o = oracle.trellis(cloud, H, W, max_dist=0.5)
    _same(g, o)


@pytest.mark.parametrize("length", [161, 300, 512])
def test_medians_of_long_beams_with_ties(ctx, length):
    """Beams of 161-512 points staged in LDS: coordinates quantised to 1 cm so
    that many values tie (index order decides the median), y crossing zero
    and the lowest beam's z around zero, with both signs of zero among the
    ties (equal under <)."""
    H, W = 24, 2048  # more than min_vertices (16) beams make a tree
    rng = np.random.default_rng(length)
    cloud = np.full((H * W, 4), np.nan, np.float32)
    for r in range(H):
        c = np.arange(100, 100 + length)
        x = np.round(6.0 + 0.0005 * c + rng.normal(0.0, 0.005, len(c)), 2)
        y = np.round(0.01 * (c - c.mean()) / 10.0, 2)
        z = np.round(rng.normal(0.1 * r, 0.02, len(c)), 2)
        y[::7] = -0.0
        if r == 0:
            z[::5] = -0.0
        idx = r * W + c
        cloud[idx, 0], cloud[idx, 1], cloud[idx, 2] = x, y, z
        cloud[idx, 3] = r
    print(f"[trellis] {H} beams of {length} points", flush=True)
    g = trellis(ctx, cloud, H, W, max_dist_to_centroid=0.5)
    o = oracle.trellis(cloud, H, W, max_dist=0.5)
    _same(g, o)
    assert len(g["tree_offsets"]) > 1


def test_empty_and_overflow(ctx):
    H, W = 8, 64
    cloud = np.full((H * W, 4), np.nan, np.float32)
    g = trellis(ctx, cloud, H, W)
    assert g["n_labels"] == 0 and len(g["tree_offsets"]) == 1 and g["overflow"] == 0
    # too-small capacities: flagged, no writes past them
    cloud, H, W = organized_cloud("still_tree_t0")
    import torch
    from sloam_amd.gpu import trellis_dev
    dev = torch.device("cuda", 0)
    x = torch.as_tensor(cloud).to(dev)
    out = {"labels": torch.empty(H * W, dtype=torch.int32, device=dev),
           "tree_offsets": torch.full((3,), -7, dtype=torch.int32, device=dev),
           "vertex_int": torch.empty((40, 4), dtype=torch.int32, device=dev),
           "vertex_f64": torch.empty((40, 4), dtype=torch.float64, device=dev),
           "point_offsets": torch.empty((41,), dtype=torch.int32, device=dev),
           "points": torch.empty((H * W, 4), dtype=torch.float32, device=dev),
           "counts": torch.empty((8,), dtype=torch.int32, device=dev)}
    # the wrapper passes capacities from the sizes it would allocate; call
    # through the C ABI with the small ones instead
    import ctypes as C
    from sloam_amd._lib import TrellisCfg, check
    from sloam_amd.gpu import TRELLIS_DEFAULTS, _ptr
    cfg = TrellisCfg(**TRELLIS_DEFAULTS)
    torch.cuda.synchronize()
    check(ctx.lib.sloam_trellis(ctx.handle, _ptr(x), H, W, C.byref(cfg), _ptr(out["labels"]),
                                2, 40, H * W, _ptr(out["tree_offsets"]), _ptr(out["vertex_int"]),
                                _ptr(out["vertex_f64"]), _ptr(out["point_offsets"]),
                                _ptr(out["points"]), _ptr(out["counts"])), ctx.handle)
    ctx.sync()
    cnt = out["counts"].cpu().numpy()
    o = oracle.trellis(cloud, H, W)
    assert cnt[4] == 1 and cnt[0] == len(o["tree_offsets"]) - 1
    # the first two trees' offsets are written, the sentinel after them is not
    t = out["tree_offsets"].cpu().numpy()
    assert np.array_equal(t[:2], o["tree_offsets"][:2]) and t[2] == -7
