"""Keyframe front end (sloam_frontend_*) replayed as a HIP graph: every
replay, on a different scan and with eager front-end work between replays,
equals the eager result bit for bit, and the eager result equals the oracle
(projection / partition, tree instances, ground cells)."""
import numpy as np
import pytest

import oracle
from sloam_amd import synth
from sloam_amd.gpu import Frontend

pytestmark = pytest.mark.gpu

H, W = 64, 2048
KEYS = ("mask", "ground", "tree", "trellis_counts", "tree_offsets", "vertex_int", "vertex_f64",
        "point_offsets", "points", "cell_info", "cell_model", "cell_features")


def _scans(k):
    forest = synth.make_forest(seed=21, n_trees=120, length=60.0, width=40.0)
    out = []
    for i in range(k):
        R, t = synth.trajectory_pose(10.0 + 3.0 * i)
        out.append(synth.scan_from(forest, R, t, seed=100 + i, H=H, W=W))
    return out


def _same(a, b):
    for k in KEYS:
        assert a[k].shape == b[k].shape, k
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def test_graph_replays_match_eager(ctx):
    scans = _scans(4)
    g = Frontend(ctx, H, W, use_graph=True)
    e = Frontend(ctx, H, W, use_graph=False)
    for rep in range(2):
        for i, sc in enumerate(scans):
            print(f"[frontend] pass {rep} scan {i}", flush=True)
            g.upload(sc["xyzi"], sc["label"])
            g.exec()
            rg = g.download()
            e.upload(sc["xyzi"], sc["label"])
            e.exec()
            re_ = e.download()
            _same(rg, re_)
            assert rg["trellis_counts"][0] > 0
    g.close()
    e.close()


def test_eager_matches_oracle(ctx):
    sc = _scans(1)[0]
    f = Frontend(ctx, H, W, use_graph=False)
    f.upload(sc["xyzi"], sc["label"])
    f.exec()
    r = f.download()
    pp = oracle.project_partition(sc["xyzi"], sc["label"], H=H, W=W)
    assert np.array_equal(r["mask"], pp["mask"].reshape(-1))
    assert np.array_equal(r["ground"], pp["ground"])
    assert np.array_equal(r["tree"], pp["tree"], equal_nan=True)
    t = oracle.trellis(pp["tree"], H, W)
    for k in ("tree_offsets", "vertex_int", "vertex_f64", "point_offsets", "points"):
        assert np.array_equal(r[k], t[k]), k
    gp = oracle.ground_planes(pp["ground"], radii_bins=2, theta_bins=18, min_dist=5.0,
                              max_dist=25.0, retain_thresh=0.05, num_features=5)
    assert np.array_equal(r["cell_info"], gp["info"])
    valid = gp["info"][:, 0] == 1
    assert np.array_equal(r["cell_model"][valid, 4:7], gp["model"][valid, 4:7])
    assert np.allclose(r["cell_model"][valid, :4], gp["model"][valid, :4], atol=1e-9, rtol=0)
    f.close()


def test_host_results_match_device_buffers(ctx):
    """With host_results the front end's last kernel writes the used part of
    the tree and ground-cell outputs into pinned host memory: after every
    graph replay the host view equals the device buffers."""
    scans = _scans(3)
    g = Frontend(ctx, H, W, use_graph=True, host_results=True)
    for rep in range(2):
        for i, sc in enumerate(scans):
            print(f"[frontend] host results pass {rep} scan {i}", flush=True)
            g.upload(sc["xyzi"], sc["label"])
            g.exec()
            hv = g.host_view()
            rd = g.download()
            for k, v in hv.items():
                assert v.shape == rd[k].shape, k
                assert np.array_equal(v, rd[k], equal_nan=True), k
    g.close()


def test_device_line_fits_match_fit_lines(ctx):
    """With fit_lines every tree's axis line is fitted inside the front end:
    per tree it equals sloam_fit_lines on that tree's vertex medians (as float),
    bit for bit, on every graph replay, and the host view carries the same."""
    from sloam_amd.gpu import fit_lines
    scans = _scans(3)
    g = Frontend(ctx, H, W, use_graph=True, host_results=True, fit_lines=True)
    for rep in range(2):
        for i, sc in enumerate(scans):
            print(f"[frontend] line fits pass {rep} scan {i}", flush=True)
            g.upload(sc["xyzi"], sc["label"])
            g.exec()
            hv = g.host_view()
            r = g.download()
            nt = int(r["trellis_counts"][0])
            assert nt > 0
            med = r["vertex_f64"][:, 1:4].astype(np.float32)
            ref = fit_lines(ctx, med, r["tree_offsets"])
            assert np.array_equal(r["line_coeff"], ref["coeff"])
            assert np.array_equal(r["line_stats"], ref["stats"])
            assert np.array_equal(hv["line_coeff"], r["line_coeff"])
            assert np.array_equal(hv["line_stats"], r["line_stats"])
    g.close()
