"""GPU ground planes (sloam_ground_planes) and batched plane RANSAC
(sloam_plane_plan_*) against the CPU oracle, on the reference's own ground
clouds (sloam/src/tests/aux/still_ground_t{0,1}.pcd) with the parameter sets of
core_test.cpp:94-119 and params/sloam.yaml."""
import os

import numpy as np
import pytest

import oracle
from sloam_amd.gpu import fit_planes, ground_planes

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

PARAMS = {
    # core_test.cpp: 1 x 18 bins, 0..30 m, retain 0.05, 60 features
    "core_test": dict(radii_bins=1, theta_bins=18, min_dist=0.0, max_dist=30.0,
                      retain_thresh=0.05, num_features=60),
    # params/sloam.yaml: 2 x 18 bins, 5..25 m, retain 0.05, 5 features
    "sloam_yaml": dict(radii_bins=2, theta_bins=18, min_dist=5.0, max_dist=25.0,
                       retain_thresh=0.05, num_features=5),
    # more cells than a workgroup has threads: the cells' offsets come from a
    # prefix kernel instead of each fit workgroup's own sum
    "many_cells": dict(radii_bins=16, theta_bins=36, min_dist=0.0, max_dist=30.0,
                       retain_thresh=0.05, num_features=5),
}


def _pts(name):
    xyz = np.load(os.path.join(GOLDEN, f"{name}.npz"))["xyz"]
    p = np.zeros((len(xyz), 4), np.float32)
    p[:, :3] = xyz
    return p


@pytest.mark.parametrize("fixture", ["still_ground_t0", "still_ground_t1"])
@pytest.mark.parametrize("pname", list(PARAMS))
def test_ground_planes_reference_fixtures(ctx, fixture, pname):
    pts = _pts(fixture)
    prm = PARAMS[pname]
    print(f"[ground] {fixture} ({len(pts)} points) with {pname}", flush=True)
    g = ground_planes(ctx, pts, **prm)
    o = oracle.ground_planes(pts, **prm)
    assert np.array_equal(g["info"], o["info"])
    valid = o["info"][:, 0] == 1
    assert valid.sum() > 0
    # float centroid summed in the same order: identical
    assert np.array_equal(g["model"][valid, 4:7], o["model"][valid, 4:7])
    # normal and offset: FP64 sums in a different order
    assert np.allclose(g["model"][valid, :4], o["model"][valid, :4], atol=1e-9, rtol=0)
    nf = prm["num_features"]
    assert np.array_equal(g["features"][valid, :nf], o["features"][valid, :nf])


def test_plane_ransac_bit_exact(ctx):
    rng = np.random.default_rng(3)
    sets = []
    for s in range(48):
        n = rng.normal(size=3) * np.array([0.05, 0.05, 1.0])
        n /= np.linalg.norm(n)
        m = int(rng.integers(30, 4000))
        xy = rng.uniform(-8, 8, (m, 2))
        z = -(n[0] * xy[:, 0] + n[1] * xy[:, 1] + rng.uniform(1, 2)) / n[2]
        P = np.c_[xy, z] + rng.normal(0, 0.02, (m, 3))
        k = int(0.3 * m)
        P[:k, 2] += rng.uniform(-2, 2, k)
        sets.append(P)
    sets.append(np.zeros((0, 3)))
    sets.append(np.ones((2, 3)))
    xyz = np.concatenate(sets).astype(np.float32)
    off = np.cumsum([0] + [len(p) for p in sets]).astype(np.int32)
    g = fit_planes(ctx, xyz, off, n_hyp=200, thr=0.05, counts=True, samples=True, mask=True,
                   best=True)
    o = oracle.plane_fit_batch(xyz, off, 200, thr=0.05, samples=True, threads=8)
    for k in ("samples", "counts", "best", "mask"):
        assert np.array_equal(g[k], o[k]), k
    assert np.allclose(g["model"][:, :7], o["model"][:, :7], atol=1e-9, rtol=0)
    assert np.all(np.abs(g["model"][:, 7] - o["model"][:, 7]) <= 1)


@pytest.mark.parametrize("n,bins", [(3000, (1, 1)), (30000, (1, 4)), (40000, (1, 1))],
                         ids=["3000-one-cell", "30000-four-cells", "40000-one-cell"])
def test_ground_planes_bottom_k_order(ctx, n, bins):
    """The bottom-k order by (z, input order) of the (cell, z) radix sort.  z
    is quantised to 1 cm so that many points tie, -0 and +0 among them (equal
    under <, so input order decides); the kept points and the features must
    come out in the oracle's order."""
    rng = np.random.default_rng(n)
    r = 25.0 * np.sqrt(rng.uniform(0.0, 1.0, n))
    th = rng.uniform(-np.pi, np.pi, n)
    pts = np.zeros((n, 4), np.float32)
    pts[:, 0] = r * np.cos(th)
    pts[:, 1] = r * np.sin(th)
    pts[:, 2] = np.round(0.02 * pts[:, 0] + rng.normal(0.0, 0.05, n), 2)
    prm = dict(radii_bins=bins[0], theta_bins=bins[1], min_dist=0.0, max_dist=30.0,
               retain_thresh=0.05, num_features=60)
    g = ground_planes(ctx, pts, **prm)
    o = oracle.ground_planes(pts, **prm)
    print(f"[ground] {n} points in {bins[0]} x {bins[1]} cells: kept "
          f"{o['info'][:, 1].tolist()}", flush=True)
    assert np.array_equal(g["info"], o["info"])
    valid = o["info"][:, 0] == 1
    assert valid.all()
    assert np.array_equal(g["model"][valid, 4:7], o["model"][valid, 4:7])
    assert np.allclose(g["model"][valid, :4], o["model"][valid, :4], atol=1e-9, rtol=0)
    assert np.array_equal(g["features"][valid, :60], o["features"][valid, :60])
