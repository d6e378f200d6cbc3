"""GPU batched line RANSAC (sloam_fit_lines) against the CPU oracle: bit-exact
coefficients, statistics and inliers on the reference's landmark fixtures and
on synthetic vertex sets with outliers and degenerate samples."""
import os

import numpy as np
import pytest

import oracle
from sloam_amd.gpu import fit_lines

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sets():
    sets = []
    for name in ("still_landmarks_t0", "still_landmarks_t1", "moving_landmarks_t0",
                 "moving_landmarks_t1"):
        f = np.load(os.path.join(GOLDEN, f"{name}.npz"))
        to = f["tree_offsets"]
        coords = f["vertex_radius_coords"][:, 1:4].astype(np.float32)
        sets += [coords[to[t]:to[t + 1]] for t in range(len(to) - 1)]
    rng = np.random.default_rng(5)
    for k in range(40):
        m = int(rng.integers(2, 300))
        t = rng.uniform(0, 6, m)
        d = np.array([rng.normal(0, 0.05), rng.normal(0, 0.05), 1.0])
        d /= np.linalg.norm(d)
        P = rng.uniform(-20, 20, 3) + t[:, None] * d + rng.normal(0, 0.05, (m, 3))
        n_out = int(0.3 * m)
        P[:n_out] += rng.uniform(-2, 2, (n_out, 3))
        sets.append(P.astype(np.float32))
    sets.append(np.zeros((0, 3), np.float32))
    sets.append(np.ones((1, 3), np.float32))
    sets.append(np.ones((5, 3), np.float32))  # no good sample: all equal
    q = np.zeros((6, 3), np.float32)
    q[:, 2] = np.arange(6)  # x and y equal: never a good sample
    sets.append(q)
    return sets


def test_line_ransac_bit_exact(ctx):
    sets = _sets()
    xyz = np.concatenate(sets)
    off = np.cumsum([0] + [len(s) for s in sets]).astype(np.int32)
    print(f"[line] {len(sets)} sets, {len(xyz)} points", flush=True)
    g = fit_lines(ctx, xyz, off)
    o = oracle.line_ransac_batch(xyz, off)
    assert np.array_equal(g["stats"], o["stats"])
    assert np.array_equal(g["coeff"], o["coeff"])
    assert np.array_equal(g["inliers"], o["inliers"])
    assert g["stats"][-4:, 0].tolist() == [0, 0, 0, 0]
