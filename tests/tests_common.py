"""Fixture loaders shared by the tests (the reference's test data converted
by tests/golden/make_fixtures.py)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def organized_cloud(name):
    """An organized tree cloud as H*W x 4 float32 (NaN where empty)."""
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    W, H = int(d["width"]), int(d["height"])
    cloud = np.full((H * W, 4), np.nan, np.float32)
    cloud[d["index"], :3] = d["xyz"]
    cloud[d["index"], 3] = 0.0
    return cloud, H, W


def load_landmarks(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))
