"""Seeded synthetic workloads (numpy) for tests, smoke and bench.

Forest clusters: each cluster is the lidar-visible half of a tree trunk
(cylinder) with radial Gaussian noise, plus uniform outliers in a box around
it.  Sizes, radii, tilts and heights follow the batched-fitting workload:
n = floor(exp(U(ln 50, ln 5000))), r ~ U(0.05, 0.6), tilt <= 5 deg,
height U(2, 8) m, noise sigma = 2 cm, 20 % outliers.
"""
import math

import numpy as np


def _perp_basis(a):
    a = a / np.linalg.norm(a)
    ref = np.array([1.0, 0.0, 0.0]) if abs(a[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = ref - ref.dot(a) * a
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(a, e1)
    return e1, e2


def trunk_points(rng, n, base, axis, radius, height, noise, outlier_frac, visible_half=True):
    """Points of one trunk cluster; returns (xyz float64 [n,3], is_inlier bool [n])."""
    axis = axis / np.linalg.norm(axis)
    e1, e2 = _perp_basis(axis)
    n_out = int(round(outlier_frac * n))
    n_in = n - n_out
    h = rng.uniform(0.0, height, n_in)
    if visible_half:
        to_sensor = -base[:2]
        phi_s = math.atan2(to_sensor[1], to_sensor[0])
        # angle measured in the (e1, e2) frame: direction to the sensor projected
        d3 = np.array([math.cos(phi_s), math.sin(phi_s), 0.0])
        phi0 = math.atan2(d3.dot(e2), d3.dot(e1))
        th = rng.uniform(phi0 - math.pi / 2, phi0 + math.pi / 2, n_in)
    else:
        th = rng.uniform(-math.pi, math.pi, n_in)
    rr = radius + rng.normal(0.0, noise, n_in)
    pin = (base[None, :] + h[:, None] * axis[None, :] +
           (rr * np.cos(th))[:, None] * e1[None, :] + (rr * np.sin(th))[:, None] * e2[None, :])
    box = radius + 1.0
    pout = np.stack([
        base[0] + rng.uniform(-box, box, n_out),
        base[1] + rng.uniform(-box, box, n_out),
        base[2] + rng.uniform(0.0, height, n_out),
    ], axis=1)
    pts = np.concatenate([pin, pout], axis=0)
    lab = np.concatenate([np.ones(n_in, bool), np.zeros(n_out, bool)])
    perm = rng.permutation(n)
    return pts[perm], lab[perm]


def make_clusters(n_clusters=2000, seed=1, n_min=50, n_max=5000, r_min=0.05, r_max=0.6,
                  outlier_frac=0.2, noise=0.02, tilt_max_deg=5.0, h_min=2.0, h_max=8.0,
                  range_min=1.0, range_max=40.0):
    """Batched-fitting workload.  Returns dict with xyz float32 [N,3], offsets
    int32 [C+1], truth arrays (base, axis, radius) and per-point inlier labels."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_clusters))).astype(
        np.int64)
    offsets = np.zeros(n_clusters + 1, np.int64)
    offsets[1:] = np.cumsum(sizes)
    N = int(offsets[-1])
    xyz = np.empty((N, 3), np.float64)
    lab = np.empty(N, bool)
    base = np.empty((n_clusters, 3))
    axis = np.empty((n_clusters, 3))
    radius = np.empty(n_clusters)
    for c in range(n_clusters):
        rad = rng.uniform(r_min, r_max)
        tilt = math.radians(rng.uniform(0.0, tilt_max_deg))
        az = rng.uniform(-math.pi, math.pi)
        a = np.array([math.sin(tilt) * math.cos(az), math.sin(tilt) * math.sin(az),
                      math.cos(tilt)])
        rng_d = rng.uniform(max(range_min, rad + 0.5), range_max)
        bearing = rng.uniform(-math.pi, math.pi)
        b = np.array([rng_d * math.cos(bearing), rng_d * math.sin(bearing),
                      rng.uniform(-0.5, 0.5)])
        H = rng.uniform(h_min, h_max)
        p, l = trunk_points(rng, int(sizes[c]), b, a, rad, H, noise, outlier_frac)
        xyz[offsets[c]:offsets[c + 1]] = p
        lab[offsets[c]:offsets[c + 1]] = l
        base[c], axis[c], radius[c] = b, a, rad
    if offsets[-1] >= 2**31:
        raise ValueError("too many points for int32 offsets")
    return {
        "xyz": xyz.astype(np.float32),
        "offsets": offsets.astype(np.int32),
        "base": base,
        "axis": axis,
        "radius": radius,
        "inlier": lab,
    }
