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


def _rotvec_to_R(v):
    th = np.linalg.norm(v)
    if th < 1e-15:
        return np.eye(3)
    k = v / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K


def R_to_quat(R):
    """Rotation matrix -> unit quaternion (x, y, z, w), w >= 0."""
    tr = R[0, 0] + R[1, 1] + R[2, 2]
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        w, x, y, z = 0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = math.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        w, x, y, z = (R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s
    elif R[1, 1] > R[2, 2]:
        s = math.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        w, x, y, z = (R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s
    else:
        s = math.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        w, x, y, z = (R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s
    q = np.array([x, y, z, w])
    q /= np.linalg.norm(q)
    return q if q[3] >= 0 else -q


def quat_to_R(q):
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def make_pose_problems(n_problems=64, n_cyl=500, n_planes=20, pts_per_cyl=60, pts_per_plane=50,
                       seed=2, t_err=0.5, r_err_deg=5.0, cyl_noise=0.02, plane_noise=0.03,
                       map_radius=30.0):
    """Batched pose workload: per problem, a map of n_cyl vertical-ish trunk
    cylinders and n_planes ground patches around a true sensor pose, features
    observed in the sensor frame with noise, and an initial guess perturbed by
    U(0, t_err) m and U(0, r_err_deg) degrees.  Landmark ids are global."""
    rng = np.random.default_rng(seed)
    pts, cyl, planes, pose_true, pose_init = [], [], [], [], []
    res_off, cyl_off, pl_off = [0], [0], [0]
    for k in range(n_problems):
        t_true = np.array([rng.uniform(-100, 100), rng.uniform(-100, 100), rng.uniform(1, 3)])
        ypr = [rng.uniform(-math.pi, math.pi), math.radians(rng.uniform(-3, 3)),
               math.radians(rng.uniform(-3, 3))]
        Rz = _rotvec_to_R(np.array([0, 0, ypr[0]]))
        Ry = _rotvec_to_R(np.array([0, ypr[1], 0]))
        Rx = _rotvec_to_R(np.array([ypr[2], 0, 0]))
        R_true = Rz @ Ry @ Rx
        ground_z = t_true[2] - rng.uniform(1.2, 2.0)
        base_id = cyl_off[-1]
        feats = []
        for j in range(n_cyl):
            rr = map_radius * math.sqrt(rng.uniform(0.01, 1.0))
            bearing = rng.uniform(-math.pi, math.pi)
            root = np.array([t_true[0] + rr * math.cos(bearing), t_true[1] + rr * math.sin(bearing),
                             ground_z])
            tilt = math.radians(rng.uniform(0, 5))
            az = rng.uniform(-math.pi, math.pi)
            ray = np.array([math.sin(tilt) * math.cos(az), math.sin(tilt) * math.sin(az),
                            math.cos(tilt)])
            rad = rng.uniform(0.1, 0.4)
            cyl.append(np.concatenate([root, ray, [rad, 0.0]]))
            p, _ = trunk_points(rng, pts_per_cyl, root - t_true + np.array([0, 0, 0]), ray, rad,
                                6.0, cyl_noise, 0.0)
            p = p + t_true  # trunk_points takes the base relative to the sensor
            ids = np.full(pts_per_cyl, base_id + j, np.int32)
            feats.append((p, ids))
        pl_base = pl_off[-1]
        for j in range(n_planes):
            slope = math.radians(rng.uniform(0, 5))
            az = rng.uniform(-math.pi, math.pi)
            n = np.array([math.sin(slope) * math.cos(az), math.sin(slope) * math.sin(az),
                          math.cos(slope)])
            rr = map_radius * math.sqrt(rng.uniform(0.02, 1.0))
            bearing = rng.uniform(-math.pi, math.pi)
            c = np.array([t_true[0] + rr * math.cos(bearing), t_true[1] + rr * math.sin(bearing),
                          ground_z])
            d = -n @ c
            planes.append(np.concatenate([n, [d]]))
            e1, e2 = _perp_basis(n)
            a = rng.uniform(0, 2 * math.pi, pts_per_plane)
            rad = 2.0 * np.sqrt(rng.uniform(0, 1, pts_per_plane))
            p = (c[None, :] + (rad * np.cos(a))[:, None] * e1 + (rad * np.sin(a))[:, None] * e2 +
                 rng.normal(0, plane_noise, pts_per_plane)[:, None] * n)
            ids = np.full(pts_per_plane, -(pl_base + j) - 1, np.int32)
            feats.append((p, ids))
        P = np.concatenate([f[0] for f in feats])
        ids = np.concatenate([f[1] for f in feats])
        ps = (P - t_true) @ R_true  # R^T (p - t)
        rec = np.empty((len(ps), 4), np.float32)
        rec[:, :3] = ps
        rec[:, 3] = ids.view(np.float32)
        pts.append(rec)
        res_off.append(res_off[-1] + len(ps))
        cyl_off.append(cyl_off[-1] + n_cyl)
        pl_off.append(pl_off[-1] + n_planes)
        # initial guess
        dt = rng.normal(size=3)
        dt *= rng.uniform(0, t_err) / np.linalg.norm(dt)
        ax = rng.normal(size=3)
        ax *= math.radians(rng.uniform(0, r_err_deg)) / np.linalg.norm(ax)
        R_init = _rotvec_to_R(ax) @ R_true
        pose_true.append(np.concatenate([t_true, R_to_quat(R_true)]))
        pose_init.append(np.concatenate([t_true + dt, R_to_quat(R_init)]))
    return {
        "points": np.concatenate(pts),
        "res_offsets": np.array(res_off, np.int32),
        "cyl": np.array(cyl, np.float64).reshape(-1, 8),
        "cyl_offsets": np.array(cyl_off, np.int32),
        "planes": np.array(planes, np.float64).reshape(-1, 4),
        "plane_offsets": np.array(pl_off, np.int32),
        "pose_true": np.array(pose_true),
        "pose_init": np.array(pose_init),
    }


def pose_error(pa, pb):
    """Translation (m) and rotation (rad) distance between poses t,q(xyzw)."""
    dt = np.linalg.norm(pa[..., :3] - pb[..., :3], axis=-1)
    dot = np.abs(np.sum(pa[..., 3:] * pb[..., 3:], axis=-1)).clip(0, 1)
    return dt, 2 * np.arccos(dot)


def make_scan(seed=0, H=64, W=2048, fov_deg=22.5, n_trees=40, r_min=0.1, r_max=0.4,
              h_min=2.0, h_max=8.0, tilt_max_deg=5.0, sensor_height=1.5, slope_max_deg=5.0,
              clutter_frac=0.3, tree_noise=0.02, ground_noise=0.03, range_min=1.0,
              range_max=40.0):
    """One organised H x W lidar scan of a synthetic forest (ray cast).

    Beam (row, col) points at the centre of range-image pixel (row, col) of
    the spherical projection (row 0 = +fov, col from yaw = -atan2(y, x)).
    Hits: the nearest of n_trees tilted cylinders and a sloped ground plane;
    a clutter_frac share of returns is replaced by understory clutter at a
    random range along the beam.  No return -> NaN.
    Returns xyzi float32 [H*W, 4], label uint8 [H*W] (0 other, 1 ground,
    2 tree), instance int32 [H*W] (tree index or -1) and the trees."""
    rng = np.random.default_rng(seed)
    fov = math.radians(fov_deg)
    rows = np.arange(H)
    cols = np.arange(W)
    pitch = fov - (rows + 0.5) / H * (2 * fov)
    yaw = (cols + 0.5) / W * 2 * math.pi - math.pi
    az = -yaw  # atan2(y, x)
    P, A = np.meshgrid(pitch, az, indexing="ij")
    d = np.stack([np.cos(P) * np.cos(A), np.cos(P) * np.sin(A), np.sin(P)], -1).reshape(-1, 3)
    n = H * W
    best = np.full(n, np.inf)
    label = np.zeros(n, np.uint8)
    inst = np.full(n, -1, np.int32)
    # ground plane n.x + d0 = 0 below the sensor
    slope = math.radians(rng.uniform(0, slope_max_deg))
    gaz = rng.uniform(-math.pi, math.pi)
    gn = np.array([math.sin(slope) * math.cos(gaz), math.sin(slope) * math.sin(gaz),
                   math.cos(slope)])
    gd = sensor_height * gn[2]
    den = d @ gn
    with np.errstate(divide="ignore", invalid="ignore"):
        tg = -gd / den
    ok = (den < -1e-6) & (tg > 0) & (tg < range_max)
    best[ok] = tg[ok]
    label[ok] = 1
    trees = []
    for k in range(n_trees):
        rr = rng.uniform(3.0, 0.8 * range_max)
        b = rng.uniform(-math.pi, math.pi)
        base_xy = np.array([rr * math.cos(b), rr * math.sin(b)])
        bz = -(gn[0] * base_xy[0] + gn[1] * base_xy[1] + gd) / gn[2]
        base = np.array([base_xy[0], base_xy[1], bz])
        tilt = math.radians(rng.uniform(0, tilt_max_deg))
        ta = rng.uniform(-math.pi, math.pi)
        ax = np.array([math.sin(tilt) * math.cos(ta), math.sin(tilt) * math.sin(ta),
                       math.cos(tilt)])
        rad = rng.uniform(r_min, r_max)
        hgt = rng.uniform(h_min, h_max)
        trees.append({"base": base, "axis": ax, "radius": rad, "height": hgt})
        # ray / infinite cylinder: |(t d - base) x ax| = rad
        w = d - (d @ ax)[:, None] * ax[None, :]
        c0 = -base - (-base @ ax) * ax
        qa = (w * w).sum(1)
        qb = 2 * (w @ c0)
        qc = c0 @ c0 - rad * rad
        disc = qb * qb - 4 * qa * qc
        hit = disc > 0
        t = np.full(n, np.inf)
        t[hit] = (-qb[hit] - np.sqrt(disc[hit])) / (2 * qa[hit])
        with np.errstate(invalid="ignore"):
            hz = (np.where(hit, t, 0.0)[:, None] * d - base) @ ax
        hit &= (t > 0) & (hz >= 0) & (hz <= hgt)
        closer = hit & (t < best)
        best[closer] = t[closer]
        label[closer] = 2
        inst[closer] = k
    noise = np.where(label == 2, tree_noise, ground_noise)
    hitm = np.isfinite(best) & (best >= range_min)
    rng_vals = best + rng.normal(0, 1, n) * noise
    # understory clutter replaces a share of the returns
    cl = hitm & (rng.uniform(0, 1, n) < clutter_frac)
    rng_vals[cl] = rng.uniform(range_min, np.minimum(best[cl], range_max))
    label[cl] = 0
    inst[cl] = -1
    xyz = rng_vals[:, None] * d
    xyz[~hitm] = np.nan
    label[~hitm] = 0
    inst[~hitm] = -1
    xyzi = np.empty((n, 4), np.float32)
    xyzi[:, :3] = xyz
    xyzi[:, 3] = rng.uniform(0, 255, n)
    return {"xyzi": xyzi, "label": label, "instance": inst, "trees": trees,
            "ground": np.concatenate([gn, [gd]]), "H": H, "W": W, "fov_deg": fov_deg}


def make_forest(seed=4, n_trees=300, length=200.0, width=60.0, r_min=0.1, r_max=0.4,
                h_min=2.0, h_max=8.0, tilt_max_deg=5.0, slope_max_deg=3.0):
    """A static forest: n_trees tilted trunks uniformly in [0, length] x
    [-width/2, width/2] (keeping 1.5 m off the path y = 0), on a plane
    z = g0 x + g1 y (slope <= slope_max_deg)."""
    rng = np.random.default_rng(seed)
    slope = math.radians(rng.uniform(0, slope_max_deg))
    gaz = rng.uniform(-math.pi, math.pi)
    g = (math.tan(slope) * math.cos(gaz), math.tan(slope) * math.sin(gaz))
    trees = []
    while len(trees) < n_trees:
        x = rng.uniform(0.0, length)
        y = rng.uniform(-width / 2, width / 2)
        if abs(y) < 1.5:
            continue
        tilt = math.radians(rng.uniform(0, tilt_max_deg))
        ta = rng.uniform(-math.pi, math.pi)
        ax = np.array([math.sin(tilt) * math.cos(ta), math.sin(tilt) * math.sin(ta),
                       math.cos(tilt)])
        trees.append({"base": np.array([x, y, g[0] * x + g[1] * y]), "axis": ax,
                      "radius": rng.uniform(r_min, r_max), "height": rng.uniform(h_min, h_max)})
    return {"trees": trees, "ground": g}


def trajectory_pose(s, sensor_height=1.5, ground=None):
    """Pose (R, t) at arc length s of the path: along x with a gentle lateral
    weave, heading along the path tangent, the sensor level (yaw only) and
    sensor_height above the ground plane z = g0 x + g1 y of `ground` (a
    forest's "ground"), or above z = 0 without one."""
    y = 1.0 * math.sin(s / 25.0)
    dy = math.cos(s / 25.0) / 25.0
    yaw = math.atan2(dy, 1.0)
    c, sn = math.cos(yaw), math.sin(yaw)
    R = np.array([[c, -sn, 0.0], [sn, c, 0.0], [0.0, 0.0, 1.0]])
    z0 = 0.0 if ground is None else ground[0] * s + ground[1] * y
    return R, np.array([s, y, z0 + sensor_height])


def scan_from(forest, R, t, seed=0, H=128, W=2048, fov_deg=22.5, clutter_frac=0.4,
              tree_noise=0.02, ground_noise=0.03, range_min=1.0, range_max=40.0,
              occlusion=0.1, clutter="understory", understory_height=1.5):
    """One organised H x W scan of `forest` from the sensor pose (R, t)
    (world = R sensor + t), points in the sensor frame.  Clutter makes up
    clutter_frac of the returns: with clutter="understory" (low vegetation)
    they replace returns whose surface point lies less than understory_height
    above the ground (ground and trunk bases), at a range in the last half of
    the ray; with clutter="uniform" they replace returns drawn at random, at a
    uniform range.  `occlusion` of each trunk's hits (a random vertical band)
    are lost to foliage.  Same layout and outputs as make_scan."""
    rng = np.random.default_rng(seed)
    fov = math.radians(fov_deg)
    pitch = fov - (np.arange(H) + 0.5) / H * (2 * fov)
    yaw = (np.arange(W) + 0.5) / W * 2 * math.pi - math.pi
    P, A = np.meshgrid(pitch, -yaw, indexing="ij")
    ds = np.stack([np.cos(P) * np.cos(A), np.cos(P) * np.sin(A), np.sin(P)], -1).reshape(-1, 3)
    dw = ds @ R.T
    n = H * W
    best = np.full(n, np.inf)
    label = np.zeros(n, np.uint8)
    inst = np.full(n, -1, np.int32)
    g0, g1 = forest["ground"]
    # ground z = g0 x + g1 y  <=>  gn . p + gd = 0
    gn = np.array([-g0, -g1, 1.0])
    den = dw @ gn
    with np.errstate(divide="ignore", invalid="ignore"):
        tg = -(gn @ t) / den
    ok = (den < -1e-6) & (tg > 0) & (tg < range_max)
    best[ok] = tg[ok]
    label[ok] = 1
    azw = np.arctan2(dw[:, 1], dw[:, 0])
    for k, tr in enumerate(forest["trees"]):
        D = np.hypot(*(tr["base"][:2] - t[:2]))
        if D > range_max + 1.0:
            continue
        z0 = rng.uniform(0, tr["height"])  # occluded band (drawn for every tree in range)
        ax, base, rad = tr["axis"], tr["base"] - t, tr["radius"]
        # only rays within the trunk's azimuth window (radius + lean margin)
        margin = rad + tr["height"] * math.sin(math.acos(min(1.0, ax[2]))) + 0.3
        half = math.asin(min(1.0, margin / D)) if D > margin else math.pi
        bearing = math.atan2(base[1], base[0])
        sel = np.nonzero(np.abs((azw - bearing + math.pi) % (2 * math.pi) - math.pi) <= half)[0]
        if sel.size == 0:
            continue
        dsel = dw[sel]
        w = dsel - (dsel @ ax)[:, None] * ax[None, :]
        c0 = -base - (-base @ ax) * ax
        qa = (w * w).sum(1)
        qb = 2 * (w @ c0)
        qc = c0 @ c0 - rad * rad
        disc = qb * qb - 4 * qa * qc
        hit = disc > 0
        tt = np.full(sel.size, np.inf)
        tt[hit] = (-qb[hit] - np.sqrt(disc[hit])) / (2 * qa[hit])
        with np.errstate(invalid="ignore"):
            hz = (np.where(hit, tt, 0.0)[:, None] * dsel - base) @ ax
        hit &= (tt > 0) & (hz >= 0) & (hz <= tr["height"])
        if occlusion > 0:
            hit &= ~((hz > z0) & (hz < z0 + occlusion * tr["height"]))
        closer = hit & (tt < best[sel])
        idx = sel[closer]
        best[idx] = tt[closer]
        label[idx] = 2
        inst[idx] = k
    noise = np.where(label == 2, tree_noise, ground_noise)
    hitm = np.isfinite(best) & (best >= range_min)
    rv = best + rng.normal(0, 1, n) * noise
    u = rng.uniform(0, 1, n)
    if clutter == "understory":
        with np.errstate(invalid="ignore"):
            pw = t[None, :] + np.where(hitm, best, 0.0)[:, None] * dw
            above = pw[:, 2] - (g0 * pw[:, 0] + g1 * pw[:, 1])
        low = hitm & (above < understory_height)
        frac = min(1.0, clutter_frac * hitm.sum() / max(low.sum(), 1))
        cl = low & (u < frac)
        lo = np.maximum(range_min, 0.5 * (best[cl] + range_min))
        rv[cl] = rng.uniform(lo, np.minimum(best[cl], range_max))
    elif clutter == "uniform":
        cl = hitm & (u < clutter_frac)
        rv[cl] = rng.uniform(range_min, np.minimum(best[cl], range_max))
    else:
        raise ValueError(f"unknown clutter model {clutter!r}")
    label[cl] = 0
    inst[cl] = -1
    xyz = rv[:, None] * ds
    xyz[~hitm] = np.nan
    label[~hitm] = 0
    inst[~hitm] = -1
    xyzi = np.empty((n, 4), np.float32)
    xyzi[:, :3] = xyz
    xyzi[:, 3] = rng.uniform(0, 255, n)
    return {"xyzi": xyzi, "label": label, "instance": inst, "H": H, "W": W, "fov_deg": fov_deg}


def _sequence_scan(job):
    forest, s, seed, H, W, clutter_frac = job
    R, t = trajectory_pose(s, ground=forest["ground"])
    sc = scan_from(forest, R, t, seed=seed, H=H, W=W, clutter_frac=clutter_frac)
    return (np.ascontiguousarray(sc["xyzi"], "<f4").tobytes(),
            np.ascontiguousarray(sc["label"], np.uint8).tobytes())


def write_keyframe_sequence(path, n_keyframes=16, H=128, W=2048, spacing=0.5, start=20.0,
                            seed=4, n_trees=300, clutter_frac=0.4, odom_sigma_t=0.05,
                            odom_sigma_yaw_deg=0.5, forest_length=200.0, workers=1,
                            log_every=0):
    """Writes the keyframe sequence file of sloam_amd/bench/keyframe_bench.cpp:
    scans of one forest (make_forest, `forest_length` metres long) from poses
    `spacing` metres apart along trajectory_pose (1.5 m above the forest's
    ground), with noisy odometry
    increments between them.  With workers > 1 the scans are cast by a pool
    of that many processes; the file is the same byte for byte.  With
    log_every > 0 a line goes to stderr after every log_every scans."""
    import struct
    forest = make_forest(seed=seed, n_trees=n_trees, length=forest_length)
    rng = np.random.default_rng(seed + 1)
    prev = None
    header = []
    for k in range(n_keyframes):
        R, t = trajectory_pose(start + k * spacing, ground=forest["ground"])
        truth = np.concatenate([t, R_to_quat(R)])
        if prev is None:
            odom = np.array([0, 0, 0, 0, 0, 0, 1.0])
        else:
            Rp, tp = prev
            Rd = Rp.T @ R
            td = Rp.T @ (t - tp) + rng.normal(0, odom_sigma_t, 3) * np.array([1, 1, 0.2])
            Rd = _rotvec_to_R(np.array([0, 0, math.radians(
                rng.normal(0, odom_sigma_yaw_deg))])) @ Rd
            odom = np.concatenate([td, R_to_quat(Rd)])
        prev = (R, t)
        header.append((truth, odom))
    jobs = [(forest, start + k * spacing, seed * 1000 + k, H, W, clutter_frac)
            for k in range(n_keyframes)]
    pool = None
    if workers > 1 and n_keyframes > 1:
        import multiprocessing
        pool = multiprocessing.get_context("fork").Pool(min(workers, n_keyframes))
        scans = pool.imap(_sequence_scan, jobs, chunksize=4)
    else:
        scans = map(_sequence_scan, jobs)
    try:
        with open(path, "wb") as f:
            f.write(struct.pack("<iii", n_keyframes, H, W))
            for k, ((truth, odom), (xyzi, label)) in enumerate(zip(header, scans)):
                f.write(np.asarray(truth, "<f8").tobytes())
                f.write(np.asarray(odom, "<f8").tobytes())
                f.write(xyzi)
                f.write(label)
                if log_every > 0 and (k + 1) % log_every == 0:
                    import sys
                    print(f"[synth] {k + 1}/{n_keyframes} scans written", file=sys.stderr,
                          flush=True)
    finally:
        if pool is not None:
            pool.close()
            pool.join()


XYYAW_MASK = 0b100011  # t_x, t_y, w_z free (sloam.cpp:75-77 fixes {2, 3, 4})
ZROLLPITCH_MASK = 0b011100  # t_z, w_x, w_y free (sloam.cpp:136-138 fixes {0, 1, 5})


def split_two_step(prob):
    """Two-step pose workload from a joint one: per problem, an XY-yaw problem
    over its cylinder features and a Z-roll-pitch problem over its plane
    features, both from the same initial pose (TwoStepOptimizePose,
    sloam/src/core/sloam.cpp:33-53)."""
    pts, off, masks, init = [], [0], [], []
    P = prob["points"]
    ro = prob["res_offsets"]
    for k in range(len(ro) - 1):
        seg = P[ro[k]:ro[k + 1]]
        ids = seg[:, 3].view(np.int32)
        for sel, m in ((ids >= 0, XYYAW_MASK), (ids < 0, ZROLLPITCH_MASK)):
            pts.append(seg[sel])
            off.append(off[-1] + int(sel.sum()))
            masks.append(m)
            init.append(prob["pose_init"][k])
    out = dict(prob)
    out["points"] = np.concatenate(pts) if pts else np.zeros((0, 4), np.float32)
    out["res_offsets"] = np.array(off, np.int32)
    out["free_mask"] = np.array(masks, np.int32)
    out["pose_init"] = np.array(init)
    # each derived problem keeps the landmark tables of its source problem
    n = len(ro) - 1
    out["cyl_offsets"] = np.repeat(prob["cyl_offsets"][:-1], 2).tolist() + [prob["cyl_offsets"][-1]]
    out["cyl_offsets"] = np.array(out["cyl_offsets"], np.int32)
    co = prob["cyl_offsets"]
    po = prob["plane_offsets"]
    cyl_off, pl_off = [0], [0]
    for k in range(n):
        # XY-yaw problem owns the cylinders, Z-roll-pitch the planes
        cyl_off += [co[k + 1], co[k + 1]]
        pl_off += [po[k], po[k + 1]]
    out["cyl_offsets"] = np.array(cyl_off, np.int32)
    out["plane_offsets"] = np.array(pl_off, np.int32)
    return out


def make_small_keyframe(seed=0, t_err=0.3, r_err_deg=3.0):
    """The small-scan keyframe workload: make_scan(seed, 64 x 2048, 40 trunks),
    its trunks as the map (root, axis, radius; the world frame is the sensor
    frame of the true pose, the identity), the ground plane (n, d) and an
    estimate t_err metres and r_err_deg degrees off the true pose, in
    directions drawn from seed 100 + seed.  Returns (scan, submap [40, 7],
    map plane [4], estimate (t, q x y z w))."""
    s = make_scan(seed=seed, H=64, W=2048)
    submap = np.array([[*t["base"], *t["axis"], t["radius"]] for t in s["trees"]])
    rng = np.random.default_rng(100 + seed)
    dt = rng.normal(size=3)
    dt *= t_err / np.linalg.norm(dt)
    ax = rng.normal(size=3)
    ax *= math.radians(r_err_deg) / np.linalg.norm(ax)
    ang = np.linalg.norm(ax)
    q = np.concatenate([ax / ang * math.sin(ang / 2), [math.cos(ang / 2)]])
    return s, submap, s["ground"], np.concatenate([dt, q])


def read_keyframe_sequence(path, frames=None):
    """Reads a keyframe sequence file (write_keyframe_sequence): the frames
    whose index is in `frames` (all when None), as a dict index -> (true pose
    [7], odometry increment [7], xyzi [H*W, 4] float32, labels [H*W] uint8),
    plus the (K, H, W) header."""
    with open(path, "rb") as f:
        K, H, W = np.frombuffer(f.read(12), "<i4")
        rec = 112 + int(H) * int(W) * 17
        out = {}
        for k in range(int(K)):
            if frames is not None and k not in frames:
                f.seek(rec, 1)
                continue
            hdr = np.frombuffer(f.read(112), "<f8")
            xyzi = np.frombuffer(f.read(int(H) * int(W) * 16), "<f4").reshape(-1, 4)
            lab = np.frombuffer(f.read(int(H) * int(W)), np.uint8)
            out[k] = (hdr[:7].copy(), hdr[7:].copy(), xyzi.copy(), lab.copy())
    return out, (int(K), int(H), int(W))


if __name__ == "__main__":
    # python -m sloam_amd.synth PATH N_KEYFRAMES [spacing forest_length odom_sigma_t
    #     odom_sigma_yaw_deg workers H]: writes a keyframe sequence file from a
    # process of its own (a pool may fork here; nothing here touches the GPU)
    import sys
    a = sys.argv[1:]
    if len(a) < 2:
        sys.exit("usage: python -m sloam_amd.synth PATH N_KEYFRAMES [spacing forest_length "
                 "odom_sigma_t odom_sigma_yaw_deg workers H]")
    kw = {}
    for i, (name, cast) in enumerate((("spacing", float), ("forest_length", float),
                                      ("odom_sigma_t", float), ("odom_sigma_yaw_deg", float),
                                      ("workers", int), ("H", int))):
        if len(a) > 2 + i:
            kw[name] = cast(a[2 + i])
    write_keyframe_sequence(a[0], n_keyframes=int(a[1]), log_every=100, **kw)
