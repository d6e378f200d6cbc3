"""TEST INFRASTRUCTURE ONLY -- ctypes binding of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker / CPU baseline.  The product code under
sloam_amd/ never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# the reference's Release flags, one thread (oracle/Makefile)
FAITHFUL_PATH = os.path.join(HERE, "liboracle_faithful.so")

_lib = None
_libs = {}
_variant = "tuned"


def build():
    subprocess.run(["make", "-C", HERE, "-j8", "-s"], check=True)


def select(variant):
    """Which build the following calls use: "tuned" (liboracle.so, the
    checker) or "faithful" (liboracle_faithful.so, the reference's flags)."""
    global _variant, _lib
    if variant not in ("tuned", "faithful"):
        raise ValueError(variant)
    _variant = variant
    _lib = None


def load():
    global _lib
    if _lib is None:
        path = LIB_PATH if _variant == "tuned" else FAITHFUL_PATH
        if not os.path.exists(path):
            build()
        if path not in _libs:
            lib = C.CDLL(path)
            _declare(lib)
            _declare_estimator(lib)
            _libs[path] = lib
        _lib = _libs[path]
    return _lib


def _declare(lib):
    P = C.c_void_p
    lib.oracle_philox4x32_10.argtypes = [P, P, P]
    lib.oracle_draw3.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P]
    lib.oracle_eig3.argtypes = [P, P, P]
    lib.oracle_cyl_seed.argtypes = [P, C.c_int, P, P]
    lib.oracle_cyl_hypothesis.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                          C.c_uint64, C.c_int, P, P, P]
    lib.oracle_cyl_fit_batch.argtypes = [P, P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                         C.c_int, C.c_uint64, C.c_int, P, P, P, P, P, P, C.c_int,
                                         C.c_int]
    lib.oracle_cyl_distances_fp64.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_float,
                                              C.c_float, C.c_uint64, C.c_int, P]
    lib.oracle_pose_residuals.argtypes = [P, C.c_int, P, P, P, P, P]
    lib.oracle_mt19937.argtypes = [C.c_uint32, C.c_int, P]
    lib.oracle_line_ransac.argtypes = [P, C.c_int, C.c_double, C.c_int, C.c_double, C.c_uint32,
                                       C.c_int, P, P, P]
    lib.oracle_ground_planes.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_double, C.c_int, P, P, P, P]
    lib.oracle_plane_fit_batch.argtypes = [P, P, C.c_int, C.c_int, C.c_float, C.c_uint64, C.c_int,
                                           P, P, P, P, P, C.c_int]
    lib.oracle_project_partition.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_float,
                                             C.c_float, C.c_int, P, P, P, P, P, P, P]
    lib.oracle_pose_lm_batch.argtypes = [P, P, C.c_int, P, P, P, P, P, P, P, P, P, P, C.c_int]
    lib.oracle_left_jacobian.argtypes = [P, P]
    lib.oracle_aa_quat.argtypes = [P, P, P]
    lib.oracle_pn_forward.argtypes = [P, P, P, C.c_int, C.c_int, P, P, P, P]
    lib.oracle_pn_loss_grad.argtypes = [P, P, P, C.c_int, P, P, C.c_int, P, P]
    lib.oracle_pn_param_count.restype = C.c_int
    lib.oracle_trellis_clusters.argtypes = [P, C.c_int, C.c_int, C.c_float, P]
    lib.oracle_trellis_clusters.restype = C.c_int
    lib.oracle_trellis_trees.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_float, C.c_int, C.c_int, C.c_double, C.c_int,
                                         C.c_int, C.c_int, P, P, P, P, P, P]
    lib.oracle_trellis_trees.restype = C.c_int


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def philox(ctr, key):
    lib = load()
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib.oracle_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def draw3(x0, x1, x2, n):
    out = np.zeros(3, np.int32)
    st = load().oracle_draw3(x0, x1, x2, n, _p(out))
    if st:
        raise ValueError("n < 3")
    return out


def eig3(a6):
    a = np.ascontiguousarray(a6, np.float64)
    w = np.zeros(3)
    V = np.zeros(9)
    load().oracle_eig3(_p(a), _p(w), _p(V))
    return w, V.reshape(3, 3)


def cyl_seed(xyz):
    x = np.ascontiguousarray(xyz, np.float32)
    fr = np.zeros(9, np.float32)
    ax = np.zeros(3)
    load().oracle_cyl_seed(_p(x), len(x), _p(fr), _p(ax))
    return fr, ax


def cyl_hypothesis(xyz, cluster, h, thr, max_radius, seed, tag=0):
    x = np.ascontiguousarray(xyz, np.float32)
    idx = np.zeros(3, np.int32)
    circ = np.zeros(4, np.float32)
    prm = np.zeros(4, np.float32)
    st = load().oracle_cyl_hypothesis(_p(x), len(x), cluster, h, thr, max_radius, seed, tag,
                                      _p(idx), _p(circ), _p(prm))
    if st:
        raise ValueError("cluster too small")
    return idx, circ, prm


def cyl_fit_batch(xyz, offsets, n_hyp, thr=0.05, max_radius=1.0, refit_rounds=3, refit_iters=2,
                  seed=0x5EED,
                  tag=0, counts=True, samples=False, mask=True, threads=1, proposal=None,
                  fp64_refit=False):
    """The batched cylinder fit (oracle_cyl_fit_batch).  The refit runs in the
    GPU's FP32 arithmetic and summation order, or in FP64 (fp64_refit)."""
    x = np.ascontiguousarray(xyz, np.float32)
    off = np.ascontiguousarray(offsets, np.int32)
    prop = None if proposal is None else np.ascontiguousarray(proposal, np.float32)
    C_ = len(off) - 1
    N = int(off[-1])
    out = {
        "model": np.zeros((C_, 8)),
        "best": np.zeros(C_, np.int32),
    }
    if counts:
        out["counts"] = np.zeros((C_, n_hyp), np.int32)
    if samples:
        out["samples"] = np.zeros((C_, n_hyp, 3), np.int32)
    if mask:
        out["mask"] = np.zeros((N + 31) // 32, np.uint32)
    st = load().oracle_cyl_fit_batch(_p(x), _p(off), C_, n_hyp, thr, max_radius, refit_rounds,
                                     refit_iters,
                                     seed, tag,
                                     _p(prop),
                                     _p(out.get("counts")), _p(out.get("samples")),
                                     _p(out.get("mask")), _p(out["best"]), _p(out["model"]),
                                     threads, int(bool(fp64_refit)))
    if st:
        raise ValueError("oracle_cyl_fit_batch failed")
    return out


def cyl_distances_fp64(xyz, cluster, h, thr, max_radius, seed, tag=0):
    x = np.ascontiguousarray(xyz, np.float32)
    d = np.zeros(len(x))
    st = load().oracle_cyl_distances_fp64(_p(x), len(x), cluster, h, thr, max_radius, seed, tag,
                                          _p(d))
    return (d if st == 0 else None)


POSE_DEFAULTS = dict(max_iters=10, stop_on_convergence=False, huber_delta=0.1,
                     function_tolerance=1e-6, gradient_tolerance=1e-10,
                     parameter_tolerance=1e-8, initial_radius=1e4, min_relative_decrease=1e-3,
                     parametrization=0)


def pose_residuals(points, cyl, planes, pose):
    pts = np.ascontiguousarray(points, np.float32)
    n = len(pts)
    r = np.zeros(n)
    J = np.zeros((n, 6))
    load().oracle_pose_residuals(_p(pts), n, _p(np.ascontiguousarray(cyl, np.float64)),
                                 _p(np.ascontiguousarray(planes, np.float64)),
                                 _p(np.ascontiguousarray(pose, np.float64)), _p(r), _p(J))
    return r, J


def pose_lm_batch(prob, threads=1, **kw):
    c = dict(POSE_DEFAULTS)
    c.update(kw)
    cfg_i = np.array([c["max_iters"], int(c["stop_on_convergence"]), c["parametrization"]],
                     np.int32)
    cfg_d = np.array([c["huber_delta"], c["function_tolerance"], c["gradient_tolerance"],
                      c["parameter_tolerance"], c["initial_radius"],
                      c["min_relative_decrease"]], np.float64)
    P = len(prob["res_offsets"]) - 1
    pose = np.zeros((P, 7))
    params = np.zeros((P, 6))
    status = np.zeros((P, 2), np.int32)
    cost = np.zeros((P, 2))
    pts = np.ascontiguousarray(prob["points"], np.float32)
    fm = prob.get("free_mask")
    fm = None if fm is None else np.ascontiguousarray(fm, np.int32)
    load().oracle_pose_lm_batch(_p(pts), _p(np.ascontiguousarray(prob["res_offsets"], np.int32)),
                                P, _p(np.ascontiguousarray(prob["cyl"], np.float64)),
                                _p(np.ascontiguousarray(prob["planes"], np.float64)),
                                _p(np.ascontiguousarray(prob["pose_init"], np.float64)), _p(fm),
                                _p(cfg_i), _p(cfg_d), _p(pose), _p(params), _p(status), _p(cost),
                                threads)
    return {"pose": pose, "params": params, "status": status, "cost": cost}


def left_jacobian(w):
    J = np.zeros(9)
    load().oracle_left_jacobian(_p(np.ascontiguousarray(w, np.float64)), _p(J))
    return J.reshape(3, 3)


def aa_quat(aa):
    q = np.zeros(4)
    back = np.zeros(3)
    load().oracle_aa_quat(_p(np.ascontiguousarray(aa, np.float64)), _p(q), _p(back))
    return q, back


def project_partition(xyzi, label, H=64, W=2048, fov_up=22.5, fov_down=-22.5, destagger=False):
    x = np.ascontiguousarray(xyzi, np.float32)
    lab = np.ascontiguousarray(label, np.uint8)
    n = len(x)
    out = {
        "proj": np.zeros((n, 2), np.int32),
        "range": np.zeros((H, W), np.float32),
        "winner": np.zeros((H, W), np.int32),
        "mask": np.zeros((H, W), np.uint8),
        "ground": np.zeros((n, 4), np.float32),
        "tree": np.zeros((n, 4), np.float32),
    }
    ng = np.zeros(1, np.int32)
    load().oracle_project_partition(_p(x), _p(lab), n, H, W, fov_up, fov_down, int(destagger),
                                    _p(out["proj"]), _p(out["range"]), _p(out["winner"]),
                                    _p(out["mask"]), _p(out["ground"]), _p(ng), _p(out["tree"]))
    out["ground"] = out["ground"][:int(ng[0])]
    out["n_ground"] = ng
    return out


def ground_planes(points, radii_bins=1, theta_bins=18, min_dist=0.0, max_dist=30.0,
                  retain_thresh=0.05, num_features=60, origin=(0.0, 0.0, 0.0)):
    x = np.ascontiguousarray(points, np.float32)
    nc = radii_bins * theta_bins
    info = np.zeros((nc, 4), np.int32)
    model = np.zeros((nc, 8))
    feats = np.zeros((nc, max(num_features, 1), 4), np.float32)
    org = np.array(origin, np.float32)
    load().oracle_ground_planes(_p(x), len(x), radii_bins, theta_bins, min_dist, max_dist,
                                retain_thresh, num_features, _p(org), _p(info), _p(model),
                                _p(feats))
    return {"info": info, "model": model, "features": feats}


def plane_fit_batch(xyz, offsets, n_hyp, thr=0.05, seed=0x5EED, tag=1, counts=True, samples=False,
                    mask=True, threads=1):
    x = np.ascontiguousarray(xyz, np.float32)
    off = np.ascontiguousarray(offsets, np.int32)
    S = len(off) - 1
    N = int(off[-1])
    out = {"model": np.zeros((S, 8)), "best": np.zeros(S, np.int32)}
    if counts:
        out["counts"] = np.zeros((S, n_hyp), np.int32)
    if samples:
        out["samples"] = np.zeros((S, n_hyp, 3), np.int32)
    if mask:
        out["mask"] = np.zeros((N + 31) // 32, np.uint32)
    load().oracle_plane_fit_batch(_p(x), _p(off), S, n_hyp, thr, seed, tag, _p(out.get("counts")),
                                  _p(out.get("samples")), _p(out.get("mask")), _p(out["best"]),
                                  _p(out["model"]), threads)
    return out


def mt19937(seed, n):
    out = np.zeros(n, np.uint32)
    load().oracle_mt19937(seed, n, _p(out))
    return out


def line_ransac(points, threshold=0.25, max_iterations=50, probability=0.99, seed=12345,
                optimize=True):
    P = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    n = len(P)
    coeff = np.zeros(6, np.float32)
    inl = np.zeros(max(n, 1), np.uint8)
    stats = np.zeros(4, np.int32)
    load().oracle_line_ransac(_p(P), n, threshold, max_iterations, probability, seed,
                              int(optimize), _p(coeff), _p(inl), _p(stats))
    return {"coeff": coeff, "inliers": inl[:n], "stats": stats}


def line_ransac_batch(xyz, offsets, **kw):
    off = np.asarray(offsets)
    res = [line_ransac(xyz[off[s]:off[s + 1]], **kw) for s in range(len(off) - 1)]
    return {"coeff": np.array([r["coeff"] for r in res]).reshape(-1, 6),
            "stats": np.array([r["stats"] for r in res]).reshape(-1, 4),
            "inliers": np.concatenate([r["inliers"] for r in res]) if res else np.zeros(0, np.uint8)}


TRELLIS_DEFAULTS = dict(dist_thr=1.0, min_label_points=80, min_row_points=3, max_dist=0.2,
                        min_vertices=16, max_vertices=56, radius_scale=1.0)


def trellis_clusters(cloud, H, W, dist_thr=1.0):
    x = np.ascontiguousarray(cloud, np.float32).reshape(H * W, 4)
    labels = np.zeros(H * W, np.int32)
    n = load().oracle_trellis_clusters(_p(x), H, W, dist_thr, _p(labels))
    return labels, n


def trellis(cloud, H, W, **kw):
    """Components and trees of an organized H x W tree cloud (x NaN = empty)."""
    c = dict(TRELLIS_DEFAULTS)
    c.update(kw)
    x = np.ascontiguousarray(cloud, np.float32).reshape(H * W, 4)
    labels, n = trellis_clusters(x, H, W, c["dist_thr"])
    cap_t = H * W // (c["min_label_points"] + 1) + 1
    cap_v = cap_t * c["max_vertices"]
    cap_p = H * W
    out = {"labels": labels, "n_labels": n,
           "tree_offsets": np.zeros(cap_t + 1, np.int32),
           "vertex_int": np.zeros((cap_v, 4), np.int32),
           "vertex_f64": np.zeros((cap_v, 4)),
           "point_offsets": np.zeros(cap_v + 1, np.int32),
           "points": np.zeros((cap_p, 4), np.float32)}
    counts = np.zeros(3, np.int32)
    rc = load().oracle_trellis_trees(_p(x), _p(labels), H, W, n, c["min_label_points"],
                                     c["min_row_points"], c["max_dist"], c["min_vertices"],
                                     c["max_vertices"], c["radius_scale"], cap_t, cap_v, cap_p,
                                     _p(out["tree_offsets"]), _p(out["vertex_int"]),
                                     _p(out["vertex_f64"]), _p(out["point_offsets"]),
                                     _p(out["points"]), _p(counts))
    assert rc == 0
    nt, nv, npt = (int(v) for v in counts)
    out["tree_offsets"] = out["tree_offsets"][:nt + 1]
    out["vertex_int"] = out["vertex_int"][:nv]
    out["vertex_f64"] = out["vertex_f64"][:nv]
    out["point_offsets"] = out["point_offsets"][:nv + 1]
    out["points"] = out["points"][:npt]
    return out


PN_LOSS_WEIGHTS = (1.0, 1.0, 1.0)


def pn_forward(params, xyz, offsets, emulate=True):
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(xyz, np.float32)
    off = np.ascontiguousarray(offsets, np.int32)
    S = len(off) - 1
    out = np.zeros((S, 7))
    f = np.zeros((S, 256))
    fidx = np.zeros((S, 256), np.int32)
    prop = np.zeros((S, 8))
    load().oracle_pn_forward(_p(p), _p(x), _p(off), S, int(emulate), _p(out), _p(f), _p(fidx),
                             _p(prop))
    return {"out": out, "f": f, "fidx": fidx, "proposal": prop}


def pn_loss_grad(params, xyz, offsets, targets, weights=PN_LOSS_WEIGHTS, emulate=True):
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(xyz, np.float32)
    off = np.ascontiguousarray(offsets, np.int32)
    t = np.ascontiguousarray(targets, np.float64)
    w = np.ascontiguousarray(weights, np.float32)
    S = len(off) - 1
    loss = np.zeros(S)
    grad = np.zeros(load().oracle_pn_param_count())
    load().oracle_pn_loss_grad(_p(p), _p(x), _p(off), S, _p(t), _p(w), int(emulate), _p(loss),
                               _p(grad))
    return loss, grad


# ---- per-keyframe estimator (oracle_estimator.cpp) -------------------------

def _declare_estimator(lib):
    P = C.c_void_p
    lib.oracle_sloam_create.argtypes = [P]
    lib.oracle_sloam_create.restype = P
    lib.oracle_sloam_destroy.argtypes = [P]
    lib.oracle_sloam_set_state.argtypes = [P, C.c_int, C.c_int, P]
    lib.oracle_sloam_run.argtypes = [P, P, C.c_int, P, P, C.c_int, P, C.c_int, P, P, P, P, P,
                                     C.c_int, P, P, P, P, C.c_int, P, P, P, P, P, P, P, P, P]
    lib.oracle_from_two_vectors_euler.argtypes = [P, P, P, P]


# FeatureModelParams defaults (sloam/params/sloam.yaml, and the node's
# hard-coded maxTreeRadius 0.3 / maxAxisTheta 10)
SLOAM_PARAMS = dict(maxLidarDist=20.0, maxTreeRadius=0.3, maxAxisTheta=10.0,
                    defaultTreeRadius=0.2, featuresPerTree=20, numGroundFeatures=5,
                    groundRadiiBins=2, groundThetaBins=18, treeMatchThresh=0.5,
                    AddNewTreeThreshDist=1.5, minTreeModels=5, minGroundModels=36,
                    twoStepOptim=1)
_PARAM_ORDER = ("maxLidarDist", "maxTreeRadius", "maxAxisTheta", "defaultTreeRadius",
                "featuresPerTree", "numGroundFeatures", "groundRadiiBins", "groundThetaBins",
                "treeMatchThresh", "AddNewTreeThreshDist", "minTreeModels", "minGroundModels",
                "twoStepOptim")


def from_two_vectors_euler(a, b):
    """Rotation matrix of the smallest rotation taking a onto b and its
    Euler angles (0, 1, 2), as the plane acceptance reads them."""
    lib = load()
    R = np.zeros(9)
    e = np.zeros(3)
    lib.oracle_from_two_vectors_euler(_p(np.ascontiguousarray(a, np.float64)),
                                      _p(np.ascontiguousarray(b, np.float64)), _p(R), _p(e))
    return R.reshape(3, 3), e


class SloamEstimator:
    """RunSloam restated on the CPU (oracle_estimator.cpp), with its state
    (first scan, previous ground planes)."""

    def __init__(self, **params):
        lib = load()
        p = dict(SLOAM_PARAMS)
        p.update(params)
        arr = np.array([float(p[k]) for k in _PARAM_ORDER], np.float64)
        self.params = p
        self.h = lib.oracle_sloam_create(_p(arr))
        self.lib = lib

    def close(self):
        if self.h:
            self.lib.oracle_sloam_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, first, prev_planes):
        """prev_planes: [n, 7] plane[4], centroid[3]."""
        pp = np.zeros((len(prev_planes), 8))
        if len(prev_planes):
            pp[:, :7] = prev_planes
        self.lib.oracle_sloam_set_state(self.h, int(first), len(pp), _p(pp))

    def run(self, rec, tree_mode=None):
        """One RunSloam call on a keyframe record (the dict of
        tests' keyframe dump reader: pose, cells, trees, line/cylinder fits,
        map).  tree_mode: 0 line fits computed here, 1 the record's line
        fits, 2 the record's cylinder fits (default: cylinder fits when the
        record has them, else its line fits, else computed)."""
        nt = len(rec["tree_off"]) - 1
        if tree_mode is None:
            tree_mode = 2 if len(rec["cyl_ok"]) == nt and nt > 0 else \
                (1 if len(rec["line_int"]) == nt and nt > 0 else 0)
        nc = len(rec["cell_model"])
        nf = rec["nf"]
        cm = np.zeros((nc, 8))
        if nc:
            cm[:, :7] = rec["cell_model"]
        i32 = lambda a: np.ascontiguousarray(a, np.int32)  # noqa: E731
        f64 = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
        f32 = lambda a: np.ascontiguousarray(a, np.float32)  # noqa: E731
        info = np.zeros(12, np.int32)
        pose = np.zeros(7)
        delta = np.zeros(7)
        tm_id = np.zeros(max(nt, 1), np.int32)
        tm = np.zeros((max(nt, 1), 7))
        matches = np.zeros(max(nt, 1), np.int32)
        valid = np.zeros(max(nt, 1), np.int32)
        planes = np.zeros((max(nc, 1), 8))
        nm = len(rec["map"])
        args = [self.h, _p(f64(rec["pose"])), nc, _p(cm), _p(i32(rec["cell_valid"])), nf,
                _p(f32(rec["cell_feat"])), nt, _p(i32(rec["tree_off"])),
                _p(i32(rec["vtx_int"])), _p(f64(rec["vtx_f64"])), _p(i32(rec["vtx_pt_off"])),
                _p(f32(rec["pts"])), tree_mode,
                _p(i32(rec["line_int"])) if tree_mode == 1 else None,
                _p(f32(rec["line_coeff"])) if tree_mode == 1 else None,
                _p(i32(rec["cyl_ok"])) if tree_mode == 2 else None,
                _p(f64(rec["cyl_fit"])) if tree_mode == 2 else None,
                nm, _p(f64(rec["map"]) if nm else np.zeros(7)), _p(info), _p(pose), _p(delta),
                _p(tm_id), _p(tm), _p(matches), _p(valid), _p(planes)]
        self.lib.oracle_sloam_run(*args)
        k, n_pl = int(info[6]), int(info[7])
        return {"ok": bool(info[0]), "early": bool(info[1]), "tree_step": bool(info[2]),
                "ground_step": bool(info[3]), "tree_matches": int(info[4]),
                "plane_matches": int(info[5]), "first": bool(info[8]),
                "converged": bool(info[9]), "tree_iters": int(info[10]),
                "ground_iters": int(info[11]), "pose": pose, "delta": delta,
                "tm_id": tm_id[:k].copy(), "tm": tm[:k].copy(), "matches": matches[:k].copy(),
                "tree_valid": valid[:nt].copy(), "planes": planes[:n_pl, :7].copy()}


# ---- small-scan plain keyframe (the GPU's sloam_amd/plain_keyframe.py) ------

def _axis_gap(root_a, ray_a, root_b, ray_b):
    """Cylinder::distance(model) (sloam/src/objects/cylinder.cpp:175-194):
    mean over z = 0, 3, 6 of |a(z) - b(z)|, a(z) = root + ((z - root_z) /
    ray_z) ray."""
    tot = 0.0
    for z in (0.0, 3.0, 6.0):
        a = root_a + ((z - root_a[2]) / ray_a[2]) * ray_a
        b = root_b + ((z - root_b[2]) / ray_b[2]) * ray_b
        tot += float(np.sqrt(np.sum((a - b) ** 2)))
    return tot / 3.0


def plain_keyframe(xyzi, label, estimate, submap, map_plane, H=64, W=2048, n_hyp=200, thr=0.05,
                   max_radius=1.0, plane_hyp=200, max_tree_radius=0.45, tree_match=1.0,
                   features_per_tree=20, ground_features=256, seed=0x5EED, threads=1):
    """The small-scan plain-method keyframe on the CPU: projection and
    partition, tree instances, per-tree cylinder RANSAC, ground RANSAC over
    the ground list, association (each valid tree with radius <=
    max_tree_radius to the map tree of smallest axis gap, strict <, if below
    tree_match; its first features_per_tree points; every k-th ground inlier
    up to ground_features) and the FP64 LM pose solve (Huber 0.1, 50
    iterations, stop at convergence)."""
    pp = project_partition(xyzi, label, H=H, W=W)
    tr = trellis(pp["tree"], H, W)
    nt = len(tr["tree_offsets"]) - 1
    offsets = tr["point_offsets"][tr["tree_offsets"]].astype(np.int32) if nt else \
        np.zeros(1, np.int32)
    txyz = np.ascontiguousarray(tr["points"][:, :3])
    gxyz = np.ascontiguousarray(pp["ground"][:, :3])
    cyl = cyl_fit_batch(txyz, offsets, n_hyp, thr=thr, max_radius=max_radius, seed=seed,
                        samples=True, threads=threads) if nt else None
    g_off = np.array([0, len(gxyz)], np.int32)
    pl = plane_fit_batch(gxyz, g_off, plane_hyp, thr=thr, seed=seed, tag=1, samples=True,
                         threads=threads)
    est = np.asarray(estimate, np.float64)
    x, y, z, w = est[3:]
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    submap = np.asarray(submap, np.float64)
    pts, ids = [], []
    matches = np.full(nt, -1, np.int32)
    for k in range(nt):
        m = cyl["model"][k]
        if not (0.0 < m[6] <= max_tree_radius) or len(submap) == 0:
            continue
        root = R @ m[:3] + est[:3]
        ray = R @ (m[:3] + m[3:6]) + est[:3] - root
        best, bj = tree_match + 100.0, -1
        for j in range(len(submap)):
            d = _axis_gap(submap[j, :3], submap[j, 3:6], root, ray)
            if d < best:
                best, bj = d, j
        if not best < tree_match:
            continue
        matches[k] = bj
        p = txyz[offsets[k]:offsets[k + 1]][:features_per_tree]
        pts.append(p)
        ids.append(np.full(len(p), bj, np.int32))
    if pl["model"][0, 7] > 0 and len(gxyz):
        inl = [i for i in range(len(gxyz)) if (pl["mask"][i >> 5] >> (i & 31)) & 1]
        if inl:
            step = max(1, -(-len(inl) // ground_features))
            g = gxyz[inl[::step]]
            pts.append(g)
            ids.append(np.full(len(g), -1, np.int32))
    P = np.concatenate(pts) if pts else np.zeros((0, 3), np.float32)
    I = np.concatenate(ids) if ids else np.zeros(0, np.int32)
    rec = np.empty((len(P), 4), np.float32)
    rec[:, :3] = P
    rec[:, 3] = I.view(np.float32)
    table = np.zeros((max(len(submap), 1), 8))
    table[:len(submap), :7] = submap
    prob = {"points": rec, "res_offsets": np.array([0, len(rec)], np.int32), "cyl": table,
            "planes": np.asarray(map_plane, np.float64).reshape(1, 4), "pose_init": est[None]}
    sol = pose_lm_batch(prob, threads=1, max_iters=50, stop_on_convergence=True)
    return {"pose": sol["pose"][0], "status": sol["status"][0], "matches": matches,
            "proj": pp, "trellis": tr, "tree_offsets": offsets, "tree_xyz": txyz,
            "ground_xyz": gxyz, "cyl": cyl, "plane": pl, "problem": prob}


def learned_fit(params, xyz, offsets, n_hyp=16, plain_hyp=1024, thr=0.05, max_radius=1.0,
                min_inlier_ratio=0.5, seed=0x5EED, threads=1):
    """The learned fit on the CPU (the CPU baseline of the GPU's
    sloam_cyl_learned_exec): proposals of the network in FP64, n_hyp
    verifying hypotheses with the proposal as hypothesis 0 and LO refit;
    the clusters whose best verifying hypothesis holds fewer than
    min_inlier_ratio x n points are fitted again by the plain
    plain_hyp-hypothesis method, and so are those (n >= 3) whose verified
    model is invalid or keeps fewer than min_inlier_ratio x n inliers after
    its refit (the rescue of sloam_cyl_learned_exec)."""
    off = np.ascontiguousarray(offsets, np.int32)
    prop = pn_forward(params, xyz, off, emulate=False)["proposal"].astype(np.float32)
    v = cyl_fit_batch(xyz, off, n_hyp, thr=thr, max_radius=max_radius, seed=seed, tag=1,
                      counts=True, mask=False, threads=threads, proposal=prop)
    n = np.diff(off)
    S = len(n)
    best_count = v["counts"][np.arange(S), v["best"]]
    decided = (n >= 3) & (best_count < min_inlier_ratio * n)
    vm = v["model"]
    rescued = ~decided & (n >= 3) & ~((vm[:, 6] > 0) & (vm[:, 7] >= min_inlier_ratio * n))
    fb = np.nonzero(decided | rescued)[0]
    model = v["model"].copy()
    if len(fb):
        pts = np.concatenate([xyz[off[c]:off[c + 1]] for c in fb])
        sub = np.concatenate([[0], np.cumsum(n[fb])]).astype(np.int32)
        p = cyl_fit_batch(pts, sub, plain_hyp, thr=thr, max_radius=max_radius, seed=seed, tag=0,
                          counts=False, mask=False, threads=threads)
        model[fb] = p["model"]
    return {"model": model, "n_fallback": len(fb)}


def keyframe_step(est, xyzi, label, pose, submap, H=128, W=2048, n_hyp=1024, thr=0.05,
                  max_radius=1.0, seed=0x5EED, ground=None, threads=1):
    """One keyframe of the keyframe stream on the CPU (the CPU baseline of the
    GPU's KeyframeRunner with the plain tree model): projection and
    partition, tree instances, ground cells (sloam.yaml: 2 x 18 bins of
    5-25 m, bottom 5 %, 5 features), the plain cylinder fit of every tree's
    points, then RunSloam (est: a SloamEstimator holding the state) against
    the submap [K, 7]."""
    g = dict(radii_bins=2, theta_bins=18, min_dist=5.0, max_dist=25.0, retain_thresh=0.05,
             num_features=5)
    if ground:
        g.update(ground)
    pp = project_partition(xyzi, label, H=H, W=W)
    tr = trellis(pp["tree"], H, W)
    cells = ground_planes(pp["ground"], **g)
    nt = len(tr["tree_offsets"]) - 1
    vo = tr["point_offsets"]
    offsets = vo[tr["tree_offsets"]].astype(np.int32) if nt else np.zeros(1, np.int32)
    if nt:
        fit = cyl_fit_batch(np.ascontiguousarray(tr["points"][:, :3]), offsets, n_hyp, thr=thr,
                            max_radius=max_radius, seed=seed, counts=False, mask=False,
                            threads=threads)
        model = fit["model"]
    else:
        model = np.zeros((0, 8))
    rec = {"pose": np.asarray(pose, np.float64), "nf": g["num_features"],
           "cell_model": cells["model"][:, :7], "cell_valid": cells["info"][:, 0].copy(),
           "cell_feat": cells["features"], "tree_off": tr["tree_offsets"],
           "vtx_int": np.ascontiguousarray(tr["vertex_int"][:, [0, 3]]),
           "vtx_f64": tr["vertex_f64"].astype(np.float64), "vtx_pt_off": vo,
           "pts": tr["points"], "line_int": np.zeros((0, 2), np.int32),
           "line_coeff": np.zeros((0, 6), np.float32),
           "cyl_ok": ((model[:, 6] > 0) & (model[:, 7] >= 3)).astype(np.int32) if nt else
           np.zeros(0, np.int32),
           "cyl_fit": np.ascontiguousarray(model[:, :7]),
           "map": np.asarray(submap, np.float64).reshape(-1, 7)}
    return est.run(rec, tree_mode=2 if nt else 0)
