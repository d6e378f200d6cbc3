"""One keyframe of the plain method on a small scan, on the GPU through the C
ABI: a labelled H x W scan is projected and partitioned
(sloam_project_partition), its trees are clustered and sliced into beam
vertices (sloam_trellis), every tree's points get a point-level cylinder
RANSAC fit (sloam_cyl_plan: n_hyp hypotheses, thr inlier band, LO refit),
the ground list a point-level plane RANSAC fit (sloam_plane_plan), and the
pose is solved against a submap of map trees and the map's ground plane
(sloam_pose_plan: Huber(0.1) Levenberg-Marquardt, quaternion, 50
iterations, stop at convergence).

The reference's per-keyframe path is SLOAMNode::run -> RunSloam
(sloam/src/core/sloamNode.cpp:186-282, sloam/src/core/sloam.cpp:453-532);
its tree model is a line through the vertex medians
(sloam/src/objects/cylinder.cpp:75-173) and its ground model a PCA plane
per cell (sloam/src/objects/plane.cpp:96-128).  This keyframe replaces both
with point-level RANSAC fits, and associates each fitted tree with the map
tree nearest by the reference's model distance (Cylinder::distance,
cylinder.cpp:175-194; matchFeatures' nearest-below-threshold rule,
sloam.cpp:258-285) from the pose estimate.  Residuals: the reference's tree
features (the first features_per_tree points of each matched tree,
cylinder.cpp:86-96,172) against the map cylinder, and ground_features
points of the ground fit's inliers (every k-th) against the map plane.
"""
import math

import numpy as np
import torch

from .gpu import CylinderFitPlan, PlaneFitPlan, PosePlan, _ptr, check, project_partition_dev, \
    trellis_dev


def quat_to_R(q):
    """q = (x, y, z, w)."""
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def model_distance(root_a, ray_a, root_b, ray_b):
    """Cylinder::distance(model): the mean over heights 0, 3 and 6 of the gap
    between the two axes' points at that height (cylinder.cpp:175-194).
    root_b / ray_b may hold many models ([k, 3])."""
    d = 0.0
    for h in (0.0, 3.0, 6.0):
        pa = root_a + ((h - root_a[2]) / ray_a[2]) * ray_a
        pb = root_b + ((h - root_b[..., 2:3]) / ray_b[..., 2:3]) * ray_b
        d = d + np.linalg.norm(pb - pa, axis=-1)
    return d / 3.0


def associate(tree_models, tree_points, tree_offsets, ground_model, ground_points, ground_mask,
              estimate, submap, map_plane, max_tree_radius=0.45, tree_match=1.0,
              features_per_tree=20, ground_features=256):
    """Builds the pose problem of one keyframe from its fits.

    tree_models [T, 8] (sloam_cyl_plan model records, sensor frame),
    tree_points [N, 3] with tree_offsets [T + 1]; ground_model [8] (plane
    fit record), ground_points [G, 3], ground_mask (bit per ground point);
    estimate (t, q x y z w); submap [K, 7] map trees (root, ray, radius);
    map_plane [4].  A tree takes part if its fit is valid with radius <=
    max_tree_radius; it is matched to the nearest map tree (strict <, first
    wins) if that distance is below tree_match.  Returns the problem in the
    sloam_pose_plan layout and the matches (map index per tree, -1)."""
    R = quat_to_R(estimate[3:])
    t = estimate[:3]
    T = len(tree_models)
    matches = np.full(T, -1, np.int32)
    pts, ids = [], []
    for k in range(T):
        m = tree_models[k]
        if not (0.0 < m[6] <= max_tree_radius) or len(submap) == 0:
            continue
        root = R @ m[:3] + t
        ray = R @ (m[:3] + m[3:6]) + t - root
        d = model_distance(root, ray, submap[:, :3], submap[:, 3:6])
        j = int(np.argmin(d))  # first minimum
        if not d[j] < tree_match:
            continue
        matches[k] = j
        p = tree_points[tree_offsets[k]:tree_offsets[k + 1]][:features_per_tree]
        pts.append(p)
        ids.append(np.full(len(p), j, np.int32))
    g = np.zeros((0, 3), np.float32)
    if ground_model[7] > 0 and len(ground_points):
        bits = np.unpackbits(ground_mask.view(np.uint8), bitorder="little")[:len(ground_points)]
        inl = np.nonzero(bits)[0]
        if len(inl):
            step = max(1, int(math.ceil(len(inl) / ground_features)))
            g = ground_points[inl[::step]]
    pts.append(g)
    ids.append(np.full(len(g), -1, np.int32))
    P = np.concatenate(pts) if pts else np.zeros((0, 3), np.float32)
    I = np.concatenate(ids) if ids else np.zeros(0, np.int32)
    rec = np.empty((len(P), 4), np.float32)
    rec[:, :3] = P
    rec[:, 3] = I.view(np.float32)
    cyl = np.zeros((max(len(submap), 1), 8))
    if len(submap):
        cyl[:len(submap), :7] = submap
    return {
        "points": rec, "res_offsets": np.array([0, len(rec)], np.int32), "cyl": cyl,
        "cyl_offsets": np.array([0, len(cyl)], np.int32),
        "planes": np.asarray(map_plane, np.float64).reshape(1, 4),
        "plane_offsets": np.array([0, 1], np.int32),
        "pose_init": np.asarray(estimate, np.float64).reshape(1, 7),
    }, matches


# the reference's OptimizePose settings (sloam.cpp:174-255): LM, Huber 0.1,
# at most 50 iterations, stop at convergence, quaternion parametrisation
POSE_CFG = dict(max_iters=50, stop_on_convergence=True, parametrization=0)


class PlainKeyframe:
    """The small-scan plain-method keyframe on one context.  Plans and
    buffers are kept across run() calls (plans are re-laid out per scan)."""

    def __init__(self, ctx, H=64, W=2048, n_hyp=200, thr=0.05, max_radius=1.0, plane_hyp=200,
                 max_tree_radius=0.45, tree_match=1.0, features_per_tree=20,
                 ground_features=256, seed=0x5EED, fov_up=22.5, fov_down=-22.5):
        self.ctx = ctx
        self.H, self.W = H, W
        self.n_hyp, self.thr, self.max_radius = n_hyp, thr, max_radius
        self.plane_hyp = plane_hyp
        self.max_tree_radius, self.tree_match = max_tree_radius, tree_match
        self.features_per_tree, self.ground_features = features_per_tree, ground_features
        self.seed = seed
        self.fov = (fov_up, fov_down)
        dev = torch.device("cuda", ctx.device)
        N = H * W
        self.dev = dev
        self.proj_out = {
            "mask": torch.empty((H, W), dtype=torch.uint8, device=dev),
            "ground": torch.empty((N, 4), dtype=torch.float32, device=dev),
            "n_ground": torch.empty((1,), dtype=torch.int32, device=dev),
            "tree": torch.empty((N, 4), dtype=torch.float32, device=dev),
        }
        self.tr_out = None
        self.cyl = None
        self.tree_xyz = torch.empty((max(N, 1), 3), dtype=torch.float32, device=dev)
        self.ground_xyz = torch.empty((max(N, 1), 3), dtype=torch.float32, device=dev)

    def close(self):
        if self.cyl is not None:
            self.cyl.close()
            self.cyl = None

    def run(self, xyzi, label, estimate, submap, map_plane, keep=False):
        """xyzi [H*W, 4] float32 and label [H*W] uint8 device tensors;
        estimate (t, q x y z w); submap [K, 7]; map_plane [4].  Returns the
        pose (t, q), the LM status, the matches and (keep=True) every
        intermediate result as numpy arrays."""
        ctx, H, W = self.ctx, self.H, self.W
        lib = ctx.lib
        project_partition_dev(ctx, xyzi, label, H=H, W=W, fov_up=self.fov[0],
                              fov_down=self.fov[1], out=self.proj_out)
        self.tr_out = trellis_dev(ctx, self.proj_out["tree"], H, W, out=self.tr_out)
        ctx.sync()
        cnt = self.tr_out["counts"].cpu().numpy()
        nt, nv = int(cnt[0]), int(cnt[1])
        n_ground = int(self.proj_out["n_ground"].cpu().numpy()[0])
        tree_off_v = self.tr_out["tree_offsets"][:nt + 1].cpu().numpy()
        pt_off = self.tr_out["point_offsets"][:nv + 1].cpu().numpy()
        # per-tree point ranges: the points of the tree's vertices
        offsets = pt_off[tree_off_v].astype(np.int32) if nt else np.zeros(1, np.int32)
        n_pts = int(offsets[-1])
        # xyz of the trellis points and of the ground list
        check(lib.sloam_xyz_from_xyzi(ctx.handle, _ptr(self.tr_out["points"]), n_pts,
                                      _ptr(self.tree_xyz)), ctx.handle)
        check(lib.sloam_xyz_from_xyzi(ctx.handle, _ptr(self.proj_out["ground"]), n_ground,
                                      _ptr(self.ground_xyz)), ctx.handle)
        g_off = np.array([0, n_ground], np.int32)
        off_d = torch.as_tensor(offsets).to(self.dev)
        g_off_d = torch.as_tensor(g_off).to(self.dev)
        torch.cuda.synchronize()  # the copies run on torch's stream
        cyl_out = None
        if nt:
            if self.cyl is None:
                self.cyl = CylinderFitPlan(ctx, offsets, n_hyp=self.n_hyp, thr=self.thr,
                                           max_radius=self.max_radius, seed=self.seed)
            else:
                self.cyl.reset(offsets)
            cyl_out = self.cyl.alloc_outputs(counts=keep, samples=keep, mask=keep, best=keep)
            self.cyl.exec(self.tree_xyz[:n_pts], off_d, cyl_out)
        plane = PlaneFitPlan(ctx, g_off, n_hyp=self.plane_hyp, thr=self.thr, seed=self.seed)
        pl_out = plane.alloc_outputs(counts=keep, samples=keep, mask=True, best=keep)
        if n_ground >= 3:
            plane.exec(self.ground_xyz[:n_ground], g_off_d, pl_out)
        ctx.keep(off_d, g_off_d)
        ctx.sync()
        models = cyl_out["model"].cpu().numpy() if nt else np.zeros((0, 8))
        gmodel = pl_out["model"].cpu().numpy()[0] if n_ground >= 3 else np.zeros(8)
        gmask = pl_out["mask"].cpu().numpy().view(np.uint32)
        tree_xyz = self.tree_xyz[:n_pts].cpu().numpy()
        ground_xyz = self.ground_xyz[:n_ground].cpu().numpy()
        plane.close()
        prob, matches = associate(models, tree_xyz, offsets, gmodel, ground_xyz, gmask,
                                  np.asarray(estimate, np.float64), np.asarray(submap),
                                  map_plane, self.max_tree_radius, self.tree_match,
                                  self.features_per_tree, self.ground_features)
        t = {k: torch.as_tensor(np.ascontiguousarray(prob[k])).to(self.dev)
             for k in ("points", "cyl", "cyl_offsets", "planes", "plane_offsets", "pose_init")}
        torch.cuda.synchronize()
        pp = PosePlan(ctx, prob["res_offsets"], len(prob["cyl"]), 1, **POSE_CFG)
        pout = pp.alloc_outputs()
        pp.exec(t["points"], t["cyl"], t["cyl_offsets"], t["planes"], t["plane_offsets"],
                t["pose_init"], pout)
        ctx.keep(*t.values(), *pout.values())
        ctx.sync()
        res = {"pose": pout["pose"].cpu().numpy()[0], "status": pout["status"].cpu().numpy()[0],
               "matches": matches, "n_trees": nt, "n_ground": n_ground}
        pp.close()
        if keep:
            res.update({
                "tree_offsets": offsets, "tree_xyz": tree_xyz, "ground_xyz": ground_xyz,
                "cyl": {k: v.cpu().numpy() for k, v in cyl_out.items()} if nt else None,
                "models": models, "ground_model": gmodel,
                "plane": {k: v.cpu().numpy() for k, v in pl_out.items()},
                "problem": prob,
            })
        return res
