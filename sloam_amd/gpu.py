"""Python front end of the C ABI: contexts and batched fitting plans.

Device buffers are torch CUDA (HIP) tensors; torch is used for device memory
only, every computation runs in the library's own kernels.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import (CYL_MODEL_STRIDE, CylFitCfg, FrontendCfg, FrontendOut, GroundCfg, LineFitCfg,
                   PlaneFitCfg, PoseCfg, ProjCfg, TrellisCfg, check)


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("device tensor expected")
    if not t.is_contiguous():
        raise ValueError("contiguous tensor expected")
    return C.c_void_p(t.data_ptr())


class Context:
    """One HIP stream + scratch on one device (sloam_ctx)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.sloam_ctx_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device
        self._keep = []

    def close(self):
        if self.handle:
            self.lib.sloam_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def keep(self, *tensors):
        """Holds tensors that enqueued work reads or writes until the next
        sync(): the context's stream runs asynchronously, and a tensor freed
        before then could be handed to another allocation by torch."""
        self._keep.extend(t for t in tensors if t is not None)

    def sync(self):
        check(self.lib.sloam_sync(self.handle), self.handle)
        self._keep.clear()

    @property
    def stream_ptr(self):
        return self.lib.sloam_ctx_stream(self.handle)

    def torch_stream(self):
        """The context's stream as a torch ExternalStream (for events)."""
        return torch.cuda.ExternalStream(self.stream_ptr, device=torch.device("cuda", self.device))

    def enable_timing(self, on=True):
        check(self.lib.sloam_ctx_enable_timing(self.handle, 1 if on else 0), self.handle)

    def reset_timing(self):
        check(self.lib.sloam_ctx_reset_timing(self.handle), self.handle)

    def kernel_time(self, name):
        ms = C.c_double()
        n = C.c_int64()
        check(self.lib.sloam_ctx_kernel_time(self.handle, name.encode(), C.byref(ms), C.byref(n)),
              self.handle)
        return ms.value, n.value

    def graph_begin(self):
        """Starts recording the library calls made on this context into a HIP
        graph (sloam_graph_begin); run the sequence once before capturing it."""
        check(self.lib.sloam_graph_begin(self.handle), self.handle)
        self._capture_from = len(self._keep)

    def graph_end(self):
        """Ends the recording; returns the Graph, which holds the tensors the
        recorded calls use."""
        h = C.c_void_p()
        check(self.lib.sloam_graph_end(self.handle, C.byref(h)), self.handle)
        return Graph(self, h, self._keep[self._capture_from:])


class Graph:
    """A recorded sequence of library calls, replayed as one launch on the
    context's stream over the same buffers (sloam_graph_*)."""

    def __init__(self, ctx, handle, refs):
        self.ctx = ctx
        self.handle = handle
        self.refs = list(refs)

    def launch(self):
        check(self.ctx.lib.sloam_graph_launch(self.handle), self.ctx.handle)
        self.ctx.keep(*self.refs)

    def close(self):
        if self.handle:
            self.ctx.lib.sloam_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_cylinder_fit(ctx, xyz, offsets_dev, offsets_host, n_hyp=1024, thr=0.05,
                          max_radius=1.0, refit_rounds=3, refit_iters=2, seed=0x5EED,
                          stream_tag=0):
    """(hyp_chunk, tile_points) of the fastest launch shape for this batch
    (sloam_cyl_autotune; cached per batch shape)."""
    off = np.ascontiguousarray(offsets_host, dtype=np.int32)
    cfg = CylFitCfg(n_hyp=n_hyp, thr=thr, max_radius=max_radius, refit_iters=refit_iters,
                    seed=seed, stream_tag=stream_tag, hyp_chunk=0, tile_points=0,
                    refit_rounds=refit_rounds)
    out = CylFitCfg()
    torch.cuda.synchronize()
    check(ctx.lib.sloam_cyl_autotune(ctx.handle, off.ctypes.data_as(C.c_void_p), len(off) - 1,
                                     C.byref(cfg), _ptr(xyz), _ptr(offsets_dev), C.byref(out)),
          ctx.handle)
    return out.hyp_chunk, out.tile_points


class CylinderFitPlan:
    """Batched point-level cylinder RANSAC over clusters (sloam_cyl_plan)."""

    def __init__(self, ctx, offsets, n_hyp=1024, thr=0.05, max_radius=1.0, refit_rounds=3, refit_iters=2,
                 seed=0x5EED, stream_tag=0, hyp_chunk=0, tile_points=0):
        self.ctx = ctx
        self.lib = ctx.lib
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        self.offsets_host = off
        self.n_clusters = len(off) - 1
        self.n_points = int(off[-1])
        self.cfg = CylFitCfg(n_hyp=n_hyp, thr=thr, max_radius=max_radius,
                             refit_iters=refit_iters, seed=seed, stream_tag=stream_tag,
                             hyp_chunk=hyp_chunk, tile_points=tile_points,
                             refit_rounds=refit_rounds)
        h = C.c_void_p()
        check(self.lib.sloam_cyl_plan_create(ctx.handle, off.ctypes.data_as(C.c_void_p),
                                             self.n_clusters, C.byref(self.cfg), C.byref(h)),
              ctx.handle)
        self.handle = h

    def close(self):
        if self.handle:
            # a plan outliving its context is left to the process's end
            # (its destroy would wait on the context's freed stream)
            if self.ctx.handle:
                self.lib.sloam_cyl_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, offsets):
        """Lays out another batch on the same plan (sloam_cyl_plan_reset:
        waits for the context stream; device buffers only grow)."""
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        st = self.lib.sloam_cyl_plan_reset(self.handle, off.ctypes.data_as(C.c_void_p),
                                           len(off) - 1)
        if st != 0:
            self.offsets_host = np.zeros(1, np.int32)
            self.n_clusters = 0
            self.n_points = 0
            check(st, self.ctx.handle)
        self.offsets_host = off
        self.n_clusters = len(off) - 1
        self.n_points = int(off[-1])

    def alloc_outputs(self, counts=False, samples=False, mask=True, best=True):
        dev = torch.device("cuda", self.ctx.device)
        out = {"model": torch.empty((self.n_clusters, CYL_MODEL_STRIDE), dtype=torch.float64,
                                    device=dev)}
        H = self.cfg.n_hyp
        if counts:
            out["counts"] = torch.empty((self.n_clusters, H), dtype=torch.int32, device=dev)
        if samples:
            out["samples"] = torch.empty((self.n_clusters, H, 3), dtype=torch.int32, device=dev)
        if mask:
            out["mask"] = torch.empty(((self.n_points + 31) // 32,), dtype=torch.int32,
                                      device=dev)
        if best:
            out["best"] = torch.empty((self.n_clusters,), dtype=torch.int32, device=dev)
        return out

    def exec(self, xyz, offsets, out, proposal=None, active=None):
        """Enqueues one batch on the context stream (asynchronous).  proposal:
        float32 [n_clusters, 8] (sloam_pn_infer); active: int32 [n_clusters]
        (clusters with 0 are skipped)."""
        if xyz.dtype != torch.float32 or xyz.shape != (self.n_points, 3):
            raise ValueError("xyz must be float32 [n_points, 3]")
        if offsets.dtype != torch.int32 or offsets.numel() != self.n_clusters + 1:
            raise ValueError("offsets must be int32 [n_clusters + 1]")
        if proposal is None and active is None:
            check(self.lib.sloam_cyl_plan_exec(
                self.handle, _ptr(xyz), _ptr(offsets), _ptr(out.get("counts")),
                _ptr(out.get("samples")), _ptr(out.get("mask")), _ptr(out.get("best")),
                _ptr(out["model"])), self.ctx.handle)
        else:
            check(self.lib.sloam_cyl_plan_exec_ex(
                self.handle, _ptr(xyz), _ptr(offsets), _ptr(proposal), _ptr(active),
                _ptr(out.get("counts")), _ptr(out.get("samples")), _ptr(out.get("mask")),
                _ptr(out.get("best")), _ptr(out["model"])), self.ctx.handle)
        self.ctx.keep(xyz, offsets, proposal, active, *out.values())
        return out


def fit_cylinders(ctx, xyz, offsets_host, **kw):
    """One-shot convenience: plan, run, sync; returns numpy outputs."""
    want = {k: kw.pop(k) for k in ("counts", "samples", "mask", "best") if k in kw}
    plan = CylinderFitPlan(ctx, offsets_host, **kw)
    dev = torch.device("cuda", ctx.device)
    xyz_d = torch.as_tensor(np.ascontiguousarray(xyz, dtype=np.float32)).to(dev)
    off_d = torch.as_tensor(np.ascontiguousarray(offsets_host, dtype=np.int32)).to(dev)
    out = plan.alloc_outputs(**want)
    torch.cuda.synchronize()
    plan.exec(xyz_d, off_d, out)
    ctx.sync()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    if "mask" in res:
        res["mask"] = res["mask"].view(np.uint32)
    plan.close()
    return res


class PosePlan:
    """Batched LM pose estimation over independent problems (sloam_pose_plan)."""

    def __init__(self, ctx, res_offsets, n_cylinders, n_planes, max_iters=10,
                 stop_on_convergence=False, huber_delta=0.1, function_tolerance=1e-6,
                 gradient_tolerance=1e-10, parameter_tolerance=1e-8, initial_radius=1e4,
                 min_relative_decrease=1e-3, parametrization=0, mode=0, precision=0):
        self.ctx = ctx
        self.lib = ctx.lib
        off = np.ascontiguousarray(res_offsets, dtype=np.int32)
        self.n_problems = len(off) - 1
        self.n_res = int(off[-1])
        self.cfg = PoseCfg(max_iters=max_iters, stop_on_convergence=int(stop_on_convergence),
                           huber_delta=huber_delta, function_tolerance=function_tolerance,
                           gradient_tolerance=gradient_tolerance,
                           parameter_tolerance=parameter_tolerance,
                           initial_radius=initial_radius,
                           min_relative_decrease=min_relative_decrease,
                           parametrization=parametrization, precision=precision)
        h = C.c_void_p()
        check(self.lib.sloam_pose_plan_create(ctx.handle, off.ctypes.data_as(C.c_void_p),
                                              self.n_problems, int(n_cylinders), int(n_planes),
                                              C.byref(self.cfg), C.byref(h)), ctx.handle)
        self.handle = h
        self.set_mode(mode)

    def set_mode(self, mode):
        """0 auto, 1 batched (two kernels per LM iteration), 2 fused (one
        workgroup per problem runs every iteration in one launch)."""
        check(self.lib.sloam_pose_plan_set_mode(self.handle, int(mode)), self.ctx.handle)

    @property
    def uses_fused(self):
        r = self.lib.sloam_pose_plan_uses_fused(self.handle)
        if r < 0:
            check(r, self.ctx.handle)
        return bool(r)

    def reset(self, res_offsets, n_cylinders, n_planes):
        """Re-targets the plan at another problem set (grow-only buffers)."""
        off = np.ascontiguousarray(res_offsets, dtype=np.int32)
        check(self.lib.sloam_pose_plan_reset(self.handle, off.ctypes.data_as(C.c_void_p),
                                             len(off) - 1, int(n_cylinders), int(n_planes),
                                             None), self.ctx.handle)
        self.n_problems = len(off) - 1
        self.n_res = int(off[-1])

    def close(self):
        if self.handle:
            # a plan outliving its context is left to the process's end
            # (its destroy would wait on the context's freed stream)
            if self.ctx.handle:
                self.lib.sloam_pose_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc_outputs(self):
        dev = torch.device("cuda", self.ctx.device)
        return {
            "pose": torch.empty((self.n_problems, 7), dtype=torch.float64, device=dev),
            "params": torch.empty((self.n_problems, 6), dtype=torch.float64, device=dev),
            "status": torch.empty((self.n_problems, 2), dtype=torch.int32, device=dev),
            "cost": torch.empty((self.n_problems, 2), dtype=torch.float64, device=dev),
        }

    def exec(self, points, cyl, cyl_off, planes, plane_off, pose_in, out, free_mask=None):
        check(self.lib.sloam_pose_plan_exec(
            self.handle, _ptr(points), _ptr(cyl), _ptr(cyl_off), _ptr(planes), _ptr(plane_off),
            _ptr(pose_in), _ptr(free_mask), _ptr(out["pose"]), _ptr(out.get("params")),
            _ptr(out.get("status")), _ptr(out.get("cost"))), self.ctx.handle)
        return out


def solve_poses(ctx, prob, **kw):
    """One-shot convenience over a dict from synth.make_pose_problems."""
    dev = torch.device("cuda", ctx.device)
    plan = PosePlan(ctx, prob["res_offsets"], len(prob["cyl"]), len(prob["planes"]), **kw)
    t = {k: torch.as_tensor(np.ascontiguousarray(prob[k])).to(dev)
         for k in ("points", "cyl", "cyl_offsets", "planes", "plane_offsets", "pose_init")}
    fm = None
    if "free_mask" in prob:
        fm = torch.as_tensor(np.ascontiguousarray(prob["free_mask"], np.int32)).to(dev)
    out = plan.alloc_outputs()
    torch.cuda.synchronize()
    plan.exec(t["points"], t["cyl"], t["cyl_offsets"], t["planes"], t["plane_offsets"],
              t["pose_init"], out, free_mask=fm)
    ctx.sync()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    plan.close()
    return res


def project_partition_dev(ctx, xyzi, label, H=64, W=2048, fov_up=22.5, fov_down=-22.5,
                          destagger=False, out=None):
    """Enqueues the projection + partition of device tensors xyzi [n, 4] float32
    and label [n] uint8; returns the dict of device outputs."""
    n = xyzi.shape[0]
    dev = xyzi.device
    if out is None:
        out = {
            "proj": torch.empty((n, 2), dtype=torch.int32, device=dev),
            "range": torch.empty((H, W), dtype=torch.float32, device=dev),
            "winner": torch.empty((H, W), dtype=torch.int32, device=dev),
            "mask": torch.empty((H, W), dtype=torch.uint8, device=dev),
            "ground": torch.empty((n, 4), dtype=torch.float32, device=dev),
            "n_ground": torch.empty((1,), dtype=torch.int32, device=dev),
            "tree": torch.empty((n, 4), dtype=torch.float32, device=dev),
        }
    cfg = ProjCfg(height=H, width=W, fov_up_deg=fov_up, fov_down_deg=fov_down,
                  destagger=int(destagger))
    check(ctx.lib.sloam_project_partition(
        ctx.handle, _ptr(xyzi), _ptr(label), n, C.byref(cfg), _ptr(out.get("proj")),
        _ptr(out.get("range")), _ptr(out.get("winner")), _ptr(out["mask"]), _ptr(out["ground"]),
        _ptr(out["n_ground"]), _ptr(out.get("tree"))), ctx.handle)
    return out


def project_partition(ctx, xyzi, label, **kw):
    """One-shot convenience on host arrays; returns numpy outputs."""
    dev = torch.device("cuda", ctx.device)
    x = torch.as_tensor(np.ascontiguousarray(xyzi, np.float32)).to(dev)
    lab = torch.as_tensor(np.ascontiguousarray(label, np.uint8)).to(dev)
    torch.cuda.synchronize()
    out = project_partition_dev(ctx, x, lab, **kw)
    ctx.sync()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["ground"] = res["ground"][:int(res["n_ground"][0])]
    return res


def ground_planes_dev(ctx, points, count=None, radii_bins=1, theta_bins=18, min_dist=0.0,
                      max_dist=30.0, retain_thresh=0.05, num_features=60, origin=(0.0, 0.0, 0.0),
                      out=None):
    """Binned per-cell ground plane fits of device points [cap, 4] (optional
    device int32 count); returns device outputs."""
    cap = points.shape[0]
    dev = points.device
    n_cells = radii_bins * theta_bins
    if out is None:
        out = {
            "info": torch.empty((n_cells, 4), dtype=torch.int32, device=dev),
            "model": torch.empty((n_cells, 8), dtype=torch.float64, device=dev),
            "features": torch.empty((n_cells, max(num_features, 1), 4), dtype=torch.float32,
                                    device=dev),
            "kept_offsets": torch.empty((n_cells + 1,), dtype=torch.int32, device=dev),
            "kept_index": torch.empty((max(cap, 1),), dtype=torch.int32, device=dev),
        }
    cfg = GroundCfg(radii_bins=radii_bins, theta_bins=theta_bins, min_dist=min_dist,
                    max_dist=max_dist, retain_thresh=retain_thresh, num_features=num_features)
    cfg.origin[0], cfg.origin[1], cfg.origin[2] = origin
    check(ctx.lib.sloam_ground_planes(ctx.handle, _ptr(points), cap, _ptr(count), C.byref(cfg),
                                      _ptr(out["info"]), _ptr(out["model"]),
                                      _ptr(out["features"]), _ptr(out.get("kept_offsets")),
                                      _ptr(out.get("kept_index"))), ctx.handle)
    return out


def ground_planes(ctx, points, **kw):
    dev = torch.device("cuda", ctx.device)
    x = torch.as_tensor(np.ascontiguousarray(points, np.float32)).to(dev)
    torch.cuda.synchronize()
    out = ground_planes_dev(ctx, x, **kw)
    ctx.sync()
    return {k: v.cpu().numpy() for k, v in out.items()}


class PlaneFitPlan:
    """Batched point-level plane RANSAC over point sets (sloam_plane_plan)."""

    def __init__(self, ctx, offsets, n_hyp=200, thr=0.05, seed=0x5EED, stream_tag=1):
        self.ctx = ctx
        self.lib = ctx.lib
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        self.n_sets = len(off) - 1
        self.n_points = int(off[-1])
        self.cfg = PlaneFitCfg(n_hyp=n_hyp, thr=thr, seed=seed, stream_tag=stream_tag, reserved=0)
        h = C.c_void_p()
        check(self.lib.sloam_plane_plan_create(ctx.handle, off.ctypes.data_as(C.c_void_p),
                                               self.n_sets, C.byref(self.cfg), C.byref(h)),
              ctx.handle)
        self.handle = h

    def close(self):
        if self.handle:
            # a plan outliving its context is left to the process's end
            # (its destroy would wait on the context's freed stream)
            if self.ctx.handle:
                self.lib.sloam_plane_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc_outputs(self, counts=False, samples=False, mask=True, best=True):
        dev = torch.device("cuda", self.ctx.device)
        out = {"model": torch.empty((self.n_sets, 8), dtype=torch.float64, device=dev)}
        H = self.cfg.n_hyp
        if counts:
            out["counts"] = torch.empty((self.n_sets, H), dtype=torch.int32, device=dev)
        if samples:
            out["samples"] = torch.empty((self.n_sets, H, 3), dtype=torch.int32, device=dev)
        if mask:
            out["mask"] = torch.empty(((self.n_points + 31) // 32,), dtype=torch.int32, device=dev)
        if best:
            out["best"] = torch.empty((self.n_sets,), dtype=torch.int32, device=dev)
        return out

    def exec(self, xyz, offsets, out):
        check(self.lib.sloam_plane_plan_exec(
            self.handle, _ptr(xyz), _ptr(offsets), _ptr(out.get("counts")),
            _ptr(out.get("samples")), _ptr(out.get("mask")), _ptr(out.get("best")),
            _ptr(out["model"])), self.ctx.handle)
        return out


def fit_planes(ctx, xyz, offsets_host, **kw):
    want = {k: kw.pop(k) for k in ("counts", "samples", "mask", "best") if k in kw}
    plan = PlaneFitPlan(ctx, offsets_host, **kw)
    dev = torch.device("cuda", ctx.device)
    x = torch.as_tensor(np.ascontiguousarray(xyz, np.float32)).to(dev)
    o = torch.as_tensor(np.ascontiguousarray(offsets_host, np.int32)).to(dev)
    out = plan.alloc_outputs(**want)
    torch.cuda.synchronize()
    plan.exec(x, o, out)
    ctx.sync()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    if "mask" in res:
        res["mask"] = res["mask"].view(np.uint32)
    plan.close()
    return res


def fit_lines(ctx, xyz, offsets_host, threshold=0.25, probability=0.99, max_iterations=50,
              optimize=True, seed=12345):
    """Batched PCL-semantics RANSAC line fits; returns numpy coeff / stats / inliers."""
    dev = torch.device("cuda", ctx.device)
    off = np.ascontiguousarray(offsets_host, np.int32)
    S = len(off) - 1
    N = int(off[-1])
    x = torch.as_tensor(np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)).to(dev)
    o = torch.as_tensor(off).to(dev)
    coeff = torch.empty((max(S, 1), 6), dtype=torch.float32, device=dev)
    stats = torch.empty((max(S, 1), 4), dtype=torch.int32, device=dev)
    inl = torch.empty((max(N, 1),), dtype=torch.uint8, device=dev)
    cfg = LineFitCfg(threshold=threshold, probability=probability, max_iterations=max_iterations,
                     optimize=int(optimize), seed=seed, reserved=0)
    torch.cuda.synchronize()
    check(ctx.lib.sloam_fit_lines(ctx.handle, _ptr(x), _ptr(o), S, C.byref(cfg), _ptr(coeff),
                                  _ptr(stats), _ptr(inl)), ctx.handle)
    ctx.sync()
    return {"coeff": coeff.cpu().numpy()[:S], "stats": stats.cpu().numpy()[:S],
            "inliers": inl.cpu().numpy()[:N]}


TRELLIS_DEFAULTS = dict(cluster_dist=1.0, max_dist_to_centroid=0.2, min_label_points=80,
                        min_row_points=3, min_vertices=16, max_vertices=56, radius_scale=1.0)


def trellis_dev(ctx, cloud, H, W, out=None, **kw):
    """Tree instances of an organized H x W cloud already on the device
    (float32 [H*W, 4]); device outputs, counts left on the device."""
    c = dict(TRELLIS_DEFAULTS)
    c.update(kw)
    dev = torch.device("cuda", ctx.device)
    N = H * W
    cap_t = N // (c["min_label_points"] + 1) + 1
    cap_v = cap_t * c["max_vertices"]
    cap_p = max(N, 1)
    if out is None:
        out = {
            "labels": torch.empty((max(N, 1),), dtype=torch.int32, device=dev),
            "tree_offsets": torch.empty((cap_t + 1,), dtype=torch.int32, device=dev),
            "vertex_int": torch.empty((cap_v, 4), dtype=torch.int32, device=dev),
            "vertex_f64": torch.empty((cap_v, 4), dtype=torch.float64, device=dev),
            "point_offsets": torch.empty((cap_v + 1,), dtype=torch.int32, device=dev),
            "points": torch.empty((cap_p, 4), dtype=torch.float32, device=dev),
            "counts": torch.empty((8,), dtype=torch.int32, device=dev),
        }
    cfg = TrellisCfg(**c)
    check(ctx.lib.sloam_trellis(ctx.handle, _ptr(cloud), H, W, C.byref(cfg), _ptr(out["labels"]),
                                cap_t, cap_v, cap_p, _ptr(out["tree_offsets"]),
                                _ptr(out["vertex_int"]), _ptr(out["vertex_f64"]),
                                _ptr(out["point_offsets"]), _ptr(out["points"]),
                                _ptr(out["counts"])), ctx.handle)
    return out


def trellis(ctx, cloud, H, W, **kw):
    """Host in / host out: numpy arrays trimmed to the counts."""
    dev = torch.device("cuda", ctx.device)
    x = torch.as_tensor(np.ascontiguousarray(cloud, np.float32).reshape(-1, 4)).to(dev)
    torch.cuda.synchronize()
    out = trellis_dev(ctx, x, H, W, **kw)
    ctx.sync()
    cnt = out["counts"].cpu().numpy()
    nt, nv, npt = int(cnt[0]), int(cnt[1]), int(cnt[2])
    return {"labels": out["labels"].cpu().numpy()[:H * W], "n_labels": int(cnt[3]),
            "overflow": int(cnt[4]), "beams": int(cnt[5]),
            "tree_offsets": out["tree_offsets"].cpu().numpy()[:nt + 1],
            "vertex_int": out["vertex_int"].cpu().numpy()[:nv],
            "vertex_f64": out["vertex_f64"].cpu().numpy()[:nv],
            "point_offsets": out["point_offsets"].cpu().numpy()[:nv + 1],
            "points": out["points"].cpu().numpy()[:npt]}


# sloam.yaml ground settings (2 x 18 bins, 5-25 m, bottom 5 %, 5 features)
GROUND_DEFAULTS = dict(radii_bins=2, theta_bins=18, min_dist=5.0, max_dist=25.0,
                       retain_thresh=0.05, num_features=5)


class Frontend:
    """Keyframe front end (sloam_frontend_*): projection + partition, tree
    instances and ground cells of one organised H x W scan on plan-owned
    buffers, optionally replayed as one HIP graph."""

    def __init__(self, ctx, H, W, use_graph=True, fov_up=22.5, fov_down=-22.5, destagger=False,
                 trellis=None, ground=None, host_results=False, fit_lines=False):
        self.ctx = ctx
        self.lib = ctx.lib
        self.H, self.W, self.N = H, W, H * W
        tc = dict(TRELLIS_DEFAULTS)
        tc.update(trellis or {})
        gc = dict(GROUND_DEFAULTS)
        gc.update(ground or {})
        self.nf = gc["num_features"]
        self.cfg = FrontendCfg(proj=ProjCfg(height=H, width=W, fov_up_deg=fov_up,
                                            fov_down_deg=fov_down, destagger=int(destagger)),
                               trellis=TrellisCfg(**tc), ground=GroundCfg(**gc),
                               use_graph=int(use_graph), host_results=int(host_results),
                               line=LineFitCfg(threshold=0.25, probability=0.99,
                                               max_iterations=50, optimize=1, seed=12345,
                                               reserved=0),
                               fit_lines=int(fit_lines), reserved=0)
        self.fit_lines = bool(fit_lines)
        h = C.c_void_p()
        check(self.lib.sloam_frontend_create(ctx.handle, C.byref(self.cfg), C.byref(h)),
              ctx.handle)
        self.handle = h
        self.out = FrontendOut()
        check(self.lib.sloam_frontend_buffers(h, C.byref(self.out)), ctx.handle)
        self._host = []

    def close(self):
        if self.handle:
            # a plan outliving its context is left to the process's end
            # (its destroy would wait on the context's freed stream)
            if self.ctx.handle:
                self.lib.sloam_frontend_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _copy(self, dst, src, nbytes):
        check(self.lib.sloam_memcpy_async(self.ctx.handle, dst, src, nbytes), self.ctx.handle)

    def upload(self, xyzi, label):
        x = np.ascontiguousarray(xyzi, np.float32).reshape(self.N, 4)
        lab = np.ascontiguousarray(label, np.uint8).reshape(self.N)
        self._host = [x, lab]  # kept until the next sync
        self._copy(self.out.xyzi, x.ctypes.data, x.nbytes)
        self._copy(self.out.label, lab.ctypes.data, lab.nbytes)

    def exec(self):
        check(self.lib.sloam_frontend_exec(self.handle), self.ctx.handle)

    def _get(self, ptr, shape, dtype):
        a = np.empty(shape, dtype)
        if a.nbytes:
            self._copy(a.ctypes.data, ptr, a.nbytes)
        return a

    def download(self):
        """Synchronises and returns the outputs as numpy arrays."""
        o = self.out
        cnt = self._get(o.trellis_counts, 8, np.int32)
        ng = self._get(o.n_ground, 1, np.int32)
        self.ctx.sync()
        nt, nv, npt = int(cnt[0]), int(cnt[1]), int(cnt[2])
        nc = o.n_cells
        r = {
            "mask": self._get(o.mask, self.N, np.uint8),
            "ground": self._get(o.ground, (int(ng[0]), 4), np.float32),
            "tree": self._get(o.tree, (self.N, 4), np.float32),
            "trellis_counts": cnt,
            "tree_offsets": self._get(o.tree_offsets, nt + 1, np.int32),
            "vertex_int": self._get(o.vertex_int, (nv, 4), np.int32),
            "vertex_f64": self._get(o.vertex_f64, (nv, 4), np.float64),
            "point_offsets": self._get(o.point_offsets, nv + 1, np.int32),
            "points": self._get(o.points, (npt, 4), np.float32),
            "cell_info": self._get(o.cell_info, (nc, 4), np.int32),
            "cell_model": self._get(o.cell_model, (nc, 8), np.float64),
            "cell_features": self._get(o.cell_features, (nc, self.nf, 4), np.float32),
        }
        if self.fit_lines:
            r["line_coeff"] = self._get(o.line_coeff, (nt, 6), np.float32)
            r["line_stats"] = self._get(o.line_stats, (nt, 4), np.int32)
        self.ctx.sync()
        self._host = []
        return r

    def host_view(self):
        """Synchronises and returns copies of the results the front end wrote
        into its pinned host buffer (plans created with host_results)."""
        self.ctx.sync()
        hv = _lib.FrontendHost()
        check(self.lib.sloam_frontend_host_view(self.handle, C.byref(hv)), self.ctx.handle)

        def arr(ptr, shape, ctype, dtype):
            n = int(np.prod(shape))
            if n == 0:
                return np.empty(shape, dtype)
            p = C.cast(ptr, C.POINTER(ctype))
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True).reshape(shape)

        cnt = arr(hv.trellis_counts, 8, C.c_int32, np.int32)
        nt, nv, npt = int(cnt[0]), int(cnt[1]), int(cnt[2])
        nc = self.out.n_cells
        return {
            "trellis_counts": cnt,
            "tree_offsets": arr(hv.tree_offsets, nt + 1, C.c_int32, np.int32),
            "vertex_int": arr(hv.vertex_int, (nv, 4), C.c_int32, np.int32),
            "vertex_f64": arr(hv.vertex_f64, (nv, 4), C.c_double, np.float64),
            "point_offsets": arr(hv.point_offsets, nv + 1, C.c_int32, np.int32),
            "points": arr(hv.points, (npt, 4), C.c_float, np.float32),
            "cell_info": arr(hv.cell_info, (nc, 4), C.c_int32, np.int32),
            "cell_model": arr(hv.cell_model, (nc, 8), C.c_double, np.float64),
            "cell_features": arr(hv.cell_features, (nc, self.nf, 4), C.c_float, np.float32),
            **({"line_coeff": arr(hv.line_coeff, (nt, 6), C.c_float, np.float32),
                "line_stats": arr(hv.line_stats, (nt, 4), C.c_int32, np.int32)}
               if self.fit_lines else {}),
        }
