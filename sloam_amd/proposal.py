"""Learned cylinder proposal: the PointNet of sloam_amd/csrc/pointnet.hip
(sloam_pn_*), its training loop and the learned fitting pipeline

    proposals (sloam_pn_infer) -> a few hypotheses seeded by them
    (sloam_cyl_plan_exec_ex with the proposal: hypothesis 0 is the proposed
    circle, the rest are samples around the proposed axis) -> LO refit ->
    clusters keeping fewer than `min_inlier_ratio` of their points fall back
    to the full plain RANSAC (sloam_cyl_fallback + exec_ex with d_active).

Everything runs on the device, on the library context's stream."""
import ctypes as C
import os

import numpy as np
import torch

from ._lib import PN_PARAMS, PN_PROPOSAL_STRIDE, PnTrainCfg, check
from .gpu import CylinderFitPlan, _ptr

# (fan_in, fan_out, offset of W, offset of b) per layer, PyTorch order
LAYERS = []
_o = 0
for _fi, _fo in ((3, 64), (64, 128), (128, 256), (256, 128), (128, 7)):
    LAYERS.append((_fi, _fo, _o, _o + _fi * _fo))
    _o += _fi * _fo + _fo
assert _o == PN_PARAMS

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                               "cylinder_proposal.npy")


def init_params(seed=0):
    """He-uniform weights for the ReLU layers, small output layer, zero
    biases (float32, numpy)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PN_PARAMS, np.float32)
    for li, (fi, fo, ow, ob) in enumerate(LAYERS):
        bound = np.sqrt(6.0 / fi) if li < len(LAYERS) - 1 else 0.5 / np.sqrt(fi)
        p[ow:ow + fi * fo] = rng.uniform(-bound, bound, fi * fo)
    return p


def load_weights(path=DEFAULT_WEIGHTS):
    w = np.load(path, allow_pickle=False)
    if w.dtype != np.float32 or w.shape != (PN_PARAMS,):
        raise ValueError(f"{path}: expected float32 [{PN_PARAMS}]")
    return w


def train_cfg(lr=1e-3, w_axis=1.0, w_centre=4.0, w_radius=16.0):
    return PnTrainCfg(lr=lr, beta1=0.9, beta2=0.999, eps=1e-8, w_axis=w_axis, w_centre=w_centre,
                      w_radius=w_radius, reserved=0)


class ProposalNet:
    """FP32 master parameters and Adam state on the device.  The calls are
    asynchronous on the library context's stream: tensors passed in must stay
    alive (not be freed to torch's allocator) until the context has been
    synchronised."""

    def __init__(self, ctx, params=None):
        self.ctx = ctx
        self.lib = ctx.lib
        dev = torch.device("cuda", ctx.device)
        p = init_params() if params is None else np.ascontiguousarray(params, np.float32)
        self.params = torch.as_tensor(p.copy()).to(dev)
        self.m = torch.zeros(PN_PARAMS, dtype=torch.float32, device=dev)
        self.v = torch.zeros(PN_PARAMS, dtype=torch.float32, device=dev)
        self.step = 0

    def infer(self, xyz, offsets, out=None):
        """Proposals [S, 8] (device) for device xyz [N, 3] / offsets [S + 1]."""
        S = offsets.numel() - 1
        if out is None:
            out = torch.empty((max(S, 1), PN_PROPOSAL_STRIDE), dtype=torch.float32,
                              device=xyz.device)
        check(self.lib.sloam_pn_infer(self.ctx.handle, _ptr(self.params), _ptr(xyz),
                                      _ptr(offsets), S, _ptr(out)), self.ctx.handle)
        self.ctx.keep(xyz, offsets, out)
        return out

    def loss_grad(self, xyz, offsets, targets, cfg=None):
        cfg = cfg or train_cfg()
        S = offsets.numel() - 1
        loss = torch.empty(S, dtype=torch.float32, device=xyz.device)
        grad = torch.empty(PN_PARAMS, dtype=torch.float32, device=xyz.device)
        check(self.lib.sloam_pn_loss_grad(self.ctx.handle, _ptr(self.params), _ptr(xyz),
                                          _ptr(offsets), S, _ptr(targets), C.byref(cfg),
                                          _ptr(loss), _ptr(grad)), self.ctx.handle)
        self.ctx.keep(xyz, offsets, targets, loss, grad)
        return loss, grad

    def train_step(self, xyz, offsets, targets, cfg=None, loss=None):
        cfg = cfg or train_cfg()
        S = offsets.numel() - 1
        self.step += 1
        check(self.lib.sloam_pn_train_step(self.ctx.handle, _ptr(self.params), _ptr(self.m),
                                           _ptr(self.v), self.step, C.byref(cfg), _ptr(xyz),
                                           _ptr(offsets), S, _ptr(targets), _ptr(loss)),
              self.ctx.handle)
        self.ctx.keep(xyz, offsets, targets, loss)
        return loss


class LearnedCylinderFit:
    """Learned proposal + `n_hyp` verifying hypotheses + LO refit, with the
    plain `plain_hyp`-hypothesis RANSAC as the fallback, for one fixed set of
    cluster offsets (plans are built once)."""

    def __init__(self, ctx, net, offsets, n_hyp=16, plain_hyp=1024, thr=0.05, max_radius=1.0,
                 min_inlier_ratio=0.5, seed=0x5EED, verify_refit_rounds=None, **kw):
        """verify_refit_rounds: LO refit rounds of the clusters that pass
        verification (default: the plain fit's); the fallback clusters always
        refit as the plain fit does."""
        self.ctx = ctx
        self.net = net
        self.min_inlier_ratio = float(min_inlier_ratio)
        vkw = dict(kw)
        if verify_refit_rounds is not None:
            vkw["refit_rounds"] = int(verify_refit_rounds)
        self.verify = CylinderFitPlan(ctx, offsets, n_hyp=n_hyp, thr=thr, max_radius=max_radius,
                                      seed=seed, stream_tag=1, **vkw)
        self.plain = CylinderFitPlan(ctx, offsets, n_hyp=plain_hyp, thr=thr,
                                     max_radius=max_radius, seed=seed, stream_tag=0, **kw)
        self.n_clusters = self.verify.n_clusters

    def alloc(self):
        dev = torch.device("cuda", self.ctx.device)
        S = max(self.n_clusters, 1)
        return {"model": torch.empty((S, 8), dtype=torch.float64, device=dev),
                "proposal": torch.empty((S, PN_PROPOSAL_STRIDE), dtype=torch.float32, device=dev),
                "active": torch.empty((S,), dtype=torch.int32, device=dev),
                "n_fallback": torch.empty((1,), dtype=torch.int32, device=dev)}

    def exec(self, xyz, offsets, out):
        """Enqueues the learned fit (asynchronous, sloam_cyl_learned_exec: one
        seed, the verifying hypotheses for every cluster, the plain ones for
        the clusters whose best verifying hypothesis holds fewer than
        min_inlier_ratio x n inliers, one refit; then the plain fit again for
        the verified clusters whose refit failed or kept fewer than
        min_inlier_ratio x n inliers); out["active"] flags the clusters that
        took the plain fit and out["n_fallback"] counts them."""
        self.net.infer(xyz, offsets, out["proposal"])
        check(self.ctx.lib.sloam_cyl_learned_exec(
            self.verify.handle, self.plain.handle, _ptr(xyz), _ptr(offsets), _ptr(out["proposal"]),
            C.c_float(self.min_inlier_ratio), None, None, _ptr(out["active"]),
            _ptr(out["n_fallback"]), None, _ptr(out["model"])), self.ctx.handle)
        self.ctx.keep(xyz, offsets, *out.values())
        return out

    def exec_two_pass(self, xyz, offsets, out):
        """The learned fit as two full passes: verify (seed, score, refit) for
        every cluster, then the plain method again for the clusters whose
        refitted model keeps fewer than min_inlier_ratio x n inliers."""
        self.net.infer(xyz, offsets, out["proposal"])
        self.verify.exec(xyz, offsets, {"model": out["model"]}, proposal=out["proposal"])
        check(self.ctx.lib.sloam_cyl_fallback(self.ctx.handle, _ptr(offsets), self.n_clusters,
                                              _ptr(out["model"]),
                                              C.c_float(self.min_inlier_ratio),
                                              _ptr(out["active"]), _ptr(out["n_fallback"])),
              self.ctx.handle)
        self.plain.exec(xyz, offsets, {"model": out["model"]}, active=out["active"])
        self.ctx.keep(xyz, offsets, *out.values())
        return out

    def close(self):
        self.verify.close()
        self.plain.close()
