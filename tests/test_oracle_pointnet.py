"""CPU tests that pin the FP64 oracle of the learned proposal network: the
analytic gradient against central differences of the loss (pure FP64), the
loss's invariance to the sign of the target axis, zero loss for an invalid
target, and the input normalisation."""
import numpy as np

import oracle
from sloam_amd import proposal, synth


def _batch(S=3, seed=3):
    d = synth.make_clusters(S, seed=seed, n_min=60, n_max=700)
    t = np.zeros((S, 8))
    for c in range(S):
        ax = d["axis"][c] / np.linalg.norm(d["axis"][c])
        t[c, :3] = d["base"][c]
        t[c, 3:6] = ax
        t[c, 6] = d["radius"][c]
    return d, t


def test_gradient_central_differences():
    d, t = _batch()
    p = proposal.init_params(seed=1)
    # non-trivial outputs: enlarge the last layer a little
    _, _, ow, ob = proposal.LAYERS[-1]
    p[ow:ob] *= 8.0
    w = (1.0, 4.0, 16.0)
    _, g = oracle.pn_loss_grad(p, d["xyz"], d["offsets"], t, weights=w, emulate=False)
    rng = np.random.default_rng(0)
    idx = []
    for fi, fo, o_w, o_b in proposal.LAYERS:
        idx += list(rng.choice(np.arange(o_w, o_b), 6, replace=False))
        idx += list(rng.choice(np.arange(o_b, o_b + fo), 2, replace=False))
    checked = 0
    for i in idx:
        h = 1e-4 * max(1.0, abs(float(p[i])))
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[i] += h
        pm[i] -= h
        # parameters are float32: use exactly representable perturbations
        pp = pp.astype(np.float32)
        pm = pm.astype(np.float32)
        hh = (float(pp[i]) - float(pm[i])) / 2
        lp, _ = oracle.pn_loss_grad(pp, d["xyz"], d["offsets"], t, weights=w, emulate=False)
        lm, _ = oracle.pn_loss_grad(pm, d["xyz"], d["offsets"], t, weights=w, emulate=False)
        num = (lp.mean() - lm.mean()) / (2 * hh)
        assert abs(num - g[i]) <= 2e-3 * max(1e-3, abs(g[i]), abs(num)), (i, num, g[i])
        checked += 1
    assert checked == len(idx)


def test_axis_sign_invariance_and_invalid_target():
    d, t = _batch(seed=5)
    p = proposal.init_params(seed=2)
    l1, g1 = oracle.pn_loss_grad(p, d["xyz"], d["offsets"], t)
    t2 = t.copy()
    t2[:, 3:6] *= -1
    l2, g2 = oracle.pn_loss_grad(p, d["xyz"], d["offsets"], t2)
    assert np.allclose(l1, l2, rtol=1e-12, atol=0)
    assert np.allclose(g1, g2, rtol=1e-9, atol=1e-15)
    t3 = t.copy()
    t3[:, 6] = -1.0
    l3, g3 = oracle.pn_loss_grad(p, d["xyz"], d["offsets"], t3)
    assert np.all(l3 == 0) and np.all(g3 == 0)


def test_proposal_frame():
    """The proposal is reported in the world frame: the normalisation is
    undone exactly for the centre and the radius."""
    d, t = _batch(S=2, seed=7)
    p = proposal.init_params(seed=3)
    o = oracle.pn_forward(p, d["xyz"], d["offsets"])
    for c in range(2):
        P = d["xyz"][d["offsets"][c]:d["offsets"][c + 1]]
        m = min(len(P), 512)
        Q = P[(np.arange(m) * len(P)) // m]
        lo, hi = Q.min(0), Q.max(0)
        cen = 0.5 * (lo + hi)
        half = 0.5 * (hi - lo)
        sh = np.float32(1) / max(np.float32(1e-3), half[:2].max())
        sz = np.float32(1) / max(np.float32(1e-3), half[2])
        sc = np.array([sh, sh, sz], np.float64)
        assert np.allclose(o["proposal"][c, 3:6], cen + o["out"][c, 3:6] / sc, rtol=1e-6)
        assert np.isclose(o["proposal"][c, 6], o["out"][c, 6] / sh, rtol=1e-6)
        v = o["out"][c, :3] / sc
        assert np.allclose(o["proposal"][c, :3], v / np.linalg.norm(v), rtol=1e-9)


def _torch_forward(params, xyz, offsets):
    """The network written again in torch, FP64, with the BF16 roundings done
    by torch's own float -> bfloat16 conversion: weights and biases, and the
    layer-1 and layer-2 activations after the ReLU; masked max-pool (channels
    whose outputs are all <= 0 pool to 0)."""
    import torch

    def bf(t):
        return t.to(torch.bfloat16).to(torch.float64)

    W = []
    p = torch.as_tensor(np.asarray(params, np.float32))
    for fi, fo, ow, ob in proposal.LAYERS:
        W.append((bf(p[ow:ow + fi * fo]).reshape(fo, fi), bf(p[ob:ob + fo])))
    outs = []
    for c in range(len(offsets) - 1):
        P = xyz[offsets[c]:offsets[c + 1]].astype(np.float32)
        n = len(P)
        m = min(n, 512)
        Q = P[(np.arange(m, dtype=np.int64) * n) // m]
        lo, hi = Q.min(0), Q.max(0)
        cen = np.float32(0.5) * (lo + hi)
        half = np.float32(0.5) * (hi - lo)
        sh = np.float32(1) / max(np.float32(1e-3), max(half[0], half[1]))
        sz = np.float32(1) / max(np.float32(1e-3), half[2])
        X = ((Q - cen) * np.array([sh, sh, sz], np.float32)).astype(np.float32)
        x = torch.as_tensor(X.astype(np.float64))
        h1 = bf(torch.relu(x @ W[0][0].T + W[0][1]))
        h2 = bf(torch.relu(h1 @ W[1][0].T + W[1][1]))
        s3 = h2 @ W[2][0].T + W[2][1]
        f = torch.clamp(s3.max(0).values, min=0.0)
        h4 = torch.relu(W[3][0] @ f + W[3][1])
        outs.append((W[4][0] @ h4 + W[4][1]).numpy())
    return np.array(outs)


def test_forward_matches_independent_torch_bf16():
    """oracle_pn_forward (emulated BF16) against the torch restatement of the
    five layers, the max-pool and the BF16 roundings: the outputs agree to
    FP64 summation-order noise, on trained and random weights."""
    d = synth.make_clusters(12, seed=21, n_min=40, n_max=2000)
    for params in (proposal.load_weights(), proposal.init_params(seed=5)):
        ref = _torch_forward(params, d["xyz"], d["offsets"])
        o = oracle.pn_forward(params, d["xyz"], d["offsets"], emulate=True)["out"]
        assert np.allclose(o, ref, rtol=1e-9, atol=1e-12), np.abs(o - ref).max()
    # the rounding matters: without it the outputs move by far more
    un = oracle.pn_forward(proposal.load_weights(), d["xyz"], d["offsets"], emulate=False)["out"]
    ref = _torch_forward(proposal.load_weights(), d["xyz"], d["offsets"])
    assert np.abs(un - ref).max() > 1e-4
