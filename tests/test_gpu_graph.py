"""Library calls recorded into a HIP graph (sloam_graph_*): the plain batched
cylinder fit and the learned fit (proposal inference + learned pass), each
replayed several times with other eager work and synchronisations between the
replays; every replay equals the eager result bit for bit."""
import numpy as np
import pytest
import torch

from sloam_amd import proposal, synth
from sloam_amd.gpu import CylinderFitPlan

pytestmark = pytest.mark.gpu


def test_plain_fit_graph_replays_match_eager(ctx):
    d = synth.make_clusters(300, seed=51)
    x = torch.as_tensor(d["xyz"]).cuda()
    o = torch.as_tensor(d["offsets"].astype(np.int32)).cuda()
    plan = CylinderFitPlan(ctx, d["offsets"], n_hyp=256)
    eager = plan.alloc_outputs(counts=True, mask=True, best=True)
    plan.exec(x, o, eager)
    ctx.sync()
    ref = {k: v.cpu().numpy() for k, v in eager.items()}
    out = plan.alloc_outputs(counts=True, mask=True, best=True)
    plan.exec(x, o, out)  # sizes every buffer before the capture
    ctx.sync()
    ctx.graph_begin()
    plan.exec(x, o, out)
    g = ctx.graph_end()
    other = synth.make_clusters(50, seed=52)
    op = CylinderFitPlan(ctx, other["offsets"], n_hyp=64)
    oo = op.alloc_outputs(mask=False, best=False)
    ox = torch.as_tensor(other["xyz"]).cuda()
    oo_off = torch.as_tensor(other["offsets"].astype(np.int32)).cuda()
    for rep in range(3):
        print(f"[graph] plain replay {rep}", flush=True)
        for v in out.values():
            v.fill_(-7)  # stale outputs must be overwritten by the replay
        torch.cuda.synchronize()
        g.launch()
        op.exec(ox, oo_off, oo)  # eager work between replays
        ctx.sync()
        for k, v in out.items():
            assert np.array_equal(v.cpu().numpy(), ref[k]), (rep, k)
    g.close()
    plan.close()
    op.close()


def test_learned_fit_graph_replays_match_eager(ctx):
    d = synth.make_clusters(200, seed=53)
    x = torch.as_tensor(d["xyz"]).cuda()
    o = torch.as_tensor(d["offsets"].astype(np.int32)).cuda()
    net = proposal.ProposalNet(ctx, proposal.load_weights())
    fit = proposal.LearnedCylinderFit(ctx, net, d["offsets"], n_hyp=16, plain_hyp=256)
    eager = fit.alloc()
    fit.exec(x, o, eager)
    ctx.sync()
    ref = {k: v.cpu().numpy() for k, v in eager.items()}
    out = fit.alloc()
    fit.exec(x, o, out)
    ctx.sync()
    ctx.graph_begin()
    fit.exec(x, o, out)
    g = ctx.graph_end()
    for rep in range(3):
        print(f"[graph] learned replay {rep}", flush=True)
        for v in out.values():
            v.fill_(-7)
        torch.cuda.synchronize()
        g.launch()
        ctx.sync()
        for k, v in out.items():
            assert np.array_equal(v.cpu().numpy(), ref[k]), (rep, k)
    g.close()
    fit.close()
