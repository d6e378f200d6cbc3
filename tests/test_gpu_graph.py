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


def test_graph_holds_its_scratch():
    """A captured graph holds the context scratch its kernels use: a later
    call on the same context that needs a larger buffer is refused (status
    -1, no memory moved under the graph) until the graph is destroyed; the
    graph's replays stay correct meanwhile."""
    from sloam_amd._lib import SloamError
    from sloam_amd.gpu import Context, project_partition_dev

    c = Context(0)
    g = None
    try:
        dev = torch.device("cuda", 0)
        small = synth.make_scan(seed=21, H=16, W=512)
        big = synth.make_scan(seed=22, H=64, W=2048)
        xs = torch.as_tensor(small["xyzi"]).to(dev)
        ls = torch.as_tensor(small["label"]).to(dev)
        xb = torch.as_tensor(big["xyzi"]).to(dev)
        lb = torch.as_tensor(big["label"]).to(dev)
        torch.cuda.synchronize()
        out = project_partition_dev(c, xs, ls, H=16, W=512)
        c.sync()
        ref = {k: v.cpu().numpy() for k, v in out.items()}
        c.graph_begin()
        project_partition_dev(c, xs, ls, H=16, W=512, out=out)
        g = c.graph_end()
        print("[graph] larger call while the graph lives", flush=True)
        with pytest.raises(SloamError, match="status -1"):
            project_partition_dev(c, xb, lb, H=64, W=2048)
        for v in out.values():
            v.fill_(-3)
        torch.cuda.synchronize()
        g.launch()
        c.sync()
        n_ground = int(ref["n_ground"][0])
        for k, v in out.items():
            got = v.cpu().numpy()
            if k == "ground":  # rows past the count are not written
                got, want = got[:n_ground], ref[k][:n_ground]
            else:
                want = ref[k]
            assert np.array_equal(got, want, equal_nan=True), k
        g.close()
        big_out = project_partition_dev(c, xb, lb, H=64, W=2048)  # the slot may move now
        c.sync()
        assert int(big_out["n_ground"].cpu()[0]) > 0
    finally:
        if g is not None:
            g.close()
        c.close()


def test_context_destroyed_before_its_graph():
    """Destroying a context releases the graphs still captured on it: the
    orphaned graph refuses to launch (status -1) and its destroy frees only
    the handle."""
    from sloam_amd._lib import SloamError
    from sloam_amd.gpu import Context, project_partition_dev

    c = Context(0)
    dev = torch.device("cuda", 0)
    s = synth.make_scan(seed=23, H=16, W=512)
    x = torch.as_tensor(s["xyzi"]).to(dev)
    lab = torch.as_tensor(s["label"]).to(dev)
    torch.cuda.synchronize()
    out = project_partition_dev(c, x, lab, H=16, W=512)
    c.sync()
    c.graph_begin()
    project_partition_dev(c, x, lab, H=16, W=512, out=out)
    g = c.graph_end()
    g.launch()
    c.sync()
    lib = c.lib
    handle = g.handle
    c.close()
    assert lib.sloam_graph_launch(handle) == -1
    g.close()
