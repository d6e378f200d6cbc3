"""GPU parity of batched cylinder fitting (C ABI sloam_cyl_plan_*) against the
CPU oracle on the same seeded inputs.

Bar: sampled indices, per-hypothesis inlier counts, best hypothesis, inlier
masks and the refitted models bit-exact against the oracle's restatement of
the GPU arithmetic (FP32 residuals and Jacobians, TwoSum sums in the GPU's
fixed order, FP64 5x5 solve).  Against the oracle's all-FP64 refit the models
agree within 1e-4 m / rad wherever both keep the same final inliers (a point
within FP32 rounding of the 5 cm band may change sides); every error is
printed."""
import numpy as np
import pytest
import torch

import oracle
from sloam_amd import synth
from sloam_amd.gpu import CylinderFitPlan, fit_cylinders

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-4


def _compare(g, o, n_points):
    assert np.array_equal(g["best"], o["best"])
    if "samples" in g:
        assert np.array_equal(g["samples"], o["samples"])
    if "counts" in g:
        assert np.array_equal(g["counts"], o["counts"])
    if "mask" in g:
        assert np.array_equal(g["mask"], o["mask"])
    gm, om = g["model"], o["model"]
    bad = np.flatnonzero(np.any(gm.view(np.uint64) != om.view(np.uint64), axis=1))
    assert bad.size == 0, (bad[:8], gm[bad[:2]], om[bad[:2]])


def _vs_fp64(g, d, n_hyp, **kw):
    """Error of the GPU's models against the oracle's FP64 refit."""
    o = oracle.cyl_fit_batch(d["xyz"], d["offsets"], n_hyp, threads=16, fp64_refit=True,
                             counts=False, mask=False, **kw)
    gm, om = g["model"], o["model"]
    both = (gm[:, 6] > 0) & (om[:, 6] > 0)
    same = both & (gm[:, 7] == om[:, 7])
    err = np.abs(gm[same, :7] - om[same, :7]).max(axis=1) if same.any() else np.zeros(1)
    print(f"[cylfit] vs FP64 refit: {int(same.sum())} of {int(both.sum())} valid fits keep the "
          f"same inliers; largest difference {err.max():.2e}, median {np.median(err):.2e}; "
          f"validity differs on {int(np.sum((gm[:, 6] > 0) != (om[:, 6] > 0)))}", flush=True)
    assert err.max() <= FP64_TOL
    assert same.sum() >= 0.99 * both.sum()


def _run_both(d, n_hyp, **kw):
    print(f"[cylfit] clusters={len(d['offsets']) - 1} points={d['offsets'][-1]} n_hyp={n_hyp} {kw}",
          flush=True)
    return d


@pytest.mark.parametrize("n_hyp", [200, 1024])
def test_small_batch_bit_exact(ctx, n_hyp):
    d = synth.make_clusters(64, seed=1)
    _run_both(d, n_hyp)
    g = fit_cylinders(ctx, d["xyz"], d["offsets"], n_hyp=n_hyp, counts=True, samples=True,
                      mask=True, best=True)
    o = oracle.cyl_fit_batch(d["xyz"], d["offsets"], n_hyp, samples=True, threads=8)
    _compare(g, o, d["offsets"][-1])
    _vs_fp64(g, d, n_hyp)


def test_edge_clusters(ctx):
    pts = [np.zeros((0, 3)), np.ones((1, 3)), np.ones((2, 3)),
           np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float),
           np.array([[0, 0, i] for i in range(10)], float),
           np.full((70, 3), 2.5)]
    rng = np.random.default_rng(4)
    p, _ = synth.trunk_points(rng, 5000, np.array([6.0, 1.0, 0.0]), np.array([0, 0, 1.0]), 0.4,
                              7.0, 0.02, 0.2)
    pts.append(p)
    xyz = np.concatenate(pts).astype(np.float32)
    off = np.cumsum([0] + [len(q) for q in pts]).astype(np.int32)
    d = {"xyz": xyz, "offsets": off}
    _run_both(d, 96)
    g = fit_cylinders(ctx, xyz, off, n_hyp=96, counts=True, samples=True, mask=True, best=True)
    o = oracle.cyl_fit_batch(xyz, off, 96, samples=True)
    _compare(g, o, off[-1])


def test_large_cluster_takes_per_step_refit(ctx):
    """A batch holding a 40000-point cluster (20 refit chunks) is refitted
    with one launch per Gauss-Newton step; without it, the same small batch
    takes the fused refit (one workgroup per cluster: at most 512 clusters of
    at most 4 chunks).  Both give the oracle's records bit for bit (the
    2000-cluster test covers the per-step form at full size)."""
    rng = np.random.default_rng(12)
    big, _ = synth.trunk_points(rng, 40000, np.array([4.0, -2.0, 0.0]), np.array([0, 0, 1.0]),
                                0.3, 8.0, 0.02, 0.2)
    d = synth.make_clusters(12, seed=13)
    xyz = np.concatenate([d["xyz"], big.astype(np.float32)])
    off = np.concatenate([d["offsets"], [d["offsets"][-1] + len(big)]]).astype(np.int32)
    for n_c in (13, 12):
        o_off = off[:n_c + 1]
        o_xyz = xyz[:o_off[-1]]
        _run_both({"xyz": o_xyz, "offsets": o_off}, 128)
        g = fit_cylinders(ctx, o_xyz, o_off, n_hyp=128, counts=True, mask=True, best=True)
        o = oracle.cyl_fit_batch(o_xyz, o_off, 128, threads=8)
        _compare(g, o, o_off[-1])


@pytest.mark.parametrize("hyp_chunk,tile", [(32, 64), (96, 512), (1024, 4096)])
def test_launch_shapes_do_not_change_results(ctx, hyp_chunk, tile):
    d = synth.make_clusters(40, seed=9)
    _run_both(d, 256, hyp_chunk=hyp_chunk, tile=tile)
    ref = fit_cylinders(ctx, d["xyz"], d["offsets"], n_hyp=256, counts=True, mask=True, best=True)
    g = fit_cylinders(ctx, d["xyz"], d["offsets"], n_hyp=256, counts=True, mask=True, best=True,
                      hyp_chunk=hyp_chunk, tile_points=tile)
    for k in ref:
        assert np.array_equal(ref[k], g[k]), k


def test_full_batch_2000_clusters(ctx):
    """The batched-fitting workload at full size: 2000 clusters, 1024 hypotheses."""
    d = synth.make_clusters(2000, seed=1)
    _run_both(d, 1024)
    g = fit_cylinders(ctx, d["xyz"], d["offsets"], n_hyp=1024, counts=True, mask=True, best=True)
    o = oracle.cyl_fit_batch(d["xyz"], d["offsets"], 1024, threads=16)
    _compare(g, o, d["offsets"][-1])
    _vs_fp64(g, d, 1024)
    # size-independent sanity: every count <= cluster size, mask popcount per
    # cluster == count of the best hypothesis
    sizes = np.diff(d["offsets"])
    assert np.all(g["counts"] <= sizes[:, None])
    bits = np.unpackbits(g["mask"].view(np.uint8), bitorder="little")[:d["offsets"][-1]]
    per = np.add.reduceat(bits.astype(np.int64), d["offsets"][:-1])
    per[sizes == 0] = 0
    best_cnt = g["counts"][np.arange(len(sizes)), np.maximum(g["best"], 0)]
    assert np.array_equal(per, np.where(g["best"] >= 0, best_cnt, 0))


def test_deterministic_repeat(ctx):
    d = synth.make_clusters(300, seed=3)
    plan = CylinderFitPlan(ctx, d["offsets"], n_hyp=512)
    dev = torch.device("cuda", 0)
    xyz = torch.as_tensor(d["xyz"]).to(dev)
    off = torch.as_tensor(d["offsets"]).to(dev)
    torch.cuda.synchronize()
    outs = []
    for _ in range(2):
        o = plan.alloc_outputs(counts=True)
        plan.exec(xyz, off, o)
        ctx.sync()
        outs.append({k: v.cpu().numpy() for k, v in o.items()})
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    plan.close()


def _exec_plan(ctx, plan, d):
    dev = torch.device("cuda", 0)
    xyz = torch.as_tensor(np.ascontiguousarray(d["xyz"], np.float32)).to(dev)
    off = torch.as_tensor(np.ascontiguousarray(d["offsets"], np.int32)).to(dev)
    o = plan.alloc_outputs(counts=True, samples=True, mask=True, best=True)
    torch.cuda.synchronize()
    plan.exec(xyz, off, o)
    ctx.sync()
    res = {k: v.cpu().numpy() for k, v in o.items()}
    res["mask"] = res["mask"].view(np.uint32)
    return res


def test_plan_reset_matches_new_plans(ctx):
    """sloam_cyl_plan_reset: a plan laid out again for a larger, then a
    smaller batch gives bit-for-bit the outputs of a plan created for that
    batch (and the oracle's); an invalid batch leaves an empty plan that runs
    as a no-op and can be reset again."""
    batches = [synth.make_clusters(40, seed=2), synth.make_clusters(220, seed=3),
               synth.make_clusters(9, seed=4)]
    n_hyp = 256
    plan = CylinderFitPlan(ctx, batches[0]["offsets"], n_hyp=n_hyp)
    for i, d in enumerate(batches):
        print(f"[cylfit] reset batch {i}: clusters={len(d['offsets']) - 1} "
              f"points={d['offsets'][-1]}", flush=True)
        if i > 0:
            plan.reset(d["offsets"])
        got = _exec_plan(ctx, plan, d)
        fresh = CylinderFitPlan(ctx, d["offsets"], n_hyp=n_hyp)
        want = _exec_plan(ctx, fresh, d)
        fresh.close()
        for k in want:
            assert np.array_equal(got[k], want[k]), (i, k)
    o = oracle.cyl_fit_batch(batches[2]["xyz"], batches[2]["offsets"], n_hyp, samples=True)
    _compare(got, o, batches[2]["offsets"][-1])
    from sloam_amd._lib import SloamError
    with pytest.raises(SloamError):
        plan.reset(np.array([3, 10, 20], np.int32))  # offsets must start at 0
    assert plan.n_clusters == 0
    plan.reset(batches[0]["offsets"])
    got = _exec_plan(ctx, plan, batches[0])
    fresh = CylinderFitPlan(ctx, batches[0]["offsets"], n_hyp=n_hyp)
    want = _exec_plan(ctx, fresh, batches[0])
    fresh.close()
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    plan.close()


def test_xyz_from_xyzi(ctx):
    """sloam_xyz_from_xyzi copies the first three floats of every point
    (bit-exact, NaN payloads included) and nothing past n."""
    from sloam_amd._lib import check
    from sloam_amd.gpu import _ptr
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(5)
    for n in (0, 1, 255, 256, 257, 70001):
        print(f"[xyz] n={n}", flush=True)
        a = rng.normal(size=(n, 4)).astype(np.float32)
        if n > 3:
            a[3, 1] = np.nan
        src = torch.as_tensor(a).to(dev)
        dst = torch.full((n + 8, 3), 7.0, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        check(ctx.lib.sloam_xyz_from_xyzi(ctx.handle, _ptr(src), n, _ptr(dst)), ctx.handle)
        ctx.sync()
        got = dst.cpu().numpy()
        assert np.array_equal(got[:n].view(np.uint32), a[:, :3].view(np.uint32))
        assert np.all(got[n:] == 7.0)


def test_autotune_choice_is_valid_and_cached(ctx):
    """sloam_cyl_autotune returns one of its candidate shapes, the same one
    again from its cache, and that shape reproduces the default results."""
    import torch
    from sloam_amd.gpu import CylinderFitPlan, autotune_cylinder_fit
    d = synth.make_clusters(300, seed=31)
    x = torch.as_tensor(d["xyz"]).cuda()
    o = torch.as_tensor(d["offsets"]).cuda()
    a = autotune_cylinder_fit(ctx, x, o, d["offsets"], n_hyp=256)
    b = autotune_cylinder_fit(ctx, x, o, d["offsets"], n_hyp=256)
    print(f"[autotune] chose hyp_chunk {a[0]}, tile {a[1]}", flush=True)
    assert a == b and a[0] in (64, 128, 256, 512) and a[1] in (1024, 2048, 4096)
    outs = []
    for chunk, tile in ((0, 0), a):
        p = CylinderFitPlan(ctx, d["offsets"], n_hyp=256, hyp_chunk=chunk, tile_points=tile)
        out = p.alloc_outputs(counts=True, mask=True, best=True)
        p.exec(x, o, out)
        ctx.sync()
        outs.append({k: v.cpu().numpy() for k, v in out.items()})
        p.close()
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


def _read_counts(ctx, plan_handle):
    """The 4 device counts of a counted plan: clusters, work items, refit
    chunks, overflow."""
    import ctypes as C
    p = C.c_void_p()
    assert ctx.lib.sloam_cyl_plan_counts(plan_handle, C.byref(p)) == 0
    host = np.zeros(4, np.int32)
    assert ctx.lib.sloam_memcpy_async(ctx.handle, host.ctypes.data_as(C.c_void_p), p, 16) == 0
    ctx.sync()
    return host


def _counted(ctx, cap, cap_points, d_n, n_hyp, tag=0, hyp_chunk=0):
    import ctypes as C
    from sloam_amd._lib import CylFitCfg, check
    from sloam_amd.gpu import _ptr
    cfg = CylFitCfg(n_hyp=n_hyp, thr=0.05, max_radius=1.0, refit_iters=2, seed=0x5EED,
                    stream_tag=tag, hyp_chunk=hyp_chunk, tile_points=0, refit_rounds=3)
    h = C.c_void_p()
    check(ctx.lib.sloam_cyl_plan_create_counted(ctx.handle, cap, cap_points, _ptr(d_n),
                                                C.byref(cfg), C.byref(h)), ctx.handle)
    return h


def test_counted_plan_matches_host_plan(ctx):
    """A counted plan (batch size read on the device, layout made on the
    device) gives a host-planned plan's counts, best hypotheses and models
    bit for bit, on two batches in a row (the second smaller); a batch past
    its capacity sets the overflow word and fits nothing."""
    import ctypes as C
    from sloam_amd._lib import check
    from sloam_amd.gpu import _ptr
    dev = torch.device("cuda", 0)
    cap, cap_points, n_hyp = 80, 300000, 512
    d_n = torch.zeros(1, dtype=torch.int32, device=dev)
    plan = _counted(ctx, cap, cap_points, d_n, n_hyp)
    try:
        for seed, n in ((41, 60), (42, 25)):
            d = synth.make_clusters(n, seed=seed)
            print(f"[cylfit] counted plan: {n} clusters, {d['offsets'][-1]} points", flush=True)
            x = torch.as_tensor(d["xyz"]).to(dev)
            off = torch.as_tensor(d["offsets"]).to(dev)
            d_n.fill_(n)
            counts = torch.full((cap, n_hyp), -1, dtype=torch.int32, device=dev)
            best = torch.full((cap,), -9, dtype=torch.int32, device=dev)
            model = torch.zeros((cap, 8), dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            check(ctx.lib.sloam_cyl_plan_exec(plan, _ptr(x), _ptr(off), _ptr(counts), None, None,
                                              _ptr(best), _ptr(model)), ctx.handle)
            ctx.sync()
            ref = fit_cylinders(ctx, d["xyz"], d["offsets"], n_hyp=n_hyp, counts=True,
                                mask=False, best=True)
            assert np.array_equal(counts.cpu().numpy()[:n], ref["counts"])
            assert np.array_equal(best.cpu().numpy()[:n], ref["best"])
            assert np.array_equal(model.cpu().numpy()[:n].view(np.uint64),
                                  ref["model"].view(np.uint64))
            k = _read_counts(ctx, plan)
            assert k[0] == n and k[1] > 0 and k[2] >= n and k[3] == 0
        # past the capacity: overflow, nothing fitted
        d = synth.make_clusters(cap + 5, seed=43)
        x = torch.as_tensor(d["xyz"]).to(dev)
        off = torch.as_tensor(d["offsets"]).to(dev)
        d_n.fill_(cap + 5)
        model = torch.full((cap, 8), 7.0, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        check(ctx.lib.sloam_cyl_plan_exec(plan, _ptr(x), _ptr(off), None, None, None, None,
                                          _ptr(model)), ctx.handle)
        ctx.sync()
        k = _read_counts(ctx, plan)
        print(f"[cylfit] counted plan past its capacity: counts {k.tolist()}", flush=True)
        assert k[3] == 1 and k[0] == 0 and k[2] == 0
        assert np.all(model.cpu().numpy() == 7.0)
        # a cluster beyond 131072 points (the host plans' bound): overflow
        rng = np.random.default_rng(44)
        big, _ = synth.trunk_points(rng, 140000, np.array([3.0, 0.0, 0.0]),
                                    np.array([0, 0, 1.0]), 0.3, 8.0, 0.02, 0.2)
        x = torch.as_tensor(big.astype(np.float32)).to(dev)
        off = torch.as_tensor(np.array([0, len(big)], np.int32)).to(dev)
        d_n.fill_(1)
        torch.cuda.synchronize()
        check(ctx.lib.sloam_cyl_plan_exec(plan, _ptr(x), _ptr(off), None, None, None, None,
                                          _ptr(model)), ctx.handle)
        ctx.sync()
        k = _read_counts(ctx, plan)
        print(f"[cylfit] counted plan, a 140000-point cluster: counts {k.tolist()}", flush=True)
        assert k[3] == 1 and k[0] == 0
        assert np.all(model.cpu().numpy() == 7.0)
    finally:
        ctx.lib.sloam_cyl_plan_destroy(plan)


def test_counted_learned_fit_matches_host_plans(ctx):
    """The learned fit on counted plans with the counted proposal inference
    (sloam_pn_infer_counted) equals the host-planned learned fit."""
    import ctypes as C
    from sloam_amd import proposal
    from sloam_amd._lib import check
    from sloam_amd.gpu import _ptr
    dev = torch.device("cuda", 0)
    d = synth.make_clusters(48, seed=44)
    n = len(d["offsets"]) - 1
    cap, cap_points = 64, 300000
    x = torch.as_tensor(d["xyz"]).to(dev)
    off = torch.as_tensor(d["offsets"]).to(dev)
    net = proposal.ProposalNet(ctx, proposal.load_weights(proposal.DEFAULT_WEIGHTS))
    lf = proposal.LearnedCylinderFit(ctx, net, d["offsets"])
    ref_out = lf.alloc()
    torch.cuda.synchronize()
    lf.exec(x, off, ref_out)
    ctx.sync()
    ref = ref_out["model"].cpu().numpy()
    d_n = torch.full((1,), n, dtype=torch.int32, device=dev)
    verify = _counted(ctx, cap, cap_points, d_n, 16, tag=1, hyp_chunk=0)
    plain = _counted(ctx, cap, cap_points, d_n, 1024, tag=0)
    try:
        prop = torch.zeros((cap, 8), dtype=torch.float32, device=dev)
        flags = torch.zeros((cap,), dtype=torch.int32, device=dev)
        n_fb = torch.zeros((1,), dtype=torch.int32, device=dev)
        model = torch.zeros((cap, 8), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        check(ctx.lib.sloam_pn_infer_counted(ctx.handle, _ptr(net.params), _ptr(x), _ptr(off),
                                             cap, _ptr(d_n), _ptr(prop)), ctx.handle)
        check(ctx.lib.sloam_cyl_learned_exec(verify, plain, _ptr(x), _ptr(off), _ptr(prop),
                                             C.c_float(0.5), None, None, _ptr(flags),
                                             _ptr(n_fb), None, _ptr(model)), ctx.handle)
        ctx.sync()
        got = model.cpu().numpy()[:n]
        print(f"[cylfit] counted learned fit: {int(n_fb.cpu()[0])} fallbacks, host "
              f"{int(ref_out['n_fallback'].cpu()[0])}", flush=True)
        assert np.array_equal(got.view(np.uint64), ref[:n].view(np.uint64))
        assert int(n_fb.cpu()[0]) == int(ref_out["n_fallback"].cpu()[0])
    finally:
        ctx.lib.sloam_cyl_plan_destroy(verify)
        ctx.lib.sloam_cyl_plan_destroy(plain)
        lf.close()
