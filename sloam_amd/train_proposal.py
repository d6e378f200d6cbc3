"""Trains the learned cylinder proposal on the GPU.

Training data: synthetic trunk clusters (sloam_amd.synth.make_clusters:
log-uniform 50-5000 points, radius U(0.05, 0.6) m, 2 cm noise, an outlier
share drawn per chunk from --outliers), labelled by the plain method (1024-hypothesis RANSAC + LO refit,
sloam_cyl_plan) on the same device.  Optimiser: Adam on the FP32 master
weights through sloam_pn_train_step, cosine-decayed step size.

    python -m sloam_amd.train_proposal \\
        --out sloam_amd/weights/cylinder_proposal.npy

(the defaults: 100 chunks of 2000 clusters = 200k clusters, 12 epochs; the
committed weights come from this command, about 25 s on one MI355X)

prints one progress line per chunk generated and every 100 steps."""
import argparse
import math
import os
import sys
import time

import numpy as np
import torch

from sloam_amd import proposal, synth
from sloam_amd.gpu import Context, CylinderFitPlan


def make_chunk(ctx, seed, n, batch, outlier_frac=0.2):
    """One chunk of labelled clusters and its training batches.  Every device
    tensor a batch uses is created here and kept for the whole run: the
    library's stream reads them asynchronously, after the host has moved on,
    so none may be freed (and reused by torch's allocator) while steps are in
    flight."""
    d = synth.make_clusters(n, seed=seed, outlier_frac=outlier_frac)
    x = torch.as_tensor(d["xyz"]).cuda()
    o = torch.as_tensor(d["offsets"].astype(np.int32)).cuda()
    plan = CylinderFitPlan(ctx, d["offsets"], n_hyp=1024, thr=0.05, max_radius=1.0)
    out = plan.alloc_outputs(mask=False, best=False)
    plan.exec(x, o, out)
    ctx.sync()
    plan.close()
    off = d["offsets"].astype(np.int64)
    batches = []
    for a in range(0, n - batch + 1, batch):
        b = a + batch
        p0, p1 = int(off[a]), int(off[b])
        ob = torch.as_tensor((off[a:b + 1] - p0).astype(np.int32)).cuda()
        batches.append((x[p0:p1], ob, out["model"][a:b]))
    torch.cuda.synchronize()
    return {"xyz": x, "model": out["model"], "batches": batches}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--chunks", type=int, default=100)
    ap.add_argument("--chunk-size", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=250)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-min", type=float, default=5e-5)
    ap.add_argument("--seed", type=int, default=13)
    ap.add_argument("--outliers", default="0.2,0.6",
                    help="outlier fraction range of the chunks (uniform)")
    ap.add_argument("--init", default=None, help="start from these weights (.npy)")
    ap.add_argument("--out", default=proposal.DEFAULT_WEIGHTS)
    args = ap.parse_args(argv)
    ctx = Context(0)
    t0 = time.time()
    chunks = []
    lo, hi = (float(v) for v in args.outliers.split(","))
    orng = np.random.default_rng(args.seed + 1)
    for k in range(args.chunks):
        chunks.append(make_chunk(ctx, args.seed * 1000 + k, args.chunk_size, args.batch,
                                 outlier_frac=float(orng.uniform(lo, hi))))
        if k % 5 == 0 or k == args.chunks - 1:
            print(f"[train] chunk {k + 1}/{args.chunks} ready ({time.time() - t0:.0f} s)",
                  flush=True)
    init = proposal.load_weights(args.init) if args.init else proposal.init_params(args.seed)
    net = proposal.ProposalNet(ctx, init)
    per_chunk = args.chunk_size // args.batch
    steps = args.epochs * args.chunks * per_chunk
    rng = np.random.default_rng(args.seed)
    loss = torch.empty(args.batch, dtype=torch.float32, device="cuda")
    run = []
    step = 0
    for ep in range(args.epochs):
        order = [(k, j) for k in range(args.chunks) for j in range(per_chunk)]
        rng.shuffle(order)
        for k, j in order:
            lr = args.lr_min + 0.5 * (args.lr - args.lr_min) * (1 + math.cos(math.pi * step / steps))
            cfg = proposal.train_cfg(lr=lr)
            x, o, t = chunks[k]["batches"][j]
            net.train_step(x, o, t, cfg, loss)
            step += 1
            if step % 100 == 0 or step == steps:
                ctx.sync()
                run.append(float(loss.mean()))
                print(f"[train] epoch {ep + 1} step {step}/{steps} lr {lr:.2e} "
                      f"loss {np.mean(run[-5:]):.5f} ({time.time() - t0:.0f} s)", flush=True)
    ctx.sync()
    w = net.params.cpu().numpy().astype(np.float32)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, w)
    print(f"[train] saved {args.out} ({w.size} parameters, {time.time() - t0:.0f} s)", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
