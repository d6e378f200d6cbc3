"""Benchmark of the batched cylinder-fitting hot path on one MI355X.

Workload: 2000 synthetic trunk clusters (n = floor(exp(U(ln 50, ln 5000))),
radius U(0.05, 0.6) m, tilt <= 5 deg, sigma 2 cm, 20 % outliers), 1024 seeded
RANSAC hypotheses per cluster, 5 cm inlier threshold, LO refit.  A step is one
batch: seed + score + refit kernels.  Inputs are resident in HBM before the
timed region.  Prints one JSON line.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cylinder fits/sec + keyframe latency p50/p99 (ms) at ≤1 cm radius err vs plain"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
PMC_SUMMARY = os.path.join(ROOT, "profiles", "cyl_score_plain_2000x1024_pmc.md")


def pmc_traffic(clusters, hyp, points, chunk, tile):
    """HBM bytes per cyl_score launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (scaled by a copy of known size), or None when that
    summary holds no line for this workload and launch shape."""
    try:
        text = open(PMC_SUMMARY).read()
    except OSError:
        return None
    m = re.search(r"^traffic_workload: (.+)$", text, re.M)
    if not m or m.group(1).strip() != f"clusters={clusters} hypotheses={hyp} points={points}":
        return None
    m = re.search(rf"^traffic_shape: hyp_chunk={chunk} tile={tile} bytes=(\d+)$", text, re.M)
    return int(m.group(1)) if m else None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--clusters", type=int, default=2000)
    ap.add_argument("--hyp", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=1)
    ap.add_argument("--cpu-pool", type=int, default=16,
                    help="threads of the second, multi-threaded CPU measurement (0: skip)")
    ap.add_argument("--hyp-chunk", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="default launch shape instead of sloam_cyl_autotune's")
    ap.add_argument("--no-pose", action="store_true", help="skip the batched pose leg")
    ap.add_argument("--no-graph", action="store_true",
                    help="time the eager launches only (no HIP-graph replay)")
    ap.add_argument("--pose-steps", type=int, default=10)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--learned-clusters", type=int, default=4096)
    ap.add_argument("--no-keyframes", action="store_true", help="skip the keyframe-stream leg")
    ap.add_argument("--keyframes", type=int, default=16)
    ap.add_argument("--kf-passes", type=int, default=9,
                    help="passes over the keyframe sequence; the first is a warm-up")
    ap.add_argument("--kf-height", type=int, default=128)
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the long keyframe stream (--stream-keyframes scans, two passes)")
    ap.add_argument("--stream-keyframes", type=int, default=1000)
    ap.add_argument("--stream-spacing", type=float, default=0.2,
                    help="metres between scans of the long stream (2 m/s at 10 Hz)")
    ap.add_argument("--stream-forest-length", type=float, default=260.0)
    ap.add_argument("--learned-cpu-clusters", type=int, default=512,
                    help="clusters of the learned leg's CPU baseline sample")
    ap.add_argument("--kf-cpu-frames", type=int, default=12,
                    help="keyframes the CPU baseline of the keyframe sequence times")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the small-scan keyframe leg (64x2048, 40 trunks)")
    ap.add_argument("--small-steps", type=int, default=50)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device to device three times and exit: the run "
                         "whose rocprofv3 FETCH_SIZE / WRITE_SIZE readings scale the byte "
                         "counters of the library's kernels")
    return ap.parse_args()


def copy_probe():
    """A device copy of known size (256 MiB read and 256 MiB written, three
    times): a counter's reading for it divided by 256 MiB is the factor that
    turns the same counter's reading for a kernel into bytes moved."""
    import torch
    n = 64 * 1024 * 1024  # float32 elements: 256 MiB
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 3 * n * 4}), flush=True)


def bench_pose(ctx, args):
    """Batched pose estimation: 64 problems x (500 cylinders + 20 ground
    patches), ~2M points, 10 LM iterations each; one step = one solve of all
    64 problems.  Reports LM iterations/s and the residual kernel's HBM rate."""
    import time as _t

    import numpy as np
    import torch

    from sloam_amd import synth
    from sloam_amd.gpu import PosePlan

    prob = synth.make_pose_problems(64, seed=2)
    dev = torch.device("cuda", 0)
    iters = 10
    plan = PosePlan(ctx, prob["res_offsets"], len(prob["cyl"]), len(prob["planes"]),
                    max_iters=iters)
    t = {k: torch.as_tensor(np.ascontiguousarray(prob[k])).to(dev)
         for k in ("points", "cyl", "cyl_offsets", "planes", "plane_offsets", "pose_init")}
    out = plan.alloc_outputs()
    torch.cuda.synchronize()
    for _ in range(2):
        plan.exec(t["points"], t["cyl"], t["cyl_offsets"], t["planes"], t["plane_offsets"],
                  t["pose_init"], out)
    ctx.sync()
    stream = ctx.torch_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    ctx.reset_timing()
    ctx.enable_timing(True)
    e0.record(stream)
    for _ in range(args.pose_steps):
        plan.exec(t["points"], t["cyl"], t["cyl_offsets"], t["planes"], t["plane_offsets"],
                  t["pose_init"], out)
    e1.record(stream)
    ctx.sync()
    ctx.enable_timing(False)
    ms = e0.elapsed_time(e1) / args.pose_steps
    ev_ms, ev_n = ctx.kernel_time("pose_eval")
    sv_ms, sv_n = ctx.kernel_time("pose_solve")
    n_res = int(prob["res_offsets"][-1])
    eval_ms = ev_ms / max(ev_n, 1)
    # algorithmic bytes per evaluation: 16 B per residual (xyz + landmark id)
    # + the FP32 landmark tables (32 B per cylinder, 16 B per plane)
    eval_bytes = 16.0 * n_res + 32.0 * len(prob["cyl"]) + 16.0 * len(prob["planes"])
    res = {k: v.cpu().numpy() for k, v in out.items()}
    dt, dr = synth.pose_error(res["pose"], prob["pose_true"])
    # the batched loop (2 launches per iteration) replayed as one HIP graph,
    # which removes the per-launch gaps between the iterations
    ctx.graph_begin()
    plan.exec(t["points"], t["cyl"], t["cyl_offsets"], t["planes"], t["plane_offsets"],
              t["pose_init"], out)
    graph = ctx.graph_end()
    graph.launch()
    ctx.sync()
    graph_same = all(np.array_equal(res[k], v.cpu().numpy()) for k, v in out.items())
    e0.record(stream)
    for _ in range(args.pose_steps):
        graph.launch()
    e1.record(stream)
    ctx.sync()
    graph_ms = e0.elapsed_time(e1) / args.pose_steps
    graph.close()
    # the same solves in the fused mode (one workgroup per problem runs all
    # iterations in one launch), for the comparison of the two modes
    plan.set_mode(2)
    plan.exec(t["points"], t["cyl"], t["cyl_offsets"], t["planes"], t["plane_offsets"],
              t["pose_init"], out)
    ctx.sync()
    e0.record(stream)
    for _ in range(args.pose_steps):
        plan.exec(t["points"], t["cyl"], t["cyl_offsets"], t["planes"], t["plane_offsets"],
                  t["pose_init"], out)
    e1.record(stream)
    ctx.sync()
    fused_ms = e0.elapsed_time(e1) / args.pose_steps
    fused_same = all(np.array_equal(res[k], v.cpu().numpy()) for k, v in out.items())
    plan.close()
    # the same 64 solves with FP64 evaluation (the keyframe solves' form)
    plan = PosePlan(ctx, prob["res_offsets"], len(prob["cyl"]), len(prob["planes"]),
                    max_iters=iters, precision=1, mode=1)
    plan.exec(t["points"], t["cyl"], t["cyl_offsets"], t["planes"], t["plane_offsets"],
              t["pose_init"], out)
    ctx.sync()
    ctx.reset_timing()
    ctx.enable_timing(True)
    e0.record(stream)
    for _ in range(args.pose_steps):
        plan.exec(t["points"], t["cyl"], t["cyl_offsets"], t["planes"], t["plane_offsets"],
                  t["pose_init"], out)
    e1.record(stream)
    ctx.sync()
    ctx.enable_timing(False)
    f64_ms = e0.elapsed_time(e1) / args.pose_steps
    ev64_ms, ev64_n = ctx.kernel_time("pose_eval")
    eval64_ms = ev64_ms / max(ev64_n, 1)
    eval64_bytes = 16.0 * n_res + 64.0 * len(prob["cyl"]) + 32.0 * len(prob["planes"])
    res64 = {k: v.cpu().numpy() for k, v in out.items()}
    dt64 = np.abs(res64["pose"][:, :3] - res["pose"][:, :3]).max()
    plan.close()
    # one keyframe-sized problem (2 x 3250 residuals, 50 iterations, stop at
    # convergence): fused against batched, FP32 and FP64 evaluation
    small = synth.make_pose_problems(2, n_cyl=50, n_planes=5, seed=36)
    ts = {k: torch.as_tensor(np.ascontiguousarray(small[k])).to(dev)
          for k in ("points", "cyl", "cyl_offsets", "planes", "plane_offsets", "pose_init")}
    small_ms = {}
    for mode, precision, name in ((1, 0, "batched"), (2, 0, "fused"), (1, 1, "batched_fp64"),
                                  (2, 1, "fused_fp64")):
        sp = PosePlan(ctx, small["res_offsets"], len(small["cyl"]), len(small["planes"]),
                      max_iters=50, stop_on_convergence=True, mode=mode, precision=precision)
        so = sp.alloc_outputs()
        torch.cuda.synchronize()
        sp.exec(ts["points"], ts["cyl"], ts["cyl_offsets"], ts["planes"], ts["plane_offsets"],
                ts["pose_init"], so)
        ctx.sync()
        e0.record(stream)
        for _ in range(args.pose_steps):
            sp.exec(ts["points"], ts["cyl"], ts["cyl_offsets"], ts["planes"],
                    ts["plane_offsets"], ts["pose_init"], so)
        e1.record(stream)
        ctx.sync()
        small_ms[name] = e0.elapsed_time(e1) / args.pose_steps
        sp.close()
    cpu = None
    if not args.no_cpu:
        import oracle
        oracle.load()
        med, tot = faithful(lambda: cpu_median(
            lambda: oracle.pose_lm_batch(prob, threads=1, max_iters=iters)))
        cpu = {"value": 64 * iters / med, "unit": "LM iterations/s", "cores": 1, "kind": "port",
               "ms_per_solve": med * 1e3,
               "sample": f"the same 64 problems x {iters} iterations through the oracle's FP64 "
                         f"LM built with the reference's flags (-O3 -DNDEBUG), 1 thread, "
                         f"median of 5 after 1 warm-up, {tot:.1f} s in total",
               "gpu_speedup": med * 1e3 / ms}
    return {
        "workload": f"64 problems x (500 cylinders + 20 ground patches), {n_res} residuals, "
                    f"{iters} LM iterations",
        "ms_per_solve": ms,
        "lm_iters_per_s": 64 * iters / (ms / 1e3),
        "residuals_per_s": n_res * (iters + 1) / (ms / 1e3),
        "eval_kernel_ms": eval_ms,
        "solve_kernel_ms": sv_ms / max(sv_n, 1),
        "eval_hbm_gbs": eval_bytes / (eval_ms / 1e3) / 1e9,
        "eval_hbm_frac": eval_bytes / (eval_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "pose_err_vs_truth": {"max_m": float(dt.max()), "max_deg": float(np.degrees(dr.max()))},
        "graph_ms_per_solve": graph_ms,
        "graph_lm_iters_per_s": 64 * iters / (graph_ms / 1e3),
        "graph_bitwise_equal": bool(graph_same),
        "fused_ms_per_solve": fused_ms,
        "fused_bitwise_equal": bool(fused_same),
        "fp64": {"ms_per_solve": f64_ms, "eval_kernel_ms": eval64_ms,
                 "eval_hbm_frac": eval64_bytes / (eval64_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                 "max_translation_diff_vs_fp32_m": float(dt64)},
        "keyframe_size_solve_ms": small_ms,
        "cpu_baseline": cpu,
    }


def host_cpu():
    """Model name (lscpu) and the number of cores this process may run on."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    return {"model": model, "cores_available": n}


def faithful(fn):
    """Runs fn with the oracle built with the reference's flags selected."""
    import oracle
    oracle.select("faithful")
    try:
        return fn()
    finally:
        oracle.select("tuned")


def cpu_median(fn, reps=5):
    """Median wall time (s) of reps runs of fn after one warm-up run, and
    the total time spent."""
    t_all = time.perf_counter()
    fn()
    times = []
    for _ in range(reps):
        c0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - c0)
    times.sort()
    return times[len(times) // 2], time.perf_counter() - t_all


def bench_small_keyframe(ctx, args):
    """The small-scan plain keyframe: one 64 x 2048 scan of 40 trunks,
    projection + partition, tree instances, per-tree cylinder RANSAC (200
    hypotheses, 5 cm), ground RANSAC (200 hypotheses) and the LM pose against
    the 40-trunk map from a 0.3 m / 3 deg perturbation.  Latency of the whole
    keyframe from device-resident inputs (it synchronises three times: for
    the tree and ground counts, for the fits the association reads, and for
    the pose); the same steps on the CPU through the oracle."""
    import numpy as np
    import torch

    import oracle
    from sloam_amd import synth
    from sloam_amd.plain_keyframe import PlainKeyframe

    s, submap, plane, est = synth.make_small_keyframe(0)
    dev = torch.device("cuda", 0)
    x = torch.as_tensor(s["xyzi"]).to(dev)
    lab = torch.as_tensor(s["label"]).to(dev)
    torch.cuda.synchronize()
    kf = PlainKeyframe(ctx)
    for _ in range(3):
        r = kf.run(x, lab, est, submap, plane)
    lat = []
    for _ in range(args.small_steps):
        t0 = time.perf_counter()
        r = kf.run(x, lab, est, submap, plane)
        lat.append((time.perf_counter() - t0) * 1e3)
    kf.close()
    lat = np.array(lat)
    et, er = synth.pose_error(r["pose"], np.array([0, 0, 0, 0, 0, 0, 1.0]))
    cpu = None
    if not args.no_cpu:
        oracle.load()
        holder = {}

        def run_cpu():
            holder["o"] = oracle.plain_keyframe(s["xyzi"], s["label"], est, submap, plane,
                                                threads=1)

        med, tot = faithful(lambda: cpu_median(run_cpu))
        o = holder["o"]
        dp = float(np.abs(o["pose"][:3] - r["pose"][:3]).max())
        cpu = {"ms": med * 1e3, "cores": 1, "kind": "port",
               "sample": f"the same keyframe through the oracle built with the reference's "
                         f"flags (-O3 -DNDEBUG), 1 thread (projection, trellis, fits, "
                         f"association in numpy, LM), median of 5 after 1 warm-up, "
                         f"{tot:.1f} s in total",
               "gpu_speedup": med * 1e3 / float(np.median(lat)),
               "pose_diff_vs_gpu_m": dp}
    return {
        "workload": "one 64x2048 scan, 40 trunks (r U(0.1,0.4) m, tilt <= 5 deg, sigma 2 cm), "
                    "ground slope <= 5 deg sigma 3 cm, 30 % clutter; per-tree cylinder RANSAC "
                    "200 hyps at 5 cm, ground RANSAC 200 hyps, LM pose vs the 40-trunk map "
                    "from 0.3 m / 3 deg",
        "latency_ms": {"p50": float(np.median(lat)), "p99": float(np.percentile(lat, 99)),
                       "min": float(lat.min())},
        "trees": int(r["n_trees"]), "matched": int(np.sum(r["matches"] >= 0)),
        "ground_points": int(r["n_ground"]),
        "pose_err_vs_truth": {"m": float(et), "deg": float(np.degrees(er))},
        "cpu_baseline": cpu,
    }


PN_FLOP_PER_POINT = 2 * (3 * 64 + 64 * 128 + 128 * 256)  # shared per-point MLP
PN_FLOP_HEAD = 2 * (256 * 128 + 128 * 7)
BF16_PEAK_TFLOPS = 2500.0


def bench_learned(ctx, args):
    """Learned proposal + 16 verifying hypotheses + LO refit, plain
    1024-hypothesis fallback below 0.5 inlier ratio, on 4096 clusters of the
    same generator (20 % outliers).  Accuracy against the plain fit of the same
    clusters; network throughput from the pn_fwd kernel time."""
    import numpy as np
    import torch

    from sloam_amd import proposal, synth
    from sloam_amd.gpu import CylinderFitPlan

    S = args.learned_clusters
    d = synth.make_clusters(S, seed=3)
    dev = torch.device("cuda", 0)
    xyz = torch.as_tensor(d["xyz"]).to(dev)
    off = torch.as_tensor(d["offsets"]).to(dev)
    n = np.diff(d["offsets"])
    m = np.minimum(n, 512)
    net_flop = float(PN_FLOP_PER_POINT * m.sum())  # the per-point MLP of pn_fwd
    plain = CylinderFitPlan(ctx, d["offsets"], n_hyp=args.hyp, stream_tag=0)
    pout = plain.alloc_outputs(mask=False, best=False)
    net = proposal.ProposalNet(ctx, proposal.load_weights())
    fit = proposal.LearnedCylinderFit(ctx, net, d["offsets"], n_hyp=16, plain_hyp=args.hyp)
    lout = fit.alloc()
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        plain.exec(xyz, off, pout)
        fit.exec(xyz, off, lout)
    ctx.sync()
    stream = ctx.torch_stream()

    def timed(fn):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        ctx.sync()
        e0.record(stream)
        for _ in range(args.steps):
            fn()
        e1.record(stream)
        ctx.sync()
        return e0.elapsed_time(e1) / args.steps

    ms_plain = timed(lambda: plain.exec(xyz, off, pout))
    ctx.reset_timing()
    ctx.enable_timing(True)
    ms_learned = timed(lambda: fit.exec(xyz, off, lout))
    ctx.enable_timing(False)
    lout2 = fit.alloc()
    fit.exec_two_pass(xyz, off, lout2)
    ms_two_pass = timed(lambda: fit.exec_two_pass(xyz, off, lout2))
    kern = {}
    for name in ("pn_prep", "pn_fwd", "pn_head", "cyl_seed", "cyl_score", "cyl_decide",
                 "cyl_refit", "cyl_rescue"):
        ms, k = ctx.kernel_time(name)
        kern[name] = ms / args.steps  # per batch (some kernels launch twice per batch)
    ref = pout["model"].cpu().numpy()
    lm = lout["model"].cpu().numpy()
    nf = int(lout["n_fallback"].cpu().numpy()[0])
    # the plain fit's own spread: the same clusters with another hypothesis
    # stream (another seed), against the first
    plain2 = CylinderFitPlan(ctx, d["offsets"], n_hyp=args.hyp, stream_tag=1)
    pout2 = plain2.alloc_outputs(mask=False, best=False)
    plain2.exec(xyz, off, pout2)
    ctx.sync()
    ref2 = pout2["model"].cpu().numpy()
    plain2.close()
    # one training step of the network on the first 250 clusters with the
    # plain fits as targets (forward saving activations, backward on MFMA,
    # batch gradient, Adam), on a copy of the weights
    B = min(250, S)
    tnet = proposal.ProposalNet(ctx, proposal.load_weights())
    txyz, toff, tgt = xyz[:int(d["offsets"][B])], off[:B + 1], pout["model"][:B].contiguous()
    tloss = torch.empty(B, dtype=torch.float32, device=dev)
    tcfg = proposal.train_cfg(lr=1e-4)
    tnet.train_step(txyz, toff, tgt, tcfg, tloss)
    ctx.reset_timing()
    ctx.enable_timing(True)
    ms_train = timed(lambda: tnet.train_step(txyz, toff, tgt, tcfg, tloss))
    ctx.enable_timing(False)
    train = {"clusters": B, "ms_per_step": ms_train}
    for name in ("pn_prep", "pn_fwd", "pn_bwd", "pn_reduce"):
        ms, k = ctx.kernel_time(name)
        train[name + "_ms"] = ms / args.steps
    bins = bench_learned_size_bins(ctx, args, net, timed)
    ok = (ref[:, 6] > 0) & (lm[:, 6] > 0)
    err = np.abs(lm[ok, 6] - ref[ok, 6]) * 100
    ok2 = (ref[:, 6] > 0) & (ref2[:, 6] > 0)
    err2 = np.abs(ref2[ok2, 6] - ref[ok2, 6]) * 100

    def vs_truth(m):
        v = m[:, 6] > 0
        e = np.abs(m[v, 6] - d["radius"][v]) * 100
        return {"median": float(np.median(e)), "p90": float(np.percentile(e, 90)),
                "frac_over_1cm": float(np.mean(e > 1.0)), "valid": int(v.sum())}

    def by_size(e, okm):
        sz = n[okm]
        return {f"{lo}-{hi - 1}": float(np.mean(e[(sz >= lo) & (sz < hi)] > 1.0))
                for lo, hi in ((50, 256), (256, 1024), (1024, 5001))}
    fwd_s = kern["pn_fwd"] / 1e3
    plain.close()
    fit.close()
    cpu = None
    if not args.no_cpu:
        import oracle
        oracle.load()
        k = min(args.learned_cpu_clusters, S)
        sub_off = d["offsets"][:k + 1]
        sub_xyz = d["xyz"][:int(sub_off[-1])]
        params = proposal.load_weights()
        med, tot = faithful(lambda: cpu_median(
            lambda: oracle.learned_fit(params, sub_xyz, sub_off, plain_hyp=args.hyp)))
        cpu = {"value": k / med, "unit": "cylinder fits/s", "cores": 1, "kind": "port",
               "sample": f"the first {k} of the {S} clusters: the network in FP64, 16 "
                         f"verifying hypotheses, plain {args.hyp}-hypothesis fallback, through "
                         f"the oracle built with the reference's flags (-O3 -DNDEBUG), 1 "
                         f"thread, median of 5 after 1 warm-up, {tot:.1f} s in total",
               "gpu_speedup": (S / (ms_learned / 1e3)) / (k / med)}
    return {
        "workload": f"{S} clusters (log-U[50,5000] points, 20 % outliers), proposal net on "
                    f"<= 512 points each, 16 verifying hypotheses, plain {args.hyp}-hypothesis "
                    f"fallback below 0.5 verified inlier ratio, one LO refit per cluster",
        "fits_per_s": S / (ms_learned / 1e3),
        "ms_per_batch": ms_learned,
        "plain_fits_per_s": S / (ms_plain / 1e3),
        "plain_ms_per_batch": ms_plain,
        "speedup_vs_plain": ms_plain / ms_learned,
        "two_pass_ms_per_batch": ms_two_pass,
        "fallback_fraction": nf / S,
        "radius_diff_vs_plain_cm": {"median": float(np.median(err)),
                                    "p90": float(np.percentile(err, 90)),
                                    "frac_over_1cm": float(np.mean(err > 1.0)),
                                    "frac_over_1cm_by_size": by_size(err, ok)},
        # the floor of the comparison: plain against plain with another seed
        "plain_vs_plain_other_seed_cm": {"median": float(np.median(err2)),
                                         "p90": float(np.percentile(err2, 90)),
                                         "frac_over_1cm": float(np.mean(err2 > 1.0)),
                                         "frac_over_1cm_by_size": by_size(err2, ok2)},
        "radius_err_vs_truth_cm": {"learned": vs_truth(lm), "plain": vs_truth(ref),
                                   "plain_other_seed": vs_truth(ref2)},
        "kernel_ms_per_batch": kern,
        "train_step": train,
        "size_bins": bins,
        "cpu_baseline": cpu,
        "net_mfma": {"achieved": net_flop / fwd_s / 1e12, "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": net_flop / fwd_s / 1e12 / BF16_PEAK_TFLOPS,
                     "flop_per_launch": net_flop},
    }


def bench_learned_size_bins(ctx, args, net, timed, sizes=(256, 1024, 4096), clusters=1024):
    """The learned fit against the plain one per cluster size: `clusters`
    clusters of about n points for each n (same generator, 20 % outliers;
    keyed by the size drawn).
    The network sees at most 512 points of a cluster, so its share of the
    work shrinks as n grows while the plain method's grows with n."""
    import numpy as np
    import torch

    from sloam_amd import proposal, synth
    from sloam_amd.gpu import CylinderFitPlan

    dev = torch.device("cuda", 0)
    out = {}
    for n in sizes:
        d = synth.make_clusters(clusters, seed=30 + n, n_min=n, n_max=n)
        xyz = torch.as_tensor(d["xyz"]).to(dev)
        off = torch.as_tensor(d["offsets"]).to(dev)
        plain = CylinderFitPlan(ctx, d["offsets"], n_hyp=args.hyp, stream_tag=0)
        pout = plain.alloc_outputs(mask=False, best=False)
        fit = proposal.LearnedCylinderFit(ctx, net, d["offsets"], n_hyp=16, plain_hyp=args.hyp)
        lout = fit.alloc()
        torch.cuda.synchronize()
        plain.exec(xyz, off, pout)
        fit.exec(xyz, off, lout)
        ms_plain = timed(lambda: plain.exec(xyz, off, pout))
        ms_learned = timed(lambda: fit.exec(xyz, off, lout))
        ref = pout["model"].cpu().numpy()
        lm = lout["model"].cpu().numpy()
        ok = (ref[:, 6] > 0) & (lm[:, 6] > 0)
        err = np.abs(lm[ok, 6] - ref[ok, 6]) * 100
        # the generator's size draw floors exp(log n): n or n - 1 points
        out[str(int(np.median(np.diff(d["offsets"]))))] = {
            "plain_fits_per_s": clusters / (ms_plain / 1e3),
            "learned_fits_per_s": clusters / (ms_learned / 1e3),
            "speedup_vs_plain": ms_plain / ms_learned,
            "fallback_fraction": int(lout["n_fallback"].cpu().numpy()[0]) / clusters,
            "radius_frac_over_1cm": float(np.mean(err > 1.0)) if err.size else None,
        }
        plain.close()
        fit.close()
    return out


def traj_diff(a, b):
    """Per-keyframe difference of two trajectories ([K, 7]: t, quaternion xyzw)."""
    import numpy as np
    dt = np.linalg.norm(a[:, :3] - b[:, :3], axis=1)
    dot = np.clip(np.abs(np.sum(a[:, 3:] * b[:, 3:], axis=1)), 0, 1)
    ang = np.degrees(2 * np.arccos(dot))
    return {"max_m": float(dt.max()), "mean_m": float(dt.mean()), "final_m": float(dt[-1]),
            "max_deg": float(ang.max()), "mean_deg": float(ang.mean())}


# estimator parameters of the keyframe streams: the forest's trunks reach
# 0.4 m radius (the reference's hard-coded 0.3 m cap would drop them) and a
# 3 deg heading error moves a trunk 0.3 m at 6 m height (the yaml's 0.5 m
# tree match threshold is marginal)
KF_PARAMS = ("--max-tree-radius", "0.45", "--tree-match-thresh", "1.0")


def run_keyframe_bench(path, passes, model, pinned, pipelined, traj_path, graph=1, shadow=None,
                       extra=KF_PARAMS):
    """One run of the C++ program sloam_amd/keyframe_bench; returns its JSON
    line and the trajectory it wrote.  shadow: a second tree model run beside
    `model` on every keyframe of the first pass, from the same estimate and
    submap (the one-keyframe pose difference of the two); extra: further
    options (estimator parameters, --dump)."""
    import subprocess

    import numpy as np
    exe = os.path.join(ROOT, "sloam_amd", "keyframe_bench")
    weights = os.path.join(ROOT, "sloam_amd", "weights", "cylinder_proposal.npy")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: build first (python -m sloam_amd.build)")
    r = subprocess.run([exe, path, str(passes), str(graph), "--tree-model", model,
                        "--weights", weights, "--traj-out", traj_path] +
                       (["--pinned"] if pinned else []) +
                       (["--pipelined"] if pipelined else []) +
                       (["--shadow", shadow] if shadow else []) + list(extra),
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"keyframe_bench failed: {r.stderr[-2000:]}")
    for line in r.stderr.splitlines():
        if line.startswith("shadow keyframe"):
            print(f"[bench] {line}", file=sys.stderr, flush=True)
    return (json.loads(r.stdout.strip().splitlines()[-1]),
            np.fromfile(traj_path, dtype=np.float64).reshape(-1, 7))


def keyframe_cpu_baseline(args, seq_path, tmp, gpu_plain):
    """The keyframe's work on the CPU: for keyframes 1..args.kf_cpu_frames of
    the sequence, the oracle's projection and partition, tree instances,
    ground cells, plain cylinder fits (1024 hypotheses) of every tree and
    RunSloam, from the estimate, submap and previous ground planes the GPU
    runner (plain tree model) had at that keyframe (keyframe_bench --dump).
    Faithful build (the reference's -O3 -DNDEBUG), one thread; each keyframe
    timed once after one warm-up keyframe.  Also reports how far the CPU's
    pose lands from the GPU runner's."""
    import numpy as np

    import oracle
    from sloam_amd import synth
    from sloam_amd import keyframe_dump as kf_dump

    dump = os.path.join(tmp, "plain.dump")
    run_keyframe_bench(seq_path, 1, "plain", 1, 0, os.path.join(tmp, "traj_dump.bin"),
                       extra=KF_PARAMS + ("--dump", dump))
    recs = [r for r in kf_dump.read(dump) if not r["first"]][:args.kf_cpu_frames + 1]
    frames, (_, H, W) = synth.read_keyframe_sequence(seq_path, {r["keyframe"] for r in recs})
    oracle.select("faithful")
    try:
        est = oracle.SloamEstimator(maxTreeRadius=0.45, treeMatchThresh=1.0)
        times, dpose, agree = [], [], 0
        for i, r in enumerate(recs):
            _, _, xyzi, lab = frames[r["keyframe"]]
            est.set_state(False, r["prev_planes"])
            t0 = time.perf_counter()
            o = oracle.keyframe_step(est, xyzi, lab, r["pose"], r["map"], H=H, W=W)
            dt = time.perf_counter() - t0
            if i == 0:
                continue  # warm-up
            times.append(dt)
            if o["ok"] and r["out"]["ok"]:
                dpose.append(float(np.abs(o["pose"][:3] - r["out"]["pose"][:3]).max()))
            agree += int(o["ok"] == r["out"]["ok"] and
                         o["tree_step"] == r["out"]["tree_step"])
    finally:
        oracle.select("tuned")
    med = float(np.median(times))
    return {"ms_per_keyframe": med * 1e3, "cores": 1, "kind": "port",
            "sample": f"keyframes {recs[1]['keyframe']}-{recs[-1]['keyframe']} of the sequence "
                      f"({len(times)} timed after one warm-up keyframe), oracle built with the "
                      f"reference's flags (-O3 -DNDEBUG), 1 thread: front end, ground cells, "
                      f"plain cylinder fits (1024 hypotheses), RunSloam",
            "gpu_plain_latency_p50_ms": gpu_plain["latency_ms"]["p50"],
            "gpu_speedup_vs_plain_p50": med * 1e3 / gpu_plain["latency_ms"]["p50"],
            "same_decisions": f"{agree} of {len(times)}",
            "pose_diff_vs_gpu_max_m": max(dpose) if dpose else None}


def bench_keyframes(args):
    """Keyframe stream: scans of a 300-tree forest (128 x 2048, 40 % understory
    clutter, partial trunk occlusion) every 0.5 m along a weaving trajectory,
    noisy odometry between them; each keyframe runs the GPU front end (one
    HIP graph: projection + partition, tree instances, ground cells), the
    tree models and the estimator (ground and tree models, association,
    two-step pose).  The C++ program sloam_amd/keyframe_bench times each
    keyframe end to end (host buffers in, pose out).  The headline runs the
    learned tree model (proposal network + 16 verifying hypotheses + plain
    fallback over each tree's points); the plain cylinder fit and the
    reference's median-line model run beside it, and the learned trajectory
    is compared with the plain one."""
    import tempfile

    import numpy as np

    from sloam_amd import synth

    # name, tree model, graph, pinned input, pipelined
    runs = (("learned", "learned", 1, 1, 0), ("learned_pipelined", "learned", 1, 1, 1),
            ("learned_pageable", "learned", 1, 0, 0), ("plain", "plain", 1, 1, 0),
            ("lines", "lines", 1, 1, 0), ("lines_eager", "lines", 0, 1, 0))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "sequence.bin")
        t0 = time.perf_counter()
        synth.write_keyframe_sequence(path, n_keyframes=args.keyframes, H=args.kf_height)
        gen_s = time.perf_counter() - t0
        res, traj = {}, {}
        for name, model, graph, pinned, pipelined in runs:
            print(f"[bench] keyframe stream: {name}", file=sys.stderr, flush=True)
            res[name], traj[name] = run_keyframe_bench(
                path, args.kf_passes, model, pinned, pipelined,
                os.path.join(d, f"traj_{name}.bin"), graph=graph,
                shadow="plain" if name == "learned" else None)
        cpu = None if args.no_cpu else keyframe_cpu_baseline(args, path, d, res["plain"])

    g = res["learned"]
    g["drift_vs_plain"] = traj_diff(traj["learned"], traj["plain"])
    g["step_vs_plain"] = g.pop("shadow")
    pip = res["learned_pipelined"]
    g["pipelined"] = {"throughput_kf_per_s": pip["throughput_kf_per_s"],
                      "latency_ms": pip["latency_ms"],
                      "trajectory_equal": bool(np.array_equal(traj["learned_pipelined"],
                                                              traj["learned"]))}
    g["tree_models"] = {
        name: {"latency_ms": res[name]["latency_ms"],
               "tree_fit_p50_ms": res[name]["tree_fit_p50_ms"],
               "throughput_kf_per_s": res[name]["throughput_kf_per_s"],
               "trees_mean": res[name]["trees_mean"],
               "drift_vs_truth": res[name]["drift"]}
        for name in ("learned", "plain", "lines")}
    g["tree_models"]["lines"]["drift_vs_plain"] = traj_diff(traj["lines"], traj["plain"])
    g["pageable_input"] = {"latency_ms": res["learned_pageable"]["latency_ms"],
                           "upload_p50_ms": res["learned_pageable"]["stage_p50_ms"]["upload"]}
    g["lines_eager_frontend_p50_ms"] = res["lines_eager"]["stage_p50_ms"]["frontend"]
    g["lines_eager_latency_p50_ms"] = res["lines_eager"]["latency_ms"]["p50"]
    g["cpu_baseline"] = cpu
    g["workload"] = (f"{args.keyframes} keyframes of {args.kf_height}x2048 scans, 0.5 m apart, "
                     f"300-tree forest, 40 % understory clutter; {args.kf_passes} passes, the "
                     f"first a warm-up reported apart; learned tree models (plain and the "
                     f"reference's median lines beside them); each scan handed over in the "
                     f"runner's pinned input buffers (pageable arrays: pageable_input); "
                     f"throughput of the sequential runner and of the pipelined one (a "
                     f"scan's front end overlapping the previous keyframe's estimator); "
                     f"sequence generated on the host in {gen_s:.1f} s (not timed)")
    return g


def bench_keyframe_stream(args):
    """The long keyframe stream: --stream-keyframes scans (128 x 2048) of a
    300-tree forest along the weaving trajectory, one every --stream-spacing
    metres (2 m/s at 10 Hz), 40 % understory clutter and partial trunk
    occlusion; odometry noise per step scaled by sqrt(spacing / 0.5 m), so
    the odometry's random walk has the same variance per metre as the short
    sequence's.  Every scan is a keyframe.  Two passes per run: the first
    warms every buffer and plan, the second is timed.  The learned tree model
    runs sequentially and pipelined, the plain cylinder fit beside it; the
    learned trajectory is compared with the plain one keyframe by keyframe."""
    import math
    import shutil
    import subprocess
    import tempfile

    import numpy as np

    n = args.stream_keyframes
    need = 12 + n * (112 + args.kf_height * 2048 * 17)
    with tempfile.TemporaryDirectory() as d:
        free = shutil.disk_usage(d).free
        if free < need + (1 << 30):
            msg = (f"the {need / 2**30:.1f} GiB sequence file needs more than the "
                   f"{free / 2**30:.1f} GiB free under {d}")
            print(f"[bench] long keyframe stream skipped: {msg}", file=sys.stderr, flush=True)
            return {"skipped": msg}
        path = os.path.join(d, "stream.bin")
        scale = math.sqrt(args.stream_spacing / 0.5)
        sigma_t, sigma_yaw = 0.05 * scale, 0.5 * scale
        workers = max(1, min(16, os.cpu_count() or 1))
        print(f"[bench] long keyframe stream: writing {n} scans ({workers} processes)",
              file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        # a process of its own: its pool forks, and this process holds the GPU
        subprocess.run([sys.executable, "-m", "sloam_amd.synth", path, str(n),
                        str(args.stream_spacing), str(args.stream_forest_length),
                        str(sigma_t), str(sigma_yaw), str(workers), str(args.kf_height)],
                       check=True, timeout=1200, cwd=ROOT)
        gen_s = time.perf_counter() - t0
        res, traj = {}, {}
        # name, tree model, pipelined, device estimator (the whole keyframe
        # as one captured graph), shadow model and options: learned against
        # plain one keyframe at a time, and plain against plain with another
        # hypothesis stream (the plain fit's own spread)
        for name, model, pipelined, device, shadow, more in (
                ("learned", "learned", 0, 0, "plain", ()),
                ("learned_pipelined", "learned", 1, 0, None, ()),
                ("plain", "plain", 0, 0, None, ()),
                ("learned_device", "learned", 0, 1, None, ("--device-estimator",)),
                ("plain_device", "plain", 0, 1, None, ("--device-estimator",)),
                ("plain_seed", "plain", 0, 0, "plain", ("--shadow-seed", "0x1234"))):
            print(f"[bench] long keyframe stream: {name}", file=sys.stderr, flush=True)
            res[name], traj[name] = run_keyframe_bench(
                path, 2, model, 1, pipelined, os.path.join(d, f"traj_{name}.bin"),
                shadow=shadow, extra=KF_PARAMS + more)

    def summary(r):
        return {k: r[k] for k in ("latency_ms", "throughput_kf_per_s", "stage_p50_ms",
                                  "estimator_p50_ms", "tree_fit_p50_ms", "trees_mean",
                                  "planes_mean", "success_rate", "discarded", "trees_fitted",
                                  "trees_fallback", "drift", "map_p50_ms",
                                  "latency_with_map_ms", "odometry_drift", "steps", "params")}

    g = summary(res["learned"])
    g["keyframes"] = n
    g["drift_vs_plain"] = traj_diff(traj["learned"], traj["plain"])
    g["step_vs_plain"] = res["learned"]["shadow"]
    # the floor of that comparison: plain against plain with another
    # hypothesis stream, one keyframe at a time the same way
    g["step_plain_vs_plain_other_seed"] = res["plain_seed"]["shadow"]
    g["pipelined"] = {"throughput_kf_per_s": res["learned_pipelined"]["throughput_kf_per_s"],
                      "latency_ms": res["learned_pipelined"]["latency_ms"],
                      "trajectory_equal": bool(np.array_equal(traj["learned_pipelined"],
                                                              traj["learned"]))}
    g["plain"] = summary(res["plain"])
    # the device estimator: the same keyframes, bit for bit, with the
    # estimator on the GPU behind the front end in one graph per keyframe
    g["device"] = {
        name: {"latency_ms": res[f"{name}_device"]["latency_ms"],
               "latency_with_map_ms": res[f"{name}_device"]["latency_with_map_ms"],
               "throughput_kf_per_s": res[f"{name}_device"]["throughput_kf_per_s"],
               "success_rate": res[f"{name}_device"]["success_rate"],
               "trajectory_equal_to_host": bool(np.array_equal(traj[f"{name}_device"],
                                                               traj[name]))}
        for name in ("learned", "plain")}
    g["workload"] = (f"{n} keyframes of {args.kf_height}x2048 scans, {args.stream_spacing} m "
                     f"apart ({n * args.stream_spacing:.0f} m), 300 trees over "
                     f"{args.stream_forest_length:.0f} x 60 m, 40 % understory clutter, "
                     f"odometry noise {sigma_t:.4f} m / {sigma_yaw:.3f} deg per step; "
                     f"two passes, the second timed; learned tree model (sequential and "
                     f"pipelined runner), plain cylinder fit beside it, each also with "
                     f"the estimator on the GPU (device: the keyframe as one HIP graph); "
                     f"pinned input buffers; latency without the map query and update "
                     f"(latency_with_map_ms with them); drift is the final/max keyframe position error against "
                     f"the true poses, drift_vs_plain against the plain trajectory "
                     f"(two maps built apart), step_vs_plain the per-keyframe pose "
                     f"difference of a plain runner fed the learned run's estimate and "
                     f"submap (first pass), step_plain_vs_plain_other_seed the same for "
                     f"a plain runner against plain with another hypothesis stream; "
                     f"sequence generated on the host in {gen_s:.1f} s (not timed)")
    return g


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only (a batch would shard by clusters, "
              "size-balanced, with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return

    import numpy as np
    import torch

    from sloam_amd import synth
    from sloam_amd.gpu import Context, CylinderFitPlan, autotune_cylinder_fit

    thr, max_radius, seed = 0.05, 1.0, 0x5EED
    d = synth.make_clusters(args.clusters, seed=1)
    n_points = int(d["offsets"][-1])
    ctx = Context(0)
    dev = torch.device("cuda", 0)
    xyz = torch.as_tensor(d["xyz"]).to(dev)
    off = torch.as_tensor(d["offsets"]).to(dev)
    chunk, tile = args.hyp_chunk, args.tile
    tuned = False
    if chunk == 0 and tile == 0 and not args.no_autotune:
        # launch shape picked by sloam_cyl_autotune on this batch (untimed)
        chunk, tile = autotune_cylinder_fit(ctx, xyz, off, d["offsets"], n_hyp=args.hyp, thr=thr,
                                            max_radius=max_radius, seed=seed)
        tuned = True
    plan = CylinderFitPlan(ctx, d["offsets"], n_hyp=args.hyp, thr=thr, max_radius=max_radius,
                           seed=seed, hyp_chunk=chunk, tile_points=tile)
    out = plan.alloc_outputs(counts=False, mask=True, best=True)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        plan.exec(xyz, off, out)
    ctx.sync()

    # timed region: HIP events on the context stream + per-kernel events
    stream = ctx.torch_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ctx.reset_timing()
    ctx.enable_timing(True)
    torch.cuda.synchronize()
    ctx.sync()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        plan.exec(xyz, off, out)
    ev1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.enable_timing(False)
    gpu_ms = ev0.elapsed_time(ev1)
    wall_ms = (t1 - t0) * 1e3
    ms_per_step = gpu_ms / args.steps
    fits_per_s = args.clusters * args.steps / (gpu_ms / 1e3)

    kern = {}
    for name in ("cyl_seed", "cyl_score", "cyl_refit"):
        ms, n = ctx.kernel_time(name)
        kern[name] = ms / max(n, 1)

    # the same batch recorded once into a HIP graph and replayed, one launch
    # per step; the replays' models must equal the eager run's bit for bit
    graph_ms = None
    if not args.no_graph:
        ref_model = out["model"].clone()
        ctx.graph_begin()
        plan.exec(xyz, off, out)
        graph = ctx.graph_end()
        graph.launch()
        ctx.sync()
        ev0.record(stream)
        for _ in range(args.steps):
            graph.launch()
        ev1.record(stream)
        ctx.sync()
        graph_ms = ev0.elapsed_time(ev1) / args.steps
        if not torch.equal(out["model"], ref_model):
            raise RuntimeError("graph replay differs from the eager fit")
        graph.close()
    eager_fits_per_s = fits_per_s
    launch = "eager"
    if graph_ms is not None and graph_ms < ms_per_step:
        launch = "graph"
        ms_per_step = graph_ms
        fits_per_s = args.clusters / (graph_ms / 1e3)

    # dominant kernel: cyl_score.  Algorithmic work per (point, hypothesis)
    # pair: 2 FMA + 2 compares = 6 FP32 ops; bytes: points once (12 B) +
    # counts (4 B per hypothesis).
    pairs = n_points * args.hyp
    score_s = kern["cyl_score"] / 1e3
    ops = 6.0 * pairs
    score_bytes = 12.0 * n_points + 4.0 * args.clusters * args.hyp
    achieved_tflops = ops / score_s / 1e12
    achieved_gbs = score_bytes / score_s / 1e9
    traffic = pmc_traffic(args.clusters, args.hyp, n_points, chunk, tile)

    # accuracy against the generator's truth (informational)
    ctx.sync()
    model = out["model"].cpu().numpy()
    rerr = np.abs(model[:, 6] - d["radius"])

    small = None
    if not args.no_small:
        small = bench_small_keyframe(ctx, args)
    pose = None
    if not args.no_pose:
        pose = bench_pose(ctx, args)
    learned = None
    if not args.no_learned:
        learned = bench_learned(ctx, args)
    keyframes = None
    stream = None
    if not args.no_keyframes:
        keyframes = bench_keyframes(args)
        if not args.no_stream and args.stream_keyframes > 0:
            try:
                stream = bench_keyframe_stream(args)
            except (OSError, RuntimeError, subprocess.SubprocessError) as e:
                # reported in the line (the keyframe headline then comes from
                # the short sequence); the other legs' figures still print
                msg = f"{type(e).__name__}: {e}"[-800:]
                print(f"[bench] long keyframe stream failed: {msg}", file=sys.stderr, flush=True)
                stream = {"skipped": f"failed: {msg}"}

    cpu = None
    cpu_threaded = None
    if not args.no_cpu:
        import oracle
        # bounded sample: the whole batch, fitted six times (one warm-up, the
        # median of five), one thread (the reference is single-threaded); then
        # the same with --cpu-pool threads over clusters
        oracle.load()

        def cpu_fit(threads):
            return cpu_median(lambda: oracle.cyl_fit_batch(
                d["xyz"], d["offsets"], args.hyp, thr=thr, max_radius=max_radius, seed=seed,
                counts=False, mask=True, threads=threads))

        med, tot = faithful(lambda: cpu_fit(1))
        host = host_cpu()
        cpu = {
            "value": args.clusters / med,
            "unit": "cylinder fits/s",
            "cores": 1,
            "kind": "port",
            "host_cpu": host,
            "sample": f"all {args.clusters} clusters ({n_points} points), {args.hyp} "
                      f"hypotheses, same seed, oracle built with the reference's flags (-O3 "
                      f"-DNDEBUG), 1 thread (the reference is single-threaded); median of 5 "
                      f"runs after 1 warm-up, {tot:.1f} s in total",
        }
        if args.cpu_pool > 1:
            med_t, tot_t = cpu_fit(args.cpu_pool)
            cpu_threaded = {"value": args.clusters / med_t, "unit": "cylinder fits/s",
                            "cores": args.cpu_pool,
                            "sample": f"the same, oracle built -O3 -mfma -fopenmp, "
                                      f"{args.cpu_pool} threads over clusters; median of 5 "
                                      f"runs after 1 warm-up, {tot_t:.1f} s in total"}

    # keyframe headline: the long stream when it ran, else the short sequence
    head, head_name = None, None
    if stream is not None and "skipped" not in stream:
        head, head_name = stream, "keyframe_stream"
    elif keyframes is not None:
        head, head_name = keyframes, "keyframe"
    line = {
        "metric": METRIC,
        "value": fits_per_s,
        "unit": "cylinder fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched cylinder fitting: {args.clusters} clusters of "
                        f"log-U[50,5000] points ({n_points} total), {args.hyp} RANSAC "
                        f"hypotheses each, 5 cm threshold, LO refit",
            "clusters": args.clusters,
            "hypotheses": args.hyp,
            "points": n_points,
            "parallelism": "single GPU",
            "launch_shape": {"hyp_chunk": chunk, "tile_points": tile, "autotuned": tuned},
            "launch": launch,
        },
        "eager_fits_per_s": eager_fits_per_s,
        "graph_ms_per_step": graph_ms,
        "roofline": {
            "bound": "valu",
            "kernel": "cyl_score",
            "achieved": achieved_tflops,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_source": None if traffic is None else os.path.relpath(PMC_SUMMARY, ROOT),
            "work_per_launch": f"{pairs} point-hypothesis pairs x 6 FP32 ops",
            "avg_launch_ms": kern["cyl_score"],
            "hbm_achieved_gbs": achieved_gbs,
            "hbm_frac": achieved_gbs / HBM_PEAK_GBS,
        },
        "kernel_ms": kern,
        "wall_ms": wall_ms,
        "radius_err_vs_truth_cm": {"median": float(np.median(rerr) * 100),
                                   "p90": float(np.percentile(rerr, 90) * 100)},
        "cpu_baseline": cpu,
        "cpu_multithreaded": cpu_threaded,
        "keyframe_latency_ms": None if head is None else
        {"p50": head["latency_ms"]["p50"], "p99": head["latency_ms"]["p99"]},
        "keyframe_throughput_kf_per_s": None if head is None else
        {"sequential": head["throughput_kf_per_s"],
         "pipelined": head["pipelined"]["throughput_kf_per_s"]},
        "keyframe_device_latency_ms": None if head is None or "device" not in head else
        {"p50": head["device"]["learned"]["latency_ms"]["p50"],
         "p99": head["device"]["learned"]["latency_ms"]["p99"],
         "trajectory_equal_to_host": head["device"]["learned"]["trajectory_equal_to_host"]},
        "keyframe_drift_vs_plain": None if head is None else head["drift_vs_plain"],
        "keyframe_step_vs_plain": None if head is None else head["step_vs_plain"],
        "keyframe_step_plain_vs_plain_other_seed": None if head is None else
        head.get("step_plain_vs_plain_other_seed"),
        "keyframe_headline": head_name,
        "small_keyframe": small,
        "pose": pose,
        "learned": learned,
        "keyframe": keyframes,
        "keyframe_stream": stream,
    }
    print(json.dumps(line), flush=True)
    plan.close()
    ctx.close()


if __name__ == "__main__":
    main()
