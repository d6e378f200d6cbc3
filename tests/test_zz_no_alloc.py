"""sloam_project_partition_noalloc (scratch in a fixed device array, no
allocation) against sloam_project_partition (scratch from the context) through
the C ABI: two calls in a row, the second on a smaller scan, every output
bit-exact; then a request beyond the array's limits fails on the host with
SLOAM_ERR_TOO_LARGE before any kernel runs."""
import ctypes as C

import numpy as np
import pytest
import torch

from sloam_amd import synth
from sloam_amd._lib import ProjCfg
from sloam_amd.gpu import _ptr

pytestmark = pytest.mark.gpu

SLOAM_ERR_TOO_LARGE = -4
OUTPUTS = ("proj", "range", "winner", "mask", "ground", "n_ground", "tree")


def _outputs(n, H, W, dev):
    return {
        "proj": torch.full((n, 2), -7, dtype=torch.int32, device=dev),
        "range": torch.full((H, W), -7.0, dtype=torch.float32, device=dev),
        "winner": torch.full((H, W), -7, dtype=torch.int32, device=dev),
        "mask": torch.full((H, W), 7, dtype=torch.uint8, device=dev),
        "ground": torch.full((n, 4), -7.0, dtype=torch.float32, device=dev),
        "n_ground": torch.full((1,), -7, dtype=torch.int32, device=dev),
        "tree": torch.full((n, 4), -7.0, dtype=torch.float32, device=dev),
    }


def _call(ctx, fn, x, lab, H, W, destagger, own_proj):
    n = x.shape[0]
    out = _outputs(n, H, W, x.device)
    cfg = ProjCfg(height=H, width=W, fov_up_deg=22.5, fov_down_deg=-22.5,
                  destagger=int(destagger))
    st = fn(ctx.handle, _ptr(x), _ptr(lab), n, C.byref(cfg),
            _ptr(out["proj"]) if own_proj else None, _ptr(out["range"]), _ptr(out["winner"]),
            _ptr(out["mask"]), _ptr(out["ground"]), _ptr(out["n_ground"]), _ptr(out["tree"]))
    assert st == 0, (st, ctx.lib.sloam_ctx_last_error(ctx.handle))
    return out


def _host(out):
    r = {k: v.cpu().numpy() for k, v in out.items()}
    r["ground"] = r["ground"][:int(r["n_ground"][0])]
    return r


@pytest.mark.parametrize("own_proj", [False, True])
def test_noalloc_matches_allocating_entry_point(ctx, own_proj):
    dev = torch.device("cuda", 0)
    # larger scan first, then a smaller one (the array keeps the first call's
    # contents; the second call must not depend on them)
    cases = [(128, 2048, False, 11), (64, 2048, True, 12), (32, 1024, False, 13)]
    for H, W, destagger, seed in cases:
        s = synth.make_scan(seed=seed, H=H, W=W)
        print(f"[noalloc] {H}x{W} scan, destagger {destagger}, caller proj {own_proj}",
              flush=True)
        x = torch.as_tensor(s["xyzi"]).to(dev)
        lab = torch.as_tensor(s["label"]).to(dev)
        torch.cuda.synchronize()
        a = _call(ctx, ctx.lib.sloam_project_partition_noalloc, x, lab, H, W, destagger,
                  own_proj)
        ctx.sync()
        b = _call(ctx, ctx.lib.sloam_project_partition, x, lab, H, W, destagger, own_proj)
        ctx.sync()
        ha, hb = _host(a), _host(b)
        assert int(ha["n_ground"][0]) > 0
        for k in OUTPUTS:
            if k == "proj" and not own_proj:
                continue
            assert np.array_equal(ha[k].view(np.uint8), hb[k].view(np.uint8)), k


def test_noalloc_rejects_sizes_beyond_its_array(ctx):
    dev = torch.device("cuda", 0)
    n = 16
    x = torch.zeros((n, 4), dtype=torch.float32, device=dev)
    lab = torch.zeros((n,), dtype=torch.uint8, device=dev)
    out = _outputs(n, 1, 1, dev)
    for H, W in ((512, 2048), (1024, 1024)):  # 1,048,576 pixels > 524,288
        cfg = ProjCfg(height=H, width=W, fov_up_deg=22.5, fov_down_deg=-22.5, destagger=0)
        st = ctx.lib.sloam_project_partition_noalloc(
            ctx.handle, _ptr(x), _ptr(lab), n, C.byref(cfg), None, None, None, _ptr(out["mask"]),
            _ptr(out["ground"]), _ptr(out["n_ground"]), None)
        assert st == SLOAM_ERR_TOO_LARGE
        assert b"524288" in ctx.lib.sloam_ctx_last_error(ctx.handle)
