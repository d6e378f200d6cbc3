"""Device copy of a known size (256 MiB, three times), for scaling the
rocprofv3 HBM byte counters (FETCH_SIZE, WRITE_SIZE) of the library's
kernels: the counter's reading for this copy divided by 256 MiB is the
factor by which a kernel's reading relates to bytes moved.

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o copy -- python3 sloam_amd/bench/copy_probe.py"""
import torch

n = 64 * 1024 * 1024  # float32 elements: 256 MiB
x = torch.ones(n, device="cuda")
y = torch.empty_like(x)
torch.cuda.synchronize()
for _ in range(3):
    y.copy_(x)
torch.cuda.synchronize()
print("copied", 3 * n * 4, "bytes", flush=True)
