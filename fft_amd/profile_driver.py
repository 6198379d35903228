"""Small driver for rocprofv3 runs: executes one workload a few times so a
profiler can attribute kernel time and counters.  Used by the profile summaries
under profiles/ (the exact command sits in each summary).

    python -m fft_amd.profile_driver c2c_2p20 [iters]
    python -m fft_amd.profile_driver copy [iters]
"""
from __future__ import annotations

import sys

import numpy as np


def main():
    import torch
    from fft_amd import _lib
    from fft_amd.afft import Plan

    work = sys.argv[1] if len(sys.argv) > 1 else "c2c_2p20"
    iters = int(sys.argv[2]) if len(sys.argv) > 2 else 3
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(1)
    print("workload", work, "iters", iters, flush=True)
    if work == "copy":
        x = torch.rand(128 << 20, device=dev, generator=g)  # 512 MiB
        y = torch.empty_like(x)
        for _ in range(iters):
            _lib.lib().afft_gpu_copy(int(x.data_ptr()), int(y.data_ptr()), x.numel() * 4, None)
    else:
        shapes = {
            "c2c_2p20": ("c", (1 << 20,), 64),
            "c2c_2p16": ("c", (1 << 16,), 1024),
            "c2c_2p12": ("c", (1 << 12,), 16384),
            "r2c_2p22": ("r", (1 << 22,), 16),
            "c2c_256c": ("c", (256, 256, 256), 1),
            "r2c_8192s": ("r", (8192, 8192), 1),
        }
        kind, dims, batch = shapes[work]
        n = int(np.prod(dims))
        if kind == "c":
            x = torch.rand(2 * n * batch, device=dev, generator=g)
            y = torch.empty_like(x)
        else:
            x = torch.randn(n * batch, device=dev, generator=g)
            hs = (dims[0] // 2 + 1) * int(np.prod(dims[1:]))
            y = torch.empty(2 * hs * batch, device=dev)
        p = Plan(kind, dims, batch)
        print("passes:", p.describe(), flush=True)
        for _ in range(iters):
            p(x, None, y, None)
    torch.cuda.synchronize()
    print("done", flush=True)


if __name__ == "__main__":
    main()
