"""Small driver for rocprofv3 runs: executes one workload a few times so a
profiler can attribute kernel time and counters.  Used by the profile summaries
under profiles/ (the exact command sits in each summary).

    python -m fft_amd.profile_driver c2c_2p20 [iters]
    python -m fft_amd.profile_driver copy [iters]
    python -m fft_amd.profile_driver queue [iters]

Workloads: c2c_2p20 (64 x 2^20), c2c_2p16 (1024 x 2^16), c2c_2p12, r2c_2p22
(16 x 2^22), r2c_2p20 (64 x 2^20; r2c_2p20_tr with the complex four-step and
its real pass), c2r_2p20 (64 x 2^20), c2r_2p22 (16 x 2^22), c2c_2p25 (2 x 2^25,
three passes), c2c_256c / r2c_256c / c2r_256c (256^3), r2c_8192s (8192^2), c2c_2p16_4096 (4096 x 2^16)
and c2c_2p16_4096_mfma (the same in the split-precision MFMA mode), the short
rows c2c_2p4 (2^21 x 16) and r2c_2p5 (2^20 x 32), the longest one-pass rows
c2c_2p15 (1024 x 2^15) and c2r_2p16 (512 x 2^16), the three-pass c2c_2p23
(8 x 2^23) and c2c_2p24 (4 x 2^24), copy (512 MiB device copy, the counter
calibration), queue (10,000 uneven jobs).
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
    elif work == "queue":
        # 10,000 jobs, N = 2^k, k ~ U{6..16}, 70% c2c / 30% real, fwd/inv at random
        from fft_amd.afft import Queue
        rng = np.random.default_rng(5)
        lay, off = [], 0
        for _ in range(10000):
            lg = int(rng.integers(6, 17))
            real = rng.random() < 0.3
            inv = rng.random() < 0.5
            n = 1 << lg
            kind = (3 if inv else 2) if real else (1 if inv else 0)
            nin = {0: 2 * n, 1: 2 * n, 2: n, 3: 2 * (n // 2 + 1)}[kind]
            nout = {0: 2 * n, 1: 2 * n, 2: 2 * (n // 2 + 1), 3: n}[kind]
            so = off
            off += (nin + 63) // 64 * 64
            do = off
            off += (nout + 63) // 64 * 64
            lay.append((kind, lg, so, do))
        buf = torch.rand(off, device=dev, generator=g)
        base = int(buf.data_ptr())
        q = Queue(10000, 16)
        q.prepare(Queue.jobs([(k, lg, base + 4 * so, base + 4 * do) for (k, lg, so, do) in lay]))
        for _ in range(iters):
            q.launch()
    else:
        shapes = {
            "c2c_2p20": ("c", (1 << 20,), 64),
            "c2c_2p16": ("c", (1 << 16,), 1024),
            "c2c_2p12": ("c", (1 << 12,), 16384),
            "r2c_2p22": ("r", (1 << 22,), 16),
            "r2c_2p20": ("r", (1 << 20,), 64),
            "r2c_2p20_tr": ("r", (1 << 20,), 64),
            "c2c_256c": ("c", (256, 256, 256), 1),
            "r2c_256c": ("r", (256, 256, 256), 1),
            "c2r_256c": ("ri", (256, 256, 256), 1),
            "r2c_8192s": ("r", (8192, 8192), 1),
            "c2r_8192s": ("ri", (8192, 8192), 1),
            "c2r_2p20": ("ri", (1 << 20,), 64),
            "c2r_2p22": ("ri", (1 << 22,), 16),
            "c2c_2p25": ("c", (1 << 25,), 2),
            "c2c_2p16_4096": ("c", (1 << 16,), 4096),
            "c2c_2p16_4096_mfma": ("c", (1 << 16,), 4096),
            "c2c_2p4": ("c", (16,), 1 << 21),
            "r2c_2p5": ("r", (32,), 1 << 20),
            "c2c_2p15": ("c", (1 << 15,), 1024),
            "c2r_2p16": ("ri", (1 << 16,), 512),
            "c2c_2p23": ("c", (1 << 23,), 8),
            "c2c_2p24": ("c", (1 << 24,), 4),
        }
        kind, dims, batch = shapes[work]
        n = int(np.prod(dims))
        if kind == "c":
            x = torch.rand(2 * n * batch, device=dev, generator=g)
            y = torch.empty_like(x)
        elif kind == "ri":
            # timing input: random half spectrum (not Hermitian-consistent)
            hs = (dims[0] // 2 + 1) * int(np.prod(dims[1:]))
            x = torch.randn(2 * hs * batch, device=dev, generator=g)
            y = torch.empty(n * batch, device=dev)
        else:
            x = torch.randn(n * batch, device=dev, generator=g)
            hs = (dims[0] // 2 + 1) * int(np.prod(dims[1:]))
            y = torch.empty(2 * hs * batch, device=dev)
        p = Plan(kind, dims, batch)
        if work.endswith("_mfma"):
            p.set_mode(1)
        if work.endswith("_tr"):
            p.set_variant(real_pair=-1)
        print("passes:", p.describe(), flush=True)
        for _ in range(iters):
            p(x, None, y, None)
    torch.cuda.synchronize()
    print("done", flush=True)


if __name__ == "__main__":
    main()
