"""Benchmark of the afft hot path on one MI355X.

One step = the batched c2c workload afft's headline config names: 64 complex
FP32 transforms of N = 2^20 (interleaved, 512 MiB), forward then inverse,
inputs uniform[-1, 1) already resident in HBM.  value = GFLOP/s with the
5*N*log2(N) convention per transform and direction.

The JSON line also carries
  roofline      the dominant kernel (four-step pass A, the 1024-point column
                pass with the W_N twiddle fused), achieved algorithmic GB/s
                against the 8 TB/s HBM peak, from HIP events on the stream the
                kernel runs on;
  cpu_baseline  afft itself (oracle/_ref, built from the reference sources) on
                this host's cores over a bounded sample of the same workload;
  extra         the rest of the afft configs measured the same way (r2c
                16 x 2^22, c2c 256^3, r2c 8192^2, the device-copy ceiling).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    return ap.parse_args()


def flops_c2c(n, count):
    return 5.0 * n * math.log2(n) * count


def time_events(torch, stream, fn, iters):
    """Average ms of fn() over iters, HIP events on `stream`."""
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record(stream)
    for _ in range(iters):
        fn()
    e.record(stream)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / iters


def cpu_baseline(seconds: float):
    """afft (reference build) on the host: 2^20 c2c forward + inverse,
    interleaved, one plan per thread, as many transforms as fit in about
    `seconds` of wall time (measured, then scaled to GFLOP/s)."""
    from oracle import oracle as O
    if not O.ref_available():
        return None
    R = O.RefAfft()
    n = 1 << 20
    threads = int(os.environ.get("AFFT_BENCH_CPU_THREADS", min(16, os.cpu_count() or 1)))
    x = O.complex_input(1, (n,))
    res = {}
    for nthr in sorted({1, threads}):
        plans = [(R.plan("c", (n,), O.AFFT_INTERLEAVED, O.AFFT_AUTO), R.plan("ci", (n,), O.AFFT_INTERLEAVED, O.AFFT_AUTO))
                 for _ in range(nthr)]
        bufs = [(x.copy(), np.empty_like(x)) for _ in range(nthr)]
        # warm-up, one pair per thread
        for t in range(nthr):
            R.run(plans[t][0], bufs[t][0], None, bufs[t][1], None)
        counts = [0] * nthr
        stop = time.perf_counter() + seconds / 2

        def work(t):
            pf, pi = plans[t]
            a, b = bufs[t]
            while time.perf_counter() < stop:
                R.run(pf, a, None, b, None)
                R.run(pi, b, None, a, None)
                counts[t] += 2
                np.multiply(a, 1.0 / n, out=a)  # keep magnitudes bounded

        t0 = time.perf_counter()
        ths = [threading.Thread(target=work, args=(t,)) for t in range(nthr)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        el = time.perf_counter() - t0
        total = sum(counts)
        res[nthr] = (flops_c2c(n, total) / el / 1e9, total, el)
    isa = "avx512f" if R.supported(O.AFFT_AVX512F) else "avx2" if R.supported(O.AFFT_AVX2) else "scalar"
    return res, threads, isa


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the afft hot path has no multi-GPU step (independent transforms shard "
              "by batch with no communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from fft_amd import _lib
    from fft_amd.afft import Plan

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    n, batch = 1 << 20, 64
    g = torch.Generator(device=dev).manual_seed(20)
    x = torch.rand(2 * n * batch, device=dev, generator=g, dtype=torch.float32).mul_(2).sub_(1)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    fwd = Plan("c", (n,), batch)
    inv = Plan("ci", (n,), batch)

    def step():
        fwd(x, None, y, None, stream)
        inv(y, None, z, None, stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms_ev = time_events(torch, stream, step, args.steps)
    wall = (time.perf_counter() - t0) / args.steps * 1e3
    ms = ms_ev
    gflops = flops_c2c(n, 2 * batch) / (ms * 1e-3) / 1e9
    hbm_frac_step = 2 * batch * 16.0 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS

    # dominant kernel: pass A of the forward plan, timed on its own
    launches = fwd.launches(0)
    ms_a = time_events(torch, stream, lambda: fwd.run_range(0, 1, x, None, y, None, stream), args.steps)
    ms_b = time_events(torch, stream, lambda: fwd.run_range(1, 2, x, None, y, None, stream), args.steps)
    per_launch_ms = ms_a / launches
    bytes_per_launch = 16.0 * n * (batch / launches)
    achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                "kernel": "fft_pass_kernel<float,1024,16,col,EPI_TW4> (four-step pass A)",
                "avg_launch_us": round(per_launch_ms * 1e3, 2),
                "algorithmic_bytes_per_launch": int(bytes_per_launch)}

    extra = {"step_wall_ms": round(wall, 3), "step_hbm_frac_algorithmic": round(hbm_frac_step, 4),
             "pass_a_ms_per_direction": round(ms_a, 4), "pass_b_ms_per_direction": round(ms_b, 4),
             "chunk": int(batch / launches)}
    # measured copy ceiling on the same buffers
    ms_copy = time_events(torch, stream, lambda: _lib.lib().afft_gpu_copy(int(x.data_ptr()), int(z.data_ptr()),
                                                                         x.numel() * 4, stream.cuda_stream), 10)
    extra["copy_ceiling_GBps"] = round(2 * x.numel() * 4 / (ms_copy * 1e-3) / 1e9, 1)
    extra["step_frac_of_copy_ceiling"] = round(hbm_frac_step * HBM_PEAK_GBPS / extra["copy_ceiling_GBps"], 4)

    if not args.no_extra:
        del z
        torch.cuda.empty_cache()
        # r2c 16 x 2^22 (standard normal)
        nr, br = 1 << 22, 16
        xr = torch.randn(nr * br, device=dev, generator=g, dtype=torch.float32)
        yr = torch.empty(2 * (nr // 2 + 1) * br, device=dev, dtype=torch.float32)
        pr = Plan("r", (nr,), br)
        pr(xr, None, yr, None, stream)
        msr = time_events(torch, stream, lambda: pr(xr, None, yr, None, stream), 10)
        extra["r2c_16x2^22_ms"] = round(msr, 4)
        extra["r2c_16x2^22_GFLOPs_2.5NlogN"] = round(2.5 * nr * 22 * br / (msr * 1e-3) / 1e9, 1)
        extra["r2c_16x2^22_hbm_frac"] = round((4 * nr + 8 * (nr // 2 + 1)) * br / (msr * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                                              4)
        del xr, yr, pr
        # c2c 256^3
        n3 = 256 ** 3
        x3 = torch.rand(2 * n3, device=dev, generator=g, dtype=torch.float32).mul_(2).sub_(1)
        y3 = torch.empty_like(x3)
        p3 = Plan("c", (256, 256, 256), 1)
        p3(x3, None, y3, None, stream)
        ms3 = time_events(torch, stream, lambda: p3(x3, None, y3, None, stream), 10)
        extra["c2c_256^3_ms"] = round(ms3, 4)
        extra["c2c_256^3_GFLOPs"] = round(flops_c2c(n3, 1) / (ms3 * 1e-3) / 1e9, 1)
        extra["c2c_256^3_hbm_frac"] = round(16 * n3 / (ms3 * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
        del x3, y3, p3
        # r2c 8192^2
        n2 = 8192 * 8192
        x2 = torch.randn(n2, device=dev, generator=g, dtype=torch.float32)
        y2 = torch.empty(2 * 4097 * 8192, device=dev, dtype=torch.float32)
        p2 = Plan("r", (8192, 8192), 1)
        p2(x2, None, y2, None, stream)
        ms2 = time_events(torch, stream, lambda: p2(x2, None, y2, None, stream), 10)
        extra["r2c_8192^2_ms"] = round(ms2, 4)
        extra["r2c_8192^2_GFLOPs_2.5NlogN"] = round(2.5 * n2 * 26 / (ms2 * 1e-3) / 1e9, 1)
        extra["r2c_8192^2_hbm_frac"] = round((4 * n2 + 8 * 4097 * 8192) / (ms2 * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
        del x2, y2, p2
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        r = cpu_baseline(args.cpu_seconds)
        if r is not None:
            res, threads, isa = r
            v1 = res[1]
            vt = res[threads]
            cpu = {"value": round(vt[0], 2), "unit": "GFLOP/s", "cores": threads, "kind": "reference",
                   "sample": f"afft ({isa}, oracle/_ref) c2c N=2^20 interleaved fwd+inv, one plan per thread, "
                             f"{vt[1]} transforms in {vt[2]:.1f} s on {threads} threads; single thread: "
                             f"{v1[0]:.2f} GFLOP/s ({v1[1]} transforms in {v1[2]:.1f} s)",
                   "single_thread_value": round(v1[0], 2)}

    out = {
        "metric": "batched c2c FFT GFLOP/s (5N*log2N), 64 x N=2^20 FP32 fwd+inv",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic uniform[-1,1) (torch generator, seed 20), resident in HBM",
        "config": {"workload": "64 x c2c N=2^20 interleaved FP32, forward then inverse (afft batched config)",
                   "batch": batch, "n": n},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
