"""Benchmark of the afft hot path on one MI355X.

One step = the batched c2c workload afft's headline config names: 64 complex
FP32 transforms of N = 2^20 (interleaved, 512 MiB), forward then inverse,
inputs uniform[-1, 1) already resident in HBM.  value = GFLOP/s with the
5*N*log2(N) convention per transform and direction.

The JSON line also carries
  roofline      the dominant kernel (the fused four-step launch, or the pass
                of the two-pass plan that takes longest), achieved algorithmic
                GB/s against the 8 TB/s HBM peak, from HIP events on the
                stream the kernel runs on, timed inside steps; beside it the
                step's own fraction and, when profiles/ holds a rocprofv3
                kernel-trace summary of this command, the profiler's average;
  cpu_baseline  afft itself (oracle/_ref, built from the reference sources) on
                this host's cores over a bounded sample of the same workload;
  extra         the rest of the afft configs measured the same way (r2c
                16 x 2^22, c2c and r2c 256^3, r2c 8192^2, FP64 c2c 16 x 2^20, the
                device-copy ceiling),
                the MFMA split-precision mode against the FP32-FMA plan on
                4096 x 2^16 with 10^u amplitudes (time and error vs FP64),
                and the 10,000-job uneven batch through the persistent work
                queue (eager and graph-replay p50/p99).

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
    ap.add_argument("--steps", type=int, default=50)
    # the first ~40 ms of steps on a fresh box run about 8 % slower (0.85 against
    # 0.78 ms per step) while the clocks settle: 60 untimed steps (~50 ms) first
    ap.add_argument("--warmup", type=int, default=60)
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


def time_median(torch, stream, fn, iters, reps=5):
    """Median over reps of time_events(fn, iters): the side workloads are
    short enough that one sample moves with the clocks by several percent."""
    return sorted(time_events(torch, stream, fn, iters) for _ in range(reps))[reps // 2]


def cpu_info():
    """(model name, logical CPUs of the host, CPUs this process may run on,
    threads the baseline uses).  The thread count is the affinity set, capped
    by OMP_NUM_THREADS when set (a shared host's CPU share for this job)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = avail
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    threads = int(os.environ.get("AFFT_BENCH_CPU_THREADS", threads))
    return model, os.cpu_count() or 1, avail, max(1, threads)


def cpu_baseline(seconds: float):
    """afft (reference build) on the host: 2^20 c2c forward + inverse,
    interleaved, one plan per thread, as many transforms as fit in about
    `seconds` of wall time (measured, then scaled to GFLOP/s)."""
    from oracle import oracle as O
    if not O.ref_available():
        return None
    R = O.RefAfft()
    n = 1 << 20
    model, nproc, avail, threads = cpu_info()
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
    return res, threads, isa, model, nproc, avail


def measured_traffic(describe: str, launch_transforms: int):
    """HBM bytes per launch of the dominant pass from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/*_pmc.json, written by
    fft_amd.profile_summary), if one matches this pass and launch size."""
    import glob
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_pmc.json"))):
        try:
            with open(f) as fh:
                js = json.load(fh)
        except (OSError, ValueError):
            continue
        if js.get("describe") == describe and int(js.get("launch_transforms", -1)) == int(launch_transforms):
            return int(js["traffic_bytes_per_launch"]), os.path.relpath(f, ROOT)
    return None, None


def rocprof_summary(describe: str):
    """Average duration of the dominant kernel from the committed rocprofv3
    kernel-trace summary of this bench command (profiles/*_ktrace.json,
    written by fft_amd.profile_summary), if one matches the kernel."""
    import glob
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_ktrace.json"))):
        try:
            with open(f) as fh:
                js = json.load(fh)
        except (OSError, ValueError):
            continue
        if js.get("describe") == describe and "avg_ns" in js:
            return {"avg_us": round(js["avg_ns"] / 1e3, 2), "source": os.path.relpath(f, ROOT)}
    return None


def bench_mfma(torch, dev, stream, g, iters=5):
    """4096 c2c transforms of N=2^16, each scaled by 10^u, u ~ U[-3, 3):
    plain FP32-FMA plan against the split-precision MFMA mode, both timed on
    the same buffers, both checked per transform against the device FP64
    plan (max rel-L2 over the 4096 transforms, bound 2*16*2^-24)."""
    from fft_amd.afft import Plan
    n, batch = 1 << 16, 4096
    u = torch.rand(batch, 1, device=dev, generator=g, dtype=torch.float64) * 6 - 3
    x = (torch.rand(batch, 2 * n, device=dev, generator=g, dtype=torch.float64) * 2 - 1) * torch.pow(10.0, u)
    x32 = x.float().contiguous()
    x64 = x32.double().contiguous()
    del x, u
    y = torch.empty_like(x32)
    ref = torch.empty_like(x64)
    p64 = Plan("c", (n,), batch, dtype=np.float64)
    p64(x64, None, ref, None, stream)
    del x64
    pf = Plan("c", (n,), batch)
    pm = Plan("c", (n,), batch)
    pm.set_mode(1)
    pb = Plan("c", (n,), batch)
    pb.set_mode(2)
    out = {}
    bound = 2.0 * 16 * 2.0 ** -24
    for tag, p in (("fp32", pf), ("mfma", pm), ("mfma_bf16", pb)):
        p(x32, None, y, None, stream)
        ms = time_events(torch, stream, lambda: p(x32, None, y, None, stream), iters)
        d = (y.double() - ref).pow_(2).sum(1)
        r = ref.pow(2).sum(1)
        err = torch.sqrt(d / r)
        out[tag] = (ms, float(err.max()), float(err.mean()))
    del y, ref, x32, pf, pm, pb, p64
    torch.cuda.empty_cache()
    (mf, ef, af), (mm, em, am), (mb, eb, ab) = out["fp32"], out["mfma"], out["mfma_bf16"]
    return {"workload": "4096 x c2c N=2^16 FP32 interleaved, amplitudes 10^u, u~U[-3,3) per transform",
            "fp32_fma_ms": round(mf, 4), "mfma_ms": round(mm, 4), "mfma_speedup": round(mf / mm, 3),
            "fp32_err_max": ef, "fp32_err_mean": af, "mfma_err_max": em, "mfma_err_mean": am,
            "mfma_bf16_ms": round(mb, 4), "mfma_bf16_speedup": round(mf / mb, 3), "mfma_bf16_err_max": eb,
            "mfma_bf16_err_mean": ab,
            "err_bound": bound, "mfma_GFLOPs": round(flops_c2c(n, batch) / (mm * 1e-3) / 1e9, 1),
            "fp32_GFLOPs": round(flops_c2c(n, batch) / (mf * 1e-3) / 1e9, 1)}


def bench_queue(torch, dev, stream, reps=40):
    """10,000 1-D FP32 jobs, N = 2^k with k ~ U{6..16}, 70% complex / 30%
    real, direction at random, in one persistent launch per batch; timed
    eagerly and as a replayed HIP graph (per-batch latency p50/p99)."""
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
        lay.append((kind, lg, so, do, nin, nout))
    buf = torch.rand(off, device=dev, dtype=torch.float32)
    base = int(buf.data_ptr())
    q = Queue(10000, 16)
    jobs = Queue.jobs([(k, lg, base + 4 * so, base + 4 * do) for (k, lg, so, do, _, _) in lay])
    t0 = time.perf_counter()
    ntask = q.prepare(jobs)
    prep_ms = (time.perf_counter() - t0) * 1e3
    # the library's queue autotuner picks groups per CU and the chain lag on
    # this batch and leaves the queue prepared with them
    q.launch(stream)  # the first launch also generates the twiddle tables of every job size
    untuned_ms = time_events(torch, stream, lambda: q.launch(stream), 10)
    tune_ms = q.tune(jobs, stream, 5)
    groups, lag = q.config
    ms = time_events(torch, stream, lambda: q.launch(stream), 10)
    algb = sum(4 * (nin + nout) for (_, _, _, _, nin, nout) in lay)
    fl = sum((5.0 if k < 2 else 2.5) * (1 << lg) * lg for (k, lg, _, _, _, _) in lay)
    s2 = torch.cuda.Stream()
    gr = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s2):
        with torch.cuda.graph(gr, stream=s2):
            q.launch(s2)
    torch.cuda.synchronize()
    lat = []
    for i in range(reps + 5):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        gr.replay()
        b.record(stream)
        torch.cuda.synchronize()
        if i >= 5:
            lat.append(a.elapsed_time(b))
    lat = np.array(lat)
    del gr
    q.close()
    del buf
    torch.cuda.empty_cache()
    return {"workload": "10000 x 1-D FP32 jobs, N=2^k k~U{6..16}, 70% c2c / 30% real, fwd/inv at random",
            "tasks": int(ntask), "prepare_host_ms": round(prep_ms, 3), "launch_ms": round(ms, 4),
            "tuned": {"groups_per_cu": groups, "lag": lag, "tune_ms": round(tune_ms, 4),
                      "untuned_launch_ms": round(untuned_ms, 4)},
            "GFLOPs": round(fl / (ms * 1e-3) / 1e9, 1), "algorithmic_GBps": round(algb / (ms * 1e-3) / 1e9, 1),
            "graph_replay_p50_ms": round(float(np.percentile(lat, 50)), 4),
            "graph_replay_p99_ms": round(float(np.percentile(lat, 99)), 4), "graph_replays": reps}


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

    # the dominant kernel, timed inside steps: the same step sequence with
    # HIP events on `stream` around every pass range (a fused plan runs its
    # passes as one launch, so its range is the whole transform)
    def ranges(p):
        return [(0, p.num_passes)] if p.describe()[0].startswith("fft[fused4]") else \
            [(i, i + 1) for i in range(p.num_passes)]
    spans = [(p, a, b, src, dst) for p, src, dst in ((fwd, x, y), (inv, y, z)) for a, b in ranges(p)]
    tot = [0.0] * len(spans)
    torch.cuda.synchronize()
    for _ in range(args.steps):
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(len(spans) + 1)]
        evs[0].record(stream)
        for k, (p, a, b, src, dst) in enumerate(spans):
            p.run_range(a, b, src, None, dst, None, stream)
            evs[k + 1].record(stream)
        torch.cuda.synchronize()
        for k in range(len(spans)):
            tot[k] += evs[k].elapsed_time(evs[k + 1])
    per_range = [t / args.steps for t in tot]
    # pass ranges of the two directions share their kernel shapes: pick the
    # range of the forward plan whose forward + inverse time is largest
    nf = len(ranges(fwd))
    comb = [per_range[k] + per_range[nf + k] for k in range(nf)]
    kdom = int(np.argmax(comb))
    a0, b0 = ranges(fwd)[kdom]
    fused = b0 - a0 > 1
    launches = 1 if fused else fwd.launches(a0)
    per_launch_ms = comb[kdom] / 2 / launches
    bytes_per_launch = 16.0 * n * (batch / launches)
    achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    desc_a = fwd.describe()[a0]
    traffic, traffic_src = measured_traffic(desc_a, batch // launches)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "kernel": ("fused four-step (both passes, one launch): " if fused else f"pass {a0}: ") + desc_a,
                "avg_launch_us": round(per_launch_ms * 1e3, 2),
                "algorithmic_bytes_per_launch": int(bytes_per_launch),
                "step_frac": round(hbm_frac_step, 4),
                "traffic_source": traffic_src}
    # every pass of the step, not only the dominant one: algorithmic bytes of
    # its launches (16 B per element read + written) over its in-step time
    roofline["passes"] = []
    for k in range(nf):
        a, b = ranges(fwd)[k]
        nl = 1 if b - a > 1 else fwd.launches(a)
        t_launch = comb[k] / 2 / nl
        bpl = 16.0 * n * (batch / nl)
        roofline["passes"].append({"kernel": fwd.describe()[a], "launches_per_direction": nl,
                                   "avg_launch_us": round(t_launch * 1e3, 2),
                                   "achieved": round(bpl / (t_launch * 1e-3) / 1e9, 1),
                                   "frac": round(bpl / (t_launch * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)})
    rp = rocprof_summary(desc_a)
    if rp:
        roofline["rocprof_avg_us"] = rp["avg_us"]
        roofline["frac_rocprof"] = round(bytes_per_launch / (rp["avg_us"] * 1e-6) / 1e9 / HBM_PEAK_GBPS, 4)
        roofline["rocprof_source"] = rp["source"]

    extra = {"step_wall_ms": round(wall, 3), "step_hbm_frac_algorithmic": round(hbm_frac_step, 4),
             "in_step_ms_per_range": {f"{'fwd' if k < nf else 'inv'} passes {spans[k][1]}-{spans[k][2] - 1}":
                                      round(per_range[k], 4) for k in range(len(spans))},
             "chunk": int(fwd.get_variant().chunk), "dominant_launch_transforms": int(batch / launches)}
    # measured copy ceiling on the same buffers
    ms_copy = time_events(torch, stream, lambda: _lib.lib().afft_gpu_copy(int(x.data_ptr()), int(z.data_ptr()),
                                                                         x.numel() * 4, stream.cuda_stream), 10)
    extra["copy_ceiling_GBps"] = round(2 * x.numel() * 4 / (ms_copy * 1e-3) / 1e9, 1)
    extra["step_frac_of_copy_ceiling"] = round(hbm_frac_step * HBM_PEAK_GBPS / extra["copy_ceiling_GBps"], 4)

    if not args.no_extra:
        # the side workloads run on their own: the headline's plans and
        # buffers are released first (with them alive, 16 x r2c 2^22 measured
        # 0.42 instead of 0.40 ms)
        fwd.close()
        inv.close()
        del z, x, y
        torch.cuda.empty_cache()
        # r2c 16 x 2^22 (standard normal)
        nr, br = 1 << 22, 16
        xr = torch.randn(nr * br, device=dev, generator=g, dtype=torch.float32)
        yr = torch.empty(2 * (nr // 2 + 1) * br, device=dev, dtype=torch.float32)
        pr = Plan("r", (nr,), br)
        pr(xr, None, yr, None, stream)
        msr = time_median(torch, stream, lambda: pr(xr, None, yr, None, stream), 10)
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
        ms3 = time_median(torch, stream, lambda: p3(x3, None, y3, None, stream), 10)
        extra["c2c_256^3_ms"] = round(ms3, 4)
        extra["c2c_256^3_GFLOPs"] = round(flops_c2c(n3, 1) / (ms3 * 1e-3) / 1e9, 1)
        extra["c2c_256^3_hbm_frac"] = round(16 * n3 / (ms3 * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
        del x3, y3, p3
        # r2c 256^3 (standard normal)
        x3 = torch.randn(n3, device=dev, generator=g, dtype=torch.float32)
        y3 = torch.empty(2 * 129 * 256 * 256, device=dev, dtype=torch.float32)
        p3 = Plan("r", (256, 256, 256), 1)
        p3(x3, None, y3, None, stream)
        ms3 = time_median(torch, stream, lambda: p3(x3, None, y3, None, stream), 10)
        extra["r2c_256^3_ms"] = round(ms3, 4)
        extra["r2c_256^3_GFLOPs_2.5NlogN"] = round(2.5 * n3 * 24 / (ms3 * 1e-3) / 1e9, 1)
        extra["r2c_256^3_hbm_frac"] = round((4 * n3 + 8 * 129 * 256 * 256) / (ms3 * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
        del x3, y3, p3
        # r2c 8192^2
        n2 = 8192 * 8192
        x2 = torch.randn(n2, device=dev, generator=g, dtype=torch.float32)
        y2 = torch.empty(2 * 4097 * 8192, device=dev, dtype=torch.float32)
        p2 = Plan("r", (8192, 8192), 1)
        p2(x2, None, y2, None, stream)
        ms2 = time_median(torch, stream, lambda: p2(x2, None, y2, None, stream), 10)
        extra["r2c_8192^2_ms"] = round(ms2, 4)
        extra["r2c_8192^2_GFLOPs_2.5NlogN"] = round(2.5 * n2 * 26 / (ms2 * 1e-3) / 1e9, 1)
        extra["r2c_8192^2_hbm_frac"] = round((4 * n2 + 8 * 4097 * 8192) / (ms2 * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
        del x2, y2, p2
        # FP64 c2c 16 x 2^20 (the afft_64_* path)
        nd, bd = 1 << 20, 16
        xd = torch.rand(2 * nd * bd, device=dev, generator=g, dtype=torch.float64).mul_(2).sub_(1)
        yd = torch.empty_like(xd)
        pd = Plan("c", (nd,), bd, dtype=np.float64)
        pd(xd, None, yd, None, stream)
        msd = time_median(torch, stream, lambda: pd(xd, None, yd, None, stream), 10)
        extra["c2c_fp64_16x2^20_ms"] = round(msd, 4)
        extra["c2c_fp64_16x2^20_GFLOPs"] = round(flops_c2c(nd, bd) / (msd * 1e-3) / 1e9, 1)
        extra["c2c_fp64_16x2^20_hbm_frac"] = round(32 * nd * bd / (msd * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
        del xd, yd, pd
        torch.cuda.empty_cache()
        # batches at both ends of the row-pass range (256 MiB of complex data
        # each): rows of 16 points on the LDS-staged tiny-row kernel, and the
        # longest one-pass rows on the half-LDS kernel; algorithmic GB/s
        rows = {}
        xs = torch.rand(1 << 26, device=dev, generator=g, dtype=torch.float32)
        ys = torch.empty(1 << 26, device=dev, dtype=torch.float32)
        for name, kind, nr_, b in (("c2c_2^4_x2^21", "c", 16, 1 << 21), ("r2c_2^5_x2^20", "r", 32, 1 << 20),
                                   ("c2c_2^15_x1024", "c", 1 << 15, 1024), ("c2r_2^16_x512", "ri", 1 << 16, 512)):
            ps = Plan(kind, (nr_,), b)
            ps(xs, None, ys, None, stream)
            mss = time_median(torch, stream, lambda: ps(xs, None, ys, None, stream), 10)
            m = nr_ // 2 + 1
            byts = 16 * nr_ * b if kind == "c" else (4 * nr_ + 8 * m) * b
            rows[name] = {"ms": round(mss, 4), "GBps": round(byts / (mss * 1e-3) / 1e9, 1),
                          "pass": ps.describe()[0].split(" in(")[0]}
            ps.close()
        extra["row_pass_ends"] = rows
        del xs, ys
        torch.cuda.empty_cache()
        extra["mfma_vs_fp32"] = bench_mfma(torch, dev, stream, g)
        extra["uneven_queue"] = bench_queue(torch, dev, stream)

    cpu = None
    if not args.no_cpu:
        r = cpu_baseline(args.cpu_seconds)
        if r is not None:
            res, threads, isa, model, nproc, avail = r
            v1 = res[1]
            vt = res[threads]
            cpu = {"value": round(vt[0], 2), "unit": "GFLOP/s", "cores": threads, "kind": "reference",
                   "sample": f"afft ({isa}, oracle/_ref) c2c N=2^20 interleaved fwd+inv, one plan per thread, "
                             f"{vt[1]} transforms in {vt[2]:.1f} s on {threads} threads; single thread: "
                             f"{v1[0]:.2f} GFLOP/s ({v1[1]} transforms in {v1[2]:.1f} s)",
                   "single_thread_value": round(v1[0], 2), "cpu_model": model, "isa": isa,
                   "host_logical_cpus": nproc, "cpus_available": avail}

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
