"""GPU tests of the uneven-batch persistent queue: every job of a mixed batch
(sizes 2^0..2^16, c / ci / r / ri) against the FP64 oracle, and HIP-graph
replays equal to the eager launch."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

C, CI, R, RI = 0, 1, 2, 3


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def make_jobs(rng, njobs, lo, hi):
    """Mixed jobs: 70 % complex / 30 % real, direction a fair coin, log2 n
    uniform in [lo, hi]; returns host inputs and the layout in one buffer."""
    jobs = []
    off = 0
    for i in range(njobs):
        lg = int(rng.integers(lo, hi + 1))
        real = rng.random() < 0.3
        inv = rng.random() < 0.5
        if real and lg == 0:
            lg = 1
        n = 1 << lg
        kind = (RI if inv else R) if real else (CI if inv else C)
        in_floats = {C: 2 * n, CI: 2 * n, R: n, RI: 2 * (n // 2 + 1)}[kind]
        out_floats = {C: 2 * n, CI: 2 * n, R: 2 * (n // 2 + 1), RI: n}[kind]
        src_off = off
        off += (in_floats + 63) // 64 * 64
        dst_off = off
        off += (out_floats + 63) // 64 * 64
        jobs.append((kind, lg, src_off, in_floats, dst_off, out_floats))
    return jobs, off


def fill_inputs(jobs, total, seed):
    host = np.zeros(total, np.float32)
    for i, (kind, lg, so, nin, do, nout) in enumerate(jobs):
        n = 1 << lg
        if kind == RI:
            h = O.hermitian_input(seed + i, (n,))
            host[so:so + nin] = np.stack([h.real, h.imag], -1).ravel()
        else:
            host[so:so + nin] = O.uniform(seed + i, nin)
    return host


def oracle_out(kind, lg, x):
    n = 1 << lg
    if kind in (C, CI):
        z = (x[0::2] + 1j * x[1::2]).astype(np.complex128)
        return O.fft(z, inverse=(kind == CI))
    if kind == R:
        return O.rfft(x.astype(np.float64))
    z = (x[0::2] + 1j * x[1::2]).astype(np.complex128)
    return O.irfft(z, (n,))


def run_queue(G, jobs, total, host):
    import torch
    from fft_amd.afft import Queue
    buf = G.to_dev(host)
    base = int(buf.data_ptr())
    q = Queue(len(jobs), 16)
    recs = [(k, lg, base + 4 * so, base + 4 * do) for (k, lg, so, nin, do, nout) in jobs]
    ntask = q.prepare(Queue.jobs(recs))
    q.launch()
    torch.cuda.synchronize()
    return q, buf, recs, ntask


@pytest.fixture(scope="module")
def G():
    import gpu_util
    gpu_util.torch_cuda()
    return gpu_util


def check_all(jobs, host, out):
    worst = 0.0
    for i, (kind, lg, so, nin, do, nout) in enumerate(jobs):
        ref = oracle_out(kind, lg, host[so:so + nin])
        y = out[do:do + nout]
        yc = (y[0::2] + 1j * y[1::2]).astype(np.complex64) if kind != RI else y
        n = 1 << lg
        err = O.rel_l2(yc, ref)
        assert err <= O.bound(n), (i, kind, lg, err)
        worst = max(worst, err / O.bound(n))
    return worst


def test_queue_all_sizes_and_kinds(G):
    rng = np.random.default_rng(5)
    jobs, total = make_jobs(rng, 400, 0, 16)
    # make sure every (kind, size) appears at least once
    extra = []
    for kind in (C, CI, R, RI):
        for lg in range(1 if kind >= R else 0, 17):
            extra.append(kind * 100 + lg)
    jobs2, total2 = [], 0
    for code in extra:
        kind, lg = divmod(code, 100)
        n = 1 << lg
        nin = {C: 2 * n, CI: 2 * n, R: n, RI: 2 * (n // 2 + 1)}[kind]
        nout = {C: 2 * n, CI: 2 * n, R: 2 * (n // 2 + 1), RI: n}[kind]
        jobs2.append((kind, lg, total + total2, nin, 0, nout))
        total2 += (nin + 63) // 64 * 64
    fixed = []
    off = total + total2
    for (kind, lg, so, nin, _, nout) in jobs2:
        fixed.append((kind, lg, so, nin, off, nout))
        off += (nout + 63) // 64 * 64
    jobs = jobs + fixed
    host = fill_inputs(jobs, off, 100)
    say("queue", len(jobs), "jobs")
    q, buf, recs, ntask = run_queue(G, jobs, off, host)
    assert q.status() == 0
    out = G.to_host(buf)
    worst = check_all(jobs, host, out)
    say("queue ok, tasks", ntask, "worst err/bound", worst)
    q.close()


def test_queue_chains_many_long_jobs(G):
    """Hundreds of long jobs (tile chains with dependency counters, longer
    than the scheduling lag) of every kind, twice in a row on one queue, with
    two and with four persistent work-groups per CU."""
    rng = np.random.default_rng(21)
    jobs, total = make_jobs(rng, 240, 13, 16)
    host = fill_inputs(jobs, total, 300)
    say("queue chains", len(jobs))
    q, buf, recs, ntask = run_queue(G, jobs, total, host)
    assert q.status() == 0
    check_all(jobs, host, G.to_host(buf))
    import torch
    for groups in (4, 2):
        say("queue chains, groups per CU", groups)
        q.set_groups_per_cu(groups)
        buf.copy_(G.to_dev(host))
        q.launch()
        torch.cuda.synchronize()
        assert q.status() == 0
        check_all(jobs, host, G.to_host(buf))
    q.close()


def test_queue_tune_and_lags(G):
    """The queue autotuner times (groups per CU, lag) candidates on a batch
    and leaves the queue prepared with the winner; every lag lists a chain
    step after the one it waits for, so outputs stay right for each."""
    import torch
    from fft_amd.afft import Queue
    rng = np.random.default_rng(33)
    jobs, total = make_jobs(rng, 200, 8, 16)
    host = fill_inputs(jobs, total, 500)
    buf = G.to_dev(host)
    base = int(buf.data_ptr())
    q = Queue(len(jobs), 16)
    recs = Queue.jobs([(k, lg, base + 4 * so, base + 4 * do) for (k, lg, so, nin, do, nout) in jobs])
    for lag in (0, 1, 32):
        say("queue lag", lag)
        q.set_lag(lag)
        q.prepare(recs)
        buf.copy_(G.to_dev(host))
        q.launch()
        torch.cuda.synchronize()
        assert q.status() == 0
        check_all(jobs, host, G.to_host(buf))
    ms = q.tune(recs, None, 2)
    groups, lag = q.config
    say("queue tuned", groups, lag, ms)
    assert groups in (2, 4) and lag in (8, 32, 128) and ms > 0
    buf.copy_(G.to_dev(host))
    q.launch()
    torch.cuda.synchronize()
    assert q.status() == 0
    check_all(jobs, host, G.to_host(buf))
    with pytest.raises(RuntimeError):
        q.set_lag(-1)
    q.close()


def test_queue_dependency_wait_is_bounded(G):
    """A zero spin limit makes a tile give up waiting at once: the launch must
    still finish (the status then reports the unmet wait, or 0 if every
    dependency happened to be complete already)."""
    import torch
    from fft_amd.afft import Queue
    rng = np.random.default_rng(3)
    jobs, total = make_jobs(rng, 4, 16, 16)
    host = fill_inputs(jobs, total, 11)
    buf = G.to_dev(host)
    base = int(buf.data_ptr())
    q = Queue(len(jobs), 16)
    q.set_spin_limit(0)
    q.prepare(Queue.jobs([(k, lg, base + 4 * so, base + 4 * do) for (k, lg, so, nin, do, nout) in jobs]))
    say("queue spin limit 0")
    q.launch()
    torch.cuda.synchronize()
    assert q.status() in (0, 1)
    q.set_spin_limit(1 << 26)
    q.launch()
    torch.cuda.synchronize()
    assert q.status() == 0
    check_all(jobs, host, G.to_host(buf))
    q.close()


def test_queue_graph_replay_matches_eager(G):
    import torch
    from fft_amd.afft import Queue
    rng = np.random.default_rng(9)
    jobs, total = make_jobs(rng, 300, 6, 16)
    host = fill_inputs(jobs, total, 7)
    say("queue graph", len(jobs))
    q, buf, recs, ntask = run_queue(G, jobs, total, host)
    eager = buf.clone()
    # clear the outputs, then capture the launch and replay it several times
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            q.launch(s)
    torch.cuda.synchronize()
    for rep in range(3):
        buf.copy_(G.to_dev(host))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(buf, eager), rep
        _ = (buf.sum() * 0).item()  # eager work between replays
    q.close()
