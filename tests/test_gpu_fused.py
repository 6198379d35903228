"""Both passes of the 1-D four-step c2c in one persistent launch (variant
fused = 1): bit-identical to the two-pass plan (the tiles are the pass
kernels' own), right against the oracle, for batches that fill every XCD
list, leave some empty or uneven, over several ring shapes, and under HIP
graph capture with replays."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="the fused four-step launch is switched off: its only run after a fix of a "
                                      "64-bit address truncation did not finish within its time limit; no cause "
                                      "shown yet, so the case stays off the GPU")]


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _cx(t):
    y = t.cpu().numpy().astype(np.float64)
    return y[0::2] + 1j * y[1::2]


@pytest.mark.parametrize("batch", [64, 13, 3, 1])
@pytest.mark.parametrize("kind", ["c", "ci"])
def test_fused_equals_two_pass(kind, batch):
    import torch
    from fft_amd.afft import Plan
    n = 1 << 20
    g = torch.Generator(device="cuda").manual_seed(31 + batch)
    x = torch.rand(2 * n * batch, device="cuda", generator=g).mul_(2).sub_(1)
    y0, y1 = torch.empty_like(x), torch.empty_like(x)
    p0 = Plan(kind, (n,), batch)
    p1 = Plan(kind, (n,), batch)
    p1.set_variant(fused=1)
    say("fused", kind, batch, p1.describe())
    assert "[fused4]" in p1.describe()[0]
    p0(x, None, y0, None)
    p1(x, None, y1, None)
    torch.cuda.synchronize()
    assert torch.equal(y0, y1)
    for t in sorted({0, batch // 2, batch - 1}):
        sl = slice(2 * n * t, 2 * n * (t + 1))
        e = O.rel_l2(_cx(y1[sl]).astype(np.complex64), O.fft(_cx(x[sl]), inverse=(kind == "ci")))
        assert e <= O.bound(n), (t, e)
    p0.close()
    p1.close()


@pytest.mark.parametrize("slots,lag", [(2, 1), (3, 2), (4, 1)])
def test_fused_ring_shapes(slots, lag):
    import torch
    from fft_amd.afft import Plan
    n, batch = 1 << 20, 40
    g = torch.Generator(device="cuda").manual_seed(7)
    x = torch.rand(2 * n * batch, device="cuda", generator=g).mul_(2).sub_(1)
    y0, y1 = torch.empty_like(x), torch.empty_like(x)
    p0 = Plan("c", (n,), batch)
    p1 = Plan("c", (n,), batch)
    p1.set_variant(fused=1, fused_slots=slots, fused_lag=lag)
    say("fused ring", slots, lag)
    p0(x, None, y0, None)
    for _ in range(3):
        y1.zero_()
        p1(x, None, y1, None)
        torch.cuda.synchronize()
        assert torch.equal(y0, y1)
    p0.close()
    p1.close()


def test_fused_graph_replay():
    """Captured once (reset kernel + persistent kernel), replayed several times
    with eager work in between: every replay equals the eager result."""
    import torch
    from fft_amd.afft import Plan
    n, batch = 1 << 20, 16
    g = torch.Generator(device="cuda").manual_seed(9)
    x = torch.rand(2 * n * batch, device="cuda", generator=g).mul_(2).sub_(1)
    y, ref = torch.empty_like(x), torch.empty_like(x)
    p = Plan("c", (n,), batch)
    p.set_variant(fused=1)
    p(x, None, ref, None)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    gr = torch.cuda.CUDAGraph()
    p(x, None, y, None, s)  # warm the pool on this stream before capture
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        with torch.cuda.graph(gr, stream=s):
            p(x, None, y, None, s)
    torch.cuda.synchronize()
    for i in range(4):
        y.zero_()
        torch.cuda.synchronize()
        gr.replay()
        torch.cuda.synchronize()
        say("replay", i)
        assert torch.equal(y, ref)
        p(x, None, ref, None)  # eager work between replays
        torch.cuda.synchronize()
    del gr
    p.close()
