"""The exact plans the benchmark times, checked against the CPU FP64 oracle at
their full sizes, and the FP64 device plans used as full-size references
pinned to the oracle as well.

- 64 x c2c 2^20 (forward, then inverse of its output): transforms 0, 7, 8,
  31, 32 and 63, which sit on both sides of the pass-A launch splits (8
  transforms) and the pass groups (32 transforms) of the default plan.
- 16 x r2c 2^22 (the real-input four-step in two groups of 8): transforms 0,
  7, 8 and 15, on both sides of the group boundary.
- c2c 256^3 and r2c 8192^2, FP32 and FP64, against the full oracle.
Bars: 2*log2(N)*2^-24 (FP32) and 16*2*log2(N)*2^-53 (FP64) relative L2."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def bound64(n):
    return 16 * 2.0 * np.log2(max(2, n)) * 2.0 ** -53


def _cx(t):
    """Interleaved device tensor slice -> complex128 host array."""
    y = t.cpu().numpy().astype(np.float64)
    return y[0::2] + 1j * y[1::2]


def test_headline_plans_against_oracle():
    import torch
    import gpu_util as G
    from fft_amd.afft import Plan
    n, batch = 1 << 20, 64
    g = torch.Generator(device="cuda").manual_seed(20)
    x = torch.rand(2 * n * batch, device="cuda", generator=g, dtype=torch.float32).mul_(2).sub_(1)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    fwd = Plan("c", (n,), batch)
    inv = Plan("ci", (n,), batch)
    say("headline plans", fwd.describe(), "chunk", fwd.get_variant().chunk, "pass-A launches", fwd.launches(0))
    assert fwd.get_variant().chunk == 32 and fwd.launches(0) == 8  # the splits the samples straddle
    fwd(x, None, y, None)
    inv(y, None, z, None)
    torch.cuda.synchronize()
    for t in (0, 7, 8, 31, 32, 63):
        sl = slice(2 * n * t, 2 * n * (t + 1))
        xt, yt, zt = _cx(x[sl]), _cx(y[sl]), _cx(z[sl])
        e_f = O.rel_l2(yt.astype(np.complex64), O.fft(xt))
        e_i = O.rel_l2(zt.astype(np.complex64), O.fft(yt, inverse=True))
        say("  transform", t, "fwd", e_f, "inv", e_i)
        assert e_f <= O.bound(n), (t, e_f)
        assert e_i <= O.bound(n), (t, e_i)
    fwd.close()
    inv.close()


def test_r2c_bench_plan_against_oracle():
    import torch
    from fft_amd.afft import Plan
    n, batch = 1 << 22, 16
    m = n // 2 + 1
    g = torch.Generator(device="cuda").manual_seed(21)
    x = torch.randn(n * batch, device="cuda", generator=g, dtype=torch.float32)
    y = torch.empty(2 * m * batch, device="cuda", dtype=torch.float32)
    p = Plan("r", (n,), batch)
    say("r2c bench plan", p.describe(), "chunk", p.get_variant().chunk, "pass-A launches", p.launches(0))
    p(x, None, y, None)
    torch.cuda.synchronize()
    assert p.describe()[0].startswith("real4a") and p.get_variant().chunk == 8
    for t in (0, 7, 8, 15):
        xt = x[n * t:n * (t + 1)].cpu().numpy().astype(np.float64)
        yt = _cx(y[2 * m * t:2 * m * (t + 1)])
        e = O.rel_l2(yt.astype(np.complex64), O.rfft(xt))
        say("  transform", t, e)
        assert e <= O.bound(n), (t, e)
    p.close()


def test_fp64_reference_plans_pinned():
    """The FP64 device plans the full-size checks compare with, against the
    oracle at 16 x 2^20 (transforms 0 and 15), 256^3 and 8192^2 r2c."""
    import torch
    from fft_amd.afft import Plan
    n, batch = 1 << 20, 16
    g = torch.Generator(device="cuda").manual_seed(22)
    x = torch.rand(2 * n * batch, device="cuda", generator=g, dtype=torch.float64).mul_(2).sub_(1)
    y = torch.empty_like(x)
    p = Plan("c", (n,), batch, dtype=np.float64)
    p(x, None, y, None)
    torch.cuda.synchronize()
    for t in (0, 15):
        sl = slice(2 * n * t, 2 * n * (t + 1))
        e = O.rel_l2(_cx(y[sl]), O.fft(_cx(x[sl])))
        say("fp64 16 x 2^20 transform", t, e)
        assert e <= bound64(n), e
    p.close()
    del x, y
    shape = (256, 256, 256)
    xc = O.complex_input(1, shape).astype(np.complex128)
    src = torch.from_numpy(xc.view(np.float64).ravel()).cuda()
    dst = torch.empty_like(src)
    p = Plan("c", shape, 1, dtype=np.float64)
    p(src, None, dst, None)
    torch.cuda.synchronize()
    e = O.rel_l2(_cx(dst).reshape(shape), O.fft(xc))
    say("fp64 256^3", e)
    assert e <= bound64(xc.size), e
    p.close()
    del src, dst, xc
    shape = (8192, 8192)
    xr = O.normal(5, 8192 * 8192).reshape(shape).astype(np.float64)
    src = torch.from_numpy(xr.ravel()).cuda()
    dst = torch.empty(2 * 4097 * 8192, device="cuda", dtype=torch.float64)
    p = Plan("r", shape, 1, dtype=np.float64)
    p(src, None, dst, None)
    torch.cuda.synchronize()
    e = O.rel_l2(_cx(dst).reshape(4097, 8192), O.rfft(xr))
    say("fp64 r2c 8192^2", e)
    assert e <= bound64(xr.size), e
    p.close()


@pytest.mark.parametrize("kind,shape", [("c", (256, 256, 256)), ("r", (8192, 8192)), ("ri", (8192, 8192))],
                         ids=["c2c_256^3", "r2c_8192^2", "c2r_8192^2"])
def test_nd_bench_plans_against_oracle(kind, shape):
    import torch
    import gpu_util as G
    from fft_amd.afft import Plan
    n = int(np.prod(shape))
    p = Plan(kind, shape, 1)
    say("nd bench plan", kind, shape, p.describe())
    if kind == "c":
        x = O.complex_input(1, shape)
        src = G.to_dev(G._il(x, np.float32))
        dst = G.empty_dev(2 * n)
        p(src, None, dst, None)
        torch.cuda.synchronize()
        e = O.rel_l2(_cx(dst).astype(np.complex64).reshape(shape), O.fft(x.astype(np.complex128)))
    elif kind == "r":
        x = O.normal(5, n).reshape(shape)
        src = G.to_dev(x.ravel())
        dst = G.empty_dev(2 * 4097 * 8192)
        p(src, None, dst, None)
        torch.cuda.synchronize()
        e = O.rel_l2(_cx(dst).astype(np.complex64).reshape(4097, 8192), O.rfft(x.astype(np.float64)))
    else:
        h = O.hermitian_input(3, shape)
        src = G.to_dev(G._il(h, np.float32))
        dst = G.empty_dev(n)
        p(src, None, dst, None)
        torch.cuda.synchronize()
        e = O.rel_l2(G.to_host(dst).reshape(shape), O.irfft(h.astype(np.complex128), shape))
    say("  rel-L2", e, "bound", O.bound(n))
    assert e <= O.bound(n), e
    p.close()
