"""The FP64 device reference (oracle/devref.hip) pinned to the CPU oracle, its
error sums pinned to a host computation, and then every transform of the
full-size plans -- the exact plans bench.py times -- checked against it: no
part of the library (plans, kernels, twiddle tables, error kernel) takes part
in the reference or the comparison."""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

IL32, SP32 = 1, 0  # afft_interleaved, afft_split


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _host_pairs(t):
    h = t.cpu().numpy()
    return h[0::2] + 1j * h[1::2]


PIN_SHAPES = [(1,), (2,), (8,), (64,), (1024,), (4096,), (1 << 16,), (16, 32), (8, 8, 8), (4, 1, 8), (2, 2, 2, 2),
              (32, 16384 // 32)]


@pytest.mark.parametrize("shape", PIN_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_devref_matches_cpu_oracle(shape):
    """c2c both ways, r2c and c2r of the device reference equal the CPU
    oracle's FP64 results to 1e-12 relative (both are exact-arithmetic
    references; a wrong index, sign or twiddle would be off by O(1))."""
    import torch
    import devref as D
    s = O.squeeze_shape(shape)
    n = int(np.prod(s))
    b = 2
    x = O.complex_input(5, (b, n))
    xd = torch.from_numpy(np.stack([x.real, x.imag], -1).astype(np.float32).ravel()).cuda()
    for inv in (False, True):
        y = _host_pairs(D.c2c(xd, None, D.IL, shape, b, inv)).reshape(b, n)
        for i in range(b):
            ref = O.fft(x[i].astype(np.complex64).astype(np.complex128).reshape(s), inverse=inv).ravel()
            e = np.linalg.norm(y[i] - ref) / max(np.linalg.norm(ref), 1e-300)
            assert e < 1e-12, (shape, inv, i, e)
    xr = O.normal(6, b * n).reshape(b, n)
    yr = _host_pairs(D.c2c(torch.from_numpy(xr).cuda(), None, D.REAL, shape, b)).reshape(b, n)
    m = int(np.prod(O.half_shape(s)))
    for i in range(b):
        ref = O.rfft(xr[i].astype(np.float64).reshape(s)).ravel()
        e = np.linalg.norm(yr[i][:m] - ref) / np.linalg.norm(ref)
        assert e < 1e-12, (shape, "r2c", i, e)
    hs = O.half_shape(s)
    z = np.stack([O.complex_input(7 + i, (m,)) for i in range(b)])
    for il in (1, 0):
        if il:
            zr, zi = torch.from_numpy(np.stack([z.real, z.imag], -1).astype(np.float32).ravel()).cuda(), None
        else:
            zr = torch.from_numpy(z.real.astype(np.float32).ravel()).cuda()
            zi = torch.from_numpy(z.imag.astype(np.float32).ravel()).cuda()
        yc = _host_pairs(D.c2r(zr, zi, il, shape, b)).reshape(b, n)
        for i in range(b):
            ref = O.irfft(z[i].astype(np.complex64).astype(np.complex128).reshape(hs), s).ravel()
            e = np.linalg.norm(yc[i].real - ref) / np.linalg.norm(ref)
            assert e < 1e-12, (shape, "c2r", il, i, e)


def test_devref_error_sums_against_host():
    """rel_l2 of the device error sums equals a host computation of the same
    norm in every layout, per transform and over the batch, and sees a
    planted 1e-3 change of one entry."""
    import torch
    import devref as D
    n, b = 4096, 3
    x = O.complex_input(9, (b, n))
    xd = torch.from_numpy(np.stack([x.real, x.imag], -1).astype(np.float32).ravel()).cuda()
    ref = D.c2c(xd, None, D.IL, (n,), b)
    rh = _host_pairs(ref).reshape(b, n)
    y = (rh * (1 + 1e-6 * O.complex_input(10, (b, n)))).astype(np.complex64)
    y[1, 17] += 1e-3 * abs(rh[1, 17]) + 1e-3
    yi = torch.from_numpy(np.stack([y.real, y.imag], -1).ravel()).cuda()
    ys_r = torch.from_numpy(np.ascontiguousarray(y.real).ravel()).cuda()
    ys_i = torch.from_numpy(np.ascontiguousarray(y.imag).ravel()).cuda()
    host = [np.linalg.norm(y[i].astype(np.complex128) - rh[i]) / np.linalg.norm(rh[i]) for i in range(b)]
    host_all = np.linalg.norm(y.astype(np.complex128) - rh) / np.linalg.norm(rh)
    for re_, im_, lay in ((yi, None, D.IL), (ys_r, ys_i, D.SP)):
        each = D.rel_l2(ref, re_, im_, lay, (n,), b, each=True)
        assert np.allclose(each, host, rtol=1e-9), (each, host)
        assert math.isclose(D.rel_l2(ref, re_, im_, lay, (n,), b), host_all, rel_tol=1e-9)
    assert each[1] > 10 * max(each[0], each[2])
    # real outputs compare with the real parts; half rows (r2c layout)
    yr = torch.from_numpy(np.ascontiguousarray(y.real).ravel()).cuda()
    host_r = np.linalg.norm(y.real - rh.real) / np.linalg.norm(rh.real)
    assert math.isclose(D.rel_l2(ref, yr, None, D.REAL, (n,), b), host_r, rel_tol=1e-9)
    m = n // 2 + 1
    yh = torch.from_numpy(np.stack([y[:, :m].real, y[:, :m].imag], -1).ravel()).cuda()
    host_h = np.linalg.norm(y[:, :m].astype(np.complex128) - rh[:, :m]) / np.linalg.norm(rh[:, :m])
    assert math.isclose(D.rel_l2(ref, yh, None, D.IL, (n,), b, half=True), host_h, rel_tol=1e-9)


def _rand(seed, count):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.rand(count, device="cuda", generator=g).mul_(2).sub_(1)


@pytest.mark.parametrize("inverse", [False, True])
def test_headline_every_transform_against_devref(inverse):
    """64 x c2c 2^20 with the plan bench.py times: all 64 transforms."""
    import torch
    import devref as D
    from fft_amd.afft import Plan
    n, b = 1 << 20, 64
    x = _rand(101 + inverse, 2 * n * b)
    p = Plan("ci" if inverse else "c", (n,), b)
    say("headline vs devref", inverse, p.describe())
    y = torch.empty_like(x)
    p(x, None, y, None)
    torch.cuda.synchronize()
    ref = D.c2c(x, None, D.IL, (n,), b, inverse)
    e = D.rel_l2(ref, y, None, D.IL, (n,), b, each=True)
    assert max(e) <= O.bound(n), (max(e), int(np.argmax(e)))
    p.close()


@pytest.mark.parametrize("fmt", [IL32, SP32])
def test_r2c_c2r_2p22x16_every_transform_against_devref(fmt):
    """16 x r2c 2^22 (the real-input four-step bench plan) and the c2r of a
    random half spectrum, all 16 transforms, both formats."""
    import torch
    import devref as D
    from fft_amd.afft import Plan
    n, b = 1 << 22, 16
    m = n // 2 + 1
    x = _rand(211, n * b)
    p = Plan("r", (n,), b, format=fmt)
    say("r2c 2^22 x 16 vs devref", fmt, p.describe())
    if fmt == IL32:
        yr, yi = torch.empty(2 * m * b, device="cuda"), None
        lay = D.IL
    else:
        yr, yi = torch.empty(m * b, device="cuda"), torch.empty(m * b, device="cuda")
        lay = D.SP
    p(x, None, yr, yi)
    torch.cuda.synchronize()
    ref = D.c2c(x, None, D.REAL, (n,), b)
    e = D.rel_l2(ref, yr, yi, lay, (n,), b, half=True, each=True)
    assert max(e) <= O.bound(n), (max(e), int(np.argmax(e)))
    del ref
    q = Plan("ri", (n,), b, format=fmt)
    say("c2r 2^22 x 16 vs devref", fmt, q.describe())
    zr = _rand(212, (2 if fmt == IL32 else 1) * m * b)
    zi = _rand(213, m * b) if fmt == SP32 else None
    out = torch.empty(n * b, device="cuda")
    q(zr, zi, out, None)
    torch.cuda.synchronize()
    ref = D.c2r(zr, zi, fmt == IL32, (n,), b)
    e = D.rel_l2(ref, out, None, D.REAL, (n,), b, each=True)
    assert max(e) <= O.bound(n), (max(e), int(np.argmax(e)))
    p.close()
    q.close()


@pytest.mark.parametrize("shape,batch", [((256, 256, 256), 1), ((8192, 8192), 1), ((1 << 23,), 2), ((1 << 24,), 1),
                                         ((4096, 4096), 2), ((64, 128, 1024), 2),
                                         # beyond two passes: the three-pass four-step
                                         ((1 << 25,), 2), ((1 << 26,), 1), ((1 << 27,), 1)],
                         ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else str(v))
def test_large_shapes_against_devref(shape, batch):
    """c2c both ways, r2c and c2r of the large N-D and 1-D shapes."""
    import torch
    import devref as D
    from fft_amd.afft import Plan
    n = math.prod(shape)
    m = (shape[0] // 2 + 1) * (n // shape[0])
    x = _rand(301, 2 * n * batch)
    for kind in ("c", "ci"):
        p = Plan(kind, shape, batch)
        say("large vs devref", shape, kind, p.describe()[:2])
        y = torch.empty_like(x)
        p(x, None, y, None)
        torch.cuda.synchronize()
        ref = D.c2c(x, None, D.IL, shape, batch, kind == "ci")
        e = D.rel_l2(ref, y, None, D.IL, shape, batch, each=True)
        assert max(e) <= O.bound(n), (kind, max(e))
        del ref, y
        p.close()
    xr = x[: n * batch]
    p = Plan("r", shape, batch)
    say("large vs devref", shape, "r", p.describe()[:2])
    y = torch.empty(2 * m * batch, device="cuda")
    p(xr, None, y, None)
    torch.cuda.synchronize()
    ref = D.c2c(xr, None, D.REAL, shape, batch)
    e = D.rel_l2(ref, y, None, D.IL, shape, batch, half=True, each=True)
    assert max(e) <= O.bound(n), ("r", max(e))
    del ref
    p.close()
    q = Plan("ri", shape, batch)
    say("large vs devref", shape, "ri", q.describe()[:2])
    z = x[: 2 * m * batch]
    out = torch.empty(n * batch, device="cuda")
    q(z, None, out, None)
    torch.cuda.synchronize()
    ref = D.c2r(z, None, True, shape, batch)
    e = D.rel_l2(ref, out, None, D.REAL, shape, batch, each=True)
    assert max(e) <= O.bound(n), ("ri", max(e))
    q.close()
