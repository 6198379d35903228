"""The 1-D FP32 r2c and c2r as a real-input four-step (afft_real4.hip) on
the GPU: every compiled split against the oracle in both formats and in
batches, bit-identical results whatever the grouping of the batch, the c2r
of the r2c giving N x back, and the other two forms of the long r2c / c2r
(complex four-step with the real pass on its own, real_pair = -1; with the
real pass fused into a four-step pass, real_pair = 1) against the same
oracle."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _run(p, x, m, b, fmt):
    import torch
    import gpu_util as G
    if fmt == G.IL:
        y = G.empty_dev(2 * m * b)
        p(x, None, y, None)
        torch.cuda.synchronize()
        h = G.to_host(y)
        return (h[0::2] + 1j * h[1::2]).reshape(b, m)
    yr, yi = G.empty_dev(m * b), G.empty_dev(m * b)
    p(x, None, yr, yi)
    torch.cuda.synchronize()
    return (G.to_host(yr) + 1j * G.to_host(yi)).reshape(b, m)


@pytest.mark.parametrize("lg", list(range(16, 24)))
def test_real4_against_oracle(lg):
    import gpu_util as G
    from fft_amd.afft import Plan
    n = 1 << lg
    m = n // 2 + 1
    b = 3 if lg <= 21 else 2
    x = O.normal(40 + lg, n * b).reshape(b, n)
    xd = G.to_dev(x.ravel())
    ref = [O.rfft(x[i].astype(np.float64)) for i in range(b)]
    for fmt in (G.IL, G.SP):
        p = Plan("r", (n,), b, format=fmt)
        p.set_variant(real_pair=2)
        say("real4", lg, fmt, p.describe())
        assert p.describe()[0].startswith("real4a")
        y = _run(p, xd, m, b, fmt)
        for i in range(b):
            e = O.rel_l2(y[i].astype(np.complex64), ref[i])
            assert e <= O.bound(n), (lg, fmt, i, e)
        p.close()


def test_real4_grouping_invariance_and_other_forms():
    import torch
    import gpu_util as G
    from fft_amd.afft import Plan
    n, b = 1 << 22, 6
    m = n // 2 + 1
    x = G.to_dev(O.normal(77, n * b))
    p = Plan("r", (n,), b)
    p.set_variant(real_pair=2)
    y0 = _run(p, x, m, b, G.IL)
    for ch in (1, 4):
        p.set_chunk(ch)
        assert np.array_equal(_run(p, x, m, b, G.IL), y0), ch
    xh = G.to_host(x).reshape(b, n)
    for rp in (-1, 1):
        q = Plan("r", (n,), b)
        q.set_variant(real_pair=rp)
        say("other form", rp, q.describe())
        assert not q.describe()[0].startswith("real4")
        y = _run(q, x, m, b, G.IL)
        for i in (0, b - 1):
            assert O.rel_l2(y[i].astype(np.complex64), O.rfft(xh[i].astype(np.float64))) <= O.bound(n)
        q.close()
    for i in (0, 3, b - 1):
        assert O.rel_l2(y0[i].astype(np.complex64), O.rfft(xh[i].astype(np.float64))) <= O.bound(n)
    p.close()


def _run_c2r(p, z, n, b, fmt):
    """z: (b, n/2 + 1) complex on the host; returns the (b, n) reals."""
    import torch
    import gpu_util as G
    if fmt == G.IL:
        zi = G.to_dev(np.stack([z.real, z.imag], -1).astype(np.float32).ravel())
        zs = None
    else:
        zi = G.to_dev(z.real.astype(np.float32).ravel())
        zs = G.to_dev(z.imag.astype(np.float32).ravel())
    y = G.empty_dev(n * b)
    p(zi, zs, y, None)
    torch.cuda.synchronize()
    return G.to_host(y).reshape(b, n)


@pytest.mark.parametrize("lg", list(range(16, 24)))
def test_real4_c2r_against_oracle(lg):
    import gpu_util as G
    from fft_amd.afft import Plan
    n = 1 << lg
    m = n // 2 + 1
    b = 3 if lg <= 21 else 2
    # arbitrary bins, imaginary parts of bins 0 and N/2 included (the c2r
    # takes only their real parts)
    z = O.complex_input(50 + lg, (b, m))
    for fmt in (G.IL, G.SP):
        p = Plan("ri", (n,), b, format=fmt)
        p.set_variant(real_pair=2)
        say("real4 c2r", lg, fmt, p.describe())
        assert p.describe()[0].startswith("real4c")
        y = _run_c2r(p, z, n, b, fmt)
        for i in range(b):
            ref = O.irfft(z[i].astype(np.complex128), (n,))
            e = O.rel_l2(y[i].astype(np.float32), ref)
            assert e <= O.bound(n), (lg, fmt, i, e)
        p.close()


def test_real4_c2r_round_trip_grouping_and_other_forms():
    import torch
    import gpu_util as G
    from fft_amd.afft import Plan
    n, b = 1 << 22, 6
    m = n // 2 + 1
    x = O.normal(78, n * b)
    xd = G.to_dev(x)
    f = Plan("r", (n,), b)
    spec = G.empty_dev(2 * m * b)
    f(xd, None, spec, None)
    p = Plan("ri", (n,), b)
    p.set_variant(real_pair=2)
    say("c2r round trip", p.describe())
    y0 = G.empty_dev(n * b)
    p(spec, None, y0, None)
    torch.cuda.synchronize()
    back = G.to_host(y0).reshape(b, n) / n
    xr = x.reshape(b, n)
    for i in range(b):
        e = np.linalg.norm(back[i] - xr[i]) / np.linalg.norm(xr[i])
        assert e <= 2 * O.bound(n), (i, e)
    for ch in (1, 4):
        p.set_chunk(ch)
        y = G.empty_dev(n * b)
        p(spec, None, y, None)
        torch.cuda.synchronize()
        assert torch.equal(y, y0), ch
    sh = G.to_host(spec).reshape(b, m, 2)
    zh = sh[..., 0] + 1j * sh[..., 1]
    for rp in (-1, 1):
        q = Plan("ri", (n,), b)
        q.set_variant(real_pair=rp)
        say("c2r other form", rp, q.describe())
        assert not q.describe()[0].startswith("real4")
        y = G.empty_dev(n * b)
        q(spec, None, y, None)
        torch.cuda.synchronize()
        yh = G.to_host(y).reshape(b, n)
        for i in (0, b - 1):
            assert O.rel_l2(yh[i].astype(np.float32), O.irfft(zh[i].astype(np.complex128), (n,))) <= O.bound(n)
        q.close()
    f.close()
    p.close()


@pytest.mark.parametrize("var", [{"rowt_tile": 16}, {"sched": 1}, {"rowt_tile": 16, "sched": 1}, {"col_tile": 16},
                                 {"col_tile": 16, "sched": 1}, {"col_tile": 8}, {"half_lds": -1},
                                 {"half_lds": -1, "rowt_tile": 16}],
                         ids=lambda v: "-".join(f"{k}{x}" for k, x in v.items()))
def test_real4_compiled_shapes_against_oracle(var):
    """Every compiled tile / schedule of the four passes (chosen through the
    variant fields) against the oracle, r2c and c2r, at the splits that use
    them (1024 x 1024, 1024 x 2048, 2048 x 2048)."""
    import gpu_util as G
    from fft_amd.afft import Plan
    for lg in (20, 21, 22):
        n, b = 1 << lg, 2
        m = n // 2 + 1
        x = O.normal(90 + lg, n * b).reshape(b, n)
        p = Plan("r", (n,), b)
        p.set_variant(real_pair=2, **var)
        say("real4 shape", var, lg, p.describe())
        y = _run(p, G.to_dev(x.ravel()), m, b, G.IL)
        for i in range(b):
            assert O.rel_l2(y[i].astype(np.complex64), O.rfft(x[i].astype(np.float64))) <= O.bound(n), (var, lg, i)
        p.close()
        z = O.complex_input(91 + lg, (b, m))
        q = Plan("ri", (n,), b)
        q.set_variant(real_pair=2, **var)
        say("real4 shape c2r", var, lg, q.describe())
        yr = _run_c2r(q, z, n, b, G.IL)
        for i in range(b):
            ref = O.irfft(z[i].astype(np.complex128), (n,))
            assert O.rel_l2(yr[i].astype(np.float32), ref) <= O.bound(n), (var, lg, i)
        q.close()
