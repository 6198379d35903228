"""afft_gpu_execute_noalloc (the entry point that allocates nothing) through the
C ABI against afft_gpu_execute on the same plan: two calls in a row, the
second with a smaller input, each bit for bit equal to the pooled execute
(the same kernels; grouping does not change results), device memory flat
across the calls, an overlapping (in-place real) call staged in the fixed
buffer, and a plan too large for the buffer rejected with a message."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(kind, dims, batch, fmt, prec=32):
    from fft_amd import _lib
    L = _lib.lib()
    mem = ctypes.create_string_buffer(L.afft_gpu_plan_memsize())
    nd, arr = _lib.dims_array(dims)
    st = L.afft_gpu_plan_create(prec, _lib.KINDS[kind], nd, arr, fmt, batch, mem)
    assert st, _lib.last_error()
    return st, mem


def _both(kind, dims, batch, fmt, args):
    """Runs the plan through the pooled execute and the no-allocation one on
    the same input, into separate outputs; returns (pooled, noalloc) outputs
    and the device memory change across the no-allocation call."""
    import torch
    from fft_amd import _lib
    L = _lib.lib()
    st, mem = _plan(kind, dims, batch, fmt)
    src_re, src_im, n_out, two_out = args
    outs = []
    grown = None
    for f in (L.afft_gpu_execute, L.afft_gpu_execute_noalloc):
        d_re = torch.empty(n_out, device="cuda")
        d_im = torch.empty(n_out, device="cuda") if two_out else None
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        r = f(st, _lib.ptr(src_re), _lib.ptr(src_im), _lib.ptr(d_re), _lib.ptr(d_im), None)
        assert r == 0, _lib.last_error()
        torch.cuda.synchronize()
        grown = free0 - torch.cuda.mem_get_info()[0]
        outs.append((d_re, d_im))
    return outs, grown, mem


def test_noalloc_equals_execute_two_calls():
    import torch
    import gpu_util as G
    g = torch.Generator(device="cuda").manual_seed(3)
    # call 1: 12 x c2c 2^20 (the pooled plan groups all 12: 96 MiB of
    # scratch; the fixed buffer takes groups of 6)
    n, b = 1 << 20, 12
    x = torch.rand(2 * n * b, device="cuda", generator=g).mul_(2).sub_(1)
    (p, q), grown, _ = _both("c", (n,), b, G.IL, (x, None, 2 * n * b, False))
    assert torch.equal(p[0], q[0])
    assert abs(grown) < (2 << 20), grown
    # call 2, smaller: 3 x r2c 2^16, split output
    n2, b2 = 1 << 16, 3
    xr = torch.randn(n2 * b2, device="cuda", generator=g)
    m = n2 // 2 + 1
    (p, q), grown, _ = _both("r", (n2,), b2, G.SP, (xr, None, m * b2, True))
    assert torch.equal(p[0], q[0]) and torch.equal(p[1], q[1])
    assert abs(grown) < (2 << 20), grown


def test_noalloc_in_place_real_is_staged():
    import torch
    import gpu_util as G
    from fft_amd import _lib
    L = _lib.lib()
    n, b = 4096, 2
    m = n // 2 + 1
    xr = torch.randn(n * b, device="cuda")
    st, mem = _plan("r", (n,), b, G.IL)
    ref = torch.empty(2 * m * b, device="cuda")
    assert L.afft_gpu_execute(st, _lib.ptr(xr), None, _lib.ptr(ref), None, None) == 0
    buf = torch.empty(2 * m * b, device="cuda")
    buf[:n * b] = xr
    assert L.afft_gpu_execute_noalloc(st, _lib.ptr(buf), None, _lib.ptr(buf), None, None) == 0, _lib.last_error()
    torch.cuda.synchronize()
    assert torch.equal(buf, ref)


def test_noalloc_rejects_plans_beyond_the_buffer():
    import torch
    import gpu_util as G
    from fft_amd import _lib
    L = _lib.lib()
    n = 1 << 24  # 128 MiB of scratch for one transform
    st, mem = _plan("c", (n,), 1, G.IL)
    x = torch.zeros(2 * n, device="cuda")
    y = torch.full((2 * n,), 7.0, device="cuda")
    assert L.afft_gpu_execute_noalloc(st, _lib.ptr(x), None, _lib.ptr(y), None, None) == -1
    assert "no-allocation buffer" in _lib.last_error()
    torch.cuda.synchronize()
    assert bool((y == 7.0).all())  # nothing ran
