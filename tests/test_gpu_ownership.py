"""afft's ownership contract on the GPU build: the plan lives in the caller's
`mem` and the caller just frees it (fft.hpp:18-45, fft_internal.hpp:24-106).
Plans own no device memory (shared twiddle cache and scratch pool), so
creating and dropping hundreds of large plans without any destroy call keeps
device memory flat, and plans that share the pool on different streams still
compute the right results."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _free_bytes():
    import torch
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


def test_500_plans_dropped_without_destroy_keep_memory_flat():
    import gpu_util as G
    from fft_amd import _lib
    L = _lib.lib()
    shape = (8192, 8192)
    nd, dims = _lib.dims_array(shape)
    size = L.afft_32_r_memsize(nd, dims, G.IL, 0)
    assert 0 < size < (1 << 20)
    src = G.empty_dev(8192 * 8192).normal_()
    dst = G.empty_dev(2 * 4097 * 8192)
    # first plan: creates the twiddle tables and grows the scratch pool
    mem = ctypes.create_string_buffer(size)
    st = L.afft_32_r_create(nd, dims, G.IL, mem, 0)
    assert st, _lib.last_error()
    L.afft_32_r_transform(st, _lib.ptr(src), _lib.ptr(dst), None)
    assert _lib.last_error() == ""
    ref = G.to_host(dst).copy()
    del mem
    base = _free_bytes()
    for i in range(500):
        mem = ctypes.create_string_buffer(size)
        st = L.afft_32_r_create(nd, dims, G.IL, mem, 0)
        assert st, _lib.last_error()
        if i % 100 == 99:
            L.afft_32_r_transform(st, _lib.ptr(src), _lib.ptr(dst), None)
            assert _lib.last_error() == ""
            assert np.array_equal(G.to_host(dst), ref)
        del mem  # no destroy: afft callers just free the plan memory
    grown = base - _free_bytes()
    print(f"device memory change over 500 plans: {grown / 2**20:.1f} MiB", flush=True)
    assert grown < (16 << 20)


def test_plan_recreated_in_same_memory():
    """afft callers may reuse `mem` for a new plan: the old one is simply gone."""
    import gpu_util as G
    from fft_amd import _lib
    L = _lib.lib()
    mem = ctypes.create_string_buffer(L.afft_32_c_memsize(*_lib.dims_array([1 << 16]), G.IL, 0))
    for n in (1 << 16, 1 << 10, 1 << 16):
        nd, dims = _lib.dims_array([n])
        st = L.afft_32_c_create(nd, dims, G.IL, mem, 0)
        assert st, _lib.last_error()
        x = O.complex_input(3, (n,))
        s = G.to_dev(G._il(x, np.float32))
        d = G.empty_dev(2 * n)
        L.afft_32_c_transform(st, _lib.ptr(s), None, _lib.ptr(d), None)
        y = G.to_host(d)
        assert O.rel_l2((y[0::2] + 1j * y[1::2]).astype(np.complex64), O.fft(x.astype(np.complex128))) <= O.bound(n)
    L.afft_32_c_destroy(st)
    L.afft_32_c_transform(st, _lib.ptr(s), None, _lib.ptr(d), None)
    assert "not a plan" in _lib.last_error()


def test_pool_shared_across_streams():
    """Two four-step plans (both use scratch from the one pool) launched on two
    streams with no host synchronisation between them: the pool orders them,
    and each output equals its single-stream result bit for bit."""
    import torch
    import gpu_util as G
    from fft_amd.afft import Plan
    n, b = 1 << 20, 8
    x = G.empty_dev(2 * n * b).uniform_(-1, 1)
    w = G.empty_dev(2 * n * b).uniform_(-1, 1)
    pf = Plan("c", (n,), b)
    pi = Plan("ci", (n,), b)
    assert pf.scratch_bytes > 0 and pi.scratch_bytes > 0
    yf, yi = torch.empty_like(x), torch.empty_like(w)
    pf(x, None, yf, None)
    pi(w, None, yi, None)
    torch.cuda.synchronize()
    want_f, want_i = yf.clone(), yi.clone()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(5):
        yf.zero_()
        yi.zero_()
        torch.cuda.synchronize()
        pf(x, None, yf, None, s1)
        pi(w, None, yi, None, s2)
        pf(x, None, yf, None, s1)
        torch.cuda.synchronize()
        assert torch.equal(yf, want_f)
        assert torch.equal(yi, want_i)
    # and the forward result is right (a sampled transform against the oracle)
    xs = G.to_host(x[2 * n * 5:2 * n * 6])
    ref = O.fft((xs[0::2] + 1j * xs[1::2]).astype(np.complex128))
    ys = G.to_host(want_f[2 * n * 5:2 * n * 6])
    assert O.rel_l2((ys[0::2] + 1j * ys[1::2]).astype(np.complex64), ref) <= O.bound(n)
    pf.close()
    pi.close()
