"""Buffer placement through the C ABI, against the oracle: in-place transforms
and pointers at any 4-byte (FP32) / 8-byte (FP64) offset, as afft accepts
them (testing.cpp:67-90 runs its harness on buffers shifted by an arbitrary
offset; in-place use is not excluded by fft.h).

- c2c / ci exactly in place (src == dst): 1-D row pass, 1-D four-step, N-D
  with column passes and a column four-step, both formats;
- r2c / c2r in place (src shares its base with the output) and c2c with
  partially overlapping buffers: the input is staged through scratch first;
- misaligned device pointers (4, 12 and 60 bytes past an allocation; 8, 24
  and 56 for FP64) for every kind."""
import ctypes
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


def _plan(kind, shape, fmt, dtype=np.float32):
    from fft_amd import _lib
    L = _lib.lib()
    p = "32" if dtype == np.float32 else "64"
    nd, dims = _lib.dims_array(shape)
    size = getattr(L, f"afft_{p}_{kind}_memsize")(nd, dims, fmt, 0)
    mem = ctypes.create_string_buffer(size)
    st = getattr(L, f"afft_{p}_{kind}_create")(nd, dims, fmt, mem, 0)
    assert st, _lib.last_error()
    fn = getattr(L, f"afft_{p}_{kind}_transform")

    def run(*ptrs):
        fn(st, *ptrs)
        err = _lib.last_error()
        assert err == "", err
    run.mem = mem
    return run


def _check(kind, shape, y, x, dtype):
    n = int(np.prod(shape))
    b = O.bound(n) if dtype == np.float32 else bound64(n)
    if kind in ("c", "ci"):
        ref = O.fft(x.astype(np.complex128), inverse=(kind == "ci"))
    elif kind == "r":
        ref = O.rfft(x.astype(np.float64))
    else:
        ref = O.irfft(x.astype(np.complex128), shape)
    yy = y.astype(np.complex64 if dtype == np.float32 else np.complex128) if np.iscomplexobj(y) else y
    e = O.rel_l2(yy, ref)
    assert e <= b, (kind, shape, e, b)
    return e


C2C_INPLACE = [(4096,), (1 << 16,), (1 << 20,), (64, 64), (8192, 16), (16, 8, 32)]


@pytest.mark.parametrize("shape", C2C_INPLACE, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("kind", ["c", "ci"])
def test_c2c_in_place(kind, shape):
    import gpu_util as G
    import torch
    n = int(np.prod(shape))
    x = O.complex_input(2, shape)
    for fmt in (G.IL, G.SP):
        say("c2c in place", kind, shape, fmt)
        run = _plan(kind, shape, fmt)
        if fmt == G.IL:
            buf = G.to_dev(G._il(x, np.float32))
            run(buf.data_ptr(), None, buf.data_ptr(), None)
            torch.cuda.synchronize()
            h = G.to_host(buf)
            y = h[0::2] + 1j * h[1::2]
        else:
            re, im = G.to_dev(x.real.ravel().astype(np.float32)), G.to_dev(x.imag.ravel().astype(np.float32))
            run(re.data_ptr(), im.data_ptr(), re.data_ptr(), im.data_ptr())
            torch.cuda.synchronize()
            y = G.to_host(re) + 1j * G.to_host(im)
        _check(kind, shape, y.reshape(shape), x, np.float32)


REAL_INPLACE = [(4096,), (1 << 16,), (1 << 22,), (64, 64), (8192, 64)]


@pytest.mark.parametrize("shape", REAL_INPLACE, ids=lambda s: "x".join(map(str, s)))
def test_real_in_place(shape):
    """r2c writing over its own input and c2r over its own spectrum (the
    output starts at the input's address)."""
    import gpu_util as G
    import torch
    n = int(np.prod(shape))
    hs = G.half(shape)
    m = int(np.prod(hs))
    x = O.normal(4, n).reshape(shape)
    h = O.hermitian_input(5, shape)
    for fmt in (G.IL, G.SP):
        say("real in place", shape, fmt)
        # r2c: dst_re at src
        buf = G.empty_dev(max(n, 2 * m) + 64)
        buf[:n] = G.to_dev(x.ravel())
        im = G.empty_dev(m)
        run = _plan("r", shape, fmt)
        run(buf.data_ptr(), buf.data_ptr(), None if fmt == G.IL else im.data_ptr())
        torch.cuda.synchronize()
        hb = G.to_host(buf)
        y = hb[0:2 * m:2] + 1j * hb[1:2 * m:2] if fmt == G.IL else hb[:m] + 1j * G.to_host(im)
        _check("r", shape, y.reshape(hs), x, np.float32)
        # c2r: dst at src_re
        run = _plan("ri", shape, fmt)
        buf = G.empty_dev(max(n, 2 * m) + 64)
        if fmt == G.IL:
            buf[:2 * m] = G.to_dev(G._il(h, np.float32))
            run(buf.data_ptr(), None, buf.data_ptr())
        else:
            buf[:m] = G.to_dev(h.real.ravel().astype(np.float32))
            im = G.to_dev(h.imag.ravel().astype(np.float32))
            run(buf.data_ptr(), im.data_ptr(), buf.data_ptr())
        torch.cuda.synchronize()
        _check("ri", shape, G.to_host(buf)[:n].reshape(shape), h, np.float32)


def test_c2c_partial_overlap():
    """Output shifted by a few elements against the input (not in place):
    staged, so the result is the transform of the original input."""
    import gpu_util as G
    import torch
    for shape in ((4096,), (1 << 18,), (32, 64)):
        n = int(np.prod(shape))
        x = O.complex_input(6, shape)
        buf = G.empty_dev(2 * n + 64)
        buf[:2 * n] = G.to_dev(G._il(x, np.float32))
        say("partial overlap", shape)
        run = _plan("c", shape, G.IL)
        run(buf.data_ptr(), None, buf.data_ptr() + 4 * 34, None)
        torch.cuda.synchronize()
        hb = G.to_host(buf)[34:34 + 2 * n]
        _check("c", shape, (hb[0::2] + 1j * hb[1::2]).reshape(shape), x, np.float32)


MIS_SHAPES = [(1024,), (1 << 16,), (1 << 21,), (64, 128), (8192, 16), (8, 4, 16)]


@pytest.mark.parametrize("shape", MIS_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_misaligned_pointers(shape, dtype):
    import gpu_util as G
    import torch
    n = int(np.prod(shape))
    hs = G.half(shape)
    m = int(np.prod(hs))
    es = np.dtype(dtype).itemsize
    offs = (1, 3, 15) if dtype == np.float32 else (1, 3, 7)  # elements: 4/12/60 bytes, 8/24/56 bytes
    x = O.complex_input(7, shape)
    xr = O.normal(8, n).reshape(shape)
    h = O.hermitian_input(9, shape)
    for k, o in enumerate(offs):
        fmt = G.IL if k != 1 else G.SP
        say("misaligned", shape, dtype.__name__, o * es, "bytes, fmt", fmt)
        big = 2 * max(n, m) + 64
        s_re, s_im, d_re, d_im = (G.empty_dev(big, dtype) for _ in range(4))
        a = lambda t: t.data_ptr() + o * es

        def put(t, v):
            t[o:o + v.size] = G.to_dev(np.ascontiguousarray(v, dtype=dtype))

        def get(t, cnt):
            torch.cuda.synchronize()
            return G.to_host(t)[o:o + cnt].astype(np.float64)
        for kind in ("c", "ci"):
            run = _plan(kind, shape, fmt, dtype)
            if fmt == G.IL:
                put(s_re, G._il(x, dtype))
                run(a(s_re), None, a(d_re), None)
                y = get(d_re, 2 * n)
                y = y[0::2] + 1j * y[1::2]
            else:
                put(s_re, x.real.ravel())
                put(s_im, x.imag.ravel())
                run(a(s_re), a(s_im), a(d_re), a(d_im))
                y = get(d_re, n) + 1j * get(d_im, n)
            _check(kind, shape, y.reshape(shape), x, dtype)
        run = _plan("r", shape, fmt, dtype)
        put(s_re, xr.ravel())
        if fmt == G.IL:
            run(a(s_re), a(d_re), None)
            y = get(d_re, 2 * m)
            y = y[0::2] + 1j * y[1::2]
        else:
            run(a(s_re), a(d_re), a(d_im))
            y = get(d_re, m) + 1j * get(d_im, m)
        _check("r", shape, y.reshape(hs), xr, dtype)
        run = _plan("ri", shape, fmt, dtype)
        if fmt == G.IL:
            put(s_re, G._il(h, dtype))
            run(a(s_re), None, a(d_re))
        else:
            put(s_re, h.real.ravel())
            put(s_im, h.imag.ravel())
            run(a(s_re), a(s_im), a(d_re))
        _check("ri", shape, get(d_re, n).reshape(shape), h, dtype)
