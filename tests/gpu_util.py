"""Helpers for the GPU tests: device buffers (torch, used only as allocator
and copier) and one-call wrappers that run a transform through the C ABI with
afft's buffer layouts."""
from __future__ import annotations

import numpy as np

from fft_amd import _lib
from fft_amd import afft

IL, SP = _lib.afft_interleaved, _lib.afft_split


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible")
    return torch


def to_dev(a: np.ndarray):
    torch = torch_cuda()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def empty_dev(n: int, dtype=np.float32):
    torch = torch_cuda()
    tdt = torch.float32 if np.dtype(dtype) == np.float32 else torch.float64
    return torch.empty(max(1, int(n)), dtype=tdt, device="cuda")


def to_host(t) -> np.ndarray:
    return t.cpu().numpy()


def _split(x, dtype):
    x = np.asarray(x).ravel()
    return np.ascontiguousarray(x.real, dtype=dtype), np.ascontiguousarray(x.imag, dtype=dtype)


def _il(x, dtype):
    x = np.asarray(x).ravel()
    return np.ascontiguousarray(np.stack([x.real, x.imag], -1).ravel(), dtype=dtype)


def gpu_c2c(x: np.ndarray, inverse=False, fmt=IL, dtype=np.float32):
    """One afft_XX_c/ci plan over x's shape; returns complex output (host)."""
    shape = x.shape
    n = int(np.prod(shape))
    cls = afft.inverse_complex_transform if inverse else afft.complex_transform
    t = cls(shape, dtype=dtype, format=fmt)
    assert t, _lib.last_error()
    if fmt == IL:
        src = to_dev(_il(x, dtype))
        dst = empty_dev(2 * n, dtype)
        t(src, None, dst, None)
        y = to_host(dst)
        out = y[0::2] + 1j * y[1::2]
    else:
        sr, si = (to_dev(a) for a in _split(x, dtype))
        dr, di = empty_dev(n, dtype), empty_dev(n, dtype)
        t(sr, si, dr, di)
        out = to_host(dr) + 1j * to_host(di)
    err = _lib.last_error()
    t.close()
    assert err == "", err
    return out.reshape(shape)


def half(shape):
    s = [int(d) for d in shape if int(d) != 1] or [1]
    return [s[0] // 2 + 1] + s[1:]


def gpu_r2c(x: np.ndarray, fmt=IL, dtype=np.float32):
    shape = x.shape
    hs = half(shape)
    m = int(np.prod(hs))
    t = afft.real_transform(shape, dtype=dtype, format=fmt)
    assert t, _lib.last_error()
    src = to_dev(np.ascontiguousarray(x, dtype=dtype).ravel())
    if fmt == IL:
        dst = empty_dev(2 * m, dtype)
        t(src, dst, None)
        y = to_host(dst)
        out = y[0::2] + 1j * y[1::2]
    else:
        off = afft.align_size(m, dtype)
        buf = empty_dev(2 * off, dtype)
        base = int(buf.data_ptr())
        t(src, base, base + off * np.dtype(dtype).itemsize)
        y = to_host(buf)
        out = y[:m] + 1j * y[off:off + m]
    err = _lib.last_error()
    t.close()
    assert err == "", err
    return out.reshape(hs)


def gpu_c2r(z: np.ndarray, shape, fmt=IL, dtype=np.float32):
    hs = half(shape)
    n = int(np.prod(shape))
    m = int(np.prod(hs))
    t = afft.inverse_real_transform(shape, dtype=dtype, format=fmt)
    assert t, _lib.last_error()
    dst = empty_dev(n, dtype)
    if fmt == IL:
        src = to_dev(_il(z, dtype))
        t(src, None, dst)
    else:
        off = afft.align_size(m, dtype)
        hbuf = np.zeros(2 * off, dtype=dtype)
        zr = np.asarray(z).ravel()
        hbuf[:m] = zr.real
        hbuf[off:off + m] = zr.imag
        buf = to_dev(hbuf)
        base = int(buf.data_ptr())
        t(base, base + off * np.dtype(dtype).itemsize, dst)
    out = to_host(dst)
    err = _lib.last_error()
    t.close()
    assert err == "", err
    return out.reshape([int(s) for s in shape])


def device_rel_l2(y_dev, ref_dev) -> float:
    """rel-L2 with the library's FP64 device reduction."""
    torch = torch_cuda()
    acc = torch.zeros(2, dtype=torch.float64, device="cuda")
    fn = _lib.lib().afft_gpu_error_sums_f32 if y_dev.dtype == torch.float32 else _lib.lib().afft_gpu_error_sums_f64
    r = fn(int(y_dev.data_ptr()), int(ref_dev.data_ptr()), y_dev.numel(), int(acc.data_ptr()), None)
    assert r == 0
    torch.cuda.synchronize()
    a = acc.cpu().numpy()
    return float(np.sqrt(a[0] / a[1])) if a[1] > 0 else float("inf")


def run_abi(kind: str, shape, x: np.ndarray, fmt=IL, dtype=np.float32, off: int = 0):
    """One afft_XX_<kind>_create / _transform through the C ABI with every
    buffer starting `off` elements past its allocation (afft's harness shifts
    its buffers the same way, testing.cpp:67-90).  x is the input in afft's
    layout for the kind (complex for c / ci / ri, real for r); returns the
    output (complex, or real for ri) on the host."""
    import ctypes
    torch = torch_cuda()
    L = _lib.lib()
    p = "32" if np.dtype(dtype) == np.float32 else "64"
    nd, dims = _lib.dims_array(shape)
    mem = ctypes.create_string_buffer(getattr(L, f"afft_{p}_{kind}_memsize")(nd, dims, fmt, 0))
    st = getattr(L, f"afft_{p}_{kind}_create")(nd, dims, fmt, mem, 0)
    assert st, _lib.last_error()
    n = int(np.prod(shape))
    m = int(np.prod(half(shape)))
    es = np.dtype(dtype).itemsize
    cnt_in = {"c": n, "ci": n, "r": n, "ri": m}[kind]
    cnt_out = {"c": n, "ci": n, "r": m, "ri": n}[kind]
    cplx_in, cplx_out = kind != "r", kind != "ri"

    def dev(cnt):
        return empty_dev(cnt + off + 16, dtype)

    def at(t):
        return t.data_ptr() + off * es
    xs = np.asarray(x).ravel()
    if cplx_in and fmt == IL:
        a = dev(2 * cnt_in)
        a[off:off + 2 * cnt_in] = to_dev(_il(xs, dtype))
        ins = (at(a), None)
    elif cplx_in:
        a, b = dev(cnt_in), dev(cnt_in)
        a[off:off + cnt_in] = to_dev(np.ascontiguousarray(xs.real, dtype=dtype))
        b[off:off + cnt_in] = to_dev(np.ascontiguousarray(xs.imag, dtype=dtype))
        ins = (at(a), at(b))
    else:
        a = dev(cnt_in)
        a[off:off + cnt_in] = to_dev(np.ascontiguousarray(xs, dtype=dtype))
        ins = (at(a),)
    if cplx_out and fmt == IL:
        o = dev(2 * cnt_out)
        outs, keep = (at(o), None), (o,)
    elif cplx_out:
        o, q = dev(cnt_out), dev(cnt_out)
        outs, keep = (at(o), at(q)), (o, q)
    else:
        o = dev(cnt_out)
        outs, keep = (at(o),), (o,)
    fn = getattr(L, f"afft_{p}_{kind}_transform")
    args = {"c": ins + outs, "ci": ins + outs, "r": ins + outs, "ri": (ins + (None,))[:2] + outs}[kind]
    fn(st, *args)
    err = _lib.last_error()
    assert err == "", err
    torch.cuda.synchronize()
    h = [to_host(t)[off:] for t in keep]
    if cplx_out and fmt == IL:
        y = h[0][0:2 * cnt_out:2] + 1j * h[0][1:2 * cnt_out:2]
    elif cplx_out:
        y = h[0][:cnt_out] + 1j * h[1][:cnt_out]
    else:
        y = h[0][:cnt_out]
    return y.reshape(half(shape) if kind == "r" else [int(s) for s in shape])
