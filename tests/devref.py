"""ctypes wrapper of the FP64 device reference (oracle/devref.hip, test
infrastructure only): radix-2 Stockham c2c / r2c / c2r in FP64 on the GPU,
independent of the library's plans and kernels, and FP64 error sums of a
library output against it.  Buffers are torch tensors; the reference arrays
are float64 tensors of interleaved (re, im) pairs, [batch][n0][inner]."""
from __future__ import annotations

import ctypes
import math
import os

from oracle import oracle as O

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LIB_PATH = os.path.join(ROOT, "oracle", "libdevref.so")
IL, SP, REAL = 0, 1, 2  # layouts of a float32 array handed to the reference

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle oracle)")
        L = ctypes.CDLL(LIB_PATH)
        sz, vp, i = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int
        L.devref_last_error.restype = ctypes.c_char_p
        L.devref_c2c.argtypes = [i, ctypes.POINTER(sz), sz, vp, vp, i, vp]
        L.devref_load.argtypes = [vp, vp, i, sz, vp, vp]
        L.devref_c2r.argtypes = [i, ctypes.POINTER(sz), sz, vp, vp, i, vp, vp, vp]
        L.devref_err.argtypes = [vp, vp, vp, i, sz, sz, sz, sz, ctypes.POINTER(ctypes.c_double), vp]
        for f in (L.devref_c2c, L.devref_load, L.devref_c2r, L.devref_err):
            f.restype = i
        _lib = L
    return _lib


def _ok(rc):
    if rc != 0:
        raise RuntimeError("devref: " + lib().devref_last_error().decode())


def _p(t, byte_off=0):
    return None if t is None else ctypes.c_void_p(t.data_ptr() + byte_off)


def _dims(shape):
    s = [int(d) for d in O.squeeze_shape(shape)]
    return len(s), (ctypes.c_size_t * len(s))(*s), s


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def c2c(re, im, layout, shape, batch, inverse=False):
    """FP64 c2c over every axis of `shape` for `batch` transforms of float32
    input (layout IL / SP, or REAL for a real input); returns the float64
    tensor of interleaved pairs [batch][n]."""
    import torch
    nd, arr, s = _dims(shape)
    n = math.prod(s)
    out = torch.empty(2 * n * batch, dtype=torch.float64, device="cuda")
    work = torch.empty_like(out)
    _ok(lib().devref_load(_p(re), _p(im), layout, n * batch, _p(out), _stream()))
    _ok(lib().devref_c2c(nd, arr, batch, _p(out), _p(work), int(inverse), _stream()))
    del work
    return out


def c2r(re, im, il, shape, batch):
    """FP64 c2r of half spectra [batch][n0/2+1][inner] (float32, interleaved
    when il, else split); returns float64 pairs [batch][n0][inner] whose real
    parts are the result."""
    import torch
    nd, arr, s = _dims(shape)
    n = math.prod(s)
    out = torch.empty(2 * n * batch, dtype=torch.float64, device="cuda")
    work = torch.empty_like(out)
    _ok(lib().devref_c2r(nd, arr, batch, _p(re), _p(im), int(bool(il)), _p(out), _p(work), _stream()))
    del work
    return out


def rel_l2(ref, re, im, layout, shape, batch, half=False, each=False):
    """rel-L2 of a library output (float32; layout IL / SP, or REAL for a c2r
    output) against the FP64 reference `ref` ([batch][n0][inner] pairs).
    half: the output holds dim-0 rows 0..n0/2 only (r2c).  each: a list with
    one value per transform instead of one over the batch."""
    _, _, s = _dims(shape)
    n0 = s[0]
    inner = math.prod(s[1:])
    rows = n0 // 2 + 1 if half else n0
    per = rows * inner  # output elements per transform
    esz = 8 if layout == IL else 4
    st = _stream()

    def one(b0, cnt):
        sums = (ctypes.c_double * 2)()
        _ok(lib().devref_err(_p(ref, b0 * n0 * inner * 16), _p(re, b0 * per * esz), _p(im, b0 * per * 4), layout,
                             cnt, n0, rows, inner, sums, st))
        return math.sqrt(sums[0] / sums[1]) if sums[1] > 0 else math.sqrt(sums[0])

    if each:
        return [one(b, 1) for b in range(batch)]
    return one(0, batch)
