"""numpy front end of the CPU oracle (liboracle.so) and of the reference build.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package fft_amd/.

- ``fft / rfft / irfft`` call the FP64 restatement in afft_oracle.c (see the
  header of that file for the reference lines each one follows).
- ``RefAfft`` drives the real afft library built from the reference sources
  into ``oracle/_ref/libafft_ref.so`` (fft.h:22-84 C API) on host buffers; it
  is used to pin the oracle, to write the golden fixtures and as the timed CPU
  baseline.
- Layout helpers mirror the reference harness wrappers (testing.cpp:259-644):
  split vs interleaved, the split imaginary offset
  afft_32_align_size((N0/2+1)*inner) (testing.cpp:548-559), and the halved
  dim 0 (HalvedDimFirst, testing.cpp:530-538).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

c_size_p = ctypes.POINTER(ctypes.c_size_t)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)


def build(force: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    path = os.path.join(HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.check_call(["make", "-s", "-C", HERE, "all"])


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])
        L = ctypes.CDLL(path)
        L.orc_fft.argtypes = [ctypes.c_size_t, c_size_p, c_double_p, c_double_p, ctypes.c_int]
        L.orc_rfft.argtypes = [ctypes.c_size_t, c_size_p, c_double_p, c_double_p]
        L.orc_irfft.argtypes = [ctypes.c_size_t, c_size_p, c_double_p, c_double_p]
        L.orc_dft_bruteforce.argtypes = [ctypes.c_size_t, c_double_p, c_double_p, ctypes.c_int]
        L.orc_rel_l2.argtypes = [ctypes.c_size_t, c_float_p, c_double_p]
        L.orc_rel_l2.restype = ctypes.c_double
        L.orc_rel_l2_d.argtypes = [ctypes.c_size_t, c_double_p, c_double_p]
        L.orc_rel_l2_d.restype = ctypes.c_double
        L.orc_harness_error.argtypes = [ctypes.c_size_t, c_float_p, c_double_p]
        L.orc_harness_error.restype = ctypes.c_double
        L.orc_uniform.argtypes = [ctypes.c_uint64, ctypes.c_size_t, c_float_p]
        L.orc_normal.argtypes = [ctypes.c_uint64, ctypes.c_size_t, c_float_p]
        _LIB = L
    return _LIB


def _dims(shape):
    shape = [int(s) for s in shape]
    return len(shape), (ctypes.c_size_t * max(1, len(shape)))(*shape)


def _dp(a):
    return a.ctypes.data_as(c_double_p)


def _fp(a):
    return a.ctypes.data_as(c_float_p)


def squeeze_shape(shape):
    """remove_redundant_dimensions (common.hpp:68-84)."""
    s = [int(d) for d in shape if int(d) != 1]
    return s if s else [1]


def half_shape(shape):
    """Shape of a real transform's complex side: dim 0 (after dropping unit
    dims) becomes N0/2+1 (fft_internal.hpp:215-219)."""
    s = squeeze_shape(shape)
    return [s[0] // 2 + 1] + s[1:]


# ---- oracle transforms (complex128 in / out) ---------------------------------

def fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """FP64 c2c of an array of any shape (unnormalised inverse)."""
    a = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(a)
    nd, dims = _dims(a.shape if a.ndim else (1,))
    lib().orc_fft(nd, dims, _dp(a.view(np.float64)), _dp(out.view(np.float64)), int(bool(inverse)))
    return out


def rfft(x: np.ndarray) -> np.ndarray:
    """FP64 afft real forward transform; returns complex of half_shape(x.shape)."""
    a = np.ascontiguousarray(x, dtype=np.float64)
    hs = half_shape(a.shape)
    out = np.empty(hs, dtype=np.complex128)
    nd, dims = _dims(a.shape)
    lib().orc_rfft(nd, dims, _dp(a), _dp(out.view(np.float64)))
    return out


def irfft(z: np.ndarray, shape) -> np.ndarray:
    """FP64 afft real inverse transform of complex data of half_shape(shape)."""
    hs = half_shape(shape)
    a = np.ascontiguousarray(z, dtype=np.complex128).reshape(hs)
    out = np.empty(squeeze_shape(shape), dtype=np.float64)
    nd, dims = _dims(shape)
    lib().orc_irfft(nd, dims, _dp(a.view(np.float64)), _dp(out))
    return out.reshape([int(s) for s in shape])


def dft_bruteforce(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).ravel()
    out = np.empty_like(a)
    lib().orc_dft_bruteforce(a.size, _dp(a.view(np.float64)), _dp(out.view(np.float64)), int(bool(inverse)))
    return out


def rel_l2(y, ref) -> float:
    """||y - ref|| / ||ref|| over all real and imaginary components."""
    y = np.asarray(y)
    ref = np.ascontiguousarray(np.asarray(ref))
    yv = np.ascontiguousarray(y).view(np.float32 if y.dtype in (np.float32, np.complex64) else np.float64).ravel()
    rv = ref.view(np.float64).ravel() if ref.dtype in (np.float64, np.complex128) else ref.astype(np.float64).ravel()
    if yv.dtype == np.float32:
        return lib().orc_rel_l2(yv.size, _fp(yv), _dp(np.ascontiguousarray(rv)))
    return lib().orc_rel_l2_d(yv.size, _dp(np.ascontiguousarray(yv)), _dp(np.ascontiguousarray(rv)))


def harness_error(y, ref) -> float:
    """The reference harness metric (testing.cpp:1113-1127)."""
    yv = np.ascontiguousarray(np.asarray(y)).view(np.float32).ravel()
    rv = np.ascontiguousarray(np.asarray(ref)).view(np.float64).ravel()
    return lib().orc_harness_error(yv.size, _fp(yv), _dp(rv))


def bound(total_elements: int) -> float:
    """2*log2(N)*2^-24, the relative-L2 bound (N = total element count)."""
    n = max(2, int(total_elements))
    return 2.0 * np.log2(n) * 2.0 ** -24


def uniform(seed: int, n: int) -> np.ndarray:
    out = np.empty(int(n), dtype=np.float32)
    lib().orc_uniform(int(seed), out.size, _fp(out))
    return out


def normal(seed: int, n: int) -> np.ndarray:
    out = np.empty(int(n), dtype=np.float32)
    lib().orc_normal(int(seed), out.size, _fp(out))
    return out


def complex_input(seed: int, shape) -> np.ndarray:
    """re, im i.i.d. uniform[-1,1): re block then im block (testing.cpp:1064-1067)."""
    n = int(np.prod(shape))
    v = uniform(seed, 2 * n)
    return (v[:n] + 1j * v[n:]).astype(np.complex64).reshape(shape)


def hermitian_input(seed: int, shape) -> np.ndarray:
    """Complex input of half_shape(shape) that is the afft-side half of a
    Hermitian-symmetric array, built like the reference harness does for
    inverse real tests (testing.cpp:1069-1094): full array re/im mirrored as
    re' = re + re[mirror], im' = im[mirror] - im, then the halved dim cut."""
    full = complex_input(seed, squeeze_shape(shape)).astype(np.complex128)
    mir = full
    for ax in range(full.ndim):
        mir = np.flip(np.roll(mir, -1, axis=ax), axis=ax)  # index (n - i) mod n
    re = full.real + mir.real
    im = mir.imag - full.imag
    herm = re + 1j * im
    hs = half_shape(shape)
    return herm[: hs[0]].astype(np.complex64)


# ---- layout packing (testing.cpp:259-644) ------------------------------------

def align_size_f32(n: int) -> int:
    """afft_32_align_size: round up to 64 bytes of floats (common.hpp:173-177)."""
    return (int(n) + 15) & ~15


def split_im_off(shape) -> int:
    """Split-format imaginary offset of a real transform's complex side
    (AlignedImOff, testing.cpp:548-559)."""
    return align_size_f32(int(np.prod(half_shape(shape))))


# ---- the reference library ----------------------------------------------------

AFFT_AUTO, AFFT_SCALAR, AFFT_SSE2, AFFT_AVX2, AFFT_AVX512F = 0, 1, 2, 3, 4
AFFT_SPLIT, AFFT_INTERLEAVED = 0, 1
ISA_NAMES = {AFFT_SCALAR: "scalar", AFFT_SSE2: "sse2", AFFT_AVX2: "avx2", AFFT_AVX512F: "avx512f"}


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libafft_ref.so"))


class RefAfft:
    """The real afft FP32 library (host memory), fft.h:22-84."""

    def __init__(self):
        global _REF
        if _REF is None:
            L = ctypes.CDLL(os.path.join(HERE, "_ref", "libafft_ref.so"))
            for k in ("c", "ci", "r", "ri"):
                ms = getattr(L, f"afft_32_{k}_memsize")
                ms.argtypes = [ctypes.c_size_t, c_size_p, ctypes.c_int, ctypes.c_int]
                ms.restype = ctypes.c_size_t
                cr = getattr(L, f"afft_32_{k}_create")
                cr.argtypes = [ctypes.c_size_t, c_size_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
                cr.restype = ctypes.c_void_p
                tr = getattr(L, f"afft_32_{k}_transform")
                n = 3 if k in ("r", "ri") else 4
                tr.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * n
                tr.restype = None
            L.afft_supported.argtypes = [ctypes.c_int]
            L.afft_supported.restype = ctypes.c_int
            _REF = L
        self.L = _REF

    def supported(self, impl: int) -> bool:
        return bool(self.L.afft_supported(impl))

    def plan(self, kind: str, shape, fmt: int = AFFT_INTERLEAVED, impl: int = AFFT_AUTO):
        nd, dims = _dims(shape)
        size = getattr(self.L, f"afft_32_{kind}_memsize")(nd, dims, fmt, impl)
        if size == 0 or size == ctypes.c_size_t(-1).value:
            raise ValueError("unsupported plan")
        mem = ctypes.create_string_buffer(size)
        st = getattr(self.L, f"afft_32_{kind}_create")(nd, dims, fmt, mem, impl)
        if not st:
            raise ValueError("create failed")
        return (kind, mem, st)

    def run(self, plan, *bufs):
        kind, mem, st = plan
        ptrs = [None if b is None else b.ctypes.data for b in bufs]
        getattr(self.L, f"afft_32_{kind}_transform")(st, *ptrs)

    # Whole-array helpers returning canonical complex64 / float32 results.
    def c2c(self, x: np.ndarray, inverse=False, fmt=AFFT_INTERLEAVED, impl=AFFT_AUTO):
        shape = x.shape
        kind = "ci" if inverse else "c"
        p = self.plan(kind, shape, fmt, impl)
        if fmt == AFFT_INTERLEAVED:
            src = np.ascontiguousarray(x, dtype=np.complex64).copy()
            dst = np.empty_like(src)
            self.run(p, src, None, dst, None)
            return dst
        sr = np.ascontiguousarray(x.real, dtype=np.float32)
        si = np.ascontiguousarray(x.imag, dtype=np.float32)
        dr = np.empty_like(sr)
        di = np.empty_like(si)
        self.run(p, sr, si, dr, di)
        return (dr + 1j * di).astype(np.complex64)

    def r2c(self, x: np.ndarray, fmt=AFFT_INTERLEAVED, impl=AFFT_AUTO):
        shape = x.shape
        hs = half_shape(shape)
        p = self.plan("r", shape, fmt, impl)
        src = np.ascontiguousarray(x, dtype=np.float32).copy()
        if fmt == AFFT_INTERLEAVED:
            dst = np.empty(hs, dtype=np.complex64)
            self.run(p, src, dst, None)
            return dst
        off = split_im_off(shape)
        buf = np.zeros(2 * off, dtype=np.float32)
        self.run(p, src, buf[:off], buf[off:])
        n = int(np.prod(hs))
        return (buf[:n] + 1j * buf[off:off + n]).astype(np.complex64).reshape(hs)

    def c2r(self, z: np.ndarray, shape, fmt=AFFT_INTERLEAVED, impl=AFFT_AUTO):
        hs = half_shape(shape)
        p = self.plan("ri", shape, fmt, impl)
        dst = np.empty([int(s) for s in shape], dtype=np.float32)
        if fmt == AFFT_INTERLEAVED:
            src = np.ascontiguousarray(z, dtype=np.complex64).reshape(hs).copy()
            self.run(p, src, None, dst)
            return dst
        off = split_im_off(shape)
        n = int(np.prod(hs))
        buf = np.zeros(2 * off, dtype=np.float32)
        zz = np.ascontiguousarray(z, dtype=np.complex64).ravel()
        buf[:n] = zz.real
        buf[off:off + n] = zz.imag
        self.run(p, buf[:off], buf[off:], dst)
        return dst
