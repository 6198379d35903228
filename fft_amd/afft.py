"""Python mirror of afft's C++ wrapper classes (fft.hpp:47-234) over the GPU
C ABI: ``complex_transform``, ``inverse_complex_transform``, ``real_transform``
and ``inverse_real_transform``, each templated in afft on the element type and
the format.  Here the element type is ``dtype`` (numpy float32 / float64) and
the buffers are device buffers (torch tensors on cuda or raw device addresses).

As in afft, a failed create leaves the object falsy (fft.hpp:68) instead of
raising; calling a falsy transform raises.  ``close()`` releases the device
resources (the reference has nothing to release).  ``Plan`` is the batched,
stream-aware extension (include/afft_gpu.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import afft_auto, afft_interleaved, afft_split  # noqa: F401


def _half_count(dims) -> int:
    """Complex elements of a real transform's complex side: dim 0 (after unit
    dims are dropped) becomes N0/2+1 (fft_internal.hpp:215-219)."""
    s = [int(d) for d in dims if int(d) != 1] or [1]
    return (s[0] // 2 + 1) * int(np.prod(s[1:], dtype=np.int64))


def required_counts(kind: str, dims, fmt: int, batch: int = 1):
    """Scalars each of (src_re, src_im, dst_re, dst_im) must hold for `batch`
    transforms back to back (0: the argument is not read or written).  Real
    kinds take (src, dst_re, dst_im) and (src_re, src_im, dst); the tuple is
    in that argument order."""
    n = int(np.prod([int(d) for d in dims], dtype=np.int64))
    m = _half_count(dims)
    il = int(fmt) == afft_interleaved
    b = int(batch)
    if kind in ("c", "ci"):
        return (2 * n * b, 0, 2 * n * b, 0) if il else (n * b, n * b, n * b, n * b)
    if kind == "r":
        return (n * b, 2 * m * b, 0) if il else (n * b, m * b, m * b)
    return (2 * m * b, 0, n * b) if il else (m * b, m * b, n * b)


def check_buffers(bufs, counts, dtype):
    """Rejects, before any launch, torch tensors that are on the host, of the
    wrong element type, not contiguous or too small for the plan, and a missing
    buffer the plan needs.  Raw integer addresses are the caller's to get
    right."""
    dt = np.dtype(dtype)
    names = ("first", "second", "third", "fourth")
    for i, (x, need) in enumerate(zip(bufs, counts)):
        if need == 0:
            continue
        if x is None:
            raise ValueError(f"the {names[i]} buffer is required ({need} elements) but is None")
        if isinstance(x, int):
            if x == 0:
                raise ValueError(f"the {names[i]} buffer is a null address")
            continue
        if not hasattr(x, "data_ptr"):
            raise TypeError(f"the {names[i]} buffer must be a device tensor or an address, not {type(x).__name__}")
        if not getattr(x, "is_cuda", False):
            raise ValueError(f"the {names[i]} buffer is not on the GPU")
        tdt = str(x.dtype).replace("torch.", "")
        if tdt != dt.name:
            raise ValueError(f"the {names[i]} buffer has dtype {tdt}, the plan computes in {dt.name}")
        if not x.is_contiguous():
            raise ValueError(f"the {names[i]} buffer is not contiguous")
        if x.numel() < need:
            raise ValueError(f"the {names[i]} buffer holds {x.numel()} elements, the plan needs {need}")


def _prefix(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "32"
    if dt == np.float64:
        return "64"
    raise TypeError(f"unsupported element type {dt}")


class _Transform:
    _kind = "c"

    def __init__(self, dims, dtype=np.float32, format=afft_split, impl=afft_auto):
        self.dims = [int(d) for d in dims]
        self.format = int(format)
        self.dtype = np.dtype(dtype)
        p = _prefix(dtype)
        L = _lib.lib()
        nd, arr = _lib.dims_array(self.dims)
        self._fn_transform = getattr(L, f"afft_{p}_{self._kind}_transform")
        self._fn_destroy = getattr(L, f"afft_{p}_{self._kind}_destroy")
        size = getattr(L, f"afft_{p}_{self._kind}_memsize")(nd, arr, self.format, int(impl))
        self.memsize = size
        self._mem = None
        self.state = None
        if size == 0 or size == ctypes.c_size_t(-1).value:
            return
        self._mem = ctypes.create_string_buffer(size)
        st = getattr(L, f"afft_{p}_{self._kind}_create")(nd, arr, self.format, self._mem, int(impl))
        self.state = st if st else None

    def __bool__(self):
        return self.state is not None

    def _check(self):
        if not self.state:
            raise RuntimeError(f"invalid transform: {_lib.last_error()}")

    def close(self):
        if self.state:
            self._fn_destroy(self.state)
        self.state = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream) -> None:
        """Run on a HIP stream (torch.cuda.Stream or raw handle)."""
        self._check()
        h = getattr(stream, "cuda_stream", stream)
        if _lib.lib().afft_gpu_set_stream(self.state, h) != 0:
            raise RuntimeError(_lib.last_error())


class complex_transform(_Transform):
    """afft::complex_transform (fft.hpp:47-118): forward c2c."""
    _kind = "c"

    def __call__(self, src_re, src_im, dst_re, dst_im):
        self._check()
        check_buffers((src_re, src_im, dst_re, dst_im), required_counts(self._kind, self.dims, self.format), self.dtype)
        self._fn_transform(self.state, _lib.ptr(src_re), _lib.ptr(src_im), _lib.ptr(dst_re), _lib.ptr(dst_im))


class inverse_complex_transform(_Transform):
    """afft::inverse_complex_transform (fft.hpp:120-164): unnormalised inverse c2c."""
    _kind = "ci"

    def __call__(self, src_re, src_im, dst_re, dst_im):
        self._check()
        check_buffers((src_re, src_im, dst_re, dst_im), required_counts(self._kind, self.dims, self.format), self.dtype)
        self._fn_transform(self.state, _lib.ptr(src_re), _lib.ptr(src_im), _lib.ptr(dst_re), _lib.ptr(dst_im))


class real_transform(_Transform):
    """afft::real_transform (fft.hpp:166-200): r2c, dim 0 halved."""
    _kind = "r"

    def __call__(self, src, dst_re, dst_im):
        self._check()
        check_buffers((src, dst_re, dst_im), required_counts("r", self.dims, self.format), self.dtype)
        self._fn_transform(self.state, _lib.ptr(src), _lib.ptr(dst_re), _lib.ptr(dst_im))


class inverse_real_transform(_Transform):
    """afft::inverse_real_transform (fft.hpp:202-234): unnormalised c2r."""
    _kind = "ri"

    def __call__(self, src_re, src_im, dst):
        self._check()
        check_buffers((src_re, src_im, dst), required_counts("ri", self.dims, self.format), self.dtype)
        self._fn_transform(self.state, _lib.ptr(src_re), _lib.ptr(src_im), _lib.ptr(dst))


def supported(impl: int) -> bool:
    return bool(_lib.lib().afft_supported(int(impl)))


def align_size(n: int, dtype=np.float32) -> int:
    p = _prefix(dtype)
    return int(getattr(_lib.lib(), f"afft_{p}_align_size")(int(n)))


# afft_gpu_mode_* (include/afft_gpu.h)
MODE_FMA, MODE_MFMA_F16, MODE_MFMA_BF16 = 0, 1, 2


class Variant(ctypes.Structure):
    """afft_gpu_variant (include/afft_gpu.h): kernel-shape choices of a plan;
    0 = default for every field."""
    _fields_ = [("col_tile", ctypes.c_int32), ("rowt_tile", ctypes.c_int32), ("stream_io", ctypes.c_int32),
                ("xcd_order", ctypes.c_int32), ("chunk", ctypes.c_int64), ("sched", ctypes.c_int32),
                ("half_lds", ctypes.c_int32), ("pipeline", ctypes.c_int32), ("real_pair", ctypes.c_int32),
                ("a_split", ctypes.c_int32), ("row_log2", ctypes.c_int32), ("fused", ctypes.c_int32),
                ("fused_slots", ctypes.c_int32), ("fused_lag", ctypes.c_int32),
                ("tiny_rows", ctypes.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Plan:
    """Batched asynchronous plan (afft_gpu_plan_create / afft_gpu_execute)."""

    def __init__(self, kind: str, dims, batch: int = 1, dtype=np.float32, format=afft_interleaved):
        L = _lib.lib()
        self.kind = kind
        self.dims = [int(d) for d in dims]
        self.batch = int(batch)
        self.format = int(format)
        prec = 32 if np.dtype(dtype) == np.float32 else 64
        self.dtype = np.dtype(np.float32 if prec == 32 else np.float64)
        nd, arr = _lib.dims_array(self.dims)
        self._mem = ctypes.create_string_buffer(L.afft_gpu_plan_memsize())
        st = L.afft_gpu_plan_create(prec, _lib.KINDS[kind], nd, arr, self.format, self.batch, self._mem)
        if not st:
            raise RuntimeError(f"plan creation failed: {_lib.last_error()}")
        self.state = st

    def _bufs(self, src_re, src_im, dst_re, dst_im):
        """Checks the buffers against the plan; the real kinds use the first
        three of the four arguments as (src, dst_re, dst_im) / (src_re, src_im,
        dst)."""
        if not self.state:
            raise RuntimeError("plan is closed")
        cnt = required_counts(self.kind, self.dims, self.format, self.batch)
        if len(cnt) == 3:
            if self.kind == "r":
                # execute's (src_re, src_im, dst_re, dst_im) with src_im unused
                check_buffers((src_re, dst_re, dst_im), cnt, self.dtype)
            else:
                check_buffers((src_re, src_im, dst_re), cnt, self.dtype)
        else:
            check_buffers((src_re, src_im, dst_re, dst_im), cnt, self.dtype)

    def __call__(self, src_re, src_im, dst_re, dst_im, stream=None):
        self._bufs(src_re, src_im, dst_re, dst_im)
        h = getattr(stream, "cuda_stream", stream)
        r = _lib.lib().afft_gpu_execute(self.state, _lib.ptr(src_re), _lib.ptr(src_im), _lib.ptr(dst_re),
                                        _lib.ptr(dst_im), h)
        if r != 0:
            raise RuntimeError(_lib.last_error())

    def run_range(self, first, last, src_re, src_im, dst_re, dst_im, stream=None):
        """Enqueues passes [first, last) only (for per-kernel timing)."""
        self._bufs(src_re, src_im, dst_re, dst_im)
        h = getattr(stream, "cuda_stream", stream)
        r = _lib.lib().afft_gpu_execute_range(self.state, int(first), int(last), _lib.ptr(src_re), _lib.ptr(src_im),
                                              _lib.ptr(dst_re), _lib.ptr(dst_im), h)
        if r != 0:
            raise RuntimeError(_lib.last_error())

    def launches(self, pass_index: int) -> int:
        return _lib.lib().afft_gpu_plan_launches(self.state, int(pass_index))

    def set_chunk(self, chunk: int):
        if _lib.lib().afft_gpu_plan_set_chunk(self.state, int(chunk)) != 0:
            raise RuntimeError(_lib.last_error())

    def get_variant(self) -> Variant:
        v = Variant()
        if _lib.lib().afft_gpu_plan_get_variant(self.state, ctypes.byref(v)) != 0:
            raise RuntimeError(_lib.last_error())
        return v

    def set_variant(self, **fields):
        """Rebuilds the plan with the given variant fields (others: default;
        stream_io defaults to 2: streaming I/O except for N-D plans whose
        arrays hold at most 128 MiB per side)."""
        v = Variant(stream_io=2)
        for k, val in fields.items():
            setattr(v, k, int(val))
        if _lib.lib().afft_gpu_plan_set_variant(self.state, ctypes.byref(v)) != 0:
            raise RuntimeError(_lib.last_error())

    def tune(self, src_re, src_im, dst_re, dst_im, stream=None, iters: int = 3) -> float:
        """Runs the autotuner on these buffers; returns the best ms."""
        self._bufs(src_re, src_im, dst_re, dst_im)
        h = getattr(stream, "cuda_stream", stream)
        best = ctypes.c_float(0)
        r = _lib.lib().afft_gpu_plan_tune(self.state, _lib.ptr(src_re), _lib.ptr(src_im), _lib.ptr(dst_re),
                                          _lib.ptr(dst_im), h, int(iters), ctypes.byref(best))
        if r != 0:
            raise RuntimeError(_lib.last_error())
        return best.value

    def set_mode(self, mode: int):
        if _lib.lib().afft_gpu_plan_set_mode(self.state, int(mode)) != 0:
            raise RuntimeError(_lib.last_error())

    @property
    def num_passes(self) -> int:
        return _lib.lib().afft_gpu_plan_num_passes(self.state)

    @property
    def scratch_bytes(self) -> int:
        return _lib.lib().afft_gpu_plan_scratch_bytes(self.state)

    def describe(self):
        out = []
        buf = ctypes.create_string_buffer(512)
        for i in range(self.num_passes):
            _lib.lib().afft_gpu_plan_describe(self.state, i, buf, 512)
            out.append(buf.value.decode())
        return out

    def close(self):
        if self.state:
            _lib.lib().afft_gpu_plan_destroy(self.state)
        self.state = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Job(ctypes.Structure):
    """afft_gpu_job (include/afft_gpu.h)."""
    _fields_ = [("kind", ctypes.c_int32), ("log2n", ctypes.c_int32), ("src", ctypes.c_uint64),
                ("dst", ctypes.c_uint64)]


class Queue:
    """Uneven batch of 1-D FP32 transforms run in one persistent launch
    (afft_gpu_queue_*)."""

    def __init__(self, max_jobs: int, max_log2n: int = 16):
        L = _lib.lib()
        self._mem = ctypes.create_string_buffer(L.afft_gpu_queue_memsize())
        q = L.afft_gpu_queue_create(int(max_jobs), int(max_log2n), self._mem)
        if not q:
            raise RuntimeError(f"queue creation failed: {_lib.last_error()}")
        self.q = q

    @staticmethod
    def jobs(records):
        """records: iterable of (kind, log2n, src_addr, dst_addr)."""
        recs = list(records)
        arr = (Job * max(1, len(recs)))()
        for i, (k, lg, s, d) in enumerate(recs):
            arr[i] = Job(int(k), int(lg), int(s), int(d))
        return arr, len(recs)

    def prepare(self, jobs) -> int:
        arr, n = jobs
        r = _lib.lib().afft_gpu_queue_prepare(self.q, arr, n)
        if r < 0:
            raise RuntimeError(_lib.last_error())
        return r

    def launch(self, stream=None):
        h = getattr(stream, "cuda_stream", stream)
        if _lib.lib().afft_gpu_queue_launch(self.q, h) != 0:
            raise RuntimeError(_lib.last_error())

    def run(self, jobs, stream=None):
        self.prepare(jobs)
        self.launch(stream)

    def status(self) -> int:
        """0: the last launch met every dependency; 1: a wait hit the spin limit."""
        r = _lib.lib().afft_gpu_queue_status(self.q)
        if r < 0:
            raise RuntimeError(_lib.last_error())
        return r

    def set_groups_per_cu(self, groups: int):
        if _lib.lib().afft_gpu_queue_set_groups_per_cu(self.q, int(groups)) != 0:
            raise RuntimeError(_lib.last_error())

    def set_spin_limit(self, limit: int):
        if _lib.lib().afft_gpu_queue_set_spin_limit(self.q, int(limit)) != 0:
            raise RuntimeError(_lib.last_error())

    def set_lag(self, lag: int):
        if _lib.lib().afft_gpu_queue_set_lag(self.q, int(lag)) != 0:
            raise RuntimeError(_lib.last_error())

    @property
    def config(self):
        """(groups per CU, lag) the next launch / prepare uses."""
        g, lag = ctypes.c_int(), ctypes.c_int()
        if _lib.lib().afft_gpu_queue_get_config(self.q, ctypes.byref(g), ctypes.byref(lag)) != 0:
            raise RuntimeError(_lib.last_error())
        return g.value, lag.value

    def tune(self, jobs, stream=None, iters: int = 3) -> float:
        """Times the (groups per CU, lag) candidates on these jobs, keeps the
        fastest and leaves the queue prepared; returns its ms per launch."""
        arr, n = jobs
        h = getattr(stream, "cuda_stream", stream)
        ms = ctypes.c_float()
        if _lib.lib().afft_gpu_queue_tune(self.q, arr, n, h, int(iters), ctypes.byref(ms)) != 0:
            raise RuntimeError(_lib.last_error())
        return ms.value

    def close(self):
        if self.q:
            _lib.lib().afft_gpu_queue_destroy(self.q)
        self.q = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
