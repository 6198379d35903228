"""ctypes binding of libafft_mi355x.so (the C ABI in include/fft.h and
include/afft_gpu.h).  Loading fails loudly when the library has not been built:
there is no CPU fallback anywhere in this package."""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libafft_mi355x.so")

afft_auto, afft_scalar, afft_sse2, afft_avx2, afft_avx512f, afft_neon = 0, 1, 2, 3, 4, 5
afft_split, afft_interleaved = 0, 1
afft_gpu = 16
KINDS = {"c": 0, "ci": 1, "r": 2, "ri": 3}

_LIB = None

c_size_p = ctypes.POINTER(ctypes.c_size_t)
vp = ctypes.c_void_p

# Every symbol include/fft.h and include/afft_gpu.h declare.
FFT_H_SYMBOLS = (
    [f"afft_{p}_{k}_{f}" for p in ("32", "64") for k in ("c", "ci", "r", "ri")
     for f in ("memsize", "create", "transform", "destroy")]
    + ["afft_32_align_size", "afft_64_align_size", "afft_supported", "afft_alignment"]
)
AFFT_GPU_H_SYMBOLS = [
    "afft_gpu_plan_memsize", "afft_gpu_plan_create", "afft_gpu_execute", "afft_gpu_plan_destroy",
    "afft_gpu_set_stream", "afft_gpu_plan_set_chunk", "afft_gpu_plan_num_passes",
    "afft_gpu_plan_scratch_bytes", "afft_gpu_plan_describe", "afft_gpu_last_error",
    "afft_gpu_error_sums_f32", "afft_gpu_error_sums_f64", "afft_gpu_copy", "afft_gpu_execute_range",
    "afft_gpu_plan_launches", "afft_gpu_plan_inspect", "afft_gpu_plan_inspect_variant",
    "afft_gpu_queue_memsize", "afft_gpu_queue_create", "afft_gpu_queue_prepare", "afft_gpu_queue_launch",
    "afft_gpu_queue_run", "afft_gpu_queue_num_tasks", "afft_gpu_queue_destroy", "afft_gpu_queue_status",
    "afft_gpu_queue_set_spin_limit", "afft_gpu_queue_set_groups_per_cu",
    "afft_gpu_plan_set_mode", "afft_gpu_mfma_selftest", "afft_gpu_plan_set_variant", "afft_gpu_plan_get_variant",
    "afft_gpu_plan_tune", "afft_gpu_plan_describe_variant", "afft_gpu_execute_noalloc",
    "afft_gpu_queue_set_lag", "afft_gpu_queue_get_config", "afft_gpu_queue_tune",
]


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m fft_amd.build` (hipcc, gfx950) first")
    # PyTorch ships its own libamdhip64.so.7.  Loading torch first makes this
    # library bind to that same HIP runtime (matched by soname), so device
    # pointers, streams and graph capture are shared; loading this library
    # first would bring in a second runtime and hide the GPU from torch.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for p, T in (("32", ctypes.c_float), ("64", ctypes.c_double)):
        for k in ("c", "ci", "r", "ri"):
            f = getattr(L, f"afft_{p}_{k}_memsize")
            f.argtypes = [ctypes.c_size_t, c_size_p, ctypes.c_int, ctypes.c_int]
            f.restype = ctypes.c_size_t
            f = getattr(L, f"afft_{p}_{k}_create")
            f.argtypes = [ctypes.c_size_t, c_size_p, ctypes.c_int, vp, ctypes.c_int]
            f.restype = vp
            f = getattr(L, f"afft_{p}_{k}_transform")
            f.argtypes = [vp] + [vp] * (3 if k in ("r", "ri") else 4)
            f.restype = None
            f = getattr(L, f"afft_{p}_{k}_destroy")
            f.argtypes = [vp]
            f.restype = None
    L.afft_32_align_size.argtypes = [ctypes.c_size_t]
    L.afft_32_align_size.restype = ctypes.c_size_t
    L.afft_64_align_size.argtypes = [ctypes.c_size_t]
    L.afft_64_align_size.restype = ctypes.c_size_t
    L.afft_supported.argtypes = [ctypes.c_int]
    L.afft_supported.restype = ctypes.c_int
    L.afft_gpu_plan_memsize.argtypes = []
    L.afft_gpu_plan_memsize.restype = ctypes.c_size_t
    L.afft_gpu_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, c_size_p, ctypes.c_int,
                                       ctypes.c_size_t, vp]
    L.afft_gpu_plan_create.restype = vp
    L.afft_gpu_execute.argtypes = [vp, vp, vp, vp, vp, vp]
    L.afft_gpu_execute.restype = ctypes.c_int
    L.afft_gpu_execute_noalloc.argtypes = [vp, vp, vp, vp, vp, vp]
    L.afft_gpu_execute_noalloc.restype = ctypes.c_int
    L.afft_gpu_execute_range.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp]
    L.afft_gpu_execute_range.restype = ctypes.c_int
    L.afft_gpu_plan_launches.argtypes = [vp, ctypes.c_int]
    L.afft_gpu_plan_launches.restype = ctypes.c_int
    L.afft_gpu_plan_destroy.argtypes = [vp]
    L.afft_gpu_plan_destroy.restype = None
    L.afft_gpu_set_stream.argtypes = [vp, vp]
    L.afft_gpu_set_stream.restype = ctypes.c_int
    L.afft_gpu_plan_set_chunk.argtypes = [vp, ctypes.c_size_t]
    L.afft_gpu_plan_set_chunk.restype = ctypes.c_int
    L.afft_gpu_plan_num_passes.argtypes = [vp]
    L.afft_gpu_plan_num_passes.restype = ctypes.c_int
    L.afft_gpu_plan_scratch_bytes.argtypes = [vp]
    L.afft_gpu_plan_scratch_bytes.restype = ctypes.c_size_t
    L.afft_gpu_plan_describe.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
    L.afft_gpu_plan_describe.restype = ctypes.c_int
    L.afft_gpu_last_error.argtypes = []
    L.afft_gpu_last_error.restype = ctypes.c_char_p
    L.afft_gpu_error_sums_f32.argtypes = [vp, vp, ctypes.c_size_t, vp, vp]
    L.afft_gpu_error_sums_f32.restype = ctypes.c_int
    L.afft_gpu_error_sums_f64.argtypes = [vp, vp, ctypes.c_size_t, vp, vp]
    L.afft_gpu_error_sums_f64.restype = ctypes.c_int
    L.afft_gpu_copy.argtypes = [vp, vp, ctypes.c_size_t, vp]
    L.afft_gpu_copy.restype = ctypes.c_int
    L.afft_gpu_queue_memsize.argtypes = []
    L.afft_gpu_queue_memsize.restype = ctypes.c_size_t
    L.afft_gpu_queue_create.argtypes = [ctypes.c_size_t, ctypes.c_int, vp]
    L.afft_gpu_queue_create.restype = vp
    L.afft_gpu_queue_prepare.argtypes = [vp, vp, ctypes.c_size_t]
    L.afft_gpu_queue_prepare.restype = ctypes.c_int
    L.afft_gpu_queue_launch.argtypes = [vp, vp]
    L.afft_gpu_queue_launch.restype = ctypes.c_int
    L.afft_gpu_queue_run.argtypes = [vp, vp, ctypes.c_size_t, vp]
    L.afft_gpu_queue_run.restype = ctypes.c_int
    L.afft_gpu_queue_num_tasks.argtypes = [vp]
    L.afft_gpu_queue_num_tasks.restype = ctypes.c_int
    L.afft_gpu_queue_destroy.argtypes = [vp]
    L.afft_gpu_queue_destroy.restype = None
    L.afft_gpu_queue_status.argtypes = [vp]
    L.afft_gpu_queue_status.restype = ctypes.c_int
    L.afft_gpu_queue_set_spin_limit.argtypes = [vp, ctypes.c_longlong]
    L.afft_gpu_queue_set_spin_limit.restype = ctypes.c_int
    L.afft_gpu_queue_set_groups_per_cu.argtypes = [vp, ctypes.c_int]
    L.afft_gpu_queue_set_groups_per_cu.restype = ctypes.c_int
    L.afft_gpu_queue_set_lag.argtypes = [vp, ctypes.c_int]
    L.afft_gpu_queue_set_lag.restype = ctypes.c_int
    L.afft_gpu_queue_get_config.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.afft_gpu_queue_get_config.restype = ctypes.c_int
    L.afft_gpu_queue_tune.argtypes = [vp, vp, ctypes.c_size_t, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    L.afft_gpu_queue_tune.restype = ctypes.c_int
    L.afft_gpu_plan_set_mode.argtypes = [vp, ctypes.c_int]
    L.afft_gpu_plan_set_mode.restype = ctypes.c_int
    L.afft_gpu_mfma_selftest.argtypes = []
    L.afft_gpu_mfma_selftest.restype = ctypes.c_int
    L.afft_gpu_plan_set_variant.argtypes = [vp, vp]
    L.afft_gpu_plan_set_variant.restype = ctypes.c_int
    L.afft_gpu_plan_get_variant.argtypes = [vp, vp]
    L.afft_gpu_plan_get_variant.restype = ctypes.c_int
    L.afft_gpu_plan_tune.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int, vp]
    L.afft_gpu_plan_tune.restype = ctypes.c_int
    L.afft_gpu_plan_describe_variant.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, c_size_p, ctypes.c_int,
                                                 ctypes.c_size_t, vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
    L.afft_gpu_plan_describe_variant.restype = ctypes.c_int
    _LIB = L
    return L


def last_error() -> str:
    return lib().afft_gpu_last_error().decode()


def dims_array(dims):
    dims = [int(d) for d in dims]
    return len(dims), (ctypes.c_size_t * max(1, len(dims)))(*dims)


def ptr(x) -> int | None:
    """Device address of a torch tensor, an int address, or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return int(x.data_ptr())
