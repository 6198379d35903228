"""Host-only: the one-line description of every pass of every plan shape the
builder makes, across the kernel-shape variants, formats, kinds and both
precisions (afft_gpu_plan_describe_variant builds the plan without a GPU and
formats pass i with the same code afft_gpu_plan_describe uses).  Every pass
type must be described at least once, with its own name, and short buffers
must be truncated safely."""
import ctypes
import itertools

import numpy as np

from fft_amd import _lib
from fft_amd.afft import Variant

import plan_model as M

NAMES = {0: "fft", 1: "real_rows", 2: "real_one", 3: "real_tr", 4: "real_pair", 5: "real_rowfft",
         6: "real_irowfft", 7: "real_colpair", 8: "real_icolpair", 9: "real4a", 10: "real4b", 11: "real4c",
         12: "real4d", 13: "tiny"}

SHAPES = [(1,), (2,), (64,), (4096,), (16384,), (1 << 15,), (1 << 16,), (1 << 20,), (1 << 21,), (1 << 22,),
          (1 << 24,), (1 << 25,), (16, 32), (8192, 64), (8192, 1024), (16384, 2048), (256, 256, 256), (8, 4, 16, 2),
          (4096, 16), (65536, 16)]
VARIANTS = [None, dict(real_pair=1), dict(real_pair=-1), dict(real_pair=2), dict(fused=1), dict(col_tile=32),
            dict(sched=-1, half_lds=-1), dict(rowt_tile=16), dict(a_split=2), dict(row_log2=9), dict(tiny_rows=-1)]


def _describe(prec, kind, shape, fmt, batch, v, i, size=512):
    L = _lib.lib()
    nd, arr = _lib.dims_array(shape)
    vv = None
    if v is not None:
        vv = Variant(stream_io=2)
        for k, val in v.items():
            setattr(vv, k, int(val))
    buf = ctypes.create_string_buffer(size)
    r = L.afft_gpu_plan_describe_variant(prec, kind, nd, arr, fmt, batch, ctypes.byref(vv) if vv else None, i,
                                         buf, size)
    return r, buf.value.decode()


def test_every_pass_of_every_plan_is_described():
    seen = set()
    count = 0
    for prec, kind, shape, fmt, v in itertools.product((32, 64), (0, 1, 2, 3), SHAPES, (0, 1), VARIANTS):
        try:
            info, passes = M.inspect("c ci r ri".split()[kind], shape, fmt, 2, prec, v)
        except ValueError:
            continue  # a shape / variant this precision does not build
        for i, p in enumerate(passes):
            r, txt = _describe(prec, kind, shape, fmt, 2, v, i)
            assert r == 0 and txt.startswith(NAMES[p.type]), (prec, kind, shape, v, i, txt)
            seen.add(p.type)
            count += 1
        r, _ = _describe(prec, kind, shape, fmt, 2, v, len(passes))
        assert r == -1
    assert seen == set(NAMES), sorted(set(NAMES) - seen)
    assert count > 500


def test_short_buffers_truncate():
    _, full = _describe(32, 0, (1 << 20,), 1, 4, None, 0)
    for size in (1, 2, 8, 17):
        r, txt = _describe(32, 0, (1 << 20,), 1, 4, None, 0, size)
        assert r == 0 and txt == full[:size - 1]
