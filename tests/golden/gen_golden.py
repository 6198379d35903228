"""Writes the golden fixtures of tests/golden from the real afft library.

Run in a container that has the reference sources (it builds
oracle/_ref/libafft_ref.so from /root/reference through oracle/Makefile):

    make -C oracle && python tests/golden/gen_golden.py

Inputs are never stored: they are regenerated from the portable counter-based
generator of the oracle (orc_uniform / orc_normal, seed recorded per entry).
Outputs are afft's FP32 results for the ISAs scalar, avx2 and avx512f (afft's
rounding differs per ISA, so each is kept):

  golden_small.npz   1-D c / ci / r / ri, N = 2^1 .. 2^12, interleaved and
                     split; 2-D 16x32 and 3-D 8x8x8 c / ci / r / ri.
  golden_large.npz   the headline shapes (c2c 2^16, 2^20; r2c 2^22; c2c 256^3;
                     r2c 8192^2): afft's rel-L2 against the FP64 oracle,
                     256 strided sample bins of afft's output, and afft's
                     output checksum (sum of re, sum of im, sum of |.|^2 in
                     FP64).

Input conventions (the reference harness's, testing.cpp:1064-1096):
  c / ci   re, im uniform[-1, 1), re block then im block, seed 1
  r        uniform[-1, 1) reals (seed 2) for the small set; standard normal
           (seed 5) for the large r2c shapes, as the benchmark configs state
  ri       Hermitian input built like testing.cpp:1069-1094 from seed 3
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

ISAS = {"scalar": O.AFFT_SCALAR, "avx2": O.AFFT_AVX2, "avx512f": O.AFFT_AVX512F}
FMTS = {"il": O.AFFT_INTERLEAVED, "sp": O.AFFT_SPLIT}


def small_cases():
    for lg in range(1, 13):
        yield (1 << lg,)
    yield (16, 32)
    yield (8, 8, 8)


def key(kind, shape, fmt, isa):
    return f"{kind}|{'x'.join(str(s) for s in shape)}|{fmt}|{isa}"


def run_small(R):
    out = {}
    for shape in small_cases():
        n = int(np.prod(shape))
        xc = O.complex_input(1, shape)
        xr = O.uniform(2, n).reshape(shape)
        hz = O.hermitian_input(3, shape)
        for fname, fmt in FMTS.items():
            if len(shape) > 1 and fname == "sp":
                continue
            for iname, impl in ISAS.items():
                if not R.supported(impl):
                    continue
                out[key("c", shape, fname, iname)] = R.c2c(xc, inverse=False, fmt=fmt, impl=impl)
                out[key("ci", shape, fname, iname)] = R.c2c(xc, inverse=True, fmt=fmt, impl=impl)
                out[key("r", shape, fname, iname)] = R.r2c(xr, fmt=fmt, impl=impl)
                out[key("ri", shape, fname, iname)] = R.c2r(hz, shape, fmt=fmt, impl=impl)
    return out


def summary(y: np.ndarray, ref: np.ndarray):
    flat = np.ascontiguousarray(y).ravel()
    idx = np.linspace(0, flat.size - 1, 256).astype(np.int64)
    yd = flat.astype(np.complex128) if np.iscomplexobj(flat) else flat.astype(np.float64)
    cs = np.array([np.sum(yd.real), np.sum(yd.imag) if np.iscomplexobj(yd) else 0.0, np.sum(np.abs(yd) ** 2)])
    return O.rel_l2(y, ref), idx, flat[idx], cs


def run_large(R):
    out = {}
    cases = [("c", (1 << 16,)), ("c", (1 << 20,)), ("r", (1 << 22,)), ("c", (256, 256, 256)), ("r", (8192, 8192))]
    for kind, shape in cases:
        n = int(np.prod(shape))
        for iname, impl in ISAS.items():
            if not R.supported(impl):
                continue
            if kind == "c":
                x = O.complex_input(1, shape)
                y = R.c2c(x, impl=impl)
                ref = O.fft(x.astype(np.complex128))
            else:
                x = O.normal(5, n).reshape(shape)
                y = R.r2c(x, impl=impl)
                ref = O.rfft(x.astype(np.float64))
            err, idx, samp, cs = summary(y, ref)
            k = key(kind, shape, "il", iname)
            out[k + "|err"] = np.array([err])
            out[k + "|idx"] = idx
            out[k + "|samples"] = samp
            out[k + "|checksum"] = cs
            print(k, "rel-L2 vs FP64:", err, "bound:", O.bound(n), flush=True)
    return out


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libafft_ref.so missing: run `make -C oracle` where the reference is present")
    R = O.RefAfft()
    small = run_small(R)
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **small)
    print("golden_small:", len(small), "arrays")
    if "--small-only" not in sys.argv:
        large = run_large(R)
        np.savez_compressed(os.path.join(HERE, "golden_large.npz"), **large)
        print("golden_large:", len(large), "arrays")


if __name__ == "__main__":
    main()
