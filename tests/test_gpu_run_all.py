"""The reference's run_all sweep (run_all.cpp:149-268) against the GPU build,
with the oracle as checker.

Sizes are lists of log2 dims (each >= 1) with at most 5 dims and at most 2^22
elements: first the fixed list (1-D 2^1..2^22, squares up to 2^11 x 2^11,
cubes up to 2^7 per side, run_all.cpp:251-254), then shapes drawn so that
every (ndim, total log2) class is equally likely (get_weights,
run_all.cpp:166-185) until a budget of processed elements is spent.  Every
size runs the four kinds r, ri, c, ci (run_all.cpp:57-66) on buffers shifted
by a random multiple of 4 bytes below 4 KiB (run_all.cpp:108-111 shifts by
any byte count; the GPU build needs 4-byte-aligned floats), alternating
interleaved and split formats.  Pass rule: the reference harness metric
<= 1.2e-7 * sqrt(log2 N) (run_all.cpp:68, testing.cpp:1321-1331) and our
rel-L2 bar 2*log2(N)*2^-24; one size in three is also run in FP64 at
16*2*log2(N)*2^-53."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

MAX_SIZE, MAX_DIM = 22, 5


def _all_sizes(s, left, max_len, out):
    if s:
        out.append(tuple(s))
    if len(s) < max_len:
        for i in range(1, left + 1):
            _all_sizes(s + [i], left - i, max_len, out)


def _sweep(budget_log2=26, seed=2024):
    fixed = [(i,) for i in range(1, MAX_SIZE + 1)] + [(i, i) for i in range(1, MAX_SIZE // 2 + 1)] + \
            [(i, i, i) for i in range(1, MAX_SIZE // 3 + 1)]
    allsz = []
    _all_sizes([], MAX_SIZE, MAX_DIM, allsz)
    cls = {}
    for s in allsz:
        cls[(len(s), sum(s))] = cls.get((len(s), sum(s)), 0) + 1
    w = np.array([1.0 / cls[(len(s), sum(s))] for s in allsz])
    w /= w.sum()
    rng = np.random.default_rng(seed)
    present = set(fixed)
    done = sum(4 << sum(s) for s in fixed)
    extra = []
    while done < (1 << budget_log2) + sum(4 << sum(s) for s in fixed):
        s = allsz[rng.choice(len(allsz), p=w)]
        if s not in present:
            present.add(s)
            extra.append(s)
            done += 4 << sum(s)
    return fixed, extra


FIXED, RANDOM = _sweep()


def _run(lg, idx, fp64):
    import gpu_util as G
    shape = tuple(1 << k for k in lg)
    n = 1 << sum(lg)
    rng = np.random.default_rng(1000 + idx)
    off = int(rng.integers(0, 1024))
    fmt = G.IL if idx % 2 == 0 else G.SP
    dtypes = [np.float32] + ([np.float64] if fp64 else [])
    for dt in dtypes:
        print("[gpu-case] run_all", shape, "offset", 4 * off, "bytes fmt", fmt, np.dtype(dt).name, flush=True)
        sys.stdout.flush()
        b = O.bound(n) if dt == np.float32 else 16 * 2.0 * np.log2(max(2, n)) * 2.0 ** -53
        for kind in ("r", "ri", "c", "ci"):
            if kind in ("c", "ci"):
                x = O.complex_input(idx, shape)
                ref = O.fft(x.astype(np.complex128), inverse=(kind == "ci"))
            elif kind == "r":
                x = O.uniform(idx, n).reshape(shape)
                ref = O.rfft(x.astype(np.float64))
            else:
                x = O.hermitian_input(idx, shape)
                ref = O.irfft(x.astype(np.complex128), shape)
            y = G.run_abi(kind, shape, x, fmt, dt, off)
            cast = (np.complex64 if dt == np.float32 else np.complex128) if np.iscomplexobj(y) else y.dtype
            yv = y.astype(cast)
            e = O.rel_l2(yv, ref)
            assert e <= b, (kind, shape, e, b)
            if dt == np.float32:
                h = O.harness_error(yv, ref)
                assert h <= 1.2e-7 * np.sqrt(max(1.0, np.log2(n))), (kind, shape, h)


@pytest.mark.gpu
@pytest.mark.parametrize("lg", FIXED, ids=lambda s: "x".join(str(1 << k) for k in s))
def test_run_all_fixed(lg):
    _run(lg, FIXED.index(lg), FIXED.index(lg) % 3 == 0)


@pytest.mark.gpu
@pytest.mark.parametrize("lg", RANDOM, ids=lambda s: "x".join(str(1 << k) for k in s))
def test_run_all_random(lg):
    i = RANDOM.index(lg)
    _run(lg, 100 + i, i % 3 == 0)


def test_sweep_covers_classes():
    """Host-side property of the sweep itself: the random part spans many
    (ndim, log2 total) classes, up to 5 dims, and stays within 2^22."""
    cls = {(len(s), sum(s)) for s in RANDOM}
    assert len(RANDOM) >= 20 and len(cls) >= 15
    assert max(len(s) for s in RANDOM) == MAX_DIM
    assert all(sum(s) <= MAX_SIZE for s in FIXED + RANDOM)
