"""Row passes at both ends of the range, through the batched plan API against
the FP64 oracle.

Rows of 2..32 points on the LDS-staged one-row-per-thread kernel (the
default up to 16 points, on request for 32)
(fft_amd/csrc/afft_tiny.hip) through the batched plan API, against the FP64
oracle: every kind, both formats, both precisions, batches that end inside a
work-group, every work-group size (variant tiny_rows = 64 / 128 / 256), N-D
shapes whose last axis is tiny, and the general row kernel (tiny_rows = -1) on
the same inputs.

The longest rows (FP32 32768, FP64 16384 points: c2c, and r2c / c2r of twice
that), which only the half-LDS row kernel holds in one work-group, against
the oracle and against the two-pass plans that variant half_lds = -1 keeps."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

IL, SP = 1, 0


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def bound(n, dtype):
    if dtype == np.float32:
        return O.bound(n)
    return 16 * 2.0 * np.log2(max(2, n)) * 2.0 ** -53


def _dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).cuda()


def _run(kind, shape, batch, fmt, dtype, x, variant=None):
    """One batched plan over x (batch rows in afft's layout for the kind);
    returns the output rows (complex, or real for ri) and the first pass's
    description."""
    import torch
    from fft_amd.afft import Plan
    n = int(np.prod(shape))
    m = (shape[0] // 2 + 1) * (n // shape[0])
    p = Plan(kind, shape, batch, dtype=dtype, format=fmt)
    if variant:
        p.set_variant(**variant)
    desc = p.describe()[0]
    cin = {"c": n, "ci": n, "r": n, "ri": m}[kind]
    cout = {"c": n, "ci": n, "r": m, "ri": n}[kind]
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    xs = x.reshape(batch, cin)
    if kind == "r":
        sre, sim = _dev(xs.ravel(), dtype), None
    elif fmt == IL:
        sre, sim = _dev(np.stack([xs.real, xs.imag], -1).ravel(), dtype), None
    else:
        sre, sim = _dev(xs.real.ravel(), dtype), _dev(xs.imag.ravel(), dtype)
    if kind == "ri":
        dre, dim = torch.empty(cout * batch, dtype=tdt, device="cuda"), None
    elif fmt == IL:
        dre, dim = torch.empty(2 * cout * batch, dtype=tdt, device="cuda"), None
    else:
        dre = torch.empty(cout * batch, dtype=tdt, device="cuda")
        dim = torch.empty(cout * batch, dtype=tdt, device="cuda")
    p(sre, sim, dre, dim)
    torch.cuda.synchronize()
    p.close()
    h = dre.cpu().numpy().astype(np.float64)
    if kind == "ri":
        y = h
    elif fmt == IL:
        y = h[0::2] + 1j * h[1::2]
    else:
        y = h + 1j * dim.cpu().numpy().astype(np.float64)
    return y.reshape(batch, cout), desc


def _inputs(kind, shape, batch, seed):
    n = int(np.prod(shape))
    if kind in ("c", "ci"):
        return O.complex_input(seed, (batch, n))
    if kind == "r":
        return O.normal(seed, batch * n).reshape(batch, n)
    return np.stack([O.hermitian_input(seed + b, shape).ravel() for b in range(batch)])


def _oracle(kind, shape, row):
    s = [int(d) for d in shape]
    if kind in ("c", "ci"):
        return O.fft(row.astype(np.complex128).reshape(s), inverse=kind == "ci").ravel()
    if kind == "r":
        return O.rfft(row.astype(np.float64).reshape(s)).ravel()
    return O.irfft(row.astype(np.complex128).reshape(O.half_shape(s)), s).ravel()


def _check(kind, shape, batch, fmt, dtype, seed, variant=None, expect_tiny=True, expect=None):
    n = int(np.prod(shape))
    x = _inputs(kind, shape, batch, seed)
    # the device computes on inputs rounded to its precision: so does the oracle
    xr = x.astype(np.complex64 if np.iscomplexobj(x) else np.float32) if dtype == np.float32 else x
    say("tiny", kind, shape, "batch", batch, "fmt", fmt, np.dtype(dtype).name, variant)
    y, desc = _run(kind, shape, batch, fmt, dtype, xr, variant)
    if expect is None:
        assert desc.startswith("tiny") == expect_tiny, desc
    else:
        assert expect in desc, (expect, desc)
    worst = 0.0
    for b in range(batch):
        ref = _oracle(kind, shape, xr[b])
        e = np.linalg.norm(y[b] - ref) / max(np.linalg.norm(ref), 1e-300)
        worst = max(worst, e)
    assert worst <= bound(n, dtype), (kind, shape, batch, fmt, worst)
    return y


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("kind", ["c", "ci", "r", "ri"])
def test_tiny_rows_against_oracle(kind, dtype):
    real = kind in ("r", "ri")
    for lg in range(2 if real else 1, 7 if real else 6):
        n = 1 << lg
        for fmt in (IL, SP):
            for batch in (1, 255, 1000):
                # 32-point rows default to the general kernel (faster there)
                L = n // 2 if real else n
                y = _check(kind, (n,), batch, fmt, dtype, 11 + lg, expect_tiny=L <= 16)
                if batch == 1000:
                    # every compiled work-group size (1000 rows end inside a group of each)
                    for rows in (64, 128, 256):
                        _check(kind, (n,), batch, fmt, dtype, 11 + lg, {"tiny_rows": rows})
                if batch == 255:
                    # the general row kernel on the same input: both within the bound
                    # of the oracle, and of each other
                    g = _check(kind, (n,), batch, fmt, dtype, 11 + lg, {"tiny_rows": -1}, expect_tiny=False)
                    e = np.linalg.norm(y - g) / np.linalg.norm(g)
                    assert e <= 2 * bound(n, dtype), (kind, n, fmt, e)


@pytest.mark.parametrize("shape", [(8, 8, 8), (64, 16), (4, 16), (2, 2)], ids=lambda s: "x".join(map(str, s)))
def test_tiny_last_axis_of_nd(shape):
    for kind in ("c", "ci"):
        for fmt in (IL, SP):
            _check(kind, shape, 3, fmt, np.float32, 21)


def test_tiny_c2c_in_place():
    """src == dst: each work-group reads all its rows before it writes them."""
    import torch
    from fft_amd.afft import Plan
    n, batch = 16, 777
    x = O.complex_input(31, (batch, n)).astype(np.complex64)
    for fmt in (IL, SP):
        p = Plan("c", (n,), batch, dtype=np.float32, format=fmt)
        assert p.describe()[0].startswith("tiny")
        say("tiny in place", fmt)
        if fmt == IL:
            re_ = _dev(np.stack([x.real, x.imag], -1).ravel(), np.float32)
            p(re_, None, re_, None)
            h = re_.cpu().numpy().astype(np.float64)
            y = (h[0::2] + 1j * h[1::2]).reshape(batch, n)
        else:
            re_, im_ = _dev(x.real.ravel(), np.float32), _dev(x.imag.ravel(), np.float32)
            p(re_, im_, re_, im_)
            y = (re_.cpu().numpy().astype(np.float64) + 1j * im_.cpu().numpy()).reshape(batch, n)
        torch.cuda.synchronize()
        p.close()
        for b in range(batch):
            ref = O.fft(x[b].astype(np.complex128))
            assert np.linalg.norm(y[b] - ref) / np.linalg.norm(ref) <= O.bound(n), (fmt, b)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("kind", ["c", "ci", "r", "ri"])
def test_longest_rows_against_oracle(kind, dtype):
    """One half-LDS row pass per transform, batch 3, both formats; the same
    inputs on the two-pass plan (half_lds = -1) agree within the bound."""
    L = 32768 if dtype == np.float32 else 16384
    n = 2 * L if kind in ("r", "ri") else L
    # FP32 r2c 2^16 defaults to the real-input four-step: real_pair = -1 asks
    # for the row pass
    v = {"real_pair": -1} if (dtype == np.float32 and kind == "r") else None
    for fmt in (IL, SP):
        y = _check(kind, (n,), 3, fmt, dtype, 41, v, expect=f"L={L} CT=1 hx row")
        g = _check(kind, (n,), 3, fmt, dtype, 41, {"half_lds": -1}, expect="")
        e = np.linalg.norm(y - g) / np.linalg.norm(g)
        assert e <= 2 * bound(n, dtype), (kind, n, fmt, e)
