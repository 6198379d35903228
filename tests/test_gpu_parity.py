"""GPU parity: the HIP path through the C ABI against the FP64 oracle and
afft's own output (golden fixtures written from the reference build).

Bar: rel-L2 <= 2*log2(N)*2^-24 against the FP64 oracle (N = total element
count) and against afft's FP32 output on the same seeded input; FP64 plans
must sit within 2*log2(N)*2^-53 * 16 of the oracle."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    import gpu_util
    gpu_util.torch_cuda()
    return gpu_util


def bound64(n):
    return 16 * 2.0 * np.log2(max(2, n)) * 2.0 ** -53


@pytest.mark.parametrize("lg", list(range(0, 23)))
def test_c2c_1d(G, lg):
    n = 1 << lg
    x = O.complex_input(1, (n,))
    ref_f = O.fft(x.astype(np.complex128))
    ref_i = O.fft(x.astype(np.complex128), inverse=True)
    for fmt in (G.IL, G.SP):
        say("c2c 1d", n, "fmt", fmt)
        y = G.gpu_c2c(x, False, fmt)
        assert O.rel_l2(y.astype(np.complex64), ref_f) <= O.bound(n), (n, fmt)
        yi = G.gpu_c2c(x, True, fmt)
        assert O.rel_l2(yi.astype(np.complex64), ref_i) <= O.bound(n), (n, fmt)


def test_c2c_against_afft_golden(G):
    g = np.load(os.path.join(GOLD, "golden_small.npz"), allow_pickle=False)
    checked = 0
    for k in g.files:
        kind, shape, fmt, isa = k.split("|")
        shape = tuple(int(s) for s in shape.split("x"))
        n = int(np.prod(shape))
        f = G.IL if fmt == "il" else G.SP
        say("golden", k)
        if kind in ("c", "ci"):
            y = G.gpu_c2c(O.complex_input(1, shape), kind == "ci", f)
        elif kind == "r":
            y = G.gpu_r2c(O.uniform(2, n).reshape(shape), f)
        else:
            y = G.gpu_c2r(O.hermitian_input(3, shape), shape, f)
        gold = g[k]
        ref = gold.astype(np.complex128) if np.iscomplexobj(gold) else gold.astype(np.float64)
        y32 = y.astype(np.complex64) if np.iscomplexobj(y) else y.astype(np.float32)
        assert O.rel_l2(y32, ref) <= O.bound(n), (k, O.rel_l2(y32, ref))
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("lg", [1, 2, 3, 4, 5, 6, 8, 10, 12, 13, 14, 15, 16, 17, 20, 22])
def test_real_1d(G, lg):
    n = 1 << lg
    x = O.normal(5, n)
    ref = O.rfft(x.astype(np.float64))
    h = O.hermitian_input(3, (n,))
    refi = O.irfft(h.astype(np.complex128), (n,))
    for fmt in (G.IL, G.SP):
        say("real 1d", n, "fmt", fmt)
        z = G.gpu_r2c(x, fmt)
        assert O.rel_l2(z.astype(np.complex64), ref) <= O.bound(n), (n, fmt)
        w = G.gpu_c2r(h, (n,), fmt)
        assert O.rel_l2(w.astype(np.float32), refi) <= O.bound(n), (n, fmt)


ND_SHAPES = [(2, 4), (4, 1, 8), (16, 32), (32, 16), (8, 8, 8), (2, 2, 2, 2), (8192, 4), (4, 32768), (64, 64, 64),
             (2, 8192), (32, 16384),
             # long strided axes and long real dim 0 (column four-step, real pass over row pairs)
             (16384, 8), (4096, 4, 4), (8192, 64),
             # pass B of the dim-0 column four-step with the real pass (64-, 128-
             # and 256-point columns)
             (16384, 64), (65536, 16), (131072, 16),
             # long real dim 0 with 1024..8192-point rows: the real pass inside the
             # column pass B (r2c) or the inverse row FFTs (c2r)
             (8192, 1024), (8192, 2048)]


@pytest.mark.parametrize("shape", ND_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_nd(G, shape):
    n = int(np.prod(shape))
    x = O.complex_input(1, shape)
    xr = O.normal(5, n).reshape(shape)
    h = O.hermitian_input(3, shape)
    ref = O.fft(x.astype(np.complex128))
    refi = O.fft(x.astype(np.complex128), inverse=True)
    refr = O.rfft(xr.astype(np.float64))
    refri = O.irfft(h.astype(np.complex128), shape)
    for fmt in (G.IL, G.SP):
        say("nd", shape, "fmt", fmt)
        assert O.rel_l2(G.gpu_c2c(x, False, fmt).astype(np.complex64), ref) <= O.bound(n)
        assert O.rel_l2(G.gpu_c2c(x, True, fmt).astype(np.complex64), refi) <= O.bound(n)
        assert O.rel_l2(G.gpu_r2c(xr, fmt).astype(np.complex64), refr) <= O.bound(n)
        assert O.rel_l2(G.gpu_c2r(h, shape, fmt).astype(np.float32), refri) <= O.bound(n)


@pytest.mark.parametrize("shape", [(8192, 64), (8192, 1024), (16384, 2048), (8192, 4, 64)],
                         ids=lambda s: "x".join(map(str, s)))
def test_nd_r2c_real_pass_forms(G, shape):
    """N-D r2c with a long dim 0: the real pass over row pairs inside the
    column pass B (default), inside the row FFTs (real_pair = 1, 2-D with
    1024..8192-point rows) and on its own (-1), each against the oracle, in
    both formats and a batch of 2."""
    import torch
    from fft_amd.afft import Plan
    n = int(np.prod(shape))
    m = int(np.prod(G.half(shape)))
    xr = O.normal(9, 2 * n).reshape((2,) + shape)
    ref = [O.rfft(xr[b].astype(np.float64)).ravel() for b in range(2)]
    src = G.to_dev(xr.ravel())
    seen = set()
    for fmt in (G.IL, G.SP):
        p = Plan("r", shape, 2, format=fmt)
        for rp in (0, 1, -1):
            p.set_variant(real_pair=rp)
            desc = tuple(p.describe())
            seen.add(desc)
            say("nd r2c real pass", shape, fmt, rp, desc)
            if fmt == G.IL:
                dst = G.empty_dev(4 * m)
                p(src, None, dst, None)
                torch.cuda.synchronize()
                y = G.to_host(dst)
                y = (y[0::2] + 1j * y[1::2]).reshape(2, m)
            else:
                dre, dim_ = G.empty_dev(2 * m), G.empty_dev(2 * m)
                p(src, None, dre, dim_)
                torch.cuda.synchronize()
                y = (G.to_host(dre) + 1j * G.to_host(dim_)).reshape(2, m)
            for b in range(2):
                assert O.rel_l2(y[b].astype(np.complex64), ref[b]) <= O.bound(n), (shape, fmt, rp, b)
        p.close()
    assert any("real_colpair" in d for s_ in seen for d in s_), seen
    assert any("real_rows" in d for s_ in seen for d in s_), seen


@pytest.mark.parametrize("shape", [(8192, 64), (8192, 1024), (16384, 2048), (8192, 4, 64)],
                         ids=lambda s: "x".join(map(str, s)))
def test_nd_c2r_real_pass_forms(G, shape):
    """N-D c2r with a long dim 0: the inverse real pass over row pairs inside
    the column pass A (default), inside the inverse row FFTs (real_pair = 1,
    2-D with 1024..8192-point rows) and on its own (-1), each against the
    oracle, in both formats and a batch of 2."""
    import torch
    from fft_amd.afft import Plan
    n = int(np.prod(shape))
    hs = G.half(shape)
    m = int(np.prod(hs))
    z = np.stack([O.hermitian_input(11 + b, shape) for b in range(2)])
    ref = [O.irfft(z[b].astype(np.complex128), shape).ravel() for b in range(2)]
    zf = z.reshape(2, m)
    seen = set()
    for fmt in (G.IL, G.SP):
        p = Plan("ri", shape, 2, format=fmt)
        for rp in (0, 1, -1):
            p.set_variant(real_pair=rp)
            desc = tuple(p.describe())
            seen.add(desc)
            say("nd c2r real pass", shape, fmt, rp, desc)
            dst = G.empty_dev(2 * n)
            if fmt == G.IL:
                src = G.to_dev(np.stack([zf.real, zf.imag], -1).astype(np.float32).ravel())
                p(src, None, dst, None)
            else:
                p(G.to_dev(zf.real.astype(np.float32).ravel()), G.to_dev(zf.imag.astype(np.float32).ravel()), dst,
                  None)
            torch.cuda.synchronize()
            y = G.to_host(dst).reshape(2, n)
            for b in range(2):
                assert O.rel_l2(y[b].astype(np.float32), ref[b]) <= O.bound(n), (shape, fmt, rp, b)
        p.close()
    assert any("real_icolpair" in d for s_ in seen for d in s_), seen
    assert any("real_rows" in d for s_ in seen for d in s_), seen


@pytest.mark.parametrize("shape", [(256, 8), (128, 16), (256, 64, 4), (128, 512), (256, 256, 2), (512, 8),
                                   (1024, 64), (512, 32, 4)],
                         ids=lambda s: "x".join(map(str, s)))
def test_nd_real_short_dim0_column_tiles(G, shape):
    """N-D r2c / c2r with a dim 0 of 128 to 1024 reals: its 64- to 512-point
    column pass with the real post-pass (r2c) or pre-pass (c2r) on the default
    tile (64 columns where compiled) and on forced 16- and 64-column tiles, in
    both formats and a batch of 2 (shapes with fewer columns than a tile
    included), each against the oracle."""
    import torch
    from fft_amd.afft import Plan
    n = int(np.prod(shape))
    hs = G.half(shape)
    m = int(np.prod(hs))
    xr = O.normal(13, 2 * n).reshape((2,) + shape)
    refr = [O.rfft(xr[b].astype(np.float64)).ravel() for b in range(2)]
    z = np.stack([O.hermitian_input(17 + b, shape) for b in range(2)])
    refi = [O.irfft(z[b].astype(np.complex128), shape).ravel() for b in range(2)]
    zf = z.reshape(2, m)
    src_r = G.to_dev(xr.ravel())
    for fmt in (G.IL, G.SP):
        pr = Plan("r", shape, 2, format=fmt)
        pi = Plan("ri", shape, 2, format=fmt)
        for ct in (0, 16, 64):
            pr.set_variant(col_tile=ct)
            pi.set_variant(col_tile=ct)
            say("nd real short dim0", shape, fmt, ct, pr.describe()[0], pi.describe()[-1])
            if fmt == G.IL:
                dst = G.empty_dev(4 * m)
                pr(src_r, None, dst, None)
                torch.cuda.synchronize()
                y = G.to_host(dst)
                y = (y[0::2] + 1j * y[1::2]).reshape(2, m)
            else:
                dre, dim_ = G.empty_dev(2 * m), G.empty_dev(2 * m)
                pr(src_r, None, dre, dim_)
                torch.cuda.synchronize()
                y = (G.to_host(dre) + 1j * G.to_host(dim_)).reshape(2, m)
            dst = G.empty_dev(2 * n)
            if fmt == G.IL:
                pi(G.to_dev(np.stack([zf.real, zf.imag], -1).astype(np.float32).ravel()), None, dst, None)
            else:
                pi(G.to_dev(zf.real.astype(np.float32).ravel()), G.to_dev(zf.imag.astype(np.float32).ravel()), dst,
                   None)
            torch.cuda.synchronize()
            w = G.to_host(dst).reshape(2, n)
            for b in range(2):
                assert O.rel_l2(y[b].astype(np.complex64), refr[b]) <= O.bound(n), (shape, fmt, ct, b)
                assert O.rel_l2(w[b].astype(np.float32), refi[b]) <= O.bound(n), (shape, fmt, ct, b)
        pr.close()
        pi.close()


@pytest.mark.parametrize("shape", [(16,), (4096,), (1 << 14,), (1 << 15,), (16, 32), (8, 8, 8), (1 << 18,), (8192, 64),
                                   (16384, 64), (65536, 16), (2, 16384)],
                         ids=lambda s: "x".join(map(str, s)))
def test_fp64_path(G, shape):
    n = int(np.prod(shape))
    x = O.complex_input(1, shape).astype(np.complex128)
    xr = O.normal(5, n).reshape(shape).astype(np.float64)
    h = O.hermitian_input(3, shape).astype(np.complex128)
    say("fp64", shape)
    y = G.gpu_c2c(x, False, G.IL, np.float64)
    assert O.rel_l2(y, O.fft(x)) <= bound64(n)
    yi = G.gpu_c2c(x, True, G.SP, np.float64)
    assert O.rel_l2(yi, O.fft(x, inverse=True)) <= bound64(n)
    z = G.gpu_r2c(xr, G.IL, np.float64)
    assert O.rel_l2(z, O.rfft(xr)) <= bound64(n)
    w = G.gpu_c2r(h, shape, G.IL, np.float64)
    assert O.rel_l2(w, O.irfft(h, shape)) <= bound64(n)


@pytest.mark.parametrize("kind", ["c", "ci"])
def test_fp64_three_pass(G, kind):
    """FP64 1-D c2c of 2^23 points (three-pass four-step) against the FP64
    oracle, both formats."""
    n = 1 << 23
    x = O.complex_input(71, (n,)).astype(np.complex128)
    ref = O.fft(x, inverse=kind == "ci")
    for fmt in (G.IL, G.SP):
        say("fp64 three-pass", kind, fmt)
        y = G.gpu_c2c(x, kind == "ci", fmt, np.float64)
        assert O.rel_l2(y, ref) <= bound64(n), (kind, fmt, O.rel_l2(y, ref))


def test_roundtrip_scales_by_n(G):
    for n in (1 << 12, 1 << 20):
        say("roundtrip", n)
        x = O.complex_input(4, (n,))
        y = G.gpu_c2c(G.gpu_c2c(x, False).astype(np.complex64), True)
        assert O.rel_l2(y.astype(np.complex64), n * x.astype(np.complex128)) <= 2 * O.bound(n)
        xr = O.normal(6, n)
        w = G.gpu_c2r(G.gpu_r2c(xr).astype(np.complex64), (n,))
        assert O.rel_l2(w.astype(np.float32), n * xr.astype(np.float64)) <= 2 * O.bound(n)


def test_batched_plan_and_chunk_invariance(G):
    import torch
    from fft_amd.afft import Plan
    n, batch = 1 << 20, 6
    say("batched plan", n, batch)
    x = O.complex_input(7, (batch, n))
    src = G.to_dev(np.stack([x.real, x.imag], -1).astype(np.float32).ravel())
    dst = G.empty_dev(2 * n * batch)
    p = Plan("c", (n,), batch)
    p(src, None, dst, None)
    torch.cuda.synchronize()
    y = G.to_host(dst)
    y = (y[0::2] + 1j * y[1::2]).reshape(batch, n)
    for b in (0, batch - 1):
        assert O.rel_l2(y[b].astype(np.complex64), O.fft(x[b].astype(np.complex128))) <= O.bound(n)
    for chunk in (1, 4):
        p.set_chunk(chunk)
        dst2 = G.empty_dev(2 * n * batch)
        p(src, None, dst2, None)
        torch.cuda.synchronize()
        assert torch.equal(dst, dst2), chunk
    p.close()


def test_headline_64x2p20_against_device_fp64(G):
    """The benchmark shape at full size: every transform of the batch against
    the device FP64 plan (itself pinned to the oracle above), with the FP64
    error reduction on the device; plus the forward/inverse round trip."""
    import torch
    from fft_amd.afft import Plan
    n, batch = 1 << 20, 64
    say("headline 64 x 2^20")
    g = torch.Generator(device="cuda").manual_seed(1234)
    x32 = torch.rand(batch * n * 2, device="cuda", generator=g, dtype=torch.float32) * 2 - 1
    x64 = x32.double()
    y32 = torch.empty_like(x32)
    y64 = torch.empty_like(x64)
    Plan("c", (n,), batch)(x32, None, y32, None)
    Plan("c", (n,), batch, dtype=np.float64)(x64, None, y64, None)
    torch.cuda.synchronize()
    err = G.device_rel_l2(y32, y64)
    assert err <= O.bound(n), err
    back = torch.empty_like(x32)
    Plan("ci", (n,), batch)(y32, None, back, None)
    torch.cuda.synchronize()
    err_rt = G.device_rel_l2(back, x64 * n)
    assert err_rt <= 2 * O.bound(n), err_rt


def test_large_nd_against_afft_summaries(G):
    """256^3 c2c and 8192^2 r2c at full size: afft's sampled output (golden)
    and the device FP64 plan."""
    import torch
    from fft_amd.afft import Plan
    g = np.load(os.path.join(GOLD, "golden_large.npz"), allow_pickle=False)
    for kind, shape in (("c", (256, 256, 256)), ("r", (8192, 8192)), ("c", (1 << 16,)), ("r", (1 << 22,))):
        say("large summary", kind, shape)
        n = int(np.prod(shape))
        k = f"{kind}|{'x'.join(map(str, shape))}|il|avx2"
        if kind == "c":
            x = O.complex_input(1, shape)
            src = G.to_dev(np.stack([x.real, x.imag], -1).astype(np.float32).ravel())
            m = n
        else:
            src = G.to_dev(O.normal(5, n))
            m = int(np.prod(G.half(shape)))
        dst = G.empty_dev(2 * m)
        dst64 = G.empty_dev(2 * m, np.float64)
        Plan(kind, shape, dtype=np.float64)(src.double(), None, dst64, None)
        if kind == "r":
            # the opt-in fused real passes (real pass inside the 8192-point
            # rows; 1-D pass B with the real post-pass) and the real pass on
            # its own, at full size
            for rp in (1, -1):
                pf = Plan(kind, shape)
                pf.set_variant(real_pair=rp)
                pf(src, None, dst, None)
                torch.cuda.synchronize()
                say("  real_pair", rp, pf.describe())
                assert G.device_rel_l2(dst, dst64) <= O.bound(n)
                pf.close()
        Plan(kind, shape)(src, None, dst, None)
        torch.cuda.synchronize()
        assert G.device_rel_l2(dst, dst64) <= O.bound(n)
        y = G.to_host(dst)
        yc = y[0::2] + 1j * y[1::2]
        idx = g[k + "|idx"]
        gold = g[k + "|samples"].astype(np.complex128)
        scale = np.sqrt(g[k + "|checksum"][2] / m)  # rms of afft's output
        assert np.abs(yc[idx] - gold).max() <= 8 * O.bound(n) * scale * np.sqrt(np.log2(n)), k


def test_errors_leave_state_clean(G):
    from fft_amd import _lib
    import ctypes
    L = _lib.lib()
    nd, dims = _lib.dims_array([24])
    mem = ctypes.create_string_buffer(64)
    assert not L.afft_32_c_create(nd, dims, 1, mem, 0)
    assert "power of two" in _lib.last_error()


VARIANTS = [dict(sched=s, half_lds=h, rowt_tile=r, col_tile=c, chunk=ch, real_pair=rp)
            for s in (0, -1) for h in (0, -1) for r in (0, 16) for c in (0, 8, 32) for ch in (1, 0)
            for rp in (1, -1)]


def _variant_case(G, kind, n, batch):
    """Seeded input on the device, the FP64 oracle's outputs and a function
    that checks one plan's device output against them."""
    if kind == "r":
        x = O.normal(31, batch * n).reshape(batch, n).astype(np.float32)
        ref = [O.rfft(x[b].astype(np.float64)) for b in range(batch)]
        src = G.to_dev(x.ravel())
        nout = 2 * (n // 2 + 1) * batch
    elif kind == "ri":
        z = np.stack([O.hermitian_input(32 + b, (n,)) for b in range(batch)])
        ref = [O.irfft(z[b].astype(np.complex128), (n,)) for b in range(batch)]
        src = G.to_dev(np.stack([z.real, z.imag], -1).astype(np.float32).ravel())
        nout = n * batch
    else:
        x = O.complex_input(33, (batch, n))
        ref = [O.fft(x[b].astype(np.complex128), inverse=(kind == "ci")) for b in range(batch)]
        src = G.to_dev(np.stack([x.real, x.imag], -1).astype(np.float32).ravel())
        nout = 2 * n * batch

    def check(p, v):
        import torch
        dst = G.empty_dev(nout)
        p(src, None, dst, None)
        torch.cuda.synchronize()
        y = G.to_host(dst)
        if kind == "ri":
            y = y.reshape(batch, n)
        elif kind == "r":
            y = (y[0::2] + 1j * y[1::2]).reshape(batch, n // 2 + 1)
        else:
            y = (y[0::2] + 1j * y[1::2]).reshape(batch, n)
        for b in range(batch):
            yb = y[b].astype(np.float32 if kind == "ri" else np.complex64)
            assert O.rel_l2(yb, ref[b]) <= O.bound(n), (kind, n, v, b)

    return check


@pytest.mark.parametrize("kind,n,batch", [("c", 1 << 20, 3), ("ci", 1 << 20, 2), ("c", 1 << 22, 1),
                                          ("c", 1 << 17, 2), ("c", 1 << 14, 3), ("ci", 1 << 13, 3),
                                          ("r", 1 << 14, 2), ("ri", 1 << 14, 2), ("r", 1 << 21, 3),
                                          ("r", 1 << 17, 3), ("ri", 1 << 21, 2)])
def test_kernel_variants_against_oracle(G, kind, n, batch):
    """Every kernel-shape variant (radix-32 schedules, half-LDS exchange,
    narrow tiles) against the FP64 oracle; the variants only change speed."""
    from fft_amd.afft import Plan
    say("variants", kind, n, batch)
    check = _variant_case(G, kind, n, batch)
    p = Plan(kind, (n,), batch)
    seen = set()
    for v in VARIANTS:
        p.set_variant(**v)
        desc = tuple(p.describe())
        if desc in seen:
            continue
        seen.add(desc)
        say("  variant", v, desc)
        check(p, v)
    if kind in ("c", "ci") and n > (1 << 14):
        assert len(seen) >= 2, seen  # four-step plans do have alternatives
    if kind == "r" and n >= (1 << 16):
        # the fused real pass and the row pass + transposing real pass
        assert any("real_pair" in d for s in seen for d in s), seen
        assert any("real_tr" in d for s in seen for d in s), seen
    p.close()


@pytest.mark.parametrize("kind,n,batch", [("c", 1 << 20, 2), ("ci", 1 << 20, 1), ("r", 1 << 20, 2),
                                          ("ri", 1 << 20, 1), ("c", 1 << 19, 2), ("c", 1 << 18, 2)])
def test_row_split_variants_against_oracle(G, kind, n, batch):
    """Other splits n = n1 * n2 of the 1-D four-step (variant row_log2) and
    the wide 256- / 512-point pass-A tiles (32 and 64 columns, half LDS)
    against the FP64 oracle."""
    from fft_amd.afft import Plan
    say("row split", kind, n, batch)
    check = _variant_case(G, kind, n, batch)
    p = Plan(kind, (n,), batch)
    seen = set()
    for rl in (0, 8, 9, 11, 12):
        for ct in (0, 16, 32, 64):
            try:
                p.set_variant(row_log2=rl, col_tile=ct)
            except RuntimeError:
                continue  # no kernel of that tile for this split
            desc = tuple(p.describe())
            if desc in seen:
                continue
            seen.add(desc)
            say("  variant", rl, ct, desc)
            check(p, (rl, ct))
    assert len(seen) >= 4, seen
    p.close()


@pytest.mark.parametrize("kind,n,batch,chunk,split", [("c", 1 << 20, 8, 8, 2), ("ci", 1 << 20, 12, 6, 3),
                                                      ("c", 1 << 16, 64, 32, 4), ("r", 1 << 20, 4, 4, 2),
                                                      ("r", 1 << 22, 8, 8, 4)])
def test_split_pass_a_matches_one_launch(G, kind, n, batch, chunk, split):
    """Pass A of each group in several launches (variant a_split) computes
    exactly what one launch does: same kernels, same scratch layout."""
    import torch
    from fft_amd.afft import Plan
    say("a_split", kind, n, batch, chunk, split)
    if kind == "r":
        src = G.to_dev(O.normal(61, batch * n))
        nout = 2 * (n // 2 + 1) * batch
    else:
        x = O.complex_input(61, (batch, n))
        src = G.to_dev(np.stack([x.real, x.imag], -1).astype(np.float32).ravel())
        nout = 2 * n * batch
    p = Plan(kind, (n,), batch)
    # the 1-D r2c on the complex four-step (the real-input four-step has no
    # split form)
    rp = {"real_pair": -1} if kind == "r" else {}
    p.set_variant(chunk=chunk, **rp)
    ref = G.empty_dev(nout)
    p(src, None, ref, None)
    p.set_variant(chunk=chunk, a_split=split, **rp)
    assert p.launches(0) == (batch // chunk) * split
    out = G.empty_dev(nout)
    p(src, None, out, None)
    torch.cuda.synchronize()
    assert torch.equal(ref, out)
    p.close()


def test_grouped_plan_graph_replay(G):
    """A four-step plan over several groups (one transform each), captured
    whole in a HIP graph and replayed with eager work and synchronisations
    between the replays, equals the eager execute bit for bit; the removed
    pipelined schedule is refused with a message."""
    import torch
    from fft_amd.afft import Plan
    n, batch = 1 << 20, 6
    say("grouped plan + graph", n, batch)
    x = O.complex_input(41, (batch, n))
    src = G.to_dev(np.stack([x.real, x.imag], -1).astype(np.float32).ravel())
    ref = G.empty_dev(2 * n * batch)
    p = Plan("c", (n,), batch)
    p.set_variant(chunk=1)
    p(src, None, ref, None)
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="pipeline"):
        p.set_variant(chunk=1, pipeline=1)
    p.set_variant(chunk=1)
    dst = G.empty_dev(2 * n * batch)
    p(src, None, dst, None)
    torch.cuda.synchronize()
    assert torch.equal(dst, ref)
    y = G.to_host(dst)
    y = (y[0::2] + 1j * y[1::2]).reshape(batch, n)
    assert O.rel_l2(y[batch - 1].astype(np.complex64), O.fft(x[batch - 1].astype(np.complex128))) <= O.bound(n)
    out = G.empty_dev(2 * n * batch)
    s2 = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s2):
        p(src, None, out, None, s2)  # warm-up on the capture stream
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s2):
            p(src, None, out, None, s2)
    torch.cuda.synchronize()
    for rep in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
    p.close()


def test_autotune_keeps_a_correct_variant_and_caches_it(G, tmp_path, monkeypatch):
    """afft_gpu_plan_tune times the candidate variants on the caller's
    buffers, keeps the fastest (which still computes the transform to the
    bar) and records it in AFFT_GPU_TUNE_CACHE; a new plan of the same shape
    starts from the recorded variant."""
    import torch
    from fft_amd.afft import Plan
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("AFFT_GPU_TUNE_CACHE", str(cache))
    for kind, shape, batch in (("c", (1 << 20,), 4), ("r", (8192, 1024), 1)):
        say("autotune", kind, shape, batch)
        n = int(np.prod(shape))
        if kind == "c":
            x = O.complex_input(71, (batch, n))
            src = G.to_dev(np.stack([x.real, x.imag], -1).astype(np.float32).ravel())
            ref = [O.fft(x[b].astype(np.complex128)).ravel() for b in range(batch)]
            m = n
        else:
            xr = O.normal(72, batch * n).reshape((batch,) + shape)
            src = G.to_dev(xr.ravel())
            ref = [O.rfft(xr[b].astype(np.float64)).ravel() for b in range(batch)]
            m = int(np.prod(G.half(shape)))
        p = Plan(kind, shape, batch)
        dst = G.empty_dev(2 * m * batch)
        best = p.tune(src, None, dst, None, iters=2)
        assert best > 0
        chosen = p.get_variant().as_dict()
        say("  best", best, chosen, p.describe())
        p(src, None, dst, None)
        torch.cuda.synchronize()
        y = G.to_host(dst)
        y = (y[0::2] + 1j * y[1::2]).reshape(batch, m)
        for b in range(batch):
            assert O.rel_l2(y[b].astype(np.complex64), ref[b]) <= O.bound(n), (kind, shape, chosen, b)
        q = Plan(kind, shape, batch)
        assert q.get_variant().as_dict() == chosen
        q.close()
        p.close()
    lines = [ln for ln in cache.read_text().splitlines() if ln.strip()]
    assert len(lines) >= 2, lines
