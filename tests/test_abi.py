"""CPU tests of the C ABI library: it loads, exports every symbol the headers
declare, keeps afft's memsize / create conventions, and every plan's pass list
(built host-only by afft_gpu_plan_inspect) computes the right transform when
executed by the numpy plan model with all accesses in bounds.  No compute call
reaches a GPU here."""
import ctypes
import os
import re

import numpy as np
import pytest

import fft_amd
from fft_amd import _lib
from oracle import oracle as O

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_symbols():
    names = set()
    for h in ("fft.h", "afft_gpu.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(afft_\w+)\s*\(", txt):
            names.add(m.group(1))
        for m in re.finditer(r"extern\s+const\s+size_t\s+(afft_\w+)", txt):
            names.add(m.group(1))
    return names


def exported_symbols():
    out = os.popen(f"nm -D --defined-only {_lib.LIB_PATH}").read()
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_library_loads_and_exports_every_declared_symbol():
    L = fft_amd.lib()
    assert L is not None
    decl = declared_symbols()
    assert len(decl) >= 40
    missing = decl - exported_symbols()
    assert not missing, missing
    for s in _lib.FFT_H_SYMBOLS + _lib.AFFT_GPU_H_SYMBOLS:
        assert s in decl, s


def test_constants_and_align():
    L = fft_amd.lib()
    assert ctypes.c_size_t.in_dll(L, "afft_alignment").value == 64
    # afft_32_align_size rounds up to 64 bytes of floats (fft.cpp:219)
    assert [L.afft_32_align_size(n) for n in (0, 1, 16, 17, 33)] == [0, 16, 16, 32, 48]
    assert [L.afft_64_align_size(n) for n in (0, 1, 8, 9)] == [0, 8, 8, 16]
    # CPU ISA ids are not provided by this library (afft: unsupported -> 0)
    for impl in (1, 2, 3, 4, 5):
        assert L.afft_supported(impl) == 0


def test_memsize_conventions():
    L = fft_amd.lib()
    nd, dims = _lib.dims_array([4096])
    for p in ("32", "64"):
        for k in ("c", "ci", "r", "ri"):
            ms = getattr(L, f"afft_{p}_{k}_memsize")
            assert ms(nd, dims, 1, 0) > 0
            assert ms(nd, dims, 0, 16) > 0  # afft_gpu
            assert ms(nd, dims, 7, 0) == ctypes.c_size_t(-1).value  # invalid format
            assert ms(nd, dims, 1, 3) == 0  # afft_avx2: not provided here
            nd3, d3 = _lib.dims_array([12])
            assert ms(nd3, d3, 1, 0) == 0  # not a power of two
    # memsize builds the plan host-only: 0 exactly where create would fail
    # (a strided axis beyond two passes), > 0 for a 1-D length beyond 2^24
    # (three passes)
    nd4, d4 = _lib.dims_array([1 << 26, 2])
    assert L.afft_32_c_memsize(nd4, d4, 1, 0) == 0
    mem = ctypes.create_string_buffer(4096)
    assert not L.afft_32_c_create(nd4, d4, 1, mem, 0)
    nd5, d5 = _lib.dims_array([1 << 25])
    for k in ("c", "ci", "r", "ri"):
        assert getattr(L, f"afft_32_{k}_memsize")(nd5, d5, 1, 0) > 0


def test_create_rejects_without_crashing():
    L = fft_amd.lib()
    nd, dims = _lib.dims_array([8])
    mem = ctypes.create_string_buffer(256)
    assert not L.afft_32_c_create(nd, dims, 5, mem, 0)  # invalid format
    assert not L.afft_32_c_create(nd, dims, 1, mem, 4)  # CPU ISA
    # a null state is a no-op for destroy
    L.afft_32_c_destroy(None)


# ---- plan model -----------------------------------------------------------------

import plan_model as M  # noqa: E402

IL, SP = _lib.afft_interleaved, _lib.afft_split


def _pack(x, fmt):
    x = np.asarray(x).reshape(-1)
    if fmt == IL:
        return np.stack([x.real, x.imag], -1).ravel(), None
    return x.real.copy(), x.imag.copy()


def _unpack(re_, im_, fmt):
    return re_[0::2] + 1j * re_[1::2] if fmt == IL else re_ + 1j * im_


def model_check(kind, dims, fmt, batch, prec=32, variant=None):
    s = O.squeeze_shape(dims)
    n = int(np.prod(s))
    if kind in ("c", "ci"):
        x = O.complex_input(1, (batch, n)).astype(np.complex128)
        re_, im_ = _pack(x, fmt)
        dre, dim_, info, passes = M.run(kind, dims, fmt, batch, re_, im_, prec, variant=variant)
        y = _unpack(dre, dim_, fmt).reshape(batch, n)
        ref = np.stack([O.fft(x[b].reshape(s), inverse=(kind == "ci")).ravel() for b in range(batch)])
    elif kind == "r":
        x = O.normal(2, batch * n).astype(np.float64).reshape(batch, n)
        dre, dim_, info, passes = M.run(kind, dims, fmt, batch, x, None, prec, variant=variant)
        y = _unpack(dre, dim_, fmt).reshape(batch, -1)
        ref = np.stack([O.rfft(x[b].reshape(s)).ravel() for b in range(batch)])
    else:
        hs = O.half_shape(s)
        z = np.stack([O.hermitian_input(3 + b, s).astype(np.complex128).ravel() for b in range(batch)])
        re_, im_ = _pack(z, fmt)
        dre, dim_, info, passes = M.run(kind, dims, fmt, batch, re_, im_, prec, variant=variant)
        y = dre.reshape(batch, n)
        ref = np.stack([O.irfft(z[b].reshape(hs), s).ravel() for b in range(batch)])
    return np.abs(y - ref).max() / max(1e-300, np.abs(ref).max()), info, passes


SHAPES = [(1,), (2,), (4,), (8,), (32,), (64,), (4096,), (16384,), (32768,), (1 << 17,), (16, 32), (8, 8, 8),
          (2, 4), (4, 1, 8), (8192, 4), (2, 2, 2, 2), (32, 16384), (4, 32768),
          # long strided axes / long real dim 0: column four-step (+ real pass)
          (8192, 8), (16384, 2), (4096, 2, 2),
          # short real dim 0 (64- / 128-point real column passes on 64-column tiles)
          (256, 8), (128, 16), (512, 8), (1024, 8)]


@pytest.mark.parametrize("kind", ["c", "ci", "r", "ri"])
@pytest.mark.parametrize("dims", SHAPES, ids=lambda d: "x".join(map(str, d)))
def test_plan_model_matches_oracle(kind, dims):
    for fmt in (IL, SP):
        for batch in (1, 3):
            err, info, passes = model_check(kind, dims, fmt, batch)
            assert err < 1e-9, (kind, dims, fmt, batch, err)


@pytest.mark.parametrize("kind", ["c", "ci", "r", "ri"])
def test_plan_model_tiny_rows(kind):
    # rows of 2..16 points (c2c N <= 16, r2c / c2r N <= 32, last axes of N-D
    # shapes) run on the LDS-staged one-row-per-thread kernel (type 13), and
    # 32-point rows on request (variant tiny_rows > 0); variant tiny_rows = -1
    # keeps the general row kernel (type 0)
    real = kind in ("r", "ri")
    for lg in range(1, 7 if real else 6):
        n = 1 << lg
        for fmt, batch, prec in ((IL, 1, 32), (SP, 300, 32), (IL, 5, 64)):
            err, info, passes = model_check(kind, (n,), fmt, batch, prec)
            L = n // 2 if real else n
            # 32-point rows default to the general kernel (faster there)
            want = 13 if 2 <= L <= 16 else 0
            assert [p.type for p in passes][-1:] == [want], (n, [(p.type, p.L) for p in passes])
            assert err < 1e-9, (kind, n, fmt, batch, err)
            err, info, passes = model_check(kind, (n,), fmt, batch, prec, variant={"tiny_rows": -1})
            assert 13 not in [p.type for p in passes] and err < 1e-9
            if L >= 2:
                # an explicit work-group size asks for the tiny kernel
                err, info, passes = model_check(kind, (n,), fmt, batch, prec, variant={"tiny_rows": 64})
                assert passes[-1].type == 13 and passes[-1].CT == 64 and err < 1e-9
    for bad in (1, 32, 100, 512):
        with pytest.raises(ValueError, match="tiny_rows"):
            M.inspect(kind, (16,), IL, 4, variant={"tiny_rows": bad})
    # N-D: the last axis's rows (c2c), later axes on the general kernels
    for dims in ((8, 8, 8), (64, 16), (4, 16)) if not real else ((8, 8, 8),):
        err, info, passes = model_check(kind, dims, SP, 2)
        assert err < 1e-9, (dims, err)
        if not real:
            assert passes[0].type == 13 and passes[0].L == dims[-1], [(p.type, p.L) for p in passes]
    # 64-point c2c rows and the one-point real transform stay on the general kernel
    info, passes = M.inspect("c", (64,), IL, 8)
    assert [p.type for p in passes] == [0]
    if real:
        info, passes = M.inspect(kind, (2,), IL, 8)
        assert 13 not in [p.type for p in passes]


def test_plan_model_r2c_transposing_real_pass_split():
    # the complex-form 1-D r2c (pass A, pass B into a second scratch, the
    # transposing real pass) takes 1024-point columns once h >= 2^21: its pass B
    # stores rows, so long rows cost nothing there
    err, info, passes = model_check("r", (1 << 22,), SP, 1, prec=64)
    assert [(p.type, p.L, p.colmap) for p in passes] == [(0, 1024, 1), (0, 2048, 0), (3, 1, 0)]
    assert passes[2].rr_m == 1024 and err < 1e-9, err
    for lg, n2 in ((24, 8192), (25, 16384)):
        info, passes = M.inspect("r", (1 << lg,), IL, 1)
        assert [(p.type, p.L) for p in passes] == [(0, 1024), (0, n2), (3, 1)], (lg, [(p.type, p.L) for p in passes])
    # an explicit row length keeps the variant's split
    info, passes = M.inspect("r", (1 << 24,), IL, 1, variant={"row_log2": 11})
    assert [(p.type, p.L) for p in passes] == [(0, 4096), (0, 2048), (3, 1)]


@pytest.mark.parametrize("kind", ["r", "ri"])
def test_plan_model_real_pass_in_rows(kind):
    # long real dim 0 with a 1024-point last axis: by default the real pass
    # over row pairs runs inside the column pass B (r2c, type 7) or A (c2r,
    # type 8); on request inside the row FFTs (type 5, or 6 for c2r)
    for fmt, batch in ((IL, 1), (SP, 3)):
        err, info, passes = model_check(kind, (8192, 1024), fmt, batch)
        assert [p.type for p in passes] == ([0, 7, 0] if kind == "r" else [0, 8, 0])
        assert err < 1e-9, (kind, fmt, batch, err)
        err, info, passes = model_check(kind, (8192, 1024), fmt, batch, variant={"real_pair": 1})
        assert [p.type for p in passes] == ([0, 0, 5] if kind == "r" else [6, 0, 0])
        assert err < 1e-9, (kind, fmt, batch, err)


@pytest.mark.parametrize("dims", [(8192, 64), (16384, 64), (8192, 2, 64)], ids=lambda d: "x".join(map(str, d)))
def test_plan_model_inverse_real_pass_in_column_pass_a(dims):
    # c2r mirror: the inverse real pass over row pairs runs in pass A of the
    # dim-0 column four-step (blocks n2 and h2 - n2 per work-group), and with
    # real_pair = -1 as a pass of its own
    for fmt in (IL, SP):
        for v in (None, {"real_pair": -1}):
            err, info, passes = model_check("ri", dims, fmt, 2, variant=v)
            assert (8 in [p.type for p in passes]) == (v is None), [(p.type, p.L) for p in passes]
            assert err < 1e-9, (dims, fmt, v, err)
    err, info, passes = model_check("ri", dims, IL, 1, prec=64)
    assert 8 in [p.type for p in passes] and err < 1e-9


@pytest.mark.parametrize("dims", [(8192, 64), (16384, 64), (65536, 16), (8192, 2, 64)],
                         ids=lambda d: "x".join(map(str, d)))
def test_plan_model_real_pass_in_column_pass_b(dims):
    # long real dim 0 whose rows take no fused row pass: the real pass over
    # row pairs runs in pass B of the dim-0 column four-step (64-, 128- and
    # 256-point columns), and with real_pair = -1 as a pass of its own
    for fmt in (IL, SP):
        for v in (None, {"real_pair": -1}):
            err, info, passes = model_check("r", dims, fmt, 2, variant=v)
            assert (7 in [p.type for p in passes]) == (v is None), [(p.type, p.L) for p in passes]
            assert err < 1e-9, (dims, fmt, v, err)
    # and in FP64 (half the columns per tile)
    err, info, passes = model_check("r", dims, IL, 1, prec=64)
    assert 7 in [p.type for p in passes] and err < 1e-9


@pytest.mark.parametrize("lg", [16, 20, 21, 22])
def test_plan_model_c2r_pre_pass_in_pass_a(lg):
    # long 1-D c2r, opt-in: the inverse real pre-pass runs in the load of
    # four-step pass A (pairing bins k and h-k across columns) where a kernel
    # exists; other shapes and variants keep the separate real pass
    for fmt in (IL, SP):
        for v in ({"real_pair": 1}, {"real_pair": 1, "sched": -1}, None):
            err, info, passes = model_check("ri", (1 << lg,), fmt, 2, variant=v)
            fused = [p.epi for p in passes] == [4, 0]
            assert fused == (lg >= 20 and v is not None and "sched" not in v), [(p.type, p.L, p.epi) for p in passes]
            assert err < 1e-9, (lg, fmt, v, err)


@pytest.mark.parametrize("lg", [17, 20, 22])
def test_plan_model_fused_real_pass(lg):
    # the opt-in fused real pass (rows k1 and n1 - k1 per work-group, 1024- or
    # 2048-point rows): slot mapping, partner bins, row 0 / row n1/2 and the
    # Nyquist bin against the oracle, both formats, a batch of 3 (2^16 splits
    # its 2^15 packed points as 64 x 512, which has no fused kernel)
    for fmt in (IL, SP):
        err, info, passes = model_check("r", (1 << lg,), fmt, 3 if lg < 22 else 1, variant={"real_pair": 1})
        assert [p.type for p in passes][-1] == 4, [p.type for p in passes]
        assert err < 1e-9, (lg, fmt, err)


@pytest.mark.parametrize("lg", [16, 17, 19, 20, 21, 22, 23])
def test_plan_model_real_input_fourstep(lg):
    # 1-D FP32 r2c above 2^15: the real-input four-step (pass A over packed
    # real columns, rows 0..n1/2 into scratch; pass B with own and mirrored
    # bins, each stored once -- the model checks both, bounds included);
    # real_pair = -1 keeps the complex four-step with the transposing real pass
    for fmt in (IL, SP):
        err, info, passes = model_check("r", (1 << lg,), fmt, 3 if lg < 22 else 1, variant={"real_pair": 2})
        assert [p.type for p in passes] == [9, 10], [(p.type, p.L) for p in passes]
        assert passes[0].L * passes[1].L == 1 << lg and passes[1].L >= passes[0].L
        assert err < 1e-9, (lg, fmt, err)
    err, info, passes = model_check("r", (1 << lg,), IL, 2, variant={"real_pair": -1})
    assert 9 not in [p.type for p in passes] and err < 1e-9
    # FP64 keeps the complex four-step
    info, passes = M.inspect("r", (1 << lg,), IL, 1, prec=64)
    assert 9 not in [p.type for p in passes]


def test_real_input_fourstep_shape_variants():
    # the tile / schedule variants pick among the compiled shapes of each
    # pass length and fall back to the default where none matches
    info, p = M.inspect("r", (1 << 22,), IL, 2, variant={"real_pair": 2})
    assert [(q.type, q.L, q.CT, q.sv) for q in p] == [(9, 2048, 16, 1), (10, 2048, 8, 0)]
    info, p = M.inspect("ri", (1 << 22,), IL, 2, variant={"real_pair": 2})
    assert [(q.type, q.L, q.CT, q.sv) for q in p] == [(11, 2048, 8, 1), (12, 2048, 8, 0)]
    # defaults by length (measured): r2c 2^16..2^23, c2r 2^17..2^22 (c2r 2^16 is
    # one half-LDS row pass)
    for kind, lg, real4 in (("r", 16, True), ("r", 17, True), ("r", 22, True), ("r", 23, True), ("r", 24, False),
                            ("ri", 16, False), ("ri", 17, True), ("ri", 20, True), ("ri", 22, True),
                            ("ri", 23, False)):
        info, p = M.inspect(kind, (1 << lg,), IL, 1)
        assert (p[0].type in (9, 11)) == real4, (kind, lg, [q.type for q in p])
    info, p = M.inspect("r", (1 << 22,), IL, 2, variant={"real_pair": 2, "rowt_tile": 16, "sched": 1, "col_tile": 16})
    assert [(q.type, q.L, q.CT, q.sv) for q in p] == [(9, 2048, 16, 1), (10, 2048, 16, 1)]
    info, p = M.inspect("ri", (1 << 21,), IL, 2, variant={"real_pair": 2, "sched": 1})
    assert [(q.type, q.L, q.CT, q.sv) for q in p] == [(11, 2048, 8, 1), (12, 1024, 16, 1)]
    info, p = M.inspect("ri", (1 << 20,), IL, 2, variant={"real_pair": 2, "rowt_tile": 4})
    assert [(q.type, q.L, q.CT, q.sv) for q in p] == [(11, 1024, 8, 1), (12, 1024, 8, 0)]
    # the c2r pass A' without the half exchange (full complex entries in LDS)
    info, p = M.inspect("ri", (1 << 22,), IL, 2, variant={"real_pair": 2, "half_lds": -1})
    assert [(q.type, q.L, q.CT, q.sv) for q in p][1] == (12, 2048, 8, 0)
    err, info, p = model_check("ri", (1 << 21,), SP, 2, variant={"real_pair": 2, "half_lds": -1, "rowt_tile": 16})
    assert (p[1].L, p[1].CT) == (1024, 16) and err < 1e-9
    err, info, p = model_check("r", (1 << 21,), SP, 2, variant={"real_pair": 2, "rowt_tile": 16, "sched": 1})
    assert p[1].CT == 16 and err < 1e-9


@pytest.mark.parametrize("lg", [16, 17, 20, 23])
def test_plan_model_real_input_fourstep_c2r(lg):
    # the 1-D FP32 c2r above 2^15 runs the same split backwards: pass B' reads
    # every bin 0..N/2 (own and mirrored, the model checks the set), pass A'
    # writes two real columns per complex column
    for fmt in (IL, SP):
        err, info, passes = model_check("ri", (1 << lg,), fmt, 3 if lg < 22 else 1, variant={"real_pair": 2})
        assert [p.type for p in passes] == [11, 12], [(p.type, p.L) for p in passes]
        assert passes[0].L * passes[1].L == 1 << lg and passes[0].L >= passes[1].L
        assert err < 1e-9, (lg, fmt, err)
    for rp in (-1, 1):
        err, info, passes = model_check("ri", (1 << lg,), IL, 2, variant={"real_pair": rp})
        assert 11 not in [p.type for p in passes] and err < 1e-9
    info, passes = M.inspect("ri", (1 << lg,), IL, 1, prec=64)
    assert 11 not in [p.type for p in passes]


@pytest.mark.parametrize("kind", ["c", "ci", "r", "ri"])
def test_plan_model_row_split_variant(kind):
    # other splits n = n1 * n2 of the 1-D four-step (variant row_log2): pass B
    # rows of 2^row_log2 points (r2c / c2r split their half-length complex
    # transform the same way)
    for lg, rl in ((17, 9), (17, 11), (20, 8), (20, 11), (20, 12)):
        err, info, passes = model_check(kind, (1 << lg,), IL, 2, variant={"row_log2": rl})
        if [p.type for p in passes] in ([9, 10], [11, 12]):
            # r2c / c2r as the real-input four-step: rows of 2^row_log2 points
            rows, cols = (passes[1], passes[0]) if kind == "r" else (passes[0], passes[1])
            assert kind in ("r", "ri") and rows.L == 1 << rl and cols.L == 1 << (lg - rl)
        else:
            ffts = [p for p in passes if p.type == 0]
            assert ffts[-1].L == 1 << rl and ffts[0].colmap == 1, [(p.type, p.L) for p in passes]
        assert err < 1e-9, (kind, lg, rl, err)
    # a row length that does not split n: the plan is refused
    for rl in (3, 20, 30):
        with pytest.raises(ValueError):
            M.inspect(kind, (1 << 20,), IL, 1, variant={"row_log2": rl})


def test_plan_model_fp64_and_chunking():
    # a batch larger than one chunk exercises the per-chunk scratch reuse
    info, passes = M.inspect("c", (1 << 20,), IL, 64)
    assert len(passes) == 2 and 1 <= info.chunk < 64
    err, info, _ = model_check("c", (1 << 15,), IL, 3, prec=64)
    assert err < 1e-9
    # FP64 real transforms around the longest FP64 row pass (8192 complex)
    for kind in ("r", "ri"):
        for lg in (14, 15, 16, 17):
            err, info, _ = model_check(kind, (1 << lg,), IL, 2, prec=64)
            assert err < 1e-9, (kind, lg, err)


def test_plan_shapes_of_headline_configs():
    # 64 x 2^20 c2c: two passes (four-step 1024 x 1024)
    info, passes = M.inspect("c", (1 << 20,), IL, 64)
    assert [(p.L, p.colmap, p.epi, p.store_t) for p in passes] == [(1024, 1, 1, 0), (1024, 0, 0, 1)]
    # 2^12 c2c: one row pass
    info, passes = M.inspect("c", (4096,), IL, 1)
    assert [(p.L, p.colmap) for p in passes] == [(4096, 0)]
    # 8192^2 r2c: dim 0 (4096 complex rows) as a 64 x 64 column four-step
    # whose pass B runs the real pass over row pairs, then the 8192-point rows
    info, passes = M.inspect("r", (8192, 8192), IL, 1)
    assert [(p.type, p.L, p.colmap, p.epi) for p in passes] == [(0, 64, 1, 1), (7, 64, 1, 0), (0, 8192, 0, 0)]
    assert passes[0].tw4_blo == 1 and passes[1].rr_h == 4096 and passes[1].rr_m == 8192 and passes[1].n_blo == 64
    # the same with the real pass on its own
    info, passes = M.inspect("r", (8192, 8192), IL, 1, variant={"real_pair": -1})
    assert [(p.type, p.L, p.colmap, p.epi) for p in passes] == [(0, 64, 1, 1), (0, 64, 1, 0), (1, 1, 0, 0),
                                                                 (0, 8192, 0, 0)]
    assert passes[2].rr_h == 4096 and passes[2].rr_m == 8192
    # 16384 x 2048 r2c: the real pass inside the 128-point column pass B, or
    # on request inside the 2048-point row FFTs
    info, passes = M.inspect("r", (16384, 2048), IL, 1)
    assert [(p.type, p.L) for p in passes] == [(0, 64), (7, 128), (0, 2048)] and passes[1].rr_h == 8192
    info, passes = M.inspect("r", (16384, 2048), IL, 1, variant={"real_pair": 1})
    assert [(p.type, p.L) for p in passes][-1] == (5, 2048) and passes[-1].rr_h == 8192
    # 8192 x 64 r2c (no fused row kernel for 64-point rows): the real pass in
    # the column pass B; 8192 x 8 (fewer columns than its tile): on its own
    info, passes = M.inspect("r", (8192, 64), IL, 1)
    assert [p.type for p in passes] == [0, 7, 0] and passes[1].rr_h == 4096 and passes[1].rr_m == 64
    info, passes = M.inspect("r", (8192, 8), IL, 1)
    assert [p.type for p in passes] == [0, 0, 1, 13] and passes[2].rr_h == 4096 and passes[2].rr_m == 8
    # 2048 x 8192 r2c: dim 0 short enough for one fused column r2c pass
    info, passes = M.inspect("r", (2048, 8192), IL, 1)
    assert [(p.L, p.colmap, p.epi) for p in passes] == [(1024, 1, 2), (8192, 0, 0)]
    # 256^3 c2c: three single passes
    info, passes = M.inspect("c", (256, 256, 256), IL, 1)
    assert len(passes) == 3


def test_plan_model_three_pass_fourstep():
    # 1-D lengths above 2048 * 2048: the four-step n = p * M whose row DFTs
    # (M = q * 2048) are a four-step too -- pass 1 p-point columns, pass 2
    # q-point columns in place, pass 3 2048-point rows storing bin
    # kp + p*(kq + q*k2); r2c / c2r put the real pass beside it
    err, info, passes = model_check("c", (1 << 25,), IL, 1)
    assert [(p.type, p.L, p.colmap) for p in passes] == [(0, 128, 1), (0, 128, 1), (0, 2048, 0)], passes
    assert passes[2].n_blo == 128 and passes[2].out.s_j == 128 * 128 and passes[2].out.s_b2 == 128
    assert err < 1e-9, err
    # (the model runs the c2c; the real plans add the real pass of the
    # shorter plans, which the other model tests execute)
    info, passes = M.inspect("ri", (1 << 26,), IL, 1)
    assert [(p.type, p.L) for p in passes] == [(1, 1), (0, 128), (0, 128), (0, 2048)]
    info, passes = M.inspect("r", (1 << 28,), SP, 1)
    assert [(p.type, p.L) for p in passes] == [(0, 256), (0, 256), (0, 2048), (1, 1)]
    # from 2^23 on (a two-pass split would need 4096-point columns): FP32 and
    # FP64, executed by the model for 2^23 in FP64 with a batch of 2
    err, info, passes = model_check("c", (1 << 23,), SP, 2, prec=64)
    assert [(p.type, p.L, p.colmap) for p in passes] == [(0, 64, 1), (0, 64, 1), (0, 2048, 0)], passes
    assert err < 1e-9, err
    info, passes = M.inspect("ci", (1 << 24,), IL, 1)
    assert [(p.type, p.L) for p in passes] == [(0, 128), (0, 64), (0, 2048)]
    info, passes = M.inspect("c", (1 << 22,), IL, 1)
    assert [(p.type, p.L) for p in passes] == [(0, 2048), (0, 2048)]


def test_removed_pipeline_variant_is_refused():
    with pytest.raises(ValueError, match="pipeline"):
        M.inspect("c", (1 << 20,), IL, 8, variant={"pipeline": 1})
    info, passes = M.inspect("c", (1 << 20,), IL, 8, variant={"pipeline": -1})
    assert len(passes) == 2


def test_unsupported_shapes_report_errors():
    with pytest.raises(ValueError):
        M.inspect("c", (1 << 26, 2), IL, 1)
    with pytest.raises(ValueError):
        M.inspect("c", (12,), IL, 1)
