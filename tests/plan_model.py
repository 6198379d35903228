"""A numpy model of the GPU plans, for CPU tests.

``afft_gpu_plan_inspect`` returns the pass list the library would launch.  This
module executes that list on host arrays using exactly the addressing the
kernels use (fft_amd/csrc/afft_pass.h, afft_aux.hip): every element address of
every pass is computed, bounds-checked against the buffer sizes the API
contract gives the caller's arrays and the plan's scratch, gathered, pushed
through the pass's arithmetic (numpy FFT in float64 plus the pass epilogue) and
scattered.  The result is compared with the oracle by the tests, so a wrong
stride, role, chunk offset or epilogue in the plan builder fails on the CPU
before any kernel runs on a GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from fft_amd import _lib


class SideInfo(ctypes.Structure):
    _fields_ = [("role", ctypes.c_int32), ("il", ctypes.c_int32), ("one_array", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("re_off", ctypes.c_int64), ("im_off", ctypes.c_int64),
                ("s_j", ctypes.c_int64), ("s_c", ctypes.c_int64), ("s_b", ctypes.c_int64),
                ("s_b2", ctypes.c_int64), ("s_u", ctypes.c_int64)]


class PassInfo(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("L", ctypes.c_int32), ("CT", ctypes.c_int32),
                ("colmap", ctypes.c_int32), ("epi", ctypes.c_int32), ("store_t", ctypes.c_int32),
                ("in_", SideInfo), ("out", SideInfo), ("n_col", ctypes.c_int64),
                ("fold_user", ctypes.c_int32), ("n_blo", ctypes.c_int32), ("b_per_user", ctypes.c_int64),
                ("swap_in", ctypes.c_int32), ("swap_out", ctypes.c_int32), ("tw4_bits", ctypes.c_int32),
                ("tw4_shift", ctypes.c_int32), ("rr_h", ctypes.c_int64), ("rr_m", ctypes.c_int64),
                ("rr_inv", ctypes.c_int32), ("tw4_blo", ctypes.c_int32), ("sv", ctypes.c_int32),
                ("nt_io", ctypes.c_int32)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("npass", ctypes.c_int32), ("fused_slots", ctypes.c_int32), ("chunk", ctypes.c_int64),
                ("scr_per_user", ctypes.c_int64 * 2)]


KIND = {"c": 0, "ci": 1, "r": 2, "ri": 3}


def inspect(kind: str, dims, fmt: int, batch: int = 1, prec: int = 32, variant: dict | None = None):
    """Pass list of the plan (variant: afft_gpu_variant fields, others 0 /
    stream_io 2 as afft.Plan.set_variant sets them)."""
    from fft_amd.afft import Variant
    L = _lib.lib()
    f = L.afft_gpu_plan_inspect_variant
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int,
                  ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(PlanInfo), ctypes.POINTER(PassInfo), ctypes.c_int]
    f.restype = ctypes.c_int
    nd, arr = _lib.dims_array(dims)
    info = PlanInfo()
    passes = (PassInfo * 200)()
    v = None
    if variant is not None:
        v = Variant(stream_io=2)
        for k, val in variant.items():
            setattr(v, k, int(val))
    n = f(prec, KIND[kind], nd, arr, fmt, batch, ctypes.byref(v) if v is not None else None, ctypes.byref(info),
          passes, 200)
    if n < 0:
        raise ValueError(_lib.last_error())
    return info, [passes[i] for i in range(n)]


def squeeze(dims):
    s = [int(d) for d in dims if int(d) != 1]
    return s or [1]


class Buffers:
    """Scalar float64 arrays standing for the device buffers of one execute."""

    def __init__(self, kind, dims, fmt, batch, info):
        d = squeeze(dims)
        n = int(np.prod(d))
        hsz = (d[0] // 2 + 1) * int(np.prod(d[1:])) if kind in ("r", "ri") else n
        il = fmt == _lib.afft_interleaved
        cx_in = kind in ("c", "ci", "ri")
        cx_out = kind in ("c", "ci", "r")
        nin = n if kind != "ri" else hsz
        nout = n if kind != "r" else hsz
        self.arrays = {}
        self.arrays["src_re"] = np.zeros((2 if (cx_in and il) else 1) * nin * batch)
        self.arrays["src_im"] = np.zeros(nin * batch if (cx_in and not il) else 0)
        self.arrays["dst_re"] = np.zeros((2 if (cx_out and il) else 1) * nout * batch)
        self.arrays["dst_im"] = np.zeros(nout * batch if (cx_out and not il) else 0)
        for i in range(2):
            self.arrays[f"scr{i}"] = np.zeros(2 * info.scr_per_user[i] * info.chunk)


def _arrays_for(side, bufs):
    r = side.role
    if r == 0:
        return bufs.arrays["src_re"], bufs.arrays["src_im"]
    if r == 1:
        return bufs.arrays["dst_re"], bufs.arrays["dst_im"]
    return bufs.arrays[f"scr{r - 2}"], None


def _addr(side, start, b, n_blo, c, j):
    """Element index (complex units when il, scalar units otherwise)."""
    uoff = start * side.s_u if side.role < 2 else 0
    return (uoff + (b // n_blo) * side.s_b + (b % n_blo) * side.s_b2 + c * side.s_c + j * side.s_j)


def _check(idx, arr, what):
    if idx.size == 0:
        return
    lo, hi = int(idx.min()), int(idx.max())
    if lo < 0 or hi >= arr.size:
        raise IndexError(f"{what}: index range [{lo}, {hi}] outside buffer of {arr.size}")


def _gather(side, bufs, e, what):
    re_arr, im_arr = _arrays_for(side, bufs)
    if side.il:
        i0 = 2 * (side.re_off + e)
        _check(i0, re_arr, what + " (re)")
        _check(i0 + 1, re_arr, what + " (im)")
        return re_arr[i0] + 1j * re_arr[i0 + 1]
    ire = side.re_off + e
    iim = side.im_off + e
    src_im = re_arr if side.one_array else im_arr
    _check(ire, re_arr, what + " (re)")
    _check(iim, src_im, what + " (im)")
    return re_arr[ire] + 1j * src_im[iim]


def _scatter(side, bufs, e, v, what):
    re_arr, im_arr = _arrays_for(side, bufs)
    if side.il:
        i0 = 2 * (side.re_off + e)
        _check(i0, re_arr, what + " (re)")
        _check(i0 + 1, re_arr, what + " (im)")
        re_arr[i0] = v.real
        re_arr[i0 + 1] = v.imag
        return
    ire = side.re_off + e
    iim = side.im_off + e
    dst_im = re_arr if side.one_array else im_arr
    _check(ire, re_arr, what + " (re)")
    _check(iim, dst_im, what + " (im)")
    re_arr[ire] = v.real
    dst_im[iim] = v.imag


def _swap(z):
    return z.imag + 1j * z.real


def _fft_pass(p, bufs, start, cnt, pi):
    L = p.L
    ncol = p.n_col * cnt if p.fold_user else p.n_col
    nb = 1 if p.fold_user else p.b_per_user * cnt
    b = np.arange(nb).reshape(nb, 1, 1)
    c = np.arange(ncol).reshape(1, ncol, 1)
    nin = L + 1 if p.epi == 3 else L
    j = np.arange(nin).reshape(1, 1, nin)
    e_in = _addr(p.in_, start, b, p.n_blo, c, j)
    x = _gather(p.in_, bufs, e_in, f"pass {pi} load")
    if p.epi == 4:  # c2r pre-pass pairing bin k with bin h - k (another column)
        h = p.rr_h
        k = c * p.in_.s_c + j * p.in_.s_j
        s0 = x
        s1 = _gather(p.in_, bufs, e_in - k + (h - k), f"pass {pi} load (pair)")
        w = np.exp(-2j * np.pi * k / (2 * h))
        z = (s0 + np.conj(s1)) - 1j * (np.conj(s1) - s0) * np.conj(w)
        z0 = (s0.real + s1.real) + 1j * (s0.real - s1.real)
        x = np.where(k == 0, z0, z)
    if p.epi == 3:  # c2r pre-pass
        s0 = x[..., :L]
        s1 = x[..., L:0:-1]  # S[L - k], k = 0..L-1
        k = np.arange(L)
        w = np.exp(-2j * np.pi * k / (2 * L))
        a = s0 + np.conj(s1)
        bb = (np.conj(s1) - s0) * np.conj(w)
        z = a - 1j * bb
        z[..., 0] = (s0[..., 0].real + s1[..., 0].real) + 1j * (s0[..., 0].real - s1[..., 0].real)
        x = z
    if p.swap_in:
        x = _swap(x)
    X = np.fft.fft(x, axis=-1)
    if p.epi in (1, 4):  # four-step twiddle
        if p.tw4_blo:
            n2 = (np.arange(nb) % p.n_blo).reshape(nb, 1, 1)
            N = L * p.n_blo
        else:
            n2 = (np.arange(ncol) >> p.tw4_shift).reshape(1, ncol, 1)
            N = L * (p.n_col >> p.tw4_shift)
        k1 = np.arange(L).reshape(1, 1, L)
        X = X * np.exp(-2j * np.pi * ((n2 * k1) % N) / N)
    if p.epi == 2:  # r2c post-pass, bins 0..L
        Z = X
        k = np.arange(L)
        Zh = Z[..., (-k) % L]
        w = np.exp(-2j * np.pi * k / (2 * L))
        a = 0.5 * (Z + np.conj(Zh))
        bb = 0.5 * (Z - np.conj(Zh)) * w
        out = a - 1j * bb
        out[..., 0] = Z[..., 0].real + Z[..., 0].imag
        last = (Z[..., 0].real - Z[..., 0].imag)[..., None]
        X = np.concatenate([out, last], axis=-1)
    if p.swap_out:
        X = _swap(X)
    kk = np.arange(X.shape[-1]).reshape(1, 1, X.shape[-1])
    e_out = _addr(p.out, start, b, p.n_blo, c, kk)
    # the kernels never write the same element twice within a pass
    flat = np.broadcast_to(e_out, X.shape).ravel()
    if np.unique(flat).size != flat.size:
        raise AssertionError(f"pass {pi}: overlapping output addresses")
    _scatter(p.out, bufs, e_out, X, f"pass {pi} store")


def _real_rows(p, bufs, start, cnt, pi):
    h, m = p.rr_h, p.rr_m
    b = np.arange(cnt).reshape(cnt, 1, 1)
    c = np.arange(m).reshape(1, m, 1)
    rows = np.arange(h + 1).reshape(1, 1, h + 1)
    sc_in = p.in_.s_c or 1
    sc_out = p.out.s_c or 1
    uin = start * p.in_.s_u if p.in_.role < 2 else 0
    uout = start * p.out.s_u if p.out.role < 2 else 0
    if p.rr_inv:
        e_in = uin + b * p.in_.s_b + c * sc_in + rows * p.in_.s_j
        S = _gather(p.in_, bufs, e_in, f"pass {pi} load")
        k = np.arange(h)
        s0 = S[..., :h]
        s1 = S[..., h - k]
        w = np.exp(-2j * np.pi * k / (2 * h))
        a = s0 + np.conj(s1)
        bb = (np.conj(s1) - s0) * np.conj(w)
        Z = a - 1j * bb
        Z[..., 0] = (S[..., 0].real + S[..., h].real) + 1j * (S[..., 0].real - S[..., h].real)
        orows = np.arange(h).reshape(1, 1, h)
        e_out = uout + b * p.out.s_b + c * sc_out + orows * p.out.s_j
        _scatter(p.out, bufs, e_out, Z, f"pass {pi} store")
    else:
        hr = np.arange(h).reshape(1, 1, h)
        e_in = uin + b * p.in_.s_b + c * sc_in + hr * p.in_.s_j
        Z = _gather(p.in_, bufs, e_in, f"pass {pi} load")
        k = np.arange(h)
        Zh = Z[..., (-k) % h]
        w = np.exp(-2j * np.pi * k / (2 * h))
        X = 0.5 * (Z + np.conj(Zh)) - 1j * 0.5 * (Z - np.conj(Zh)) * w
        X[..., 0] = Z[..., 0].real + Z[..., 0].imag
        last = (Z[..., 0].real - Z[..., 0].imag)[..., None]
        X = np.concatenate([X, last], axis=-1)
        e_out = uout + b * p.out.s_b + c * sc_out + rows * p.out.s_j
        _scatter(p.out, bufs, e_out, X, f"pass {pi} store")


def _real_colpair(p, bufs, start, cnt, pi):
    """Column pass B of the dim-0 four-step with the forward real pass over
    row pairs (afft_realpair.hip real_colpair_kernel): Y[k1][j][c] (h1 x L x
    m) -> Z[k1 + h1*k2][c], the L-point DFT over j; then rows 0..h of X as
    in _real_rows."""
    h, m, h1, L = p.rr_h, p.rr_m, p.n_blo, p.L
    assert h == h1 * L
    b = np.arange(cnt).reshape(cnt, 1, 1, 1)
    k1 = np.arange(h1).reshape(1, h1, 1, 1)
    j = np.arange(L).reshape(1, 1, L, 1)
    c = np.arange(m).reshape(1, 1, 1, m)
    uin = start * p.in_.s_u if p.in_.role < 2 else 0
    e_in = uin + b * p.in_.s_b + k1 * p.in_.s_b2 + j * p.in_.s_j + c * (p.in_.s_c or 1)
    Y = _gather(p.in_, bufs, e_in, f"pass {pi} load")
    Z = np.fft.fft(Y, axis=2).transpose(0, 2, 1, 3).reshape(cnt, h, m)  # row k1 + h1*k2
    k = np.arange(h).reshape(1, h, 1)
    Zh = Z[:, (-np.arange(h)) % h, :]
    w = np.exp(-2j * np.pi * k / (2 * h))
    X = 0.5 * (Z + np.conj(Zh)) - 1j * 0.5 * (Z - np.conj(Zh)) * w
    X[:, 0, :] = Z[:, 0, :].real + Z[:, 0, :].imag
    last = (Z[:, 0, :].real - Z[:, 0, :].imag)[:, None, :]
    X = np.concatenate([X, last], axis=1)
    uout = start * p.out.s_u if p.out.role < 2 else 0
    rows = np.arange(h + 1).reshape(1, h + 1, 1)
    bo = np.arange(cnt).reshape(cnt, 1, 1)
    co = np.arange(m).reshape(1, 1, m)
    e_out = uout + bo * p.out.s_b + rows * p.out.s_j + co * (p.out.s_c or 1)
    _scatter(p.out, bufs, e_out, X, f"pass {pi} store")


def _real_icolpair(p, bufs, start, cnt, pi):
    """Inverse real pass over row pairs fused into pass A of the dim-0 column
    four-step (afft_realpair.hip real_icolpair_kernel): S rows 0..h -> Z rows
    0..h-1 as in _real_rows (inverse); then, for each block n2, the L-point
    DFT over j of swap(Z[j*h2 + n2]) times W_h^(n2*k1), stored at entry
    (k1, n2, c) = k1*out.s_j + n2*m + c."""
    h, m, h2, L = p.rr_h, p.rr_m, p.n_blo, p.L
    assert h == L * h2
    b = np.arange(cnt).reshape(cnt, 1, 1)
    rows = np.arange(h + 1).reshape(1, h + 1, 1)
    c = np.arange(m).reshape(1, 1, m)
    uin = start * p.in_.s_u if p.in_.role < 2 else 0
    e_in = uin + b * p.in_.s_b + rows * p.in_.s_j + c * (p.in_.s_c or 1)
    S = _gather(p.in_, bufs, e_in, f"pass {pi} load")
    k = np.arange(h)
    s0 = S[:, :h, :]
    s1 = S[:, h - k, :]
    w = np.exp(-2j * np.pi * k / (2 * h)).reshape(1, h, 1)
    Z = (s0 + np.conj(s1)) - 1j * (np.conj(s1) - s0) * np.conj(w)
    Z[:, 0, :] = (S[:, 0, :].real + S[:, h, :].real) + 1j * (S[:, 0, :].real - S[:, h, :].real)
    X = np.fft.fft(_swap(Z).reshape(cnt, L, h2, m), axis=1)  # row j*h2 + n2 -> bin k1
    k1 = np.arange(L).reshape(1, L, 1, 1)
    n2 = np.arange(h2).reshape(1, 1, h2, 1)
    X = X * np.exp(-2j * np.pi * ((n2 * k1) % h) / h)
    uout = start * p.out.s_u if p.out.role < 2 else 0
    bo = np.arange(cnt).reshape(cnt, 1, 1, 1)
    co = np.arange(m).reshape(1, 1, 1, m)
    e_out = uout + bo * p.out.s_b + k1 * p.out.s_j + n2 * m + co * (p.out.s_c or 1)
    _scatter(p.out, bufs, e_out, X, f"pass {pi} store")


def _real_tr(p, bufs, start, cnt, pi):
    """Transposing forward real pass (afft_aux.hip real_tr_kernel): input
    Y[b][k1][k2] (n1 x n2) holds Z[k1 + n1*k2]; output bins 0..h natural."""
    h, n1 = p.rr_h, p.rr_m
    n2 = h // n1
    b = np.arange(cnt).reshape(cnt, 1)
    uin = start * p.in_.s_u if p.in_.role < 2 else 0
    uout = start * p.out.s_u if p.out.role < 2 else 0
    k = np.arange(h).reshape(1, h)
    k1, k2 = k % n1, k // n1
    e_in = uin + b * p.in_.s_b + k1 * n2 + k2
    Z = _gather(p.in_, bufs, e_in, f"pass {pi} load")
    kk = np.arange(h)
    Zh = Z[..., (-kk) % h]
    w = np.exp(-2j * np.pi * kk / (2 * h))
    X = 0.5 * (Z + np.conj(Zh)) - 1j * 0.5 * (Z - np.conj(Zh)) * w
    X[..., 0] = Z[..., 0].real + Z[..., 0].imag
    last = (Z[..., 0].real - Z[..., 0].imag)[..., None]
    X = np.concatenate([X, last], axis=-1)
    e_out = uout + b * p.out.s_b + np.arange(h + 1).reshape(1, h + 1) * p.out.s_j
    _scatter(p.out, bufs, e_out, X, f"pass {pi} store")


def _rp_row(j, c, R, n1):
    """Row held by slot c of tile j of real_pair_kernel (afft_realpair.hip)."""
    if c < R:
        return j * R + c
    c2 = c - R
    if j == 0 and c2 == 0:
        return n1 // 2
    return n1 - j * R - c2


def _real_pair(p, bufs, start, cnt, pi):
    """Pass B of the four-step r2c fused with the real pass (afft_realpair.hip):
    tile j of each transform transforms rows _rp_row(j, c) (c < 2R) of
    Y[b][k1][:] (n2 points each), pairs slot c with slot c +- R (rows 0 and
    n1/2 with themselves) and writes bins row + n1*k2, plus bin h."""
    h, n1, n2 = p.rr_h, p.rr_m, p.L
    R = p.CT // 2
    uin = start * p.in_.s_u if p.in_.role < 2 else 0
    uout = start * p.out.s_u if p.out.role < 2 else 0
    tiles = n1 // p.CT
    covered = np.zeros((cnt, n1), dtype=np.int64)
    for b in range(cnt):
        for j in range(tiles):
            rows = [_rp_row(j, c, R, n1) for c in range(p.CT)]
            covered[b, rows] += 1
            e_in = (uin + b * p.in_.s_b + np.array(rows).reshape(-1, 1) * p.in_.s_c
                    + np.arange(n2).reshape(1, -1) * p.in_.s_j)
            Z = np.fft.fft(_gather(p.in_, bufs, e_in, f"pass {pi} load"), axis=-1)
            k2 = np.arange(n2)
            X = np.empty_like(Z)
            for c in range(p.CT):
                row = rows[c]
                self_pair = j == 0 and c in (0, R)
                pc = c if self_pair else (c + R if c < R else c - R)
                mirror = ((n2 - k2) % n2) if (j == 0 and c == 0) else (n2 - 1 - k2)
                zp = Z[pc, mirror]
                k = row + n1 * k2
                w = np.exp(-2j * np.pi * k / (2 * h))
                X[c] = 0.5 * (Z[c] + np.conj(zp)) - 1j * 0.5 * (Z[c] - np.conj(zp)) * w
            bins = np.array(rows).reshape(-1, 1) + n1 * k2.reshape(1, -1)
            _scatter(p.out, bufs, uout + b * p.out.s_b + bins * p.out.s_j, X, f"pass {pi} store")
            if j == 0:
                z0 = Z[0, 0]
                _scatter(p.out, bufs, np.array([uout + b * p.out.s_b + h * p.out.s_j]),
                         np.array([z0.real - z0.imag + 0j]), f"pass {pi} store")
    if not np.all(covered == 1):
        raise AssertionError(f"pass {pi}: rows not covered exactly once")


def _real_rowfft(p, bufs, start, cnt, pi):
    """N-D r2c real pass over row pairs fused into the row FFTs
    (afft_realpair.hip real_rowfft_kernel), in place: work-group q reads rows q
    and h-q (q = 0: row 0 twice), forms X for its output rows (q, h-q; q = 0:
    rows 0 and h; q = h/2: row h/2 once) and writes their L-point FFTs."""
    h, L = p.rr_h, p.L
    uoff = start * p.out.s_u if p.out.role < 2 else 0
    q = np.arange(h // 2 + 1)
    # (q, slot) pairs in work-group order; slot 1 of q = h/2 stores nothing
    qq = np.concatenate([q, q[q * 2 != h]])
    c = np.concatenate([np.zeros(q.size, dtype=np.int64), np.ones(qq.size - q.size, dtype=np.int64)])
    own_row = np.where(c == 0, qq, np.where(qq == 0, 0, h - qq))
    oth_row = np.where(c == 0, np.where(qq == 0, 0, h - qq), qq)
    out_row = np.where(c == 0, qq, np.where(qq == 0, h, h - qq))
    if np.bincount(out_row, minlength=h + 1).tolist() != [1] * (h + 1):
        raise AssertionError(f"pass {pi}: rows not written exactly once")
    wq = np.exp(-2j * np.pi * qq / (2 * h))
    w = np.where(c == 0, wq, -np.conj(wq)).reshape(1, -1, 1)
    b = np.arange(cnt).reshape(cnt, 1, 1)
    j = np.arange(L).reshape(1, 1, L)
    base = uoff + b * p.out.s_b
    own = _gather(p.in_, bufs, base + own_row.reshape(1, -1, 1) * p.in_.s_c + j * p.in_.s_j, f"pass {pi} load")
    oth = _gather(p.in_, bufs, base + oth_row.reshape(1, -1, 1) * p.in_.s_c + j * p.in_.s_j, f"pass {pi} load")
    x = 0.5 * (own + np.conj(oth)) - 1j * 0.5 * w * (own - np.conj(oth))
    y = np.fft.fft(x, axis=-1)
    # every work-group reads its rows before writing them (rows are disjoint)
    _scatter(p.out, bufs, base + out_row.reshape(1, -1, 1) * p.out.s_c + j * p.out.s_j, y, f"pass {pi} store")


def _real_irowfft(p, bufs, start, cnt, pi):
    """N-D c2r inverse real pass over row pairs fused into the inverse row
    FFTs (afft_realpair.hip real_irowfft_kernel): work-group q reads input rows
    q and h-q (the partner at the mirrored column), writes rows q and h-q of Z
    (q = 0: row 0 from the Hermitian parts of rows 0 and h; q = h/2: one row)."""
    h, L = p.rr_h, p.L
    uin = start * p.in_.s_u if p.in_.role < 2 else 0
    uout = start * p.out.s_u if p.out.role < 2 else 0
    q = np.arange(1, h // 2 + 1)
    qq = np.concatenate([q, q[q * 2 != h]])
    c = np.concatenate([np.zeros(q.size, dtype=np.int64), np.ones(qq.size - q.size, dtype=np.int64)])
    own_row = np.where(c == 0, qq, h - qq)
    oth_row = np.where(c == 0, h - qq, qq)
    rows = np.concatenate([[0], own_row])
    if np.bincount(rows, minlength=h).tolist() != [1] * h:
        raise AssertionError(f"pass {pi}: rows not written exactly once")
    wq = np.exp(-2j * np.pi * qq / (2 * h))
    u = np.where(c == 0, np.conj(wq), -wq).reshape(1, -1, 1)
    b = np.arange(cnt).reshape(cnt, 1, 1)
    j = np.arange(L).reshape(1, 1, L)
    jm = (L - j) % L
    ib = uin + b * p.in_.s_b
    own = _gather(p.in_, bufs, ib + own_row.reshape(1, -1, 1) * p.in_.s_c + j * p.in_.s_j, f"pass {pi} load")
    oc = np.conj(_gather(p.in_, bufs, ib + oth_row.reshape(1, -1, 1) * p.in_.s_c + jm * p.in_.s_j,
                         f"pass {pi} load"))
    pre = (own + oc) + 1j * u * (own - oc)
    x0 = _gather(p.in_, bufs, ib + 0 * p.in_.s_c + j * p.in_.s_j, f"pass {pi} load")
    x0m = _gather(p.in_, bufs, ib + 0 * p.in_.s_c + jm * p.in_.s_j, f"pass {pi} load")
    xh = _gather(p.in_, bufs, ib + h * p.in_.s_c + j * p.in_.s_j, f"pass {pi} load")
    xhm = _gather(p.in_, bufs, ib + h * p.in_.s_c + jm * p.in_.s_j, f"pass {pi} load")
    h0, hh = 0.5 * (x0 + np.conj(x0m)), 0.5 * (xh + np.conj(xhm))
    pre0 = (h0 + hh) + 1j * (h0 - hh)
    y = np.fft.ifft(np.concatenate([pre0, pre], axis=1), axis=-1) * L  # unnormalised inverse
    ob = uout + b * p.out.s_b
    _scatter(p.out, bufs, ob + rows.reshape(1, -1, 1) * p.out.s_c + j * p.out.s_j, y, f"pass {pi} store")


def _real4a(p, bufs, start, cnt, pi):
    """Pass A of the real-input four-step (afft_real4.hip real4a_kernel):
    packed column c (reals 2c, 2c+1 of each row j1 of the n1 x n2 matrix)
    -> n1-point DFT Z -> A = (Z + conj Z[-k1]) / 2, B = (Z - conj Z[-k1]) / 2i
    -> times W_N^(2c*k1), W_N^((2c+1)*k1) -> rows k1 = 0..n1/2 of the scratch,
    columns 2c and 2c+1 (row k1 at k1*out.s_j, column at c*out.s_c)."""
    n1, n2, N = p.L, p.rr_m, 2 * p.rr_h
    npc = n2 // 2
    b = np.arange(cnt).reshape(cnt, 1, 1)
    c = np.arange(npc).reshape(1, npc, 1)
    j1 = np.arange(n1).reshape(1, 1, n1)
    uin = start * p.in_.s_u if p.in_.role < 2 else 0
    e_in = uin + b * p.in_.s_b + c * p.in_.s_c + j1 * p.in_.s_j
    z = _gather(p.in_, bufs, e_in, f"pass {pi} load")
    Z = np.fft.fft(z, axis=-1)
    Zm = np.conj(Z[..., (-np.arange(n1)) % n1])
    A = 0.5 * (Z + Zm)
    B = (Z - Zm) / 2j
    k1 = np.arange(n1 // 2 + 1).reshape(1, 1, -1)
    A = A[..., : n1 // 2 + 1] * np.exp(-2j * np.pi * ((2 * c * k1) % N) / N)
    B = B[..., : n1 // 2 + 1] * np.exp(-2j * np.pi * (((2 * c + 1) * k1) % N) / N)
    uout = start * p.out.s_u if p.out.role < 2 else 0
    e_out = uout + b * p.out.s_b + c * p.out.s_c + k1 * p.out.s_j
    both = np.concatenate([e_out.ravel(), (e_out + 1).ravel()])
    if np.unique(both).size != both.size:
        raise AssertionError(f"pass {pi}: overlapping output addresses")
    _scatter(p.out, bufs, e_out, A, f"pass {pi} store A")
    _scatter(p.out, bufs, e_out + 1, B, f"pass {pi} store B")


def _real4b(p, bufs, start, cnt, pi):
    """Pass B of the real-input four-step (real4b_kernel): rows k1 = 0..n1/2
    of the scratch -> n2-point DFT R -> X[k1 + n1*k2] = R[k1][k2] for the bins
    <= N/2, and X[(n1 - k1) + n1*k2] = conj R[k1][n2 - 1 - k2] for
    0 < k1 < n1/2, k2 < n2/2; every bin 0..N/2 once."""
    n2, n1, N = p.L, p.rr_m, 2 * p.rr_h
    rows = n1 // 2 + 1
    b = np.arange(cnt).reshape(cnt, 1, 1)
    r = np.arange(rows).reshape(1, rows, 1)
    j2 = np.arange(n2).reshape(1, 1, n2)
    uin = start * p.in_.s_u if p.in_.role < 2 else 0
    Y = _gather(p.in_, bufs, uin + b * p.in_.s_b + r * p.in_.s_c + j2 * p.in_.s_j, f"pass {pi} load")
    R = np.fft.fft(Y, axis=-1)
    uout = start * p.out.s_u if p.out.role < 2 else 0
    bb = np.arange(cnt).reshape(cnt, 1, 1)
    k1 = np.arange(rows).reshape(1, rows, 1)
    k2 = np.arange(n2).reshape(1, 1, n2)
    own = (k1 + n1 * k2 <= N // 2) & ((k2 < n2 // 2) | ((k1 == 0) & (k2 == n2 // 2)))
    own = np.broadcast_to(own, R.shape)
    e_own = np.broadcast_to(uout + bb * p.out.s_b + k1 * p.out.s_c + k2 * p.out.s_j, R.shape)
    mir = np.broadcast_to((k1 > 0) & (k1 < n1 // 2) & (k2 < n2 // 2), R.shape)
    e_mir = np.broadcast_to(uout + bb * p.out.s_b + (n1 - k1) * p.out.s_c + k2 * p.out.s_j, R.shape)
    Rm = np.conj(R[..., ::-1])  # conj R[k1][n2 - 1 - k2]
    e_all = np.concatenate([e_own[own], e_mir[mir]])
    if np.unique(e_all).size != e_all.size or e_all.size != cnt * (N // 2 + 1):
        raise AssertionError(f"pass {pi}: {e_all.size} stores, {np.unique(e_all).size} distinct")
    _scatter(p.out, bufs, e_own[own], R[own], f"pass {pi} store")
    _scatter(p.out, bufs, e_mir[mir], Rm[mir], f"pass {pi} store (mirror)")


def _real4c(p, bufs, start, cnt, pi):
    """c2r pass B' of the real-input four-step (real4c_kernel): for columns
    k1 = 0..n1/2 of the spectrum (bin k1 + n1*k2 at k1*in.s_c + k2*in.s_j),
    R[k1][k2] = X[k1 + n1*k2] where that bin is stored (k2 < n2/2, column 0
    up to k2 = n2/2), else conj X[(n1 - k1) + n1*(n2 - 1 - k2)] (column 0:
    conj X[n1*(n2 - k2)]); imaginary parts of bins 0 and N/2 dropped; the
    unnormalised inverse n2-point DFT times W_N^-(j2*k1) stored as Dt[j2][k1]
    (at j2*out.s_j + k1*out.s_c)."""
    n2, n1, N = p.L, p.rr_m, 2 * p.rr_h
    rows = n1 // 2 + 1
    b = np.arange(cnt).reshape(cnt, 1, 1)
    r = np.arange(rows).reshape(1, rows, 1)
    k2 = np.arange(n2).reshape(1, 1, n2)
    own = (k2 < n2 // 2) | ((r == 0) & (k2 == n2 // 2))
    row_bin = np.where(own, r, np.where(r == 0, 0, n1 - r))
    col_bin = np.where(own, k2, np.where(r == 0, n2 - k2, n2 - 1 - k2))
    bins = np.broadcast_to(row_bin + n1 * col_bin, (1, rows, n2))
    if bins.min() < 0 or bins.max() > N // 2 or np.unique(bins).size != N // 2 + 1:
        raise AssertionError(f"pass {pi}: the rows do not read every bin 0..N/2")
    uin = start * p.in_.s_u if p.in_.role < 2 else 0
    e = uin + b * p.in_.s_b + row_bin * p.in_.s_c + col_bin * p.in_.s_j
    R = _gather(p.in_, bufs, e, f"pass {pi} load")
    R = np.where(np.broadcast_to(own, R.shape), R, np.conj(R))
    dc = np.broadcast_to((r == 0) & ((k2 == 0) | (k2 == n2 // 2)), R.shape)
    R = np.where(dc, R.real + 0j, R)
    D = np.fft.ifft(R, axis=-1) * n2
    j2 = np.arange(n2).reshape(1, 1, n2)
    D = D * np.exp(2j * np.pi * ((j2 * r) % N) / N)
    uout = start * p.out.s_u if p.out.role < 2 else 0
    e_out = uout + b * p.out.s_b + r * p.out.s_c + j2 * p.out.s_j
    if np.unique(e_out).size != e_out.size:
        raise AssertionError(f"pass {pi}: overlapping output addresses")
    _scatter(p.out, bufs, e_out, D, f"pass {pi} store")


def _real4d(p, bufs, start, cnt, pi):
    """c2r pass A' (real4d_kernel): packed column c: Z[k1] = Dt[2c][k1] +
    i Dt[2c+1][k1] (k1 > n1/2 from the conjugates at n1 - k1; row 2c at
    c*in.s_c, row 2c+1 in.s_c/2 further, k1 at k1*in.s_j), unnormalised
    inverse n1-point DFT, stored as the reals 2c, 2c+1 of rows j1 (complex
    pairs at c*out.s_c + j1*out.s_j)."""
    n1, n2 = p.L, p.rr_m
    npc = n2 // 2
    b = np.arange(cnt).reshape(cnt, 1, 1)
    c = np.arange(npc).reshape(1, npc, 1)
    k1 = np.arange(n1).reshape(1, 1, n1)
    lo = k1 <= n1 // 2
    kk = np.where(lo, k1, n1 - k1)
    uin = start * p.in_.s_u if p.in_.role < 2 else 0
    e = uin + b * p.in_.s_b + c * p.in_.s_c + kk * p.in_.s_j
    da = _gather(p.in_, bufs, e, f"pass {pi} load")
    db = _gather(p.in_, bufs, e + p.in_.s_c // 2, f"pass {pi} load")
    lo = np.broadcast_to(lo, da.shape)
    da = np.where(lo, da, np.conj(da))
    db = np.where(lo, db, np.conj(db))
    z = np.fft.ifft(da + 1j * db, axis=-1) * n1
    uout = start * p.out.s_u if p.out.role < 2 else 0
    e_out = uout + b * p.out.s_b + c * p.out.s_c + k1 * p.out.s_j
    if np.unique(e_out).size != e_out.size:
        raise AssertionError(f"pass {pi}: overlapping output addresses")
    _scatter(p.out, bufs, e_out, z, f"pass {pi} store")


def _tiny(p, bufs, start, cnt, pi):
    """Rows of 2..32 points, one per thread, staged through LDS (tiny_rows_kernel,
    fft_amd/csrc/afft_tiny.hip): the kernel addresses row g of the launch
    (g < n_col * cnt, every user transform folded into the row index) at
    g * s_c and its entries at +j, so the pass must be a folded row pass with
    adjacent entries; its arithmetic is the row pass's (plain, r2c post-pass
    or c2r pre-pass, re / im exchanged on each side as asked)."""
    if not (p.fold_user and p.in_.s_j == 1 and p.out.s_j == 1 and p.epi in (0, 2, 3) and p.L in (2, 4, 8, 16, 32)
            and not p.colmap and not p.store_t and p.n_blo == 1):
        raise AssertionError(f"pass {pi}: tiny pass outside the kernel's contract")
    if p.epi == 3 and not (p.swap_in and p.swap_out):
        raise AssertionError(f"pass {pi}: the tiny c2r kernel always exchanges re / im on both sides")
    _fft_pass(p, bufs, start, cnt, pi)


def run(kind, dims, fmt, batch, src_re, src_im=None, prec=32, edit=None, variant=None):
    """Executes the plan model; returns the dst arrays (float64).  edit
    (optional) maps the inspected pass list to the one executed; variant
    (optional) selects the plan a kernel-shape variant builds."""
    info, passes = inspect(kind, dims, fmt, batch, prec, variant)
    if edit is not None:
        passes = edit(passes)
    bufs = Buffers(kind, dims, fmt, batch, info)
    if bufs.arrays["src_re"].size != np.asarray(src_re).size:
        raise ValueError("src_re size does not match the API contract")
    bufs.arrays["src_re"][:] = np.asarray(src_re, dtype=np.float64).ravel()
    if src_im is not None:
        bufs.arrays["src_im"][:] = np.asarray(src_im, dtype=np.float64).ravel()
    for start in range(0, batch, info.chunk):
        cnt = min(info.chunk, batch - start)
        for pi, p in enumerate(passes):
            if p.type == 0:
                _fft_pass(p, bufs, start, cnt, pi)
            elif p.type == 1:
                _real_rows(p, bufs, start, cnt, pi)
            elif p.type == 3:
                _real_tr(p, bufs, start, cnt, pi)
            elif p.type == 4:
                _real_pair(p, bufs, start, cnt, pi)
            elif p.type == 5:
                _real_rowfft(p, bufs, start, cnt, pi)
            elif p.type == 6:
                _real_irowfft(p, bufs, start, cnt, pi)
            elif p.type == 7:
                _real_colpair(p, bufs, start, cnt, pi)
            elif p.type == 8:
                _real_icolpair(p, bufs, start, cnt, pi)
            elif p.type == 9:
                _real4a(p, bufs, start, cnt, pi)
            elif p.type == 10:
                _real4b(p, bufs, start, cnt, pi)
            elif p.type == 11:
                _real4c(p, bufs, start, cnt, pi)
            elif p.type == 12:
                _real4d(p, bufs, start, cnt, pi)
            elif p.type == 13:
                _tiny(p, bufs, start, cnt, pi)
            else:
                for u in range(start, start + cnt):
                    if not p.rr_inv:
                        v = bufs.arrays["src_re"][u]
                        if p.out.il:
                            bufs.arrays["dst_re"][2 * u:2 * u + 2] = [v, 0.0]
                        else:
                            bufs.arrays["dst_re"][u] = v
                            bufs.arrays["dst_im"][u] = 0.0
                    else:
                        s = bufs.arrays["src_re"]
                        bufs.arrays["dst_re"][u] = s[2 * u] if p.in_.il else s[u]
    return bufs.arrays["dst_re"], bufs.arrays["dst_im"], info, passes
