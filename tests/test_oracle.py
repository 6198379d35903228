"""Pins the CPU oracle (oracle/afft_oracle.c) before anything is checked
against it.

Each check is chosen so a plausible oracle bug fails it: a flipped sign
(impulse / tone closed forms, inverse round trip), a wrong twiddle index or
stage stride (brute-force DFT), a transposed axis in N-D (separability against
per-axis brute force, 16x32 vs 32x16), wrong real-pass pairing or a dropped
1/2 factor (r2c against the full c2c, c2r round trip), wrong dim-0 halving
(N-D real against the c2c restricted to k0 <= N0/2), and afft's own FP32
output (golden fixtures from the reference build) within 2*log2(N)*2^-24.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_bruteforce_1d():
    for n in [1, 2, 4, 8, 16, 64, 256, 1024]:
        x = O.complex_input(11, (n,)).astype(np.complex128)
        for inv in (False, True):
            y = O.fft(x, inverse=inv)
            ref = O.dft_bruteforce(x, inverse=inv)
            assert np.abs(y - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (n, inv)


def test_closed_forms():
    n = 16
    imp = np.zeros(n, np.complex128)
    imp[1] = 1.0
    y = O.fft(imp)
    # impulse at j0 = 1: X[k] = exp(-2 pi i k / n)
    k = np.arange(n)
    assert np.allclose(y, np.exp(-2j * np.pi * k / n), atol=1e-15)
    assert abs(y[1] - (0.9238795325112867 - 0.3826834323650898j)) < 1e-15
    # constant -> n * delta
    y = O.fft(np.ones(64, np.complex128))
    assert abs(y[0] - 64) < 1e-12 and np.abs(y[1:]).max() < 1e-12
    # single tone at k0 -> n * delta(k - k0); inverse tone likewise
    k0 = 5
    tone = np.exp(2j * np.pi * k0 * np.arange(128) / 128)
    y = O.fft(tone)
    assert abs(y[k0] - 128) < 1e-10 and np.abs(np.delete(y, k0)).max() < 1e-10
    y = O.fft(tone, inverse=True)
    assert abs(y[(-k0) % 128] - 128) < 1e-10


def test_linearity_parseval_roundtrip():
    a = O.complex_input(3, (512,)).astype(np.complex128)
    b = O.complex_input(4, (512,)).astype(np.complex128)
    fa, fb = O.fft(a), O.fft(b)
    assert np.allclose(O.fft(2.5 * a - 1.5j * b), 2.5 * fa - 1.5j * fb, atol=1e-11)
    assert abs(np.sum(np.abs(fa) ** 2) - 512 * np.sum(np.abs(a) ** 2)) < 1e-8
    assert np.allclose(O.fft(fa, inverse=True), 512 * a, atol=1e-10)


def test_nd_separable_against_bruteforce():
    x = O.complex_input(7, (16, 32)).astype(np.complex128)
    y = O.fft(x)
    # brute force along axis 1 then axis 0
    t = np.stack([O.dft_bruteforce(r) for r in x])
    t = np.stack([O.dft_bruteforce(c) for c in t.T]).T
    assert np.abs(y - t).max() < 1e-11
    # shape matters (a transposed-axis bug would pass a square shape)
    assert not np.allclose(O.fft(x.reshape(32, 16)).ravel(), y.ravel())
    x3 = O.complex_input(8, (4, 8, 2)).astype(np.complex128)
    assert np.allclose(O.fft(x3), np.fft.fftn(x3), atol=1e-12)


def test_unit_dims_dropped():
    x = O.complex_input(9, (4, 1, 8)).astype(np.complex128)
    assert np.allclose(O.fft(x).ravel(), O.fft(x.reshape(4, 8)).ravel(), atol=1e-13)


@pytest.mark.parametrize("shape", [(2,), (4,), (8,), (64,), (1024,), (2, 4), (4, 8), (16, 32), (8, 8, 8), (2, 1, 8)])
def test_rfft_is_restricted_c2c(shape):
    x = O.normal(5, int(np.prod(shape))).reshape(shape).astype(np.float64)
    s = O.squeeze_shape(shape)
    full = O.fft(x.reshape(s).astype(np.complex128))
    z = O.rfft(x)
    assert z.shape == tuple(O.half_shape(shape))
    assert np.abs(z - full[: s[0] // 2 + 1]).max() < 1e-11 * max(1.0, np.abs(full).max())


@pytest.mark.parametrize("shape", [(2,), (4,), (16,), (1024,), (2, 4), (16, 32), (8, 8, 8)])
def test_irfft_roundtrip_and_hermitian(shape):
    n = int(np.prod(shape))
    x = O.uniform(6, n).reshape(shape).astype(np.float64)
    back = O.irfft(O.rfft(x), shape)
    assert np.abs(back - n * x).max() < 1e-10 * n
    # on Hermitian input c2r equals the (real) full inverse c2c
    h = O.hermitian_input(3, shape).astype(np.complex128)
    s = O.squeeze_shape(shape)
    full = O.fft(_expand(h, s), inverse=True)
    assert np.abs(O.irfft(h, shape).reshape(s) - full.real).max() < 1e-9 * n


def _expand(h, s):
    """Hermitian extension of a dim-0-halved array to the full shape s."""
    full = np.zeros(s, np.complex128)
    n0 = s[0]
    full[: n0 // 2 + 1] = h
    for k0 in range(n0 // 2 + 1, n0):
        src = h[n0 - k0]
        # X[k0, k'] = conj(X[n0-k0, -k'])
        idx = tuple(np.ix_(*[(-np.arange(d)) % d for d in s[1:]])) if len(s) > 1 else ()
        full[k0] = np.conj(src[idx]) if len(s) > 1 else np.conj(src)
    return full


def test_planted_sign_error_is_caught():
    """A sign mistake of the kind the closed forms guard against: an oracle
    that computes the inverse where the forward is asked must fail the impulse
    check.  The mistake is planted here, on the CPU, in a copy of the call."""
    n = 16
    imp = np.zeros(n, np.complex128)
    imp[1] = 1.0
    wrong = O.fft(imp, inverse=True)
    assert not np.allclose(wrong, np.exp(-2j * np.pi * np.arange(n) / n), atol=1e-6)


def _gold_small():
    return np.load(os.path.join(GOLD, "golden_small.npz"), allow_pickle=False)


def _parse(k):
    kind, shape, fmt, isa = k.split("|")
    return kind, tuple(int(s) for s in shape.split("x")), fmt, isa


def test_oracle_matches_afft_golden():
    g = _gold_small()
    assert len(g.files) > 100
    for k in g.files:
        kind, shape, fmt, isa = _parse(k)
        n = int(np.prod(shape))
        if kind in ("c", "ci"):
            x = O.complex_input(1, shape).astype(np.complex128)
            ref = O.fft(x, inverse=(kind == "ci"))
        elif kind == "r":
            ref = O.rfft(O.uniform(2, n).reshape(shape).astype(np.float64))
        else:
            ref = O.irfft(O.hermitian_input(3, shape).astype(np.complex128), shape)
        err = O.rel_l2(g[k], ref)
        assert err <= O.bound(n), (k, err)
        # afft sits well inside the bound; a broken oracle would not
        if n >= 64:
            assert err <= 0.5 * O.bound(n), (k, err)


def test_oracle_matches_afft_large_summaries():
    g = np.load(os.path.join(GOLD, "golden_large.npz"), allow_pickle=False)
    keys = sorted({k.rsplit("|", 1)[0] for k in g.files})
    assert keys
    for k in keys:
        kind, shape, fmt, isa = _parse(k)
        n = int(np.prod(shape))
        if n > (1 << 20):
            continue  # larger shapes are compared on the GPU against the device FP64 path
        if kind == "c":
            ref = O.fft(O.complex_input(1, shape).astype(np.complex128))
        else:
            ref = O.rfft(O.normal(5, n).reshape(shape).astype(np.float64))
        idx = g[k + "|idx"]
        samp = g[k + "|samples"]
        flat = ref.ravel()
        e = np.abs(samp - flat[idx]).max() / np.abs(flat).max()
        assert e < 4 * O.bound(n), (k, e)


@pytest.mark.skipif(not O.ref_available(), reason="reference build not present")
def test_oracle_against_reference_build_ndim_and_split():
    R = O.RefAfft()
    for shape in [(4, 8), (2, 16), (8, 4, 2), (32,), (2,)]:
        n = int(np.prod(shape))
        x = O.normal(5, n).reshape(shape)
        for fmt in (O.AFFT_INTERLEAVED, O.AFFT_SPLIT):
            z = R.r2c(x, fmt=fmt, impl=O.AFFT_AVX2)
            assert O.rel_l2(z, O.rfft(x.astype(np.float64))) <= O.bound(n)
            h = O.hermitian_input(3, shape)
            w = R.c2r(h, shape, fmt=fmt, impl=O.AFFT_AVX2)
            assert O.rel_l2(w, O.irfft(h.astype(np.complex128), shape)) <= O.bound(n)


# ---- the error metrics --------------------------------------------------------
# rel-L2 = ||y - ref|| / ||ref|| and the reference harness metric
# sqrt(sum (d1 - d0)^2 / sum (d0^2 + d1^2)) (testing.cpp:1113-1127) decide
# every pass / fail below, so they are pinned by cases whose value the
# mathematics fixes exactly, and each check is shown to reject the plausible
# slips (a dropped square root, numerator and denominator swapped, a
# difference that is never accumulated, a denominator over the candidate
# instead of the reference).


def _exact_pair(n=4096):
    """ref: small integers (exact in float32 and float64); y = ref * (1 + 2^-10),
    also exact in float32.  rel-L2(y, ref) is then exactly 2^-10."""
    rng = np.random.default_rng(17)
    ref = rng.integers(-1000, 1000, size=2 * n).astype(np.float64)
    ref[ref == 0] = 7.0
    y = (ref * (1.0 + 2.0 ** -10)).astype(np.float32)
    assert np.array_equal(y.astype(np.float64), ref * (1.0 + 2.0 ** -10))
    c = lambda a: a.view(np.complex64 if a.dtype == np.float32 else np.complex128)
    return c(y), c(ref)


def _rel_l2_checks(rel):
    """Properties of rel-L2 that fix its value; `rel(y, ref)` takes a complex64
    or complex128 candidate and a complex128 reference."""
    y, ref = _exact_pair()
    assert abs(rel(y, ref) - 2.0 ** -10) <= 1e-15
    assert rel(ref.astype(np.complex64), ref) == 0.0
    assert abs(rel(np.zeros_like(y), ref) - 1.0) <= 1e-15
    assert abs(rel((-ref).astype(np.complex64), ref) - 2.0) <= 1e-15
    assert abs(rel((3 * ref).astype(np.complex64), ref) - 2.0) <= 1e-15
    # one component off by d: sqrt(d^2 / sum ref^2)
    z = ref.astype(np.complex64).copy()
    z[5] += np.float32(0.5)
    assert abs(rel(z, ref) - 0.5 / np.sqrt(np.sum(np.abs(ref) ** 2))) <= 1e-15
    # scale invariance in the reference's magnitude
    assert abs(rel((1024 * y).astype(np.complex64), 1024 * ref) - 2.0 ** -10) <= 1e-15
    # random data against numpy's norms
    rng = np.random.default_rng(3)
    r = rng.standard_normal(3000) + 1j * rng.standard_normal(3000)
    c = (r + 1e-3 * (rng.standard_normal(3000) + 1j * rng.standard_normal(3000))).astype(np.complex64)
    want = np.linalg.norm(c.astype(np.complex128) - r) / np.linalg.norm(r)
    assert abs(rel(c, r) - want) <= 1e-12 * want


def test_rel_l2_pinned():
    _rel_l2_checks(O.rel_l2)
    # the FP64-candidate entry point (orc_rel_l2_d) on the same cases
    _rel_l2_checks(lambda y, ref: O.rel_l2(np.asarray(y).astype(np.complex128), ref))


def test_rel_l2_checks_reject_planted_mistakes():
    """Each plausible slip, planted on the CPU around the oracle's own metric,
    fails the checks above."""
    ident = lambda y, ref: O.rel_l2(y, ref)
    mutants = {
        "dropped sqrt": lambda y, ref: ident(y, ref) ** 2,
        "num/den swapped": lambda y, ref: 1.0 / ident(y, ref) if ident(y, ref) else np.inf,
        "never accumulated": lambda y, ref: 0.0,
        "normalised by the candidate": lambda y, ref: (
            ident(y, ref) * np.linalg.norm(ref) / max(np.linalg.norm(np.asarray(y, np.complex128)), 1e-300)),
        "abs error (no denominator)": lambda y, ref: ident(y, ref) * np.linalg.norm(ref),
    }
    for name, m in mutants.items():
        with pytest.raises(AssertionError):
            _rel_l2_checks(m)
        assert name


def _harness_checks(h):
    """testing.cpp:1113-1127: sqrt(sum (d1-d0)^2 / sum (d0^2 + d1^2)) with d0
    the reference output and d1 the candidate."""
    y, ref = _exact_pair()
    e = 2.0 ** -10
    # d1 = d0 (1 + e): e / sqrt(1 + (1 + e)^2)
    assert abs(h(y, ref) - e / np.sqrt(1.0 + (1.0 + e) ** 2)) <= 1e-15
    assert h(ref.astype(np.complex64), ref) == 0.0
    assert abs(h(np.zeros_like(y), ref) - 1.0) <= 1e-15
    # d1 = -d0: sqrt(4 / 2)
    assert abs(h((-ref).astype(np.complex64), ref) - np.sqrt(2.0)) <= 1e-15
    # d1 = 3 d0: sqrt(4 / 10)
    assert abs(h((3 * ref).astype(np.complex64), ref) - np.sqrt(0.4)) <= 1e-15
    # small errors: the harness metric is rel-L2 / sqrt(2) to first order
    rng = np.random.default_rng(4)
    r = rng.standard_normal(2000) + 1j * rng.standard_normal(2000)
    c = (r + 1e-4 * (rng.standard_normal(2000) + 1j * rng.standard_normal(2000))).astype(np.complex64)
    assert abs(h(c, r) * np.sqrt(2.0) / O.rel_l2(c, r) - 1.0) < 1e-3


def test_harness_error_pinned():
    _harness_checks(O.harness_error)


def test_harness_checks_reject_planted_mistakes():
    ident = lambda y, ref: O.harness_error(y, ref)
    mutants = {
        "dropped sqrt": lambda y, ref: ident(y, ref) ** 2,
        "rel-L2 instead": lambda y, ref: O.rel_l2(y, ref),
        "never accumulated": lambda y, ref: 0.0,
        "denominator without d1": lambda y, ref: O.rel_l2(y, ref),
        "num/den swapped": lambda y, ref: 1.0 / ident(y, ref) if ident(y, ref) else np.inf,
    }
    for name, m in mutants.items():
        with pytest.raises(AssertionError):
            _harness_checks(m)
        assert name
