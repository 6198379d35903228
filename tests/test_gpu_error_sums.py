"""The device FP64 error reduction (afft_gpu_error_sums_f32 / _f64) behind every
full-size check, pinned against the CPU oracle's rel-L2 and against sums the
mathematics fixes exactly.

acc[0] = sum (y - ref)^2 and acc[1] = sum ref^2; the checks use cases where
both are known in closed form (a uniform relative error of 2^-10, one planted
difference in a large array that spans every block of the reduction), and
random data against the oracle's long-double rel-L2."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _sums(y_dev, ref_dev):
    import gpu_util as G
    torch = G.torch_cuda()
    from fft_amd import _lib
    acc = torch.full((2,), 123.0, dtype=torch.float64, device="cuda")  # must be cleared by the call
    fn = _lib.lib().afft_gpu_error_sums_f32 if y_dev.dtype == torch.float32 else _lib.lib().afft_gpu_error_sums_f64
    assert fn(int(y_dev.data_ptr()), int(ref_dev.data_ptr()), y_dev.numel(), int(acc.data_ptr()), None) == 0
    torch.cuda.synchronize()
    return acc.cpu().numpy()


@pytest.mark.parametrize("ydt", [np.float32, np.float64])
def test_uniform_relative_error_exact(ydt):
    import gpu_util as G
    rng = np.random.default_rng(17)
    ref = rng.integers(-1000, 1000, size=1 << 16).astype(np.float64)
    ref[ref == 0] = 7.0
    y = (ref * (1.0 + 2.0 ** -10)).astype(ydt)
    a = _sums(G.to_dev(y), G.to_dev(ref))
    # every term is (ref * 2^-10)^2 and ref^2: integers times powers of two,
    # summed exactly in FP64 at these magnitudes
    assert a[1] == np.sum(ref * ref)
    assert a[0] == np.sum((ref * 2.0 ** -10) ** 2)
    assert np.sqrt(a[0] / a[1]) == 2.0 ** -10


@pytest.mark.parametrize("ydt", [np.float32, np.float64])
def test_one_planted_difference_large(ydt):
    """2^25 values (more than the reduction's 4096 x 256 threads cover in one
    sweep): zero error except one value, at the far end, off by exactly 0.5."""
    import gpu_util as G
    torch = G.torch_cuda()
    n = 1 << 25
    ref = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.ones(n, dtype=torch.float32 if ydt == np.float32 else torch.float64, device="cuda")
    y[n - 3] = 1.5
    a = _sums(y, ref)
    assert a[0] == 0.25
    assert a[1] == float(n)


def test_random_against_oracle():
    import gpu_util as G
    x = O.complex_input(9, (1 << 18,)).astype(np.complex128)
    y = (x + 1e-6 * O.complex_input(10, (1 << 18,))).astype(np.complex64)
    a = _sums(G.to_dev(y.view(np.float32)), G.to_dev(x.view(np.float64)))
    want = O.rel_l2(y, x)
    assert abs(np.sqrt(a[0] / a[1]) - want) <= 1e-12 * want
    assert abs(G.device_rel_l2(G.to_dev(y.view(np.float32)), G.to_dev(x.view(np.float64))) - want) <= 1e-12 * want
