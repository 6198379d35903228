"""GPU tests of the split-precision MFMA modes (FP16 hi+lo and BF16
hi+mid+lo operands, FP32 accumulation) against the FP64 oracle / device FP64
plans, at the same bound as the FP32 path: rel-L2 <= 2*log2(N)*2^-24 per
transform."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def G():
    import gpu_util
    gpu_util.torch_cuda()
    return gpu_util


def test_mfma_operand_layout(G):
    from fft_amd import _lib
    say("mfma selftest")
    assert _lib.lib().afft_gpu_mfma_selftest() == 0


MODES = {1: "[mfma]", 2: "[mfma-bf16]"}


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("n,batch", [(256, 64), (65536, 8)])
def test_mfma_c2c_scaled_amplitudes(G, n, batch, mode):
    import torch
    from fft_amd.afft import Plan
    say("mfma c2c", n, batch, "mode", mode)
    x = O.complex_input(11, (batch, n)).astype(np.complex128)
    amp = 10.0 ** (6 * O.uniform(12, batch).astype(np.float64) / 2)  # 10^u, u in [-3, 3)
    x = x * amp[:, None]
    src = G.to_dev(np.stack([x.real, x.imag], -1).astype(np.float32).ravel())
    x32 = (src.cpu().numpy()[0::2] + 1j * src.cpu().numpy()[1::2]).reshape(batch, n).astype(np.complex128)
    for inverse in (False, True):
        dst = G.empty_dev(2 * n * batch)
        p = Plan("ci" if inverse else "c", (n,), batch)
        p.set_mode(mode)
        assert any(MODES[mode] in d for d in p.describe())
        p(src, None, dst, None)
        torch.cuda.synchronize()
        y = G.to_host(dst)
        y = (y[0::2] + 1j * y[1::2]).reshape(batch, n)
        for b in range(batch):
            ref = O.fft(x32[b], inverse=inverse)
            err = O.rel_l2(y[b].astype(np.complex64), ref)
            assert err <= O.bound(n), (n, b, inverse, err)
        p.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_mfma_3d_256_cube(G, mode):
    import torch
    from fft_amd.afft import Plan
    say("mfma 256^3 mode", mode)
    shape = (256, 256, 256)
    n = 256 ** 3
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(2 * n, device="cuda", generator=g) * 2 - 1
    y = torch.empty_like(x)
    y64 = torch.empty(2 * n, device="cuda", dtype=torch.float64)
    p = Plan("c", shape)
    p.set_mode(mode)
    assert sum(MODES[mode] in d for d in p.describe()) == 3
    p(x, None, y, None)
    Plan("c", shape, dtype=np.float64)(x.double(), None, y64, None)
    torch.cuda.synchronize()
    assert G.device_rel_l2(y, y64) <= O.bound(n)


def test_mfma_mode_rejected_where_it_cannot_apply(G):
    from fft_amd.afft import Plan
    p = Plan("c", (1024,), 4)
    with pytest.raises(RuntimeError):
        p.set_mode(1)
    with pytest.raises(RuntimeError):
        p.set_mode(2)
    with pytest.raises(RuntimeError):
        p.set_mode(3)
