"""Host-side argument checks of the Python wrappers (fft_amd/afft.py): buffer
sizes per kind and format, and rejection of host, mistyped, strided or
undersized tensors before anything reaches the GPU.  CPU only."""
import numpy as np
import pytest
import torch

from fft_amd import afft
from fft_amd._lib import afft_interleaved as IL, afft_split as SP


def test_required_counts():
    # c / ci: N complex in and out per transform
    assert afft.required_counts("c", (8, 16), IL, 3) == (2 * 128 * 3, 0, 2 * 128 * 3, 0)
    assert afft.required_counts("ci", (8, 16), SP, 2) == (256, 256, 256, 256)
    # r: N reals in, (N0/2+1)*inner complex out (dim 0 halved, unit dims dropped)
    assert afft.required_counts("r", (1, 8, 16), IL, 1) == (128, 2 * 5 * 16, 0)
    assert afft.required_counts("r", (1024,), SP, 2) == (2048, 2 * 513, 2 * 513)
    assert afft.required_counts("ri", (8, 16), IL, 1) == (2 * 5 * 16, 0, 128)
    assert afft.required_counts("ri", (8, 16), SP, 1) == (80, 80, 128)


class _FakeDev:
    """Stands in for a CUDA tensor (no GPU here)."""

    def __init__(self, n, dtype=torch.float32, contiguous=True):
        self._n, self.dtype, self._c, self.is_cuda = n, dtype, contiguous, True

    def numel(self):
        return self._n

    def is_contiguous(self):
        return self._c

    def data_ptr(self):
        return 0x1000


def test_check_buffers_accepts_and_rejects():
    cnt = afft.required_counts("c", (64,), IL, 1)
    ok = (_FakeDev(128), None, _FakeDev(128), None)
    afft.check_buffers(ok, cnt, np.float32)
    # raw addresses are the caller's to get right, but not null
    afft.check_buffers((0x2000, None, 0x3000, None), cnt, np.float32)
    with pytest.raises(ValueError, match="null"):
        afft.check_buffers((0, None, 0x3000, None), cnt, np.float32)
    with pytest.raises(ValueError, match="not on the GPU"):
        afft.check_buffers((torch.zeros(128), None, _FakeDev(128), None), cnt, np.float32)
    with pytest.raises(ValueError, match="dtype"):
        afft.check_buffers((_FakeDev(128, torch.float64), None, _FakeDev(128), None), cnt, np.float32)
    with pytest.raises(ValueError, match="contiguous"):
        afft.check_buffers((_FakeDev(128), None, _FakeDev(128, contiguous=False), None), cnt, np.float32)
    with pytest.raises(ValueError, match="needs 128"):
        afft.check_buffers((_FakeDev(127), None, _FakeDev(128), None), cnt, np.float32)
    with pytest.raises(ValueError, match="None"):
        afft.check_buffers((_FakeDev(128), None, None, None), cnt, np.float32)
    with pytest.raises(TypeError):
        afft.check_buffers((np.zeros(128, np.float32), None, _FakeDev(128), None), cnt, np.float32)
    # split format needs the imaginary arrays too
    cs = afft.required_counts("c", (64,), SP, 1)
    with pytest.raises(ValueError, match="second"):
        afft.check_buffers((_FakeDev(64), None, _FakeDev(64), _FakeDev(64)), cs, np.float32)
    afft.check_buffers((_FakeDev(64), _FakeDev(64), _FakeDev(64), _FakeDev(64)), cs, np.float32)
