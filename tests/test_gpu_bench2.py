"""The libbench2 adapter (fft_amd/harness/bench2_afft.cpp) run through its
stand-alone driver: libbench2's can_do / setup / doit / done sequence on host
buffers, in place and out of place, forward and backward, 1-D and N-D; the
output it hands back to the harness is checked against the oracle.  The
driver's input is a fixed xorshift sequence, regenerated here."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
EXE = os.path.join(ROOT, "fft_amd", "afft_bench2")


def _input(n):
    st = np.uint64(88172645463325252)
    out = np.empty(2 * n, np.float32)
    m = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        for j in range(2 * n):
            st ^= (st << np.uint64(13)) & m
            st ^= st >> np.uint64(7)
            st ^= (st << np.uint64(17)) & m
            out[j] = np.float32(float(st >> np.uint64(11)) / 9007199254740992.0 * 2.0 - 1.0)
    return out[0::2].astype(np.float64) + 1j * out[1::2]


def test_cannot_do_non_power_of_two():
    if not os.path.exists(EXE):
        pytest.skip("afft_bench2 not built")
    r = subprocess.run([EXE, "-n", "1", "ocf12"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot do" in r.stdout
    r = subprocess.run([EXE, "ocx16"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


@pytest.mark.gpu
@pytest.mark.parametrize("prob", ["ocf1024", "icf1024", "ocb4096", "icb65536", "ocf64x128", "icf16x8x32"])
def test_bench2_adapter_against_oracle(prob, tmp_path):
    out = tmp_path / "out.bin"
    r = subprocess.run([EXE, "-n", "3", "-w", str(out), prob], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "MFLOPS" in r.stdout
    dims = [int(v) for v in prob[3:].split("x")]
    n = int(np.prod(dims))
    x = _input(n).reshape(dims)
    y = np.fromfile(out, dtype=np.float32)
    y = (y[0::2] + 1j * y[1::2]).astype(np.complex64).reshape(dims)
    ref = O.fft(x, inverse=(prob[2] == "b"))
    assert O.rel_l2(y, ref) <= O.bound(n)
