"""The afft_run harness (the reference's run / run_all option set) on the
GPU: test mode at p = 1.2e-7 (run_all.cpp:68) for complex / real, forward /
inverse, both formats, misaligned buffers, N-D shapes, bench mode output, and
a short random sweep."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(__file__), "..", "fft_amd", "afft_run")


def run(*args, timeout=240):
    print("[gpu-case] afft_run", *args, flush=True)
    sys.stdout.flush()
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("flags", [[], ["-i"], ["-r"], ["-r", "-i"]], ids=["c", "ci", "r", "ri"])
@pytest.mark.parametrize("impl", ["fft", "fft-interleaved"])
def test_run_1d_sweep(flags, impl):
    rc, out = run(impl, "1-17", *flags, "-p", "1.2e-7")
    assert rc == 0, out
    assert len(out.strip().splitlines()) == 16


def test_run_nd_double_and_misaligned():
    rc, out = run("fft", "3", "4", "5", "-p", "1.2e-7")
    assert rc == 0, out
    rc, out = run("fft-interleaved", "4", "6", "-r", "-i", "-p", "1.2e-7", "--misaligned-by", "12")
    assert rc == 0, out
    rc, out = run("fft", "10", "-d", "-p", "1e-15")
    assert rc == 0, out
    rc, out = run("fft", "10", "--misaligned-by", "3")
    assert rc == 2


def test_run_bench_prints_gflops():
    rc, out = run("fft-interleaved", "12", "-b", "-n", "1e9")
    assert rc == 0, out
    assert "GFLOPS" in out and "ns" in out


def test_run_all_short_sweep():
    rc, out = run("--all", "fft-interleaved", "-t", "3e7", timeout=600)
    assert rc == 0, out[-3000:]
    assert "0 failures" in out
