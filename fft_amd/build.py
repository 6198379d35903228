"""Builds fft_amd/libafft_mi355x.so from fft_amd/csrc with hipcc for gfx950.

    python -m fft_amd.build [--incremental]       (or fft_amd.build.build())

The four translation units compile in parallel; the FP32 and FP64 pass-kernel
instantiations dominate the build time.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_obj")
LIB = os.path.join(PKG, "libafft_mi355x.so")
SOURCES = ["afft_kernels_f32.hip", "afft_kernels_f64.hip", "afft_aux.hip", "afft_plan.hip", "afft_queue.hip", "afft_mfma.hip",
           "afft_realpair.hip", "afft_fused.hip", "afft_real4.hip", "afft_tiny.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wno-unused-command-line-argument",
    # packed FP32 instructions stay off (see DESIGN notes on numerics)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _up_to_date(src: str, out: str) -> bool:
    """True when the object exists and is newer than every file its last
    compile read (the dependency list hipcc wrote next to it)."""
    dep = out[:-2] + ".d"
    if not (os.path.exists(out) and os.path.exists(dep)):
        return False
    txt = open(dep).read().replace("\\\n", " ")
    files = txt.split(":", 1)[1].split() if ":" in txt else []
    t = os.path.getmtime(out)
    return all(os.path.exists(f) and os.path.getmtime(f) < t for f in files + [os.path.join(CSRC, src)])


def _compile(src: str, incremental: bool = False) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if incremental and _up_to_date(src, out):
        return out
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(PKG, "..", "include"), "-MD", "-MF", out[:-2] + ".d", "-c",
           os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return out


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every translation unit (incremental=True: only those older than
    their source or a header) and links the library and the harness."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    # the afft_run harness (host code; links the library, rpath to its folder)
    harness = os.path.join(PKG, "afft_run")
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wno-unused-result", "-I", os.path.join(PKG, "..", "include"),
           os.path.join(PKG, "harness", "afft_run.cpp"), "-L", PKG, "-lafft_mi355x", "-Wl,-rpath,$ORIGIN",
           "-o", harness + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"afft_run build failed:\n{r.stderr[-4000:]}")
    os.replace(harness + ".tmp", harness)
    # the libbench2 adapter with its stand-alone driver (fft_amd/harness/bench2_afft.cpp)
    bench2 = os.path.join(PKG, "afft_bench2")
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wno-unused-result", "-I", os.path.join(PKG, "..", "include"),
           os.path.join(PKG, "harness", "bench2_afft.cpp"), "-L", PKG, "-lafft_mi355x", "-Wl,-rpath,$ORIGIN",
           "-o", bench2 + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"afft_bench2 build failed:\n{r.stderr[-4000:]}")
    os.replace(bench2 + ".tmp", bench2)
    if verbose:
        print("built", LIB, "and", harness, "and", bench2)
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
