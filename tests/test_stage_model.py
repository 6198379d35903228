"""CPU model of one L-point tile of the pass kernel (afft_pass.h
run_stages) and of the stage-twiddle table the device generator writes
(afft_aux.hip gen_stage_tw_kernel).

Thread t of a transform holds elements j = t + m*TPT; stage s of radix R with
Ns = R0*..*R(s-1) multiplies by table[off_s + k*(R-1) + r-1] = W_{Ns*R}^{r*k}
(k = j mod Ns), runs a radix-R DFT and writes entry (j/Ns)*Ns*R + k + r*Ns.
The model replays exactly that, table included, against numpy's FFT, for every
schedule the kernels are compiled with."""
import re
import os

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def schedules():
    txt = open(os.path.join(ROOT, "fft_amd", "csrc", "afft_device.h")).read()
    out = {}
    for m in re.finditer(r"AFGPU_SCHED\((\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*([\d,\s]+)\)", txt):
        L, sv, E, nst = (int(m.group(i)) for i in range(1, 5))
        radix = [int(x) for x in m.group(5).split(",")]
        assert len(radix) == nst and int(np.prod(radix)) == L
        assert all(E % r == 0 for r in radix)
        out[(L, sv)] = (E, radix)
    assert len(out) >= 15
    return out


def stage_table(radix):
    """What gen_stage_tw_kernel writes: stages s >= 1 only, [k][r-1]."""
    tab, offs, ns = [], [], 1
    for s, R in enumerate(radix):
        offs.append(len(tab))
        if s >= 1:
            for k in range(ns):
                for r in range(1, R):
                    tab.append(np.exp(-2j * np.pi * ((r * k) % (ns * R)) / (ns * R)))
        ns *= R
    return np.array(tab, np.complex128), offs


def tile_model(x, E, radix):
    L = x.size
    TPT = L // E
    tab, offs = stage_table(radix)
    v = np.zeros((TPT, E), np.complex128)
    for t in range(TPT):
        for m in range(E):
            v[t, m] = x[t + m * TPT]
    ns = 1
    for s, R in enumerate(radix):
        NB = E // R
        last = s == len(radix) - 1
        lds = np.full(L, np.nan + 0j)
        for t in range(TPT):
            for b in range(NB):
                j = t + b * TPT
                a = np.array([v[t, b + r * NB] for r in range(R)])
                if ns > 1:
                    k = j & (ns - 1)
                    for r in range(1, R):
                        a[r] *= tab[offs[s] + k * (R - 1) + r - 1]
                a = np.fft.fft(a)
                for r in range(R):
                    v[t, b + r * NB] = a[r]
                if not last:
                    base = (j // ns) * ns * R + (j & (ns - 1))
                    for r in range(R):
                        assert np.isnan(lds[base + r * ns].real), "two writes to one LDS entry"
                        lds[base + r * ns] = v[t, b + r * NB]
        if not last:
            assert not np.isnan(lds.real).any(), "LDS entry never written"
            for t in range(TPT):
                for m in range(E):
                    v[t, m] = lds[t + m * TPT]
        ns *= R
    out = np.zeros(L, np.complex128)
    for t in range(TPT):
        for m in range(E):
            out[t + m * TPT] = v[t, m]
    return out


@pytest.mark.parametrize("key", sorted(schedules()), ids=lambda k: f"L{k[0]}_sv{k[1]}")
def test_tile_stages_and_table(key):
    L = key[0]
    E, radix = schedules()[key]
    rng = np.random.default_rng(L)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    y = tile_model(x, E, radix)
    assert np.abs(y - np.fft.fft(x)).max() < 1e-9 * max(1.0, np.abs(x).sum())


def test_table_offsets_match_kernel_formula():
    # StageTwOff<L,S> = sum over 1 <= s' < S of Ns(s') * (R(s') - 1)
    for L, (E, radix) in schedules().items():
        tab, offs = stage_table(radix)
        ns, acc = 1, 0
        for s, R in enumerate(radix):
            assert offs[s] == acc, (L, s)
            if s >= 1:
                acc += ns * (R - 1)
            ns *= R
        assert len(tab) == acc
