"""Turns rocprofv3 CSV output into the short summaries kept under profiles/.

    python -m fft_amd.profile_summary stats <dir>/<name>_kernel_stats.csv
        markdown table of the kernel-trace statistics (name, calls, avg/min/max ns, %)

    python -m fft_amd.profile_summary pmc --fetch D1 --write D2 --copy-fetch D3 --copy-write D4
        --kernel REGEX --describe "pass description" --launch-transforms K
        --algo-bytes B --copy-bytes C --command "..." --out profiles/<name>.json
        HBM traffic per launch of the kernel matching REGEX from a FETCH_SIZE run
        (D1) and a WRITE_SIZE run (D2), each scaled by the factor that makes
        the same counter read the known size C of a plain device copy (D3, D4:
        C bytes read and C bytes written per call).

    python -m fft_amd.profile_summary ktrace <dir>/<name>_kernel_stats.csv --kernel REGEX
        --describe "pass description" --command "..." --out profiles/<name>_ktrace.json
        average duration of the kernel matching REGEX in a kernel-trace run

The pmc summary is what bench.py reads to fill roofline.traffic when the
dominant pass of the current plan matches `describe` and `launch_transforms`;
the ktrace summary gives roofline.rocprof_avg_us for the same kernel.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
import sys


def _counter_rows(d: str):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {d}")
    rows = []
    for f in files:
        with open(f) as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def counter_mean(d: str, counter: str, kernel_regex: str):
    vals = [float(r["Counter_Value"]) for r in _counter_rows(d)
            if r["Counter_Name"] == counter and re.search(kernel_regex, r["Kernel_Name"])]
    if not vals:
        raise ValueError(f"no {counter} rows for /{kernel_regex}/ in {d}")
    return sum(vals) / len(vals), len(vals)


def stats_table(path: str) -> str:
    with open(path) as fh:
        rows = list(csv.DictReader(fh))
    out = ["| kernel | calls | avg ns | min ns | max ns | % time |", "|---|---|---|---|---|---|"]
    for r in rows:
        name = r["Name"]
        name = name if len(name) <= 110 else name[:107] + "..."
        out.append(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']):.0f} | {float(r['MinNs']):.0f} | "
                   f"{float(r['MaxNs']):.0f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def ktrace_summary(a) -> dict:
    with open(a.csv) as fh:
        rows = [r for r in csv.DictReader(fh) if re.search(a.kernel, r["Name"])]
    if len(rows) != 1:
        raise ValueError(f"{len(rows)} kernels match /{a.kernel}/ in {a.csv}")
    r = rows[0]
    return {"kernel_regex": a.kernel, "kernel": r["Name"], "describe": a.describe, "calls": int(r["Calls"]),
            "avg_ns": float(r["AverageNs"]), "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"]),
            "command": a.command}


def pmc_summary(a) -> dict:
    kib = 1024.0
    fetch, nf = counter_mean(a.fetch, "FETCH_SIZE", a.kernel)
    write, nw = counter_mean(a.write, "WRITE_SIZE", a.kernel)
    cf, _ = counter_mean(a.copy_fetch, "FETCH_SIZE", "copy16")
    cw, _ = counter_mean(a.copy_write, "WRITE_SIZE", "copy16")
    ff = a.copy_bytes / (cf * kib)
    wf = a.copy_bytes / (cw * kib)
    traffic = fetch * kib * ff + write * kib * wf
    return {
        "kernel_regex": a.kernel,
        "describe": a.describe,
        "launch_transforms": a.launch_transforms,
        "dispatches": {"fetch": nf, "write": nw},
        "fetch_kib_per_dispatch": round(fetch, 1),
        "write_kib_per_dispatch": round(write, 1),
        "copy_bytes": a.copy_bytes,
        "copy_fetch_kib": round(cf, 1),
        "copy_write_kib": round(cw, 1),
        "fetch_factor": round(ff, 4),
        "write_factor": round(wf, 4),
        "traffic_bytes_per_launch": int(traffic),
        "algorithmic_bytes_per_launch": int(a.algo_bytes),
        "traffic_over_algorithmic": round(traffic / a.algo_bytes, 4),
        "command": a.command,
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    p = sub.add_parser("pmc")
    for k in ("fetch", "write", "copy_fetch", "copy_write", "kernel", "describe", "command", "out"):
        p.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    p.add_argument("--launch-transforms", type=int, required=True)
    p.add_argument("--algo-bytes", type=float, required=True)
    p.add_argument("--copy-bytes", type=float, required=True)
    k = sub.add_parser("ktrace")
    k.add_argument("csv")
    for f in ("kernel", "describe", "command", "out"):
        k.add_argument("--" + f, dest=f, required=True)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats_table(a.csv))
    elif a.cmd == "ktrace":
        js = ktrace_summary(a)
        with open(a.out, "w") as fh:
            json.dump(js, fh, indent=1)
        print(json.dumps(js, indent=1))
    else:
        js = pmc_summary(a)
        with open(a.out, "w") as fh:
            json.dump(js, fh, indent=1)
        print(json.dumps(js, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
