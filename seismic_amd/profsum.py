"""Summarise rocprofv3 CSV output into short markdown tables.

    python -m seismic_amd.profsum trace  <kernel_trace.csv>            per (kernel, grid) durations
    python -m seismic_amd.profsum pmc    <counter_collection.csv> NAME per (kernel, grid) counter means
"""
from __future__ import annotations

import csv
import sys
from collections import defaultdict


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "")
    return name.split("(")[0][:60]


def trace(path):
    """Per (kernel, grid, duration octave): persistent kernels keep one grid
    for every batch size, so launches of different plans are told apart by
    their duration band."""
    rows = list(csv.DictReader(open(path)))
    agg = defaultdict(list)
    for r in rows:
        grid = r.get("Grid_Size_X") or r.get("Grid_Size") or r.get("Grid_X") or "?"
        d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        band = max(0, d).bit_length()
        agg[(_short(r["Kernel_Name"]), grid, band)].append(d)
    total = sum(sum(v) for v in agg.values())
    print("| kernel | grid (threads) | duration band (us) | calls | avg us | min us | total ms | share |")
    print("|---|---|---|---|---|---|---|---|")
    for (k, gx, b), v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        if sum(v) < 0.002 * total:
            continue
        lo, hi = (1 << (b - 1)) / 1e3 if b else 0.0, (1 << b) / 1e3
        print(f"| {k} | {gx} | {lo:.1f}-{hi:.1f} | {len(v)} | {sum(v)/len(v)/1e3:.2f} | {min(v)/1e3:.2f} | "
              f"{sum(v)/1e6:.2f} | {sum(v)/total:.1%} |")


def pmc(path, counter):
    rows = list(csv.DictReader(open(path)))
    agg = defaultdict(list)
    for r in rows:
        if r.get("Counter_Name") != counter:
            continue
        grid = r.get("Grid_Size") or r.get("Grid_Size_X") or "?"
        agg[(_short(r["Kernel_Name"]), grid)].append(float(r["Counter_Value"]))
    print(f"| kernel | grid (threads) | dispatches | mean {counter} (KiB) |")
    print("|---|---|---|---|")
    for (k, gx), v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        print(f"| {k} | {gx} | {len(v)} | {sum(v)/len(v):.1f} |")


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        trace(sys.argv[2])
    else:
        pmc(sys.argv[2], sys.argv[3])
