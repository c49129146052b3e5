"""Summarise rocprofv3 CSV output into short markdown tables.

    python -m seismic_amd.profsum trace   <kernel_trace.csv>            per (kernel, grid) durations
    python -m seismic_amd.profsum pmc     <counter_collection.csv> NAME per (kernel, grid) counter means
    python -m seismic_amd.profsum traffic <probe.log> <FETCH_SIZE csv> <WRITE_SIZE csv>
                                          HBM bytes per launch of each kernel_probe pass (JSON)
    python -m seismic_amd.profsum l2      <probe.log> <TCC_REQ_sum+TCC_HIT_sum csv> <TCC_READ_sum+TCC_WRITE_sum csv>
                                          L2 bytes and hit rate per launch of each pass (JSON)
    python -m seismic_amd.profsum counters <probe.log> <counter csv>
                                          mean of every counter per dispatch of each pass (JSON)
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


# algorithmic bytes per cell-update of each pass (the same table as bench.py)
ALG = {"linesearch_forward": 20, "saving_forward": 24, "tangent": 40, "born": 32, "adjoint": 28}


def _pattern(name, e):
    """Kernel-name pattern of a kernel_probe pass with config entry e."""
    rx, rz = e["window"]
    if e["strips"] == 2:   # row-streamed kernels: k_bns / k_ads<threads, levels>, k_fws<threads, levels, tan, hist>
        k = e["k"]
        return {"linesearch_forward": rf"k_fws<{rx}, {k}, false, false>",
                "saving_forward": rf"k_fws<{rx}, {k}, false, true>",
                "tangent": rf"k_fws<{rx}, {k}, true, false>",
                "born": rf"k_bns<{rx}, {k}>",
                "adjoint": rf"k_ads<{rx}, {k}>"}[name]
    st = "true" if e["strips"] else "false"
    return {"linesearch_forward": rf"k_fwd<{rx}, {rz}, \d+, false, {st}, false>",
            "saving_forward": rf"k_fwd<{rx}, {rz}, \d+, false, {st}, true>",
            "tangent": rf"k_fwd<{rx}, {rz}, \d+, true, {st}, false>",
            "born": rf"k_born<{rx}, {rz}, \d+, {st}>",
            "adjoint": rf"k_adj<{rx}, {rz}, \d+, {st}>"}[name]


def _probe_info(probe_log):
    import json
    return json.loads([line for line in open(probe_log) if line.startswith("{")][-1])


def traffic(probe_log, fetch_csv, write_csv):
    """Per pass of seismic_amd.kernel_probe: mean FETCH_SIZE / WRITE_SIZE per
    dispatch of the pass's kernel (matched by name from the probe's configs),
    scaled to bytes by the probe's 1 GiB calibration copy (the largest
    copyBuffer dispatch of each run)."""
    import json
    import re
    info = _probe_info(probe_log)
    alg = ALG
    pattern = _pattern

    def load(path):
        rows = list(csv.DictReader(open(path)))
        copy = max(float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"])
        return rows, copy

    fr, fcopy = load(fetch_csv)
    wr, wcopy = load(write_csv)
    gib_kib = info["calibration_copy_bytes"] / 1024.0
    out = {"read_factor": round(gib_kib / fcopy, 4), "write_factor": round(gib_kib / wcopy, 4)}
    for name in alg:
        e = dict(info[name])
        pat = re.compile(pattern(name, e))
        fv = [float(r["Counter_Value"]) for r in fr if pat.search(r["Kernel_Name"])]
        wv = [float(r["Counter_Value"]) for r in wr if pat.search(r["Kernel_Name"])]
        if not fv or not wv:
            continue
        e["fetch_kib"] = sum(fv) / len(fv)
        e["write_kib"] = sum(wv) / len(wv)
        e["dispatches"] = len(fv)
        e["bytes_per_launch"] = int(round((e["fetch_kib"] * out["read_factor"] + e["write_kib"] * out["write_factor"])
                                          * 1024))
        e["alg_bytes_per_launch"] = alg[name] * e["cell_updates_per_launch"]
        e["traffic_over_alg"] = round(e["bytes_per_launch"] / e["alg_bytes_per_launch"], 3)
        out[name] = e
    print(json.dumps(out, indent=1))


def l2(probe_log, req_csv, rw_csv):
    """Per pass of seismic_amd.kernel_probe: L2 (TCC) read and write
    requests, bytes and hit rate per dispatch.  Bytes per read and per write
    request come from the 1 GiB calibration copy, whose reads and writes all
    go through the L2 (TCC_READ_sum and TCC_WRITE_sum of its dispatch); the
    hit rate is TCC_HIT_sum / TCC_REQ_sum from the other run."""
    import json
    import re
    info = _probe_info(probe_log)

    def load(path):
        rows = list(csv.DictReader(open(path)))
        by = defaultdict(list)
        for r in rows:
            by[r.get("Counter_Name", "")].append(r)
        return by

    req, rw = load(req_csv), load(rw_csv)
    copy = lambda rows: max(float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"])
    nbytes = info["calibration_copy_bytes"]
    rb, wb = nbytes / copy(rw["TCC_READ_sum"]), nbytes / copy(rw["TCC_WRITE_sum"])
    out = {"bytes_per_read_request": round(rb, 2), "bytes_per_write_request": round(wb, 2)}
    for name in ALG:
        if name not in info:
            continue
        e = dict(info[name])
        pat = re.compile(_pattern(name, e))
        mean = lambda rows: (lambda v: sum(v) / len(v) if v else None)(
            [float(r["Counter_Value"]) for r in rows if pat.search(r["Kernel_Name"])])
        rd, wr = mean(rw["TCC_READ_sum"]), mean(rw["TCC_WRITE_sum"])
        rq, ht = mean(req.get("TCC_REQ_sum", [])), mean(req.get("TCC_HIT_sum", []))
        if rd is None or wr is None:
            continue
        e["l2_read_requests"], e["l2_write_requests"] = rd, wr
        e["l2_bytes_per_launch"] = int(round(rd * rb + wr * wb))
        if rq and ht is not None:
            e["l2_hit_rate"] = round(ht / rq, 4)
        out[name] = e
    print(json.dumps(out, indent=1))


def counters(probe_log, path):
    """Per pass of seismic_amd.kernel_probe: the mean of every counter in
    the run per dispatch of the pass's kernel."""
    import json
    import re
    info = _probe_info(probe_log)
    rows = list(csv.DictReader(open(path)))
    names = sorted({r["Counter_Name"] for r in rows})
    out = {}
    for name in ALG:
        if name not in info:
            continue
        e = dict(info[name])
        pat = re.compile(_pattern(name, e))
        for c in names:
            v = [float(r["Counter_Value"]) for r in rows if r["Counter_Name"] == c and pat.search(r["Kernel_Name"])]
            if v:
                e[c] = sum(v) / len(v)
                e["dispatches"] = len(v)
        out[name] = e
    print(json.dumps(out, indent=1))


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
    elif sys.argv[1] == "traffic":
        traffic(sys.argv[2], sys.argv[3], sys.argv[4])
    elif sys.argv[1] == "l2":
        l2(sys.argv[2], sys.argv[3], sys.argv[4])
    elif sys.argv[1] == "counters":
        counters(sys.argv[2], sys.argv[3])
    else:
        pmc(sys.argv[2], sys.argv[3])
