"""Helpers for the rocprofv3 evidence of the discord engine.

    python3 -m paddi_amd.profiles copy                # 1 GiB device-to-device copy (calibration)
    python3 -m paddi_amd.profiles stats CSV OUT.md CMD  # kernel-trace stats -> markdown table
    python3 -m paddi_amd.profiles traffic OUT.json KERNEL FETCH.csv WRITE.csv COPYF.csv COPYW.csv CMD
    python3 -m paddi_amd.profiles util OUT.md TRAFFIC.json BUSY.csv CMD
    python3 -m paddi_amd.profiles counters OUT.md CMD PMC.csv [PMC.csv ...]

`util` combines the traffic summary with a `rocprofv3 --pmc VALUBusy SALUBusy MfmaUtil` run
(the derived metrics, per dispatch) into per-kernel HBM GB/s against the 8 TB/s peak and
duration-weighted VALU / SALU / MFMA busy percentages.

`traffic` turns two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs (KiB per dispatch)
into HBM bytes per dispatch of KERNEL, each counter scaled by the factor that makes the
copy kernel's reading equal to the bytes it is known to move (1 GiB read, 1 GiB written).
bench.py reads the JSON for its roofline `traffic`.
"""
from __future__ import annotations

import csv
import json
import os
import sys
from collections import defaultdict

COPY_BYTES = 1 << 30


def copy():
    import torch
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print("copied", COPY_BYTES, "bytes", flush=True)


def _counter_rows(path, counter):
    """Per-dispatch values of `counter` by kernel (short name)."""
    out = defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"].split("(")[0].replace("void ", "").strip()
            out[name].append(float(r["Counter_Value"]))
    return out


def _mean(v):
    return sum(v) / len(v) if v else 0.0


def _copy_kernel(rows):
    """The largest dispatch of the calibration run (the copy)."""
    best = max(rows.items(), key=lambda kv: max(kv[1]))
    return best[0], max(best[1])


def traffic(out, kernel, fetch_csv, write_csv, copyf_csv, copyw_csv, cmd):
    fetch = _counter_rows(fetch_csv, "FETCH_SIZE")
    write = _counter_rows(write_csv, "WRITE_SIZE")
    cname_f, cf = _copy_kernel(_counter_rows(copyf_csv, "FETCH_SIZE"))
    cname_w, cw = _copy_kernel(_counter_rows(copyw_csv, "WRITE_SIZE"))
    sf = COPY_BYTES / 1024.0 / cf if cf > 0 else None
    sw = COPY_BYTES / 1024.0 / cw if cw > 0 else None
    res = {"command": cmd, "copy_kernel": [cname_f, cname_w],
           "copy_kib": {"FETCH_SIZE": cf, "WRITE_SIZE": cw},
           "scale": {"FETCH_SIZE": sf, "WRITE_SIZE": sw}, "kernels": {}}
    for k in sorted(set(fetch) | set(write)):
        f_kib, w_kib = _mean(fetch.get(k, [])), _mean(write.get(k, []))
        res["kernels"][k] = {
            "dispatches": max(len(fetch.get(k, [])), len(write.get(k, []))),
            "fetch_kib": f_kib, "write_kib": w_kib,
            "bytes_per_dispatch": (f_kib * (sf or 0) + w_kib * (sw or 0)) * 1024.0}
    res["kernel"] = kernel
    with open(out, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    md = os.path.splitext(out)[0] + ".md"
    with open(md, "w") as f:
        f.write(f"# HBM traffic per dispatch (rocprofv3 PMC)\n\nCommand:\n\n```\n{cmd}\n```\n\n")
        f.write(f"Calibration: a {COPY_BYTES >> 20} MiB device copy reads FETCH_SIZE = {cf:.0f} KiB "
                f"and WRITE_SIZE = {cw:.0f} KiB; scale factors {sf:.3f} (fetch), {sw:.3f} (write).\n\n")
        f.write("| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | scaled MB / dispatch |\n")
        f.write("|---|---:|---:|---:|---:|\n")
        for k, v in sorted(res["kernels"].items(), key=lambda kv: -kv[1]["bytes_per_dispatch"]):
            f.write(f"| {k} | {v['dispatches']} | {v['fetch_kib']:.1f} | {v['write_kib']:.1f} | "
                    f"{v['bytes_per_dispatch'] / 1e6:.3f} |\n")
    print("wrote", out, md)


HBM_PEAK_GBS = 8000.0


def util(out, traffic_json, busy_csv, cmd):
    with open(traffic_json) as f:
        tj = json.load(f)
    busy = defaultdict(lambda: defaultdict(float))
    dur = defaultdict(lambda: defaultdict(float))
    calls = defaultdict(int)
    with open(busy_csv) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"].split("(")[0].replace("void ", "").strip()
            d = (float(r["End_Timestamp"]) - float(r["Start_Timestamp"])) * 1e-9
            c = r["Counter_Name"]
            busy[name][c] += float(r["Counter_Value"]) * d
            dur[name][c] += d
            calls[name] += c == "VALUBusy"
    with open(out, "w") as f:
        f.write("# HBM bandwidth and ALU busy per kernel (rocprofv3)\n\nCommands:\n\n```\n"
                f"{tj['command']}\n{cmd}\n```\n\n")
        f.write("HBM GB/s = scaled FETCH_SIZE + WRITE_SIZE bytes per dispatch / mean dispatch "
                "time of the busy run; busy = duration-weighted mean of the derived metrics "
                "VALUBusy, SALUBusy, MfmaUtil (100 = every CU's unit busy for the whole "
                "dispatch).\n\n")
        f.write("| kernel | dispatches | mean ms | HBM MB / dispatch | HBM GB/s | % of 8 TB/s | "
                "VALUBusy % | SALUBusy % | MfmaUtil % |\n|---|---:|---:|---:|---:|---:|---:|---:|---:|\n")
        rows = []
        for k, v in tj["kernels"].items():
            if k not in dur:
                continue
            rows.append((k, v, dur[k].get("VALUBusy", 0.0)))
        for k, v, t in sorted(rows, key=lambda r: -r[2]):
            nd = calls[k]
            mean_s = t / nd if nd else 0.0
            gbs = v["bytes_per_dispatch"] / mean_s / 1e9 if mean_s > 0 else 0.0
            b = {c: busy[k][c] / dur[k][c] if dur[k][c] > 0 else 0.0
                 for c in ("VALUBusy", "SALUBusy", "MfmaUtil")}
            f.write(f"| {k} | {nd} | {mean_s * 1e3:.3f} | {v['bytes_per_dispatch'] / 1e6:.1f} | "
                    f"{gbs:.0f} | {100 * gbs / HBM_PEAK_GBS:.1f} | {b['VALUBusy']:.1f} | "
                    f"{b['SALUBusy']:.1f} | {b['MfmaUtil']:.1f} |\n")
    print("wrote", out)


def counters(out, cmd, *csvs):
    """Per-kernel means per dispatch of the SQ_ counters in one or more `rocprofv3 --pmc`
    runs, and each instruction count per wave (SQ_INSTS_* / SQ_WAVES of the same run)."""
    table = defaultdict(dict)
    for path in csvs:
        per = defaultdict(lambda: defaultdict(list))
        with open(path) as f:
            for r in csv.DictReader(f):
                name = r["Kernel_Name"].split("(")[0].replace("void ", "").strip()
                per[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
        for k, cs in per.items():
            waves = _mean(cs.get("SQ_WAVES", []))
            for c, v in cs.items():
                table[k][c] = (_mean(v), len(v))
                if c.startswith("SQ_INSTS_") and waves > 0:
                    table[k][c + " / wave"] = (_mean(v) / waves, len(v))
    cols = sorted({c for v in table.values() for c in v})
    with open(out, "w") as f:
        f.write("# SQ counters per dispatch (rocprofv3 --pmc)\n\nCommands:\n\n```\n"
                f"{cmd}\n```\n\nMean per dispatch; `/ wave` divides by SQ_WAVES of the same "
                "run.\n\n")
        for k in sorted(table, key=lambda k: -table[k].get("SQ_WAVES", (0, 0))[0]):
            f.write(f"## {k}\n\n| counter | mean per dispatch | dispatches |\n|---|---:|---:|\n")
            for c in cols:
                if c in table[k]:
                    v, n = table[k][c]
                    f.write(f"| {c} | {v:.4g} | {n} |\n")
            f.write("\n")
    print("wrote", out)


def stats(path, out, cmd):
    rows = list(csv.DictReader(open(path)))
    with open(out, "w") as f:
        f.write(f"# Kernel time (rocprofv3 --kernel-trace --stats)\n\nCommand:\n\n```\n{cmd}\n```\n\n")
        f.write("| kernel | calls | total ms | average us | min us | max us | % |\n")
        f.write("|---|---:|---:|---:|---:|---:|---:|\n")
        for r in rows:
            name = r["Name"].split("(")[0].replace("void ", "")
            f.write(f"| {name} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                    f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                    f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |\n")
    print("wrote", out)


if __name__ == "__main__":
    what = sys.argv[1]
    if what == "copy":
        copy()
    elif what == "stats":
        stats(sys.argv[2], sys.argv[3], sys.argv[4])
    elif what == "traffic":
        traffic(*sys.argv[2:9])
    elif what == "util":
        util(*sys.argv[2:6])
    elif what == "counters":
        counters(*sys.argv[2:])
    else:
        raise SystemExit(__doc__)
