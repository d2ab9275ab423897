"""Throughput of the discord engine on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3] [--lengths L]

A step is one complete discord search over the workload (every length minL..maxL, top-k
per length), from the host call to the returned discord set, with the series already
resident in HBM.  The default workload is the 8,388,608-point random walk with 20 planted
bursts, m = 128..256 (129 lengths), top-3.  Both GPU forms run: the plain form (per-length
launches, host r-descent as src/main.cu:136-233) and the primary form (batched launches
with the learned r proposal when built); `value` is the primary form's throughput and
both must return the identical discord set.

Printed: one JSON line with the metric, the roofline of the dominant kernel (algorithmic
bytes per launch / hipEvent launch time, against 8 TB/s HBM3E) and the CPU baseline: the
oracle (the reference CPU path restated) on the workload's own series at six sampled
lengths, each at its final range of the GPU run, on a declared sample of rows (cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_OPS = 256 * 4 * 16 * 2.4e9  # CUs x SIMDs x lanes x clock: one VALU op per lane
OPS_PER_PAIR = 5.5                     # recurrence + compare + max per pair (scan inner loop)

WORKLOAD_NAMES = {
    "c1": "rw65k_m128_top3",
    "c2": "ecg1m_m64-128_top3",
    "c3": "rw8m_m128-256_top3",
    "c4": "ar2mix2m_m256-512s4_top3",
    "c5": "uneven1024_m64-96_top1",
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="c3", choices=sorted(WORKLOAD_NAMES))
    p.add_argument("--lengths", type=int, default=0, help="limit the number of lengths (debug)")
    p.add_argument("--primary", default="auto", help="plain|batched|learned|auto")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-plain", action="store_true", help="skip the plain form (profiling)")
    p.add_argument("--no-tuned", action="store_true", help="ignore the autotuner's table")
    p.add_argument("--no-gain", action="store_true",
                   help="skip the batched form without the learned proposal")
    return p.parse_args()


def same_set(a, b, rtol=1e-4):
    if len(a) != len(b):
        return False
    for x, y in zip(a, b):
        if (x.index, x.length, x.series) != (y.index, y.length, y.series):
            return False
        if abs(x.nn_dist2 - y.nn_dist2) > rtol * max(abs(y.nn_dist2), 1e-30):
            return False
    return True


def run_form(ctx, handles, w, maxL, mode, steps, warmup):
    for _ in range(warmup):
        ctx.discords_series(handles, w["minL"], maxL, w["step"], w["k"], mode=mode)
    times, stats, recs = [], [], None
    for _ in range(steps):
        t0 = time.perf_counter()
        recs, st = ctx.discords_series(handles, w["minL"], maxL, w["step"], w["k"], mode=mode,
                                       with_stats=True)
        times.append(time.perf_counter() - t0)
        stats.append(st)
    return recs, times, stats


# CPU baseline sample per workload: (row-run stride, number of series sampled or None = all)
CPU_SAMPLE = {"c1": (1, None), "c2": (1, None), "c3": (4, None), "c4": (8, None), "c5": (1, 8)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def sample_lengths(w):
    """The first three lengths and three more evenly spaced over the rest (the last one
    included)."""
    ms = list(range(w["minL"], w["maxL"] + 1, w["step"]))
    if len(ms) <= 6:
        return ms
    rest = ms[3:]
    pick = [rest[round((q + 1) * (len(rest) - 1) / 3)] for q in range(3)]
    return ms[:3] + pick


def cpu_baseline(w, recs, workload):
    """The oracle (C restatement of find_discords_cpu, src/cpu_kernels/cpu_kernels.cpp:7-78,
    FP64, reference scan order, OpenMP over rows) on the workload's own series: per sampled
    length, the statistics of every row and the range pass at that length's final range of
    the GPU run (the primary form's, read from its records), the pass on every stride-th run
    of 256 rows against all columns and its time scaled by the sampled fraction of rows."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    stride, nser = CPU_SAMPLE[workload]
    series = w["series"]
    sel = list(range(len(series)))
    if nser is not None and nser < len(series):
        import numpy as np
        sel = sorted(int(q) for q in np.random.default_rng(5).choice(len(series), nser,
                                                                     replace=False))
    final_r2 = {}
    for r in recs:
        final_r2[(r.series, r.length)] = r.range2
    per, units, est = [], 0.0, 0.0
    for m in sample_lengths(w):
        for q in sel:
            x = series[q]
            r2 = final_r2.get((q, m), 0.0)
            rows, cand, nn, ts, tscan = O.range_discords_sampled(x, m, r2, 256, stride, threads)
            frac = len(rows) / (len(x) - m + 1)
            e = ts + tscan / frac
            units += len(x)
            est += e
            per.append({"series": q, "m": m, "r2": r2, "rows_sampled": int(len(rows)),
                        "stats_s": round(ts, 3), "scan_s": round(tscan, 3),
                        "estimated_s": round(e, 3)})
            print(f"[bench] cpu baseline: series {q} m={m} r2={r2:.5g}: scan of "
                  f"{len(rows)} rows {tscan:.2f} s -> {e:.2f} s for all rows", file=sys.stderr,
                  flush=True)
    ms = sample_lengths(w)
    return {"value": units / est, "unit": "points*lengths/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "nproc": os.cpu_count(), "seconds_estimated": est,
            "sample": f"{len(sel)} of {len(series)} series; lengths {ms}; final attempt only "
                      f"(each length's final range of the GPU run; the reference's own "
                      f"r-descent would also run its earlier attempts, and at a lower range); "
                      f"statistics of all rows, range pass on every {stride}th run of 256 rows "
                      f"against all columns, extrapolated to all rows; FP64, reference scan "
                      f"order, {threads} threads",
            "per_length": per}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); "
              "multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)
    from paddi_amd import synth
    from paddi_amd.engine import Context, load_library

    w = synth.WORKLOADS[a.workload]()
    series = w["series"]
    nlen_full = (w["maxL"] - w["minL"]) // w["step"] + 1
    nlen = min(nlen_full, a.lengths) if a.lengths > 0 else nlen_full
    maxL = w["minL"] + (nlen - 1) * w["step"]
    points = sum(len(x) for x in series)
    units = points * nlen

    L = load_library()
    ctx = Context(0)
    handles = [ctx.upload(x) for x in series]
    if not a.no_tuned:
        ctx.apply_tuned(w["minL"], nseries=len(w["series"]))
    tunables = ctx.get_tunables()

    primary = a.primary
    if primary == "auto":
        primary = "learned" if getattr(L, "paddi_has_mode", None) and L.paddi_has_mode(2) else \
            ("batched" if getattr(L, "paddi_has_mode", None) and L.paddi_has_mode(1) else "plain")

    print(f"[bench] workload {a.workload}: {len(series)} series, {points} points, "
          f"{nlen} lengths, top-{w['k']}; primary form: {primary}", file=sys.stderr, flush=True)
    if primary != "plain":
        recs, times, stats = run_form(ctx, handles, w, maxL, primary, a.steps, a.warmup)
        print(f"[bench] {primary}: {[round(t, 3) for t in times]} s", file=sys.stderr,
              flush=True)
    if a.no_plain and primary != "plain":
        plain_recs, plain_t, plain_st = recs, [float("nan")], [stats[-1]]
    else:
        plain_recs, plain_t, plain_st = run_form(ctx, handles, w, maxL, "plain", a.steps if
                                                 primary == "plain" else 1,
                                                 a.warmup if primary == "plain" else 0)
        print(f"[bench] plain: {[round(t, 3) for t in plain_t]} s", file=sys.stderr,
              flush=True)
    if primary == "plain":
        recs, times, stats = plain_recs, plain_t, plain_st
    identical = same_set(recs, plain_recs)
    gain = None
    if primary == "learned" and not a.no_gain:
        # what the trained network gains: the same batched form with the probe descent
        b_recs, b_t, b_st = run_form(ctx, handles, w, maxL, "batched", 1, 0)
        gain = {"batched_ms_per_step": 1e3 * b_t[0], "batched_attempts": b_st[0]["attempts"],
                "learned_attempts": stats[-1]["attempts"],
                "identical_discord_set": same_set(recs, b_recs)}
        print(f"[bench] batched (no learned proposal): {b_t[0]:.3f} s", file=sys.stderr,
              flush=True)

    t_step = sum(times) / len(times)
    value = units / t_step
    plain_value = units / (sum(plain_t) / len(plain_t))
    # The first length: the plain form starts at the reference's r = 2 sqrt(m) and halves it
    # (src/main.cu:38,155), the fast forms from the learned proposal or a probe descent; so
    # the ratio is also given without it (each form run once more on the first length alone).
    first = None
    if not (a.no_plain and primary != "plain") and nlen > 1:
        _, fp, _ = run_form(ctx, handles, w, w["minL"], "plain", 1, 0)
        _, fq, _ = run_form(ctx, handles, w, w["minL"], primary, 1, 0)
        rest_plain = sum(plain_t) / len(plain_t) - fp[0]
        rest_primary = t_step - fq[0]
        first = {"plain_ms": 1e3 * fp[0], "primary_ms": 1e3 * fq[0],
                 "speedup_vs_plain_after_first_length":
                     rest_plain / rest_primary if rest_primary > 0 else None}
        print(f"[bench] first length alone: plain {fp[0]:.3f} s, {primary} {fq[0]:.3f} s",
              file=sys.stderr, flush=True)

    # Per-kernel times come from hipEvents around eager launches.  With the length loop on
    # the device the timed steps replay graphs, which carry no per-kernel events, so one more
    # step of the same form runs with the host length loop (the same kernels on the same
    # lengths) and gives the kernel figures.
    kstats, kernel_timing = stats, "the timed steps"
    if primary != "plain" and tunables.get("device_lengths", 0) > 0:
        ctx.set_tunables(device_lengths=0)
        _, kt, kstats = run_form(ctx, handles, w, maxL, primary, 1, 0)
        ctx.set_tunables(device_lengths=tunables["device_lengths"])
        kernel_timing = (f"one extra step with the host length loop ({1e3 * kt[0]:.1f} ms): "
                         "graph replays carry no per-kernel events")
        print(f"[bench] {primary} with the host length loop (kernel times): {kt[0]:.3f} s",
              file=sys.stderr, flush=True)
    # dominant kernel: the one with the most summed time across those steps
    names = ("derive", "select", "refine", "exact")
    tot = {k: sum(s["kernel_ms"][k] for s in kstats) for k in names}
    nl = {k: sum(s["launches"][k] for s in kstats) for k in names}
    dom = max(names, key=lambda k: tot[k])
    avg_ms = tot[dom] / max(nl[dom], 1)
    kname = {"select": "paddi::k_scan_select", "refine": "paddi::k_scan_refine",
             "derive": "paddi::k_derive", "exact": "paddi::k_exact"}[dom]
    # algorithmic bytes: 16 B per (point x length) for each pass over a length, shared by the
    # kernel's launches of that pass (k_scan_refine runs as three round launches per pass)
    passes = sum(s["attempts"] for s in kstats)
    bytes_per_launch = 16.0 * points * passes / max(nl[dom], 1)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{a.workload}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        for k, v in tj.get("kernels", {}).items():
            if k.startswith(kname):
                traffic = v["bytes_per_dispatch"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    pairs = sum(s["pairs"] for s in kstats) / len(kstats)
    # the share of each scan's lane-steps spent on rows still alive when their tile started
    useful = {k: (sum(s["live_pairs"][k] for s in kstats) /
                  max(sum(s["scan_pairs"][k] for s in kstats), 1)) for k in ("select", "refine")}
    scan_ms = (tot["select"] + tot["refine"]) / len(kstats)
    valu = (pairs * OPS_PER_PAIR) / (scan_ms * 1e-3) / VALU_PEAK_OPS if scan_ms > 0 else 0.0

    out = {
        "metric": "points×lengths/sec, batched+learned vs plain GPU PALMAD, identical discord set",
        "value": value,
        "unit": "points*lengths/s",
        "n_gpus": 1,
        "steps": len(times),
        "warmup": a.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD_NAMES[a.workload], "points": points,
                   "series": len(series), "lengths": nlen, "m": f"{w['minL']}..{maxL}",
                   "step": w["step"], "top_k": w["k"], "form": primary},
        "plain": {"value": plain_value, "ms_per_step": 1e3 * sum(plain_t) / len(plain_t),
                  "attempts": plain_st[-1]["attempts"]},
        "speedup_vs_plain": value / plain_value,
        "first_length": first,
        "learned_gain": gain,
        "identical_discord_set": identical,
        "attempts": stats[-1]["attempts"],
        "survivors": stats[-1]["survivors"],
        "exact_refines": stats[-1]["exact_refines"],
        "kernel_ms_per_step": {k: tot[k] / len(kstats) for k in names},
        "kernel_timing": kernel_timing,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "avg_launch_ms": avg_ms, "launches": nl[dom],
                     "bytes_per_launch": bytes_per_launch},
        "tunables": tunables,
        # where the length loop runs: on the device, every length after the first one
        # replayed hipGraph (top-k and next range computed there), or on the host
        "length_loop": ({"where": "device", "graph_per_length": True,
                         "host_sync_every_lengths": tunables["device_lengths"]}
                        if primary != "plain" and tunables.get("device_lengths", 0) > 0
                        else {"where": "host", "graph_per_pass": bool(tunables.get("use_graph"))}),
        "valu": {"pairs_per_step": pairs, "ops_per_pair": OPS_PER_PAIR,
                 "frac_of_peak": valu, "live_row_share": useful,
                 "scan_pairs_per_step": {k: sum(s["scan_pairs"][k] for s in kstats) / len(kstats)
                                         for k in ("select", "refine")}},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, recs, a.workload)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
