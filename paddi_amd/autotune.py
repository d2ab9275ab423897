"""Autotuner of the engine's schedule (select window, refine chunking, witness tiles,
exact-refine speculation, hipGraph use).

    python -m paddi_amd.autotune [--workload c3] [--lengths 0] [--mode learned]

Coordinate descent: starting from the stored entry (or the built-in defaults), each tunable in turn is set to each
of its candidates; a candidate is timed as the median of `--reps` complete searches over
the sample (the workload's first `--lengths` lengths, series resident in HBM) and kept if it
is faster.  Every configuration must return the same discord set as the starting one (the
schedule never changes results; a mismatch aborts).  The winner is stored in
paddi_amd/model/tuned.json under the key device/length-bucket[/batch] (engine.tuned_key), which
Context.apply_tuned() and bench.py read.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from paddi_amd import synth  # noqa: E402
from paddi_amd.engine import TUNED_PATH, Context, tuned_key  # noqa: E402

GRID = {
    "witness_tiles": [0, 1, 2, 4, 8, 12, 16, 24, 32, 48],
    "select_window": [256, 512, 1024, 2048, 4096, 8192, 16384],
    "refine_items": [4096, 8192, 16384, 32768, 65536],
    "min_chunk": [512, 1024, 2048, 4096, 8192],
    "refine_chunk": [16384, 32768, 65536, 131072],
    "speculate": [1, 2, 3, 4, 8],
    "use_graph": [0, 1],
    "prev_shift": [0, 1, 2],
    "bootstrap": [0, 1],
    "bound_candidates": [16, 32, 64],
    "direct_budget": [0, 400, 550, 700, 850, 1000],
    "direct_rows": [1, 2, 3, 4],
    "device_lengths": [0, 8, 32],
    "device_rounds": [1, 2, 3],
    "refine_rounds": [3, 4],
    "seed_witnesses": [0, 1],
    "refine_order": [0, 1, 2],
}


def key_of(recs):
    return [(r.series, r.index, r.length) for r in recs]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c3", choices=sorted(synth.WORKLOADS))
    ap.add_argument("--lengths", type=int, default=0, help="0 = every length")
    ap.add_argument("--mode", default="learned")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--out", default=TUNED_PATH)
    ap.add_argument("--fresh", action="store_true", help="start from the built-in defaults")
    a = ap.parse_args()

    w = synth.WORKLOADS[a.workload]()
    ctx = Context(0)
    hs = [ctx.upload(x) for x in w["series"]]
    maxL = w["maxL"] if a.lengths <= 0 else \
        min(w["maxL"], w["minL"] + (a.lengths - 1) * w["step"])
    if not a.fresh:
        # continue from the stored entry, if any
        ctx.apply_tuned(w["minL"], path=a.out, nseries=len(w["series"]))
    base = ctx.get_tunables()

    def run(tun):
        ctx.set_tunables(**tun)
        ts, recs = [], None
        for _ in range(a.reps):
            t0 = time.perf_counter()
            recs = ctx.discords_series(hs, w["minL"], maxL, w["step"], w["k"], mode=a.mode)
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts), recs

    run(base)  # warm-up (allocations, code objects)
    best_t, ref = run(base)
    best = dict(base)
    log = [dict(tunables=dict(best), seconds=best_t)]
    print(f"[tune] start {best_t * 1e3:.1f} ms {best}", flush=True)
    for sweep in range(a.sweeps):
        changed = False
        for name, values in GRID.items():
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best)
                cand[name] = v
                if cand["min_chunk"] > cand["refine_chunk"]:
                    continue
                t, recs = run(cand)
                if key_of(recs) != key_of(ref):
                    raise SystemExit(f"schedule {cand} changed the discord set")
                log.append(dict(tunables=dict(cand), seconds=t))
                if t < best_t * 0.98:
                    best_t, best, changed = t, cand, True
                    print(f"[tune] {name}={v}: {t * 1e3:.1f} ms (new best)", flush=True)
                else:
                    print(f"[tune] {name}={v}: {t * 1e3:.1f} ms", flush=True)
        if not changed:
            break
    table = {}
    if os.path.exists(a.out):
        with open(a.out) as f:
            table = json.load(f)
    table[tuned_key(w["minL"], nseries=len(w["series"]))] = dict(
        tunables=best, seconds=best_t, baseline_seconds=log[0]["seconds"],
        sample=f"{a.workload}: {len(w['series'])} series, m={w['minL']}..{maxL} "
               f"step {w['step']}, top-{w['k']}, mode {a.mode}, median of {a.reps}",
        evaluated=len(log))
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(f"[tune] best {best_t * 1e3:.1f} ms (from {log[0]['seconds'] * 1e3:.1f} ms): {best}",
          flush=True)


if __name__ == "__main__":
    main()
