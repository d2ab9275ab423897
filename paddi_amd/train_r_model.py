"""Train the learned range proposal on the GPU and write paddi_amd/model/r_mlp.bin.

    python -m paddi_amd.train_r_model [--series 256] [--steps 4000]

1. Traces: the engine's batched form (model-free) runs over the training series
   (random walks, ECG-like, AR(2); synth.training_series; lengths log-uniform in
   [20,000, 524,288]) at m = 64..512 step 32, top-3, and yields d_1, d_2, d_3 (series, m).
2. Features of every (series, m) come from the device (paddi_features); a sample is
   (features with k, log(d_k / 2m)) for k = 1, 2, 3.
3. The MLP is trained on the device with the library's own forward/backward/Adam kernels
   (paddi_mlp_train); the last 32 series are held out.
4. margin = 0.9-quantile of (prediction - t) over all samples, so that about nine proposals
   in ten fall below the k-th distance and succeed in one pass.
A JSON report (losses, margin, coverage) is written next to the model.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from paddi_amd import model as M  # noqa: E402
from paddi_amd import synth  # noqa: E402
from paddi_amd.engine import Context  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--series", type=int, default=256)
    ap.add_argument("--nmin", type=int, default=20000)
    ap.add_argument("--nmax", type=int, default=524288)
    ap.add_argument("--minL", type=int, default=64)
    ap.add_argument("--maxL", type=int, default=512)
    ap.add_argument("--step", type=int, default=32)
    ap.add_argument("--k", type=int, default=3)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--wd", type=float, default=1e-4)
    ap.add_argument("--holdout", type=int, default=32)
    ap.add_argument("--out", default=M.DEFAULT_PATH)
    a = ap.parse_args()

    t0 = time.time()
    rng = np.random.default_rng(77)
    lens = np.exp(rng.uniform(np.log(a.nmin), np.log(a.nmax), a.series)).astype(int)
    series = [synth.training_series(i, int(lens[i])) for i in range(a.series)]
    ctx = Context(0)
    hs = [ctx.upload(x) for x in series]
    print(f"[train] {a.series} series, {int(lens.sum())} points generated in "
          f"{time.time() - t0:.1f}s", flush=True)

    t1 = time.time()
    recs = ctx.discords_series(hs, a.minL, a.maxL, a.step, a.k, mode="batched")
    lengths = list(range(a.minL, a.maxL + 1, a.step))
    dk = {}
    for r in recs:
        dk.setdefault((r.series, r.length), []).append(r.nn_dist2)
    print(f"[train] traces: {len(recs)} discords in {time.time() - t1:.1f}s", flush=True)

    t2 = time.time()
    X, T, S = [], [], []
    for m in lengths:
        F = ctx.features(hs, m, top_k=1)
        for q in range(a.series):
            d = sorted(dk.get((q, m), []), reverse=True)
            for kk in range(1, min(len(d), a.k) + 1):
                f = F[q].copy()
                f[35] = kk / 4.0
                X.append(f)
                T.append(np.log(d[kk - 1] / (2.0 * m)))
                S.append(q)
    X = np.array(X, np.float32)
    T = np.array(T, np.float32)
    S = np.array(S)
    print(f"[train] features of {len(X)} samples in {time.time() - t2:.1f}s", flush=True)

    tr = S < a.series - a.holdout
    fmu = X[tr].mean(0)
    fsd = X[tr].std(0) + 1e-3
    p0 = M.init_params(X.shape[1], 64, seed=0)
    t3 = time.time()
    params, losses = ctx.mlp_train(p0, fmu, fsd, X[tr], T[tr], a.steps, a.lr, a.wd)
    print(f"[train] {a.steps} Adam steps in {time.time() - t3:.1f}s: loss "
          f"{losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)

    pred = ctx.mlp_forward(params, fmu, fsd, X)
    res = pred - T
    mse_tr = float(np.mean(res[tr] ** 2))
    mse_va = float(np.mean(res[~tr] ** 2)) if (~tr).any() else float("nan")
    base = float(np.mean((T[~tr] - T[tr].mean()) ** 2)) if (~tr).any() else float("nan")
    margin = float(np.quantile(res, 0.9))
    cover_va = float(np.mean(pred[~tr] - margin <= T[~tr])) if (~tr).any() else float("nan")
    M.save(a.out, fmu, fsd, params, margin)
    rep = dict(samples=int(len(X)), train=int(tr.sum()), holdout=int((~tr).sum()),
               lengths=lengths, k=a.k, steps=a.steps, lr=a.lr, weight_decay=a.wd,
               loss_first=float(losses[0]), loss_last=float(losses[-1]), mse_train=mse_tr,
               mse_holdout=mse_va, mse_holdout_constant_predictor=base, margin=margin,
               holdout_coverage=cover_va,
               holdout_tightness_median=float(np.median(np.exp(T[~tr] - (pred[~tr] - margin))))
               if (~tr).any() else None)
    with open(os.path.splitext(a.out)[0] + "_report.json", "w") as f:
        json.dump(rep, f, indent=1)
    print("[train] " + json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
