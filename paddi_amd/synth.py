"""Seeded synthetic series for the discord workloads.

Every generator draws from ``numpy.random.default_rng(seed)`` (PCG64) and
works in float64 before one final cast to float32, so a given seed produces
the same bytes on every host with this numpy.  ``series_sha256`` fingerprints
a series; the test suite pins the fingerprints of the small workloads.

Workloads (points x lengths are the throughput units):

  rw_planted      random walk x0 = 0, N(0,1) increments, with sine-burst
                  anomalies: a window of ``alen`` points is replaced by
                  mu_loc + amp*sigma_loc*sin(2*pi*t/period), where mu_loc and
                  sigma_loc are the mean and std of the replaced window.
                  Anomaly starts are uniform, non-overlapping, >= 1024 apart.
  ecg_like        beats of a 250-sample template (three Gaussian bumps, widths
                  5/12/20, amplitudes 1.0/-0.3/0.4, centres 60/95/160) with
                  per-beat period jitter U(-5%, +5%), additive N(0, 0.05^2);
                  anomalous beats are amplitude x1.6 or time-warped x0.7.
  ar2_mix         AR(2) with phi = (0.6, -0.2) plus 0.8 sin(2 pi t/97) +
                  0.5 sin(2 pi t/41 + 1).
"""
from __future__ import annotations

import hashlib
import math

import numpy as np

ECG_PERIOD = 250
_ECG_BUMPS = ((60.0, 5.0, 1.0), (95.0, 12.0, -0.3), (160.0, 20.0, 0.4))


def series_sha256(x: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(x, dtype=np.float32).tobytes()).hexdigest()


def _anomaly_starts(rng, n, alen, count, min_gap=1024):
    # >= 1024 apart as the workloads specify; short test series scale the gap down
    min_gap = min(min_gap, max(alen, (n - alen) // (2 * count)))
    starts = []
    lo, hi = 0, n - alen
    tries = 0
    while len(starts) < count:
        p = int(rng.integers(lo, hi + 1))
        tries += 1
        if all(abs(p - q) >= max(min_gap, alen) for q in starts):
            starts.append(p)
        if tries > 100000:
            raise ValueError("cannot place anomalies")
    return sorted(starts)


def _plant_sine(x, starts, alen, period, amp):
    t = np.arange(alen, dtype=np.float64)
    for p in starts:
        w = x[p:p + alen]
        mu, sd = float(w.mean()), float(w.std())
        x[p:p + alen] = mu + amp * sd * np.sin(2.0 * math.pi * t / period)


def random_walk(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(n)
    e[0] = 0.0
    return np.cumsum(e), rng


def rw_planted(n: int, seed: int, n_anom: int, alen: int, period: float, amp: float = 3.0,
               return_starts: bool = False):
    x, rng = random_walk(n, seed)
    starts = _anomaly_starts(rng, n, alen, n_anom) if n_anom else []
    _plant_sine(x, starts, alen, period, amp)
    out = x.astype(np.float32)
    return (out, starts) if return_starts else out


def _ecg_beat(length: int, amp: float = 1.0) -> np.ndarray:
    t = np.arange(length, dtype=np.float64) * (ECG_PERIOD / length)
    b = np.zeros(length)
    for c, w, a in _ECG_BUMPS:
        b += a * np.exp(-0.5 * ((t - c) / w) ** 2)
    return amp * b


def ecg_like(n: int, seed: int, n_amp: int = 0, n_warp: int = 0, amp_factor: float = 1.6,
             warp: float = 0.7, return_beats: bool = False):
    rng = np.random.default_rng(seed)
    nb_est = n // int(ECG_PERIOD * 0.95) + 4
    lens = np.rint(ECG_PERIOD * (1.0 + rng.uniform(-0.05, 0.05, nb_est))).astype(np.int64)
    n_special = n_amp + n_warp
    special = {}
    if n_special:
        usable = max(1, (n // ECG_PERIOD) - 4)
        if n_special > usable:
            raise ValueError("series too short for the requested anomalous beats")
        gap = 4 if usable >= 4 * n_special else 1
        picks = []
        while len(picks) < n_special:
            b = int(rng.integers(2, 2 + usable))
            if all(abs(b - q) >= gap for q in picks):
                picks.append(b)
        for q, b in enumerate(picks):
            special[b] = "amp" if q < n_amp else "warp"
    pieces, starts, total, b = [], [], 0, 0
    while total < n:
        L = int(lens[b % nb_est])
        kind = special.get(b)
        if kind == "warp":
            L = max(8, int(round(L * warp)))
        beat = _ecg_beat(L, amp_factor if kind == "amp" else 1.0)
        if kind:
            starts.append((total, kind))
        pieces.append(beat)
        total += L
        b += 1
    x = np.concatenate(pieces)[:n]
    x += rng.normal(0.0, 0.05, n)
    out = x.astype(np.float32)
    return (out, starts) if return_beats else out


def ar2_mix(n: int, seed: int) -> np.ndarray:
    from scipy.signal import lfilter
    rng = np.random.default_rng(seed)
    # x_t = 0.6 x_{t-1} - 0.2 x_{t-2} + e_t, as an IIR filter over the noise.
    x = lfilter([1.0], [1.0, -0.6, 0.2], rng.standard_normal(n))
    t = np.arange(n, dtype=np.float64)
    x += 0.8 * np.sin(2.0 * math.pi * t / 97.0) + 0.5 * np.sin(2.0 * math.pi * t / 41.0 + 1.0)
    return x.astype(np.float32)


# --- the five workloads --------------------------------------------------

def workload_c1():
    """65,536-point random walk (seed 0), 3 planted period-32 bursts of 128; m=128, top-3."""
    return dict(series=[rw_planted(65536, 0, 3, 128, 32.0)], minL=128, maxL=128, step=1, k=3)


def workload_c2():
    """1,048,576-point ECG-like series (seed 1), 5 amplitude + 5 warp beats; m=64..128, top-3."""
    return dict(series=[ecg_like(1048576, 1, 5, 5)], minL=64, maxL=128, step=1, k=3)


def workload_c3():
    """8,388,608-point random walk (seed 2), 20 period-24 bursts of 192; m=128..256, top-3."""
    return dict(series=[rw_planted(8388608, 2, 20, 192, 24.0)], minL=128, maxL=256, step=1, k=3)


def workload_c4_eval():
    """2,097,152-point AR(2)+sine mixture (seed 3); m=256..512 step 4, top-3."""
    return dict(series=[ar2_mix(2097152, 3)], minL=256, maxL=512, step=4, k=3)


def training_series(i: int, n: int = 262144) -> np.ndarray:
    """Training series i of the learned r-proposal: RW / ECG-like / AR(2) in rotation."""
    kind = i % 3
    seed = 1000 + i
    if kind == 0:
        return rw_planted(n, seed, 4, 128, 24.0)
    if kind == 1:
        return ecg_like(n, seed, 2, 2)
    rng = np.random.default_rng(seed)
    from scipy.signal import lfilter
    return lfilter([1.0], [1.0, -0.6, 0.2], rng.standard_normal(n)).astype(np.float32)


def workload_c5(count: int = 1024, seed: int = 4):
    """`count` uneven series, n ~ logU[20,000, 500,000]; RW or ECG-like; 1-4 anomalies of 80."""
    rng = np.random.default_rng(seed)
    lens = np.exp(rng.uniform(math.log(20000), math.log(500000), count)).astype(np.int64)
    kinds = rng.integers(0, 2, count)
    nanom = rng.integers(1, 5, count)
    out = []
    for i in range(count):
        s = int(rng.integers(0, 2**31 - 1))
        if kinds[i] == 0:
            out.append(rw_planted(int(lens[i]), s, int(nanom[i]), 80, 20.0))
        else:
            x = ecg_like(int(lens[i]), s).astype(np.float64)
            r2 = np.random.default_rng(s + 1)
            starts = _anomaly_starts(r2, len(x), 80, int(nanom[i]))
            _plant_sine(x, starts, 80, 20.0, 3.0)
            out.append(x.astype(np.float32))
    return dict(series=out, minL=64, maxL=96, step=1, k=1)


WORKLOADS = {
    "c1": workload_c1,
    "c2": workload_c2,
    "c3": workload_c3,
    "c4": workload_c4_eval,
    "c5": workload_c5,
}
