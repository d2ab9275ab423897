"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (built by ``make -C oracle``).  Only tests,
the smoke check and the bench's CPU-baseline leg import this module; the
product path never does.  Also holds ``brute_profile``, an independent
brute-force nearest-neighbour profile in numpy (direct z-normalised vectors,
not the QT identity the restatement uses) that pins the restatement on small
inputs.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        P = C.c_void_p
        i64, i32, f64 = C.c_int64, C.c_int32, C.c_double
        L.oracle_stats_f64.argtypes = [P, i64, i32, P, P]
        L.oracle_stats_f32.argtypes = [P, i64, i32, P, P]
        L.oracle_find_discords_f64.argtypes = [P, P, P, P, P, i64, f64, i32, i32]
        L.oracle_find_discords_hint_f64.argtypes = [P, P, P, P, P, i64, f64, i32, i32]
        L.oracle_find_discords_runs_f64.argtypes = [P, P, P, P, P, i64, f64, i32, i32, i64, i64]
        L.oracle_find_discords_f32.argtypes = [P, P, P, P, P, i64, C.c_float, i32]
        L.oracle_profile_f64.argtypes = [P, P, P, P, i64, i32, i32]
        L.oracle_topk.argtypes = [P, P, i64, f64, i32, i32, P, P, P]
        L.oracle_topk.restype = i32
        L.oracle_palmad.argtypes = [P, i64, i32, i32, i32, i32, i32, i32, P, P, P, P, P]
        L.oracle_palmad.restype = i32
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def stats(T, m, fp32=False):
    T = np.ascontiguousarray(T, dtype=np.float32)
    N = len(T) - m + 1
    dt = np.float32 if fp32 else np.float64
    mu, sd = np.empty(N, dt), np.empty(N, dt)
    (lib().oracle_stats_f32 if fp32 else lib().oracle_stats_f64)(_p(T), N, m, _p(mu), _p(sd))
    return mu, sd


def range_discords(T, m, r2, fp32=False, nthreads=0, hint=False):
    """find_discords_cpu at squared range r2; returns (cand int32[N], nnDist[N]).
    hint=True runs the same scan with each row first trying the column one past its
    predecessor's pruning column (same cand and, where cand == 1, the same nnDist; the
    nnDist of pruned rows is unspecified)."""
    T = np.ascontiguousarray(T, dtype=np.float32)
    N = len(T) - m + 1
    mu, sd = stats(T, m, fp32)
    cand = np.ones(N, np.int32)
    if fp32:
        nn = np.full(N, np.finfo(np.float32).max, np.float32)
        lib().oracle_find_discords_f32(_p(T), _p(mu), _p(sd), _p(cand), _p(nn), N,
                                       C.c_float(r2), m)
    else:
        nn = np.full(N, np.finfo(np.float64).max, np.float64)
        fn = lib().oracle_find_discords_hint_f64 if hint else lib().oracle_find_discords_f64
        fn(_p(T), _p(mu), _p(sd), _p(cand), _p(nn), N, float(r2), m, nthreads)
    return cand, nn


def range_discords_sampled(T, m, r2, run=256, stride=8, nthreads=0):
    """find_discords_cpu at squared range r2 on every stride-th run of `run` rows (reference
    scan order).  Returns (rows, cand, nnDist, stats_seconds, scan_seconds) for the sampled
    rows; the statistics are computed for all rows, as the reference does per length."""
    import time
    T = np.ascontiguousarray(T, dtype=np.float32)
    N = len(T) - m + 1
    t0 = time.perf_counter()
    mu, sd = stats(T, m)
    t1 = time.perf_counter()
    cand = np.ones(N, np.int32)
    nn = np.full(N, np.finfo(np.float64).max, np.float64)
    lib().oracle_find_discords_runs_f64(_p(T), _p(mu), _p(sd), _p(cand), _p(nn), N, float(r2),
                                        m, nthreads, run, stride)
    t2 = time.perf_counter()
    rows = np.arange(N)
    rows = rows[(rows // run) % stride == 0]
    return rows, cand[rows], nn[rows], t1 - t0, t2 - t1


def profile(T, m, nthreads=0):
    """Exact NN ED^2 of every row under the oracle's exclusion rule (O(N^2 m))."""
    T = np.ascontiguousarray(T, dtype=np.float32)
    N = len(T) - m + 1
    mu, sd = stats(T, m)
    nn = np.empty(N, np.float64)
    lib().oracle_profile_f64(_p(T), _p(mu), _p(sd), _p(nn), N, m, nthreads)
    return nn


def topk(cand, nn, r2, m, k):
    cand = np.ascontiguousarray(cand, np.int32)
    nn = np.ascontiguousarray(nn, np.float64)
    idx = np.zeros(k, np.int64)
    d = np.zeros(k, np.float64)
    nd = np.zeros(1, np.int64)
    got = lib().oracle_topk(_p(cand), _p(nn), len(cand), float(r2), m, k, _p(idx), _p(d), _p(nd))
    return idx[:got], d[:got], int(nd[0])


def palmad(T, minL, maxL, step=1, k=1, max_attempts=64, nthreads=0):
    """The PALMAD length loop with the main.cu r-schedule.  Returns a list of
    per-length dicts {m, idx, d2, attempts, r2}."""
    T = np.ascontiguousarray(T, dtype=np.float32)
    nlen = (maxL - minL) // step + 1
    idx = np.zeros(nlen * k, np.int64)
    d = np.zeros(nlen * k, np.float64)
    cnt = np.zeros(nlen, np.int32)
    att = np.zeros(nlen, np.int32)
    rr = np.zeros(nlen, np.float64)
    rc = lib().oracle_palmad(_p(T), len(T), minL, maxL, step, k, max_attempts, nthreads,
                             _p(idx), _p(d), _p(cnt), _p(att), _p(rr))
    if rc < 0:
        raise ValueError("invalid arguments")
    out = []
    for j in range(nlen):
        c = int(cnt[j])
        out.append(dict(m=minL + j * step, idx=idx[j * k:j * k + c].copy(),
                        d2=d[j * k:j * k + c].copy(), attempts=int(att[j]), r2=float(rr[j])))
    return out


# --- independent brute force (numpy, small inputs) -------------------------

def znorm_windows(T, m):
    """z-normalised windows, computed two-pass per window in float64."""
    T = np.asarray(T, dtype=np.float64)
    W = np.lib.stride_tricks.sliding_window_view(T, m)
    mu = W.mean(axis=1, keepdims=True)
    c = W - mu
    sd = np.sqrt((c * c).mean(axis=1, keepdims=True))
    z = np.where(sd > 0, c / np.where(sd > 0, sd, 1.0), 0.0)
    return z, sd[:, 0]


def brute_profile(T, m, symmetric=False):
    """NN squared z-ED of every row by direct vector differences, with the
    exclusion j - i >= m-1 or i - j >= m+1 (src/cpu_kernels/cpu_kernels.cpp:57), or with
    symmetric=True the reference GPU kernels' |i - j| >= m-1 (src/PD3_kernels.cuh:70).
    Degenerate (sd == 0) pairs are 2m."""
    z, sd = znorm_windows(T, m)
    N = z.shape[0]
    nn = np.full(N, np.inf)
    arg = np.full(N, -1, np.int64)
    left = m - 1 if symmetric else m + 1
    for i in range(N):
        d = ((z - z[i]) ** 2).sum(axis=1)
        deg = (sd == 0) | (sd[i] == 0)
        d[deg] = 2.0 * m
        j = np.arange(N)
        ok = (j - i >= m - 1) | (i - j >= left)
        d = np.where(ok, d, np.inf)
        a = int(np.argmin(d))
        nn[i], arg[i] = d[a], a
    return nn, arg
