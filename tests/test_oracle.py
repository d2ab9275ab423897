"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the parity reference.

Each part of the restatement is checked against something it does not share code or
formula with: closed forms, the brute-force profile on explicitly z-normalised windows
(tests/golden/brute_profiles.npz), invariants, and a planted mistake that must be caught.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from paddi_amd import synth

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def brute_cases():
    z = np.load(os.path.join(GOLD, "brute_profiles.npz"))
    names = sorted({k.split("__")[0] for k in z.files})
    return [(n, z[f"{n}__x"], int(z[f"{n}__m"][0]), z[f"{n}__nn"]) for n in names]


# --- sliding statistics (cpu_kernels.cpp:7-31) --------------------------------------

def test_stats_closed_form_ramp():
    # T = 0..n-1: mean of window i is i + (m-1)/2, std is sqrt((m^2-1)/12).
    m = 37
    T = np.arange(500, dtype=np.float32)
    mu, sd = O.stats(T, m)
    np.testing.assert_allclose(mu, np.arange(len(mu)) + (m - 1) / 2, rtol=0, atol=1e-9)
    np.testing.assert_allclose(sd, np.sqrt((m * m - 1) / 12.0), rtol=1e-12)


def test_stats_shift_scale_invariants():
    x = synth.ar2_mix(3000, 5).astype(np.float64)
    m = 64
    mu, sd = O.stats(x.astype(np.float32), m)
    y = (2.0 * x + 8.0).astype(np.float32)  # exact in float32 for these magnitudes? close
    mu2, sd2 = O.stats(y, m)
    # check against the two-pass float64 definition on the actually stored float32 values
    for T, a, b in ((x.astype(np.float32), mu, sd), (y, mu2, sd2)):
        W = np.lib.stride_tricks.sliding_window_view(T.astype(np.float64), m)
        np.testing.assert_allclose(a, W.mean(1), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(b, W.std(1), rtol=1e-9)


def test_stats_fp32_literal_differs_but_close_on_small_values():
    x = synth.ar2_mix(2000, 6)
    mu64, sd64 = O.stats(x, 50)
    mu32, sd32 = O.stats(x, 50, fp32=True)
    assert np.max(np.abs(sd32 - sd64) / sd64) < 1e-4


# --- range discords and the profile (cpu_kernels.cpp:33-78) --------------------------

@pytest.mark.parametrize("case", brute_cases(), ids=lambda c: c[0])
def test_profile_matches_bruteforce(case):
    name, x, m, nn_ref = case
    nn = O.profile(x, m)
    big = nn_ref > 1e-6
    np.testing.assert_allclose(nn[big], nn_ref[big], rtol=1e-8)
    assert np.all(nn[~big] < 1e-6)


@pytest.mark.parametrize("case", brute_cases(), ids=lambda c: c[0])
def test_range_discords_match_bruteforce(case):
    """cand[i] == 1 iff NN(i) > r (prune at ED^2 <= r); nnDist exact for discords."""
    name, x, m, nn_ref = case
    for q in (0.999, 0.99, 0.9, 0.5):
        r2 = float(np.quantile(nn_ref, q))
        cand, nn = O.range_discords(x, m, r2)
        exp = nn_ref > r2
        # rows whose NN is within rounding of r2 may go either way
        near = np.abs(nn_ref - r2) <= 1e-9 * max(r2, 1.0)
        assert np.array_equal(cand[~near] == 1, exp[~near]), (name, q)
        sel = (cand == 1) & (nn_ref > 1e-6)
        np.testing.assert_allclose(nn[sel], nn_ref[sel], rtol=1e-8)


def test_exclusion_rule_is_asymmetric():
    """|i-1-j| >= m: j-i >= m-1 or i-j >= m+1 (cpu_kernels.cpp:57).  On a short series
    some rows have no valid neighbour at all (nnDist stays at its initial value)."""
    m = 8
    y = np.concatenate([np.zeros(3), np.arange(1, 2 * m + 3, dtype=np.float64) ** 1.5])
    y = y.astype(np.float32)
    N = len(y) - m + 1
    prof = O.profile(y, m)
    ref, arg = O.brute_profile(y, m)
    none = ~np.isfinite(ref)
    assert np.all(prof[none] == np.finfo(np.float64).max)
    ok = (~none) & (ref > 1e-6)
    np.testing.assert_allclose(prof[ok], ref[ok], rtol=1e-7)
    for i in range(N):
        js = [j for j in range(N) if (j - i >= m - 1) or (i - j >= m + 1)]
        assert none[i] == (len(js) == 0)
        if js:
            assert arg[i] in js


def test_affine_copy_gives_zero_distance():
    rng = np.random.default_rng(9)
    m = 32
    p = rng.standard_normal(m)
    x = np.concatenate([p, rng.standard_normal(200), 3.0 * p - 7.0]).astype(np.float32)
    nn = O.profile(x, m)
    assert nn[0] < 1e-8
    assert nn[m + 200] < 1e-8


def test_planted_mistake_is_caught():
    """Plant the GPU kernels' symmetric exclusion (|i-j| >= m-1, PD3_kernels.cuh:70) in a
    copy of the brute force: on a series whose only close pair sits at offset -(m-1) the
    planted rule reports ~0 where the oracle's rule does not, and the oracle agrees with
    the unplanted brute force.  So the brute-force check can see that class of error."""
    m = 16
    rng = np.random.default_rng(5)
    x = np.cumsum(rng.standard_normal(300))
    a = 100
    x[a + m - 1] = x[a]  # the shared point, so the copy leaves window a unchanged
    src = x[a:a + m].copy()
    x[a + m - 1:a + 2 * m - 1] = src  # window a+m-1 copies window a
    x = x.astype(np.float32)
    i = a + m - 1
    z, sd = O.znorm_windows(x, m)
    N = z.shape[0]
    d = ((z - z[i]) ** 2).sum(1)
    j = np.arange(N)
    planted = d[np.abs(i - j) >= m - 1].min()
    nn_ref, _ = O.brute_profile(x, m)
    nn = O.profile(x, m)
    assert planted < 1e-8
    assert nn_ref[i] > 0.1 and nn[i] == pytest.approx(nn_ref[i], rel=1e-8)
    # and the earlier row sees the copy at +(m-1) under both rules
    assert nn[a] < 1e-8
    # the brute force's symmetric rule (the engine's alternative) finds the planted pair
    nn_sym, arg_sym = O.brute_profile(x, m, symmetric=True)
    assert nn_sym[i] < 1e-8 and arg_sym[i] == a


def test_symmetric_rule_differs_only_at_two_offsets():
    """|i-j| >= m-1 (src/PD3_kernels.cuh:70) pairs everything |i-1-j| >= m does plus
    i-j = m-1 and i-j = m: its profile is never larger, and where it is smaller the new
    nearest neighbour sits at one of those two offsets."""
    rng = np.random.default_rng(17)
    x = np.cumsum(rng.standard_normal(500)).astype(np.float32)
    for m in (8, 16):
        nn, _ = O.brute_profile(x, m)
        nn_sym, arg_sym = O.brute_profile(x, m, symmetric=True)
        assert np.all(nn_sym <= nn)
        lower = nn_sym < nn
        assert lower.any()
        rows = np.nonzero(lower)[0]
        assert set((rows - arg_sym[rows]).tolist()) <= {m - 1, m}
        np.testing.assert_array_equal(nn_sym[~lower], nn[~lower])


# --- top-k (sets_processing.hpp:11-46) and r-independence ----------------------------

def greedy_topk_reference(nn, m, k):
    order = sorted(range(len(nn)), key=lambda i: (-nn[i], i))
    acc = []
    for i in order:
        if all(abs(i - a) >= m for a in acc):
            acc.append(i)
        if len(acc) == k:
            break
    return acc


@pytest.mark.parametrize("case", brute_cases(), ids=lambda c: c[0])
def test_topk_is_r_independent(case):
    name, x, m, nn_ref = case
    k = 3
    exp = greedy_topk_reference(nn_ref, m, k)
    kth = nn_ref[exp[-1]]
    for r2 in (0.0, 0.5 * kth, 0.99 * kth, 0.999999 * kth):
        cand, nn = O.range_discords(x, m, r2)
        idx, d, nd = O.topk(cand, nn, r2, m, k)
        assert list(idx) == exp, (name, r2)
        np.testing.assert_allclose(d, nn_ref[exp], rtol=1e-8)
    # above the k-th distance fewer than k survive
    cand, nn = O.range_discords(x, m, kth * 1.000001)
    idx, d, nd = O.topk(cand, nn, kth * 1.000001, m, k)
    assert len(idx) < k


def test_top1_matches_find_top1_rule():
    """k = 1 is find_top1_discord: argmax nnDist over nnDist > r, first index on ties."""
    cand = np.array([1, 1, 0, 1, 1], np.int32)
    nn = np.array([3.0, 5.0, 9.0, 5.0, 1.0])
    idx, d, nd = O.topk(cand, nn, 2.0, 1, 1)
    assert list(idx) == [1] and d[0] == 5.0 and nd == 3


# --- PALMAD driver (main.cu:136-233) ---------------------------------------------------

def test_palmad_schedule_first_length():
    """First length: r = 2*sqrt(minL), halved per failed attempt (main.cu:38,155)."""
    x = synth.rw_planted(5000, 21, 2, 48, 12.0)
    res = O.palmad(x, 40, 46, 1, 2)
    r0 = res[0]
    exp_r = 2.0 * np.sqrt(40) * 0.5 ** (r0["attempts"] - 1)
    assert r0["r2"] == pytest.approx(exp_r ** 2, rel=1e-12)
    # lengths 1..4 start from 0.99*sqrt(d_k of the previous length) and shrink by 0.99
    for j in range(1, 5):
        prev = res[j - 1]["d2"][-1]
        exp_r = 0.99 * np.sqrt(prev) * 0.99 ** (res[j]["attempts"] - 1)
        assert res[j]["r2"] == pytest.approx(exp_r ** 2, rel=1e-9)
    # j >= 5: r = M - 2S, then r -= S
    d = np.sqrt([r["d2"][-1] for r in res[:5]])
    M, S = d.mean(), d.std()
    exp_r = M - 2 * S - S * (res[5]["attempts"] - 1)
    assert res[5]["r2"] == pytest.approx(max(exp_r, 0) ** 2, rel=1e-9)


def test_palmad_result_equals_bruteforce_topk():
    x = synth.rw_planted(2500, 22, 2, 40, 10.0)
    res = O.palmad(x, 30, 34, 1, 3)
    for r in res:
        nn, _ = O.brute_profile(x, r["m"])
        exp = greedy_topk_reference(nn, r["m"], 3)
        assert list(r["idx"]) == exp
        np.testing.assert_allclose(r["d2"], nn[exp], rtol=1e-8)


def test_workload_fingerprints():
    with open(os.path.join(GOLD, "series_sha256.json")) as f:
        shas = json.load(f)
    assert synth.series_sha256(synth.workload_c1()["series"][0]) == shas["c1"]
    assert synth.series_sha256(synth.workload_c2()["series"][0]) == shas["c2"]


def test_mutual_neighbours_tie_to_lower_index():
    """Two exact copies of an anomaly far apart are each other's nearest neighbour at
    exactly the same distance; the lower index ranks first (sets_processing.hpp:28)."""
    rng = np.random.default_rng(17)
    x = np.cumsum(rng.standard_normal(3000))
    burst = 4.0 * np.sin(np.arange(40) * 2 * np.pi / 7.0) + rng.standard_normal(40)
    x[700:740] = x[700] + burst
    x[2200:2240] = x[2200] + burst
    x = x.astype(np.float32)
    m = 40
    nn = O.profile(x, m)
    cand, nnd = O.range_discords(x, m, 0.0)
    idx, d, _ = O.topk(cand, nnd, 0.0, m, 2)
    assert nn[700] == nn[2200]  # bitwise tie
    if d[0] == nn[700]:
        assert idx[0] == 700


# --- the column-blocked scan and the certificate scan order ------------------------------

def _scalar_profile(x, m):
    """NN ED^2 of every row by the one-pair-at-a-time loop of cpu_kernels.cpp:47-75 in
    double precision, in plain Python floats (IEEE double: each product rounded, then each
    sum, ascending k), on the oracle's own window statistics."""
    mu, sd = O.stats(x, m)
    xs = [float(v) for v in np.asarray(x, np.float32)]
    N = len(xs) - m + 1
    out = []
    for i in range(N):
        best = float(np.finfo(np.float64).max)
        for j in range(N):
            if abs(i - 1 - j) < m:
                continue
            dot = 0.0
            for k in range(m):
                dot += xs[i + k] * xs[j + k]
            den = float(m) * (float(sd[i]) * float(sd[j]))
            ed = 2.0 * m if not den > 0.0 else \
                2.0 * m * (1.0 - (dot - m * (float(mu[i]) * float(mu[j]))) / den)
            best = min(best, ed)
        out.append(best)
    return np.array(out)


def test_blocked_scan_is_bitwise_the_scalar_loop():
    """The scan evaluates eight columns at a time; every pair value, and so every minimum,
    must be bitwise the scalar loop's (N = 293 rows: blocks and a ragged tail)."""
    x = synth.ar2_mix(300, 31)
    m = 8
    assert np.array_equal(O.profile(x, m), _scalar_profile(x, m))


@pytest.mark.parametrize("kind", ["rw", "ecg", "ar2"])
def test_hinted_scan_order_gives_the_same_range_set(kind):
    """oracle_find_discords_hint_f64 changes only the order in which a row meets its columns:
    the same flags, and for the rows flagged bitwise the same NN distances, as the
    reference-order scan, at ranges from 0 to above every NN distance."""
    x, m = {"rw": (synth.rw_planted(6000, 5, 2, 64, 12.0), 64),
            "ecg": (synth.ecg_like(5000, 6, 1, 1), 80),
            "ar2": (synth.ar2_mix(4000, 7), 40)}[kind]
    nn_ref = O.profile(x, m)
    for q in (0.0, 0.5, 0.9, 0.999, 1.0):
        r2 = float(np.quantile(nn_ref, q)) if q > 0 else 0.0
        c1, n1 = O.range_discords(x, m, r2)
        c2, n2 = O.range_discords(x, m, r2, hint=True)
        assert np.array_equal(c1, c2), (kind, q)
        assert np.array_equal(n1[c1 == 1], n2[c2 == 1]), (kind, q)
        assert np.array_equal(n1[c1 == 1], nn_ref[c1 == 1]), (kind, q)


def test_sampled_rows_scan_is_the_full_scan_on_those_rows():
    """The CPU baseline's row sample (every stride-th run of rows, reference scan order) gives
    those rows exactly the full pass's flags and, where flagged, NN distances."""
    x = synth.ar2_mix(20000, 3)
    m = 50
    nn_ref = O.profile(x, m)
    for q in (0.5, 0.95):
        r2 = float(np.quantile(nn_ref, q))
        c, n = O.range_discords(x, m, r2)
        rows, cs, ns, _, _ = O.range_discords_sampled(x, m, r2, run=256, stride=3)
        assert len(rows) > 0 and np.all((rows // 256) % 3 == 0)
        assert np.array_equal(cs, c[rows])
        assert np.array_equal(ns[cs == 1], n[rows][cs == 1])
