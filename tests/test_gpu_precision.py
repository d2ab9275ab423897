"""Series far outside the usual level-to-spread ratios: noisy square waves whose plateaus are
10^5 to 10^6 times their noise (levels +-100 and +-1000 on noise 1e-3, +-10 on noise 1e-5).

Windows on a plateau are near-flat and far from the level of the prefix block they start
in, so window statistics from prefix sums cancel almost every digit there, and the raw pair
formula (dot - m mu_i mu_j, src/cpu_kernels/cpu_kernels.cpp:55) cancels ten or more of
double precision's sixteen.  The reference's own one-pass variance (cpu_kernels.cpp:7-31)
fails here too, so the check is the brute force on explicitly z-normalised windows
(oracle.brute_profile: two-pass statistics per window, direct vector differences): the
engine must return its discords, never near-flat windows at the degenerate distance 2m.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-4
CASES = [(100.0, 1e-3, 11), (1000.0, 1e-3, 12), (10.0, 1e-5, 13)]
MINL, MAXL, STEP, K = 32, 40, 4, 3


def square(n, amp, noise, period, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    x = np.where((t // (period // 2)) % 2 == 0, amp, -amp) + noise * rng.standard_normal(n)
    return x.astype(np.float32)


def brute_topk(nn, m, k):
    order = sorted(range(len(nn)), key=lambda i: (-nn[i], i))
    acc = []
    for i in order:
        if len(acc) == k:
            break
        if np.isfinite(nn[i]) and all(abs(i - a) >= m for a in acc):
            acc.append(i)
    return acc, [float(nn[i]) for i in acc]


_REF = {}


def reference(amp, noise, seed):
    key = (amp, noise, seed)
    if key not in _REF:
        x = square(3000, amp, noise, 200, seed)
        _REF[key] = (x, {m: O.brute_profile(x, m)[0] for m in range(MINL, MAXL + 1, STEP)})
    return _REF[key]


@pytest.mark.parametrize("mode", ["plain", "batched", "learned"])
@pytest.mark.parametrize("amp,noise,seed", CASES, ids=[f"{a:g}on{n:g}" for a, n, _ in CASES])
def test_near_flat_plateaus_match_the_brute_force(ctx, amp, noise, seed, mode):
    x, nn = reference(amp, noise, seed)
    got = ctx.discords(x, MINL, MAXL, STEP, K, mode=mode)
    for m in range(MINL, MAXL + 1, STEP):
        idx, d = brute_topk(nn[m], m, K)
        mine = [g for g in got if g.length == m]
        assert [g.index for g in mine] == idx, (m, [(g.index, g.nn_dist2) for g in mine], idx, d)
        for g, v in zip(mine, d):
            assert g.nn_dist2 != 2.0 * m          # no near-flat window taken as degenerate
            assert abs(g.nn_dist2 - v) <= RTOL * v, (m, g, v)


@pytest.mark.parametrize("amp,noise,seed", CASES, ids=[f"{a:g}on{n:g}" for a, n, _ in CASES])
def test_near_flat_plateaus_range_flags(ctx, amp, noise, seed):
    x, nn = reference(amp, noise, seed)
    m = MINL
    ref = nn[m]
    for q in (0.5, 0.9, 0.99):
        r2 = float(np.quantile(ref, q))
        flag, got = ctx.range_discords(x, m, r2)
        near = np.abs(ref - r2) <= 1e-6 * r2
        assert np.array_equal((flag == 1)[~near], (ref > r2)[~near]), q
        sel = (flag == 1) & (ref > r2)
        assert (np.abs(got[sel] - ref[sel]) / ref[sel]).max() <= RTOL


@pytest.mark.parametrize("amp,noise,seed", CASES, ids=[f"{a:g}on{n:g}" for a, n, _ in CASES])
def test_near_flat_statistics(ctx, amp, noise, seed):
    """Window mean and standard deviation against the two-pass definition."""
    x, _ = reference(amp, noise, seed)
    for m in (MINL, MAXL):
        mu, sd = ctx.compute_statistics(x, m)
        _, sd_ref = O.znorm_windows(x, m)
        W = np.lib.stride_tricks.sliding_window_view(x.astype(np.float64), m)
        np.testing.assert_allclose(mu, W.mean(axis=1), rtol=1e-7)
        np.testing.assert_allclose(sd, sd_ref, rtol=1e-6)
