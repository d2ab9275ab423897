"""The entry point that allocates nothing (paddi_range_discords_noalloc), called through the
C ABI: the results of paddi_range_discords, bit for bit, on two calls in a row (the second
with a smaller series, so it runs on what the first left in the scratch array), at the
largest series it accepts, and with enough rows near the range that its exact refinement
runs in several launches; the oracle's results on a series it can check; and the refusal of
a series longer than its limit.
"""
import numpy as np
import pytest

from oracle import oracle as O
from paddi_amd import synth
from paddi_amd.engine import PaddiError

pytestmark = pytest.mark.gpu
MAX_POINTS = 300000


def delta_ed2(m):
    return 2.0 * m * (4.0e-5 + 2.0e-7 * m)


def _same(a, b):
    assert np.array_equal(a[0], b[0])
    sel = a[0] == 1
    assert np.array_equal(a[1][sel], b[1][sel])


def test_two_calls_match_the_allocating_entry_point(ctx):
    # white noise: at m = 2048 its NN distances crowd around their median, so thousands of
    # rows lie within the single-precision margin of a range there
    big = np.random.default_rng(61).standard_normal(MAX_POINTS).astype(np.float32)
    small = synth.rw_planted(50000, 62, 2, 96, 16.0)
    for x, m in ((big, 2048), (small, 96)):
        _, nn0 = ctx.range_discords(x, m, 0.0)
        r2 = float(np.median(nn0))
        ref = ctx.range_discords(x, m, r2)
        got = ctx.range_discords_noalloc(x, m, r2)
        _same(got, ref)
        flag, nn = got
        near = (flag == 1) & (nn > r2) & (nn <= r2 + delta_ed2(m))
        print(f"n={len(x)} m={m} r2={r2:.4g}: {int(flag.sum())} range discords, "
              f"{int(near.sum())} of them within the margin of r2 (refined exactly)")
        if len(x) == MAX_POINTS:
            # more rows than one launch of the call's exact refinement takes (4096)
            assert near.sum() > 4096


def test_matches_the_oracle(ctx):
    x = synth.ecg_like(6000, 63, 1, 1)
    m = 80
    nn_ref = O.profile(x, m)
    for q in (0.0, 0.5, 0.99):
        r2 = float(np.quantile(nn_ref, q)) if q > 0 else 0.0
        flag, nn = ctx.range_discords_noalloc(x, m, r2)
        cand, nn_o = O.range_discords(x, m, r2)
        near = np.abs(nn_ref - r2) <= 1e-7 * max(r2, 1e-3)
        assert np.array_equal((flag == 1)[~near], (cand == 1)[~near]), q
        sel = (flag == 1) & (cand == 1) & (nn_o < 1e300) & (nn_o > 1e-3)
        assert (np.abs(nn[sel] - nn_o[sel]) / nn_o[sel]).max() <= 1e-4


def test_longer_series_is_refused(ctx):
    x = synth.ar2_mix(MAX_POINTS + 1, 64)
    with pytest.raises(PaddiError) as e:
        ctx.range_discords_noalloc(x, 64, 10.0)
    assert e.value.status == 5
    flag, _ = ctx.range_discords_noalloc(x[:MAX_POINTS], 64, 10.0)
    assert len(flag) == MAX_POINTS - 63
