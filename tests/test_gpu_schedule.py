"""The engine's schedule tunables (the autotuner's search space) change speed, never results.

Every configuration below must return the discord set the oracle returns on c1 (indices and
lengths bit-exact, squared distances within 1e-4 relative), and the C ABI must reject
out-of-range values with PADDI_ERR_INVALID and leave the previous values in place.
"""
import json
import os

import pytest

from oracle import oracle as O
from paddi_amd import synth
from paddi_amd.engine import PaddiError

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
RTOL = 1e-4

SCHEDULES = [
    dict(witness_tiles=0),
    dict(witness_tiles=8, select_window=2048),
    dict(select_window=32768, refine_chunk=16384, min_chunk=512, refine_items=4096),
    dict(refine_chunk=131072, min_chunk=8192, refine_items=65536, speculate=1),
    dict(speculate=16, witness_tiles=1),
    dict(use_graph=1),
    dict(use_graph=1, witness_tiles=0, refine_items=4096),
    dict(prev_shift=0, witness_tiles=48),
    dict(prev_shift=4, speculate=2, bootstrap=0),
    dict(direct_budget=0),
    dict(direct_budget=1 << 20, direct_rows=4, refine_items=4096),
    dict(refine_rounds=4),
    dict(refine_rounds=4, refine_chunk=8192, min_chunk=512, use_graph=1),
    dict(seed_witnesses=0),
    dict(seed_witnesses=1, witness_tiles=48, select_window=256),
    dict(refine_order=1),
    dict(refine_order=1, refine_items=4096, min_chunk=512, use_graph=1),
    dict(refine_order=2),
]


@pytest.fixture
def restore(ctx):
    saved = ctx.get_tunables()
    yield
    ctx.set_tunables(**saved)


def test_tunables_roundtrip_and_validation(ctx, restore):
    t = ctx.get_tunables()
    assert t["select_window"] % 64 == 0 and t["min_chunk"] <= t["refine_chunk"]
    ctx.set_tunables(select_window=4000, witness_tiles=2)
    u = ctx.get_tunables()
    assert u["select_window"] == 4032 and u["witness_tiles"] == 2   # rounded up to 64
    for bad in (dict(select_window=0), dict(min_chunk=1 << 20, refine_chunk=1 << 16),
                dict(witness_tiles=-1), dict(witness_tiles=65), dict(speculate=0),
                dict(prev_shift=5), dict(bootstrap=2), dict(direct_budget=-1),
                dict(direct_rows=5), dict(refine_rounds=2), dict(refine_rounds=5),
                dict(seed_witnesses=2), dict(refine_order=3)):
        with pytest.raises(PaddiError):
            ctx.set_tunables(**bad)
        assert ctx.get_tunables() == u


@pytest.mark.parametrize("mode", ["batched", "learned"])
@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: ",".join(f"{k}={v}" for k, v in s.items()))
def test_schedule_does_not_change_discords(ctx, restore, sched, mode):
    w = synth.workload_c1()
    with open(os.path.join(GOLD, "palmad_c1.json")) as f:
        ref = json.load(f)
    exp = [(int(i), r["m"], float(d)) for r in ref for i, d in zip(r["idx"], r["d2"])]
    ctx.set_tunables(**sched)
    got = ctx.discords(w["series"][0], w["minL"], w["maxL"], w["step"], w["k"], mode=mode)
    assert [(g.index, g.length) for g in got] == [(i, m) for i, m, _ in exp]
    for g, (_, _, d) in zip(got, exp):
        assert abs(g.nn_dist2 - d) <= RTOL * d


def test_graph_replays_match_eager_and_oracle(ctx, restore):
    """Captured pass graphs, replayed over lengths of two prefetch classes (m <= 128 and
    m > 128) and across calls, give the eager launches' discords, which are the oracle's."""
    series = [synth.rw_planted(20000, 31, 2, 48, 12.0), synth.ecg_like(30000, 32, 1, 1)]
    minL, maxL, step, k = 120, 136, 4, 2
    ctx.set_tunables(use_graph=0)
    eager = ctx.discords_batch(series, minL, maxL, step, k, mode="batched")
    ctx.set_tunables(use_graph=1)
    runs = [ctx.discords_batch(series, minL, maxL, step, k, mode=mode)
            for mode in ("batched", "batched", "learned", "plain")]
    key = [(g.series, g.index, g.length, g.nn_dist2) for g in eager]
    for r in runs:
        assert [(g.series, g.index, g.length, g.nn_dist2) for g in r] == key
    for q, x in enumerate(series):
        ref = O.palmad(x, minL, maxL, step, k)
        exp = [(int(i), r["m"], float(d)) for r in ref for i, d in zip(r["idx"], r["d2"])]
        mine = [g for g in eager if g.series == q]
        assert [(g.index, g.length) for g in mine] == [(i, m) for i, m, _ in exp]
        for g, (_, _, d) in zip(mine, exp):
            assert abs(g.nn_dist2 - d) <= RTOL * d


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("mode", ["batched", "learned"])
@pytest.mark.parametrize("step", [1, 4])
def test_next_length_range_from_the_exact_refinements(ctx, restore, mode, step, graph):
    """With prev_shift = 0 a length's r comes from its previous discords' distances at this
    length, computed by the previous length's exact refinements (a second length in the same
    launches) instead of a launch of its own: over several lengths, steps of 1 and 4, two
    uneven series, eager or replayed passes, the discords stay the oracle's."""
    series = [synth.rw_planted(9000, 61, 2, 64, 12.0), synth.ar2_mix(7000, 62)]
    minL, maxL, k = 48, 64, 3
    ctx.set_tunables(prev_shift=0, use_graph=graph)
    got = ctx.discords_batch(series, minL, maxL, step, k, mode=mode)
    for q, x in enumerate(series):
        ref = O.palmad(x, minL, maxL, step, k)
        exp = [(int(i), r["m"], float(d)) for r in ref for i, d in zip(r["idx"], r["d2"])]
        mine = [g for g in got if g.series == q]
        assert [(g.index, g.length) for g in mine] == [(i, m) for i, m, _ in exp], (q, step)
        for g, (_, _, d) in zip(mine, exp):
            assert abs(g.nn_dist2 - d) <= RTOL * d


def test_next_length_range_with_large_groups(ctx, restore):
    """Top-8 with prev_shift = 0 and speculate = 8: exact launches of up to eight candidates
    that also evaluate them at the next length (the widest two-length kernels)."""
    x = synth.rw_planted(16000, 93, 4, 40, 10.0)
    ctx.set_tunables(prev_shift=0, speculate=8)
    ref = O.palmad(x, 40, 46, 1, 8)
    exp = [(int(i), r["m"], float(d)) for r in ref for i, d in zip(r["idx"], r["d2"])]
    got = ctx.discords(x, 40, 46, 1, 8, mode="learned")
    assert [(g.index, g.length) for g in got] == [(i, m) for i, m, _ in exp]
    for g, (_, _, d) in zip(got, exp):
        assert abs(g.nn_dist2 - d) <= RTOL * d


@pytest.mark.parametrize("mode", ["batched", "learned"])
def test_first_length_sax_witnesses_change_nothing(ctx, restore, mode):
    """The first length's witnesses from SAX words (calls of 2^18 rows or more) only order the
    scans: one long series and a batch of uneven ones (whose words must stay within their
    series) return the same records with and without them (the pairs each evaluated are
    printed)."""
    cases = [[synth.rw_planted(400_000, 71, 3, 96, 12.0)],
             [synth.rw_planted(60_000 + 17_000 * q, 80 + q, 1, 64, 10.0) if q % 2 else
              synth.ecg_like(50_000 + 13_000 * q, 80 + q, 1, 1) for q in range(6)]]
    for series in cases:
        out = {}
        for sw in (0, 1):
            ctx.set_tunables(seed_witnesses=sw)
            recs, st = ctx.discords_batch(series, 64, 68, 1, 2, mode=mode, with_stats=True)
            out[sw] = ([(r.series, r.index, r.length, r.nn_dist2, r.range2) for r in recs], st)
        assert out[1][0] == out[0][0]
        print(f"{len(series)} series: pairs without {out[0][1]['pairs']:.4g}, "
              f"with {out[1][1]['pairs']:.4g}", flush=True)
