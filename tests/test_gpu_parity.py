"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

Bar: discord indices and lengths bit-exact; squared NN distances within 1e-4 relative of
the double-precision oracle (the engine reports the exact double-precision value for
accepted discords, so the observed error is ~1e-7); range-discord flags bit-exact except
for rows whose exact distance is within 1e-7 relative of r2, where two double-precision
evaluations in different operation orders may legitimately disagree.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from paddi_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
RTOL = 1e-4


def brute_cases():
    z = np.load(os.path.join(GOLD, "brute_profiles.npz"))
    names = sorted({k.split("__")[0] for k in z.files})
    return [(n, z[f"{n}__x"], int(z[f"{n}__m"][0]), z[f"{n}__nn"]) for n in names]


def expected(ref):
    return [(int(i), r["m"], float(d)) for r in ref for i, d in zip(r["idx"], r["d2"])]


def assert_same_discords(got, exp, tag=""):
    assert len(got) == len(exp), (tag, [(g.index, g.length) for g in got], exp)
    worst = 0.0
    for g, (i, m, d) in zip(got, exp):
        assert (g.index, g.length) == (i, m), (tag, g, i, m)
        if d < 3e38:
            rel = abs(g.nn_dist2 - d) / max(d, 1e-30)
            worst = max(worst, rel)
            assert rel <= RTOL, (tag, g, d, rel)
    return worst


def test_statistics_match_oracle(ctx):
    for x, m in ((synth.ar2_mix(5000, 31), 50),
                 (synth.rw_planted(1 << 20, 2, 4, 192, 24.0), 256),
                 (synth.ecg_like(20000, 32, 1, 1), 100)):
        mu, sd = ctx.compute_statistics(x, m)
        mu_r, sd_r = O.stats(x, m)
        np.testing.assert_allclose(mu, mu_r, rtol=2e-7, atol=1e-6)
        np.testing.assert_allclose(sd, sd_r, rtol=2e-6)


@pytest.mark.parametrize("case", brute_cases(), ids=lambda c: c[0])
def test_range_discords_match_oracle(ctx, case):
    name, x, m, nn_ref = case
    for q in (0.0, 0.5, 0.9, 0.99, 0.999):
        r2 = float(np.quantile(nn_ref, q)) if q > 0 else 0.0
        flag, nn = ctx.range_discords(x, m, r2)
        cand, nn_o = O.range_discords(x, m, r2)
        near = np.abs(nn_ref - r2) <= 1e-7 * max(r2, 1e-3)
        assert np.array_equal((flag == 1)[~near], (cand == 1)[~near]), (name, q)
        sel = (flag == 1) & (cand == 1) & (nn_o < 1e300) & (nn_o > 1e-3)
        if sel.any():
            rel = np.abs(nn[sel] - nn_o[sel]) / nn_o[sel]
            assert rel.max() <= RTOL, (name, q, rel.max())


def delta_ed2(m):
    """The engine's pruning margin on squared distances (engine.cpp, delta_ed2)."""
    return 2.0 * m * (4.0e-5 + 2.0e-7 * m)


@pytest.mark.parametrize("m", [16, 100, 241, 242, 300])
def test_single_precision_error_within_margin(ctx, m):
    """The single-precision distances (matrix-core seeds + recurrence) stay within a quarter
    of the margin that makes pruning safe, on a series far from zero (level 3000), at lengths
    on both sides of the seed kernel's window-size steps (m + 15 <= 256 < m + 15)."""
    x = (synth.random_walk(5000, 77)[0] + 3000.0).astype(np.float32)
    flag, nn = ctx.range_discords(x, m, 0.0)
    _, nn_o = O.range_discords(x, m, 0.0)
    sel = (flag == 1) & (nn_o < 1e300)
    assert sel.sum() > 1000
    err = np.abs(nn[sel].astype(np.float64) - nn_o[sel]).max()
    assert err <= delta_ed2(m) / 4, (m, err, delta_ed2(m))


@pytest.mark.parametrize("kind,n,minL,maxL,step,k", [
    ("rw", 4000, 40, 44, 1, 2),
    ("rw", 9000, 64, 70, 3, 3),
    ("ecg", 12000, 60, 66, 2, 3),
    ("ar2", 8000, 30, 36, 1, 2),
    ("rw", 2000, 100, 100, 1, 3),
    ("ar2", 5000, 8, 12, 1, 4),
])
def test_discords_match_oracle(ctx, kind, n, minL, maxL, step, k):
    x = {"rw": lambda: synth.rw_planted(n, 40 + n % 97, 2, max(minL, 16), 12.0),
         "ecg": lambda: synth.ecg_like(n, 41, 1, 1),
         "ar2": lambda: synth.ar2_mix(n, 42)}[kind]()
    ref = O.palmad(x, minL, maxL, step, k)
    got = ctx.discords(x, minL, maxL, step, k, mode="plain")
    worst = assert_same_discords(got, expected(ref), kind)
    print(f"{kind} n={n} m={minL}..{maxL}/{step} k={k}: max rel err vs FP64 = {worst:.2e}")
    # the plain form follows the same r-schedule as the oracle driver
    for r in ref:
        gs = [g for g in got if g.length == r["m"]]
        if gs:
            assert gs[0].range2 == pytest.approx(r["r2"], rel=1e-5, abs=1e-6)


def test_c1_matches_cached_oracle(ctx):
    w = synth.workload_c1()
    with open(os.path.join(GOLD, "palmad_c1.json")) as f:
        ref = json.load(f)
    got = ctx.discords(w["series"][0], w["minL"], w["maxL"], w["step"], w["k"], mode="plain")
    worst = assert_same_discords(got, expected(ref), "c1")
    print(f"c1: max rel err vs FP64 = {worst:.2e}")


def test_edge_cases(ctx):
    # fewer windows than one 64-row block; ragged last block; k larger than possible
    for n, m, k in ((60, 8, 2), (200, 16, 3), (131, 10, 20), (1000, 31, 5)):
        x = synth.ar2_mix(n, n)
        ref = O.palmad(x, m, m, 1, k)
        got = ctx.discords(x, m, m, 1, k, mode="plain")
        assert_same_discords(got, expected(ref), f"n={n} m={m} k={k}")


def test_degenerate_windows(ctx):
    x = synth.rw_planted(3000, 50, 1, 32, 8.0).astype(np.float64)
    x[1000:1200] = 2.5  # flat: zero standard deviation, ED^2 = 2m to everything
    x[2000:2100] = 0.0
    x = x.astype(np.float32)
    for m in (24, 40):
        ref = O.palmad(x, m, m, 1, 3)
        got = ctx.discords(x, m, m, 1, 3, mode="plain")
        assert_same_discords(got, expected(ref), f"flat m={m}")


def ill_conditioned_series():
    """Series whose window scales change by orders of magnitude within a few windows."""
    t = np.arange(12000.0)
    rng = np.random.default_rng(4)
    square = np.where((t // 150) % 2 == 0, 10.0, -10.0) + rng.normal(0.0, 0.01, t.size)
    spikes = synth.ar2_mix(12000, 7).astype(np.float64)
    spikes[3000] += 1e5
    spikes[8000] -= 1e5 / 3
    offset = synth.ar2_mix(12000, 5).astype(np.float64) + 1e4
    return [("square", square.astype(np.float32), 40, 44),
            ("spikes", spikes.astype(np.float32), 48, 52),
            ("offset", offset.astype(np.float32), 40, 44)]


@pytest.mark.parametrize("mode", ["plain", "batched", "learned"])
def test_ill_conditioned_series_match_oracle(ctx, mode):
    """A noisy square wave (plateaus of spread 0.01 between jumps of 20), spikes of 1e5 on
    unit-scale noise, and an offset of 1e4: a single-precision chain that crosses a jump
    carries rounding of the jump's scale into cells of the plateau's, thousands of times
    smaller, so the pruning margin has to grow with that ratio (a fixed margin misses
    discords on the square wave at m = 42 and 44)."""
    for name, x, lo, hi in ill_conditioned_series():
        ref = O.palmad(x, lo, hi, 2, 3)
        got = ctx.discords(x, lo, hi, 2, 3, mode=mode)
        assert_same_discords(got, expected(ref), f"{name} {mode}")


def test_ill_conditioned_range_flags(ctx):
    name, x, m, _ = ill_conditioned_series()[0]
    nn_ref = O.profile(x, m)
    for q in (0.5, 0.99, 0.999):
        r2 = float(np.quantile(nn_ref, q))
        flag, _ = ctx.range_discords(x, m, r2)
        cand, _ = O.range_discords(x, m, r2)
        near = np.abs(nn_ref - r2) <= 1e-7 * r2
        assert np.array_equal((flag == 1)[~near], (cand == 1)[~near]), (name, q)


@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_long_windows_match_oracle(ctx, mode):
    """m near 1000: the scans' windows and seed products span more than 1000 values, and the
    exact refinement's LDS bounds its group size (groups of up to four here)."""
    x = synth.rw_planted(6000, 77, 2, 1000, 12.0)
    ref = O.palmad(x, 1016, 1020, 4, 2)
    got = ctx.discords(x, 1016, 1020, 4, 2, mode=mode)
    assert_same_discords(got, expected(ref), mode)


@pytest.mark.parametrize("mode", ["plain", "batched", "learned"])
def test_mutual_neighbours_tie_to_the_lower_index(ctx, mode):
    """Rows 426 and 2691 (m = 24) are each other's nearest neighbour: the oracle's distances
    tie exactly and the greedy top-k takes the lower index first.  The engine's double-
    precision values of the two rows differ by rounding only (its pair value centres the
    candidate's side), and its top-k ranks distances within 1e-9 of each other as a tie, by
    index, so it reports the tie in the same order (ranking by the raw values would order
    the two by rounding noise)."""
    x = synth.rw_planted(4000, 196, 2, 48, 10.0)
    ref = O.palmad(x, 24, 24, 1, 3)
    assert list(ref[0]["idx"][1:]) == [426, 2691] and ref[0]["d2"][1] == ref[0]["d2"][2]
    got = ctx.discords(x, 24, 24, 1, 3, mode=mode)
    assert_same_discords(got, expected(ref), mode)
    assert abs(got[1].nn_dist2 - got[2].nn_dist2) <= 1.2e-7 * got[1].nn_dist2  # one float ulp


def test_deterministic_replay(ctx):
    x = synth.rw_planted(30000, 60, 3, 64, 16.0)
    a = ctx.discords(x, 60, 66, 1, 3)
    b = ctx.discords(x, 60, 66, 1, 3)
    assert [(d.index, d.length, d.nn_dist2) for d in a] == \
        [(d.index, d.length, d.nn_dist2) for d in b]


def test_certificate_at_scale(ctx):
    """262,144-point random walk: the top-3 per length is certified by the oracle run at
    r2' = (1 - 1e-3) * d_3: if the oracle's range set at r2' yields the same greedy top-3,
    the engine's answer is the exact one (the top-k is r-independent below d_k)."""
    x = synth.rw_planted(262144, 70, 6, 160, 24.0)
    got = ctx.discords(x, 128, 130, 1, 3)
    for m in (128, 130):
        gs = [g for g in got if g.length == m]
        assert len(gs) == 3
        r2 = (1 - 1e-3) * gs[-1].nn_dist2
        cand, nn = O.range_discords(x, m, r2)
        idx, d, nd = O.topk(cand, nn, r2, m, 3)
        assert [g.index for g in gs] == list(idx), (m, gs, idx)
        np.testing.assert_allclose([g.nn_dist2 for g in gs], d, rtol=RTOL)


# --- the batched form (all series in every launch, tight per-length r, witness tiles) ------

@pytest.mark.parametrize("kind,n,minL,maxL,step,k", [
    ("rw", 9000, 64, 72, 1, 3),
    ("ecg", 12000, 60, 66, 2, 3),
    ("ar2", 5000, 8, 12, 1, 4),
    ("rw", 30000, 100, 110, 5, 2),
])
def test_batched_matches_oracle(ctx, kind, n, minL, maxL, step, k):
    x = {"rw": lambda: synth.rw_planted(n, 40 + n % 97, 2, max(minL, 16), 12.0),
         "ecg": lambda: synth.ecg_like(n, 41, 1, 1),
         "ar2": lambda: synth.ar2_mix(n, 42)}[kind]()
    ref = O.palmad(x, minL, maxL, step, k)
    got, st = ctx.discords(x, minL, maxL, step, k, mode="batched", with_stats=True)
    worst = assert_same_discords(got, expected(ref), kind)
    print(f"batched {kind} n={n}: max rel err {worst:.2e}, attempts {st['attempts']}")


@pytest.mark.parametrize("kind,n,minL,maxL,step,k", [
    ("rw", 20000, 64, 70, 2, 3),
    ("ecg", 15000, 100, 104, 1, 2),
    ("ar2", 12000, 256, 264, 4, 3),
])
def test_learned_matches_oracle(ctx, kind, n, minL, maxL, step, k):
    """The learned proposal changes only how many passes run, never the discord set."""
    x = {"rw": lambda: synth.rw_planted(n, 50 + n % 89, 2, 64, 12.0),
         "ecg": lambda: synth.ecg_like(n, 43, 1, 1),
         "ar2": lambda: synth.ar2_mix(n, 44)}[kind]()
    ref = O.palmad(x, minL, maxL, step, k)
    got, st = ctx.discords(x, minL, maxL, step, k, mode="learned", with_stats=True)
    assert_same_discords(got, expected(ref), f"learned {kind}")


def test_learned_proposal_saves_passes(ctx):
    """What the trained network gains: on 24 series it never saw (random walks, ECG-like and
    AR(2) mixtures of 20k-60k points, seeds outside the training set), the learned first-length
    proposal replaces the probe descent from r2 = 2m, so the call runs fewer passes, while the
    discord set is identical to the batched form's (and the batched form is the oracle's,
    tested above)."""
    rng = np.random.default_rng(2024)
    series = []
    for q in range(24):
        n = int(rng.integers(20000, 60000))
        kind = q % 3
        series.append(synth.rw_planted(n, 5000 + q, 2, 64, 16.0) if kind == 0 else
                      synth.ecg_like(n, 6000 + q, 1, 1) if kind == 1 else
                      synth.ar2_mix(n, 7000 + q))
    minL, maxL, k = 64, 67, 2
    b, sb = ctx.discords_batch(series, minL, maxL, 1, k, mode="batched", with_stats=True)
    l, sl = ctx.discords_batch(series, minL, maxL, 1, k, mode="learned", with_stats=True)
    assert [(g.series, g.index, g.length, g.nn_dist2) for g in l] == \
        [(g.series, g.index, g.length, g.nn_dist2) for g in b]
    print(f"passes: batched {sb['attempts']}, learned {sl['attempts']}; "
          f"pairs {sb['pairs']:.3g} vs {sl['pairs']:.3g}")
    assert sl["attempts"] < sb["attempts"]


def test_learned_proposal_too_high_falls_back(ctx, tmp_path):
    """A model whose margin is hugely negative proposes r far above every discord distance.
    Without the candidate bootstrap the first pass runs at that proposal and the descent must
    fall back; with it, the k discords found among the candidate blocks fix the first range
    and the proposal is not needed.  Both return the oracle's set."""
    from paddi_amd import model as Mdl
    d = Mdl.load()
    bad = tmp_path / "bad.bin"
    Mdl.save(str(bad), d["fmu"], d["fsd"], d["params"], -3.0)
    x = synth.rw_planted(20000, 77, 2, 64, 12.0)
    ref = O.palmad(x, 64, 66, 1, 2)
    saved = ctx.get_tunables()
    try:
        ctx.set_tunables(bootstrap=0)
        got, st = ctx.discords(x, 64, 66, 1, 2, mode="learned", model_path=str(bad),
                               with_stats=True)
        assert_same_discords(got, expected(ref), "fallback")
        assert st["attempts"] > 3
        ctx.set_tunables(bootstrap=1)
        got, st = ctx.discords(x, 64, 66, 1, 2, mode="learned", model_path=str(bad),
                               with_stats=True)
        assert_same_discords(got, expected(ref), "bootstrap")
        assert st["attempts"] == 3
    finally:
        ctx.set_tunables(**saved)


def test_batched_uneven_series_match_oracle(ctx):
    """Many uneven series in one call: each series' discords equal the oracle's on it."""
    rng = np.random.default_rng(8)
    series = []
    for q in range(12):
        n = int(rng.integers(2000, 12000))
        series.append(synth.rw_planted(n, 100 + q, 1, 40, 10.0) if q % 2 == 0
                      else synth.ecg_like(n, 200 + q, 1, 0))
    minL, maxL, k = 40, 44, 1
    got = ctx.discords_batch(series, minL, maxL, 1, k, mode="batched")
    plain = ctx.discords_batch(series, minL, maxL, 1, k, mode="plain")
    assert [(g.series, g.index, g.length) for g in got] == \
        [(g.series, g.index, g.length) for g in plain]
    for q, x in enumerate(series):
        ref = O.palmad(x, minL, maxL, 1, k)
        mine = [g for g in got if g.series == q]
        assert_same_discords(mine, expected(ref), f"series {q}")


def test_batched_c1_matches_cached_oracle(ctx):
    w = synth.workload_c1()
    with open(os.path.join(GOLD, "palmad_c1.json")) as f:
        ref = json.load(f)
    got = ctx.discords(w["series"][0], w["minL"], w["maxL"], w["step"], w["k"], mode="batched")
    assert_same_discords(got, expected(ref), "c1 batched")


def test_non_finite_series_is_rejected(ctx):
    """A NaN or infinity anywhere in a series is an invalid argument (status 1), found while
    the prefix sums are built; the context stays usable."""
    from paddi_amd.engine import PaddiError
    x = synth.ar2_mix(3000, 5)
    for bad in (np.nan, np.inf, -np.inf):
        y = x.copy()
        y[1234] = bad
        with pytest.raises(PaddiError) as e:
            ctx.discords(y, 32, 34, 1, 2)
        assert e.value.status == 1
        with pytest.raises(PaddiError) as e:
            ctx.range_discords(y, 32, 10.0)
        assert e.value.status == 1
    ref = O.palmad(x, 32, 34, 1, 2)
    assert_same_discords(ctx.discords(x, 32, 34, 1, 2), expected(ref), "after rejection")


@pytest.mark.parametrize("mode", ["batched", "learned"])
def test_batched_edge_series_match_oracle(ctx, mode):
    """One call mixing a series with a single window at the longest length, series shorter
    than one 64-row block, a ragged one and normal ones, with k larger than some of them
    can hold: each series' discords equal the oracle's on it alone."""
    minL, maxL, step, k = 20, 26, 2, 4
    series = [synth.ar2_mix(maxL + 1, 11), synth.ar2_mix(70, 12), synth.ar2_mix(131, 13),
              synth.rw_planted(3000, 14, 1, 24, 8.0), synth.ecg_like(2500, 15, 1, 0)]
    got = ctx.discords_batch(series, minL, maxL, step, k, mode=mode)
    for q, x in enumerate(series):
        ref = O.palmad(x, minL, maxL, step, k)
        mine = [g for g in got if g.series == q]
        assert_same_discords(mine, expected(ref), f"series {q} n={len(x)}")


@pytest.mark.parametrize("mode", ["batched", "learned"])
def test_large_k_matches_oracle(ctx, mode):
    """Top-12 per length: the greedy top-k takes many exact rounds and drops many trivial
    matches, and the next length's range comes from twelve previous discords."""
    x = synth.rw_planted(20000, 91, 4, 48, 10.0)
    ref = O.palmad(x, 48, 52, 1, 12)
    got = ctx.discords(x, 48, 52, 1, 12, mode=mode)
    assert_same_discords(got, expected(ref), mode)
