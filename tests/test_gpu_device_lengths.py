"""The length loop on the device (paddi_tunables.device_lengths).

Every length after the first runs as one captured graph: its range from the previous length's
discords (their distances at the new length from the previous length's exact rounds, or, with
prev_shift > 0, from a launch of its own on the shifted rows), the pass, the exact top-k
(k_chain_topk, k_exact_chain); the host only replays graphs and synchronises every
`device_lengths` lengths.  The loop
must return what the host loop (src/main.cu:136-233's length loop as Engine::fast runs it)
returns, bit for bit: the same kernels compute every exact value, the greedy top-k and the range
rule are the same.  A length the device cannot finish (more exact rounds than its graph holds,
or a pass with fewer than k discords above its range) is finished by the host, without changing
the answer.  The oracle checks a small case end to end.
"""
import numpy as np
import pytest

from oracle import oracle as O
from paddi_amd import synth

pytestmark = pytest.mark.gpu
RTOL = 1e-4


def uneven(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        n = int(rng.integers(6_000, 40_000))
        if q % 2:
            out.append(synth.ecg_like(n, seed + q, 1, 0))
        else:
            out.append(synth.rw_planted(n, seed + q, 1, 48, 8.0))
    return out


CASES = {
    # name: (series, minL, maxL, step, k)
    "rw_k3": (lambda: [synth.rw_planted(150_000, 5, 3, 96, 12.0)], 64, 100, 1, 3),
    "ecg_k2_step2": (lambda: [synth.ecg_like(120_000, 6, 2, 2)], 60, 96, 2, 2),
    "ar2_k3_step4": (lambda: [synth.ar2_mix(100_000, 7)], 128, 192, 4, 3),
    "uneven_k1": (lambda: uneven(40, 11), 40, 60, 1, 1),
    "uneven_k2_step3": (lambda: uneven(12, 12), 48, 84, 3, 2),
    # a series too short for two discords at the longer lengths beside a long one: the
    # lengths where it falls short go back to the host (three attempts per length: its
    # descents would otherwise run to the default 64)
    "short_series_k2": (lambda: [synth.rw_planted(60_000, 23, 2, 48, 8.0),
                                 synth.rw_planted(190, 24, 0, 0, 0.0)], 56, 100, 2, 2,
                        dict(max_attempts=3)),
}
_HOST = {}


def key(recs):
    return [(r.index, r.length, r.series, r.nn_dist2, r.range2) for r in recs]


@pytest.fixture
def restore(ctx):
    saved = ctx.get_tunables()
    yield
    ctx.set_tunables(**saved)


def run(ctx, name, mode, shift=0, **tun):
    series, minL, maxL, step, k, *call = CASES[name]
    ctx.set_tunables(prev_shift=shift, **tun)
    return ctx.discords_batch(series(), minL, maxL, step, k, mode=mode, with_stats=True,
                              **(call[0] if call else {}))


def host(ctx, name, mode, shift=0):
    if (name, mode, shift) not in _HOST:
        _HOST[(name, mode, shift)] = run(ctx, name, mode, shift, device_lengths=0)
    return _HOST[(name, mode, shift)]


@pytest.mark.parametrize("shift", [0, 2])
@pytest.mark.parametrize("mode", ["batched", "learned"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_device_loop_equals_host_loop(ctx, restore, name, mode, shift):
    ref, st0 = host(ctx, name, mode, shift)
    print(f"device loop {name} {mode} shift {shift}: {len(ref)} discords", flush=True)
    for every in (1, 5, 64):
        got, st = run(ctx, name, mode, shift, device_lengths=every, device_rounds=2)
        assert key(got) == key(ref), (name, mode, shift, every)
        # the lengths the device ran were no host passes: fewer eager scan launches
        assert st["launches"]["select"] < st0["launches"]["select"], (st, st0)
    # twice in a row: the graphs captured by the first call are replayed by the second
    again, _ = run(ctx, name, mode, shift, device_lengths=5, device_rounds=2)
    assert key(again) == key(ref)


@pytest.mark.parametrize("shift", [0, 1])
@pytest.mark.parametrize("name", ["rw_k3", "uneven_k2_step3"])
def test_lengths_the_device_cannot_finish_go_back_to_the_host(ctx, restore, name, shift):
    """One exact round per length and one candidate per round: every length whose top-k needs
    a second round stops the device loop there; the host finishes it and restarts the loop."""
    ref, _ = host(ctx, name, "batched", shift)
    got, st = run(ctx, name, "batched", shift, device_lengths=3, device_rounds=1, speculate=1)
    assert key(got) == key(ref)
    got4, _ = run(ctx, name, "batched", shift, device_lengths=3, device_rounds=4, speculate=4)
    assert key(got4) == key(ref)


@pytest.mark.parametrize("shift", [0, 2])
@pytest.mark.parametrize("name", ["ecg_k2_step2", "uneven_k2_step3"])
def test_four_refine_rounds(ctx, restore, name, shift):
    """A fourth refine round (offsets split again at 32 x refine_chunk) with chunk-major work
    items, on the host and on the device loop (whose lengths set the rounds' bounds and order on
    the device): the same results."""
    ref, _ = host(ctx, name, "learned", shift)
    for every in (0, 5):
        got, _ = run(ctx, name, "learned", shift, device_lengths=every, device_rounds=2,
                     refine_rounds=4, refine_chunk=4096, min_chunk=512, refine_order=1)
        assert key(got) == key(ref), (name, shift, every)


@pytest.mark.parametrize("shift", [0, 3])
def test_device_loop_matches_the_oracle(ctx, restore, shift):
    x = synth.rw_planted(8000, 9, 2, 64, 16.0)
    minL, maxL, k = 40, 52, 2
    ref = O.palmad(x, minL, maxL, 1, k)
    ctx.set_tunables(prev_shift=shift, device_lengths=4, device_rounds=2)
    got = ctx.discords(x, minL, maxL, 1, k, mode="batched")
    exp = [(int(i), r["m"], float(d)) for r in ref for i, d in zip(r["idx"], r["d2"])]
    assert [(g.index, g.length) for g in got] == [(i, m) for i, m, _ in exp]
    for g, (_, _, d) in zip(got, exp):
        assert abs(g.nn_dist2 - d) <= RTOL * d, (g, d)


def test_device_tunables_validation(ctx, restore):
    from paddi_amd.engine import PaddiError
    t = ctx.get_tunables()
    for bad in (dict(device_lengths=-1), dict(device_rounds=0), dict(device_rounds=5)):
        with pytest.raises(PaddiError):
            ctx.set_tunables(**bad)
        assert ctx.get_tunables() == t
