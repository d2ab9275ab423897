"""Limits and error paths of the C ABI on a device: every failure comes back as a status
code (include/paddi.h), never as an abort, and lengths up to the largest supported m keep
parity with the oracle.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O
from paddi_amd import synth
from paddi_amd.engine import Context, Discord, Opts, PaddiError, Stats, load_library

pytestmark = pytest.mark.gpu
INVALID, DEVICE, UNSUPPORTED = 1, 2, 5


def test_features_of_a_non_finite_series_is_invalid(ctx):
    """paddi_features builds the prefix sums of its series, which flag NaN and infinity:
    status PADDI_ERR_INVALID, and the context stays usable."""
    x = synth.ar2_mix(5000, 9)
    y = x.copy()
    y[777] = np.nan
    h = ctx.upload(y)
    with pytest.raises(PaddiError) as e:
        ctx.features([h], 64, 1)
    assert e.value.status == INVALID
    f = ctx.features([ctx.upload(x)], 64, 1)
    assert np.isfinite(f).all()


def test_rows_beyond_32_bits_are_refused_before_any_read(ctx):
    """A call whose row space reaches 2^31 rows is PADDI_ERR_UNSUPPORTED, decided from the
    sizes alone: the buffers passed here are far shorter than the sizes claimed, so any read
    of them would fault the host."""
    L = load_library()
    small = np.zeros(64, np.float32)
    p = small.ctypes.data_as(C.c_void_p)
    big = 1 << 31
    h = C.c_void_p()
    assert L.paddi_series_upload(ctx._c, p, big, C.byref(h)) == UNSUPPORTED
    flag = np.zeros(4, np.int32)
    nn = np.zeros(4, np.float32)
    assert L.paddi_range_discords(ctx._c, p, big, 64, C.c_float(1.0), flag.ctypes.data,
                                  nn.ctypes.data) == UNSUPPORTED
    mu = np.zeros(4, np.float32)
    assert L.paddi_compute_statistics(ctx._c, p, big, 64, mu.ctypes.data,
                                      mu.ctypes.data) == UNSUPPORTED
    # two series of 1.1e9 points: each fits, together they do not
    ptrs = (C.c_void_p * 2)(small.ctypes.data, small.ctypes.data)
    ns = np.array([1_100_000_000, 1_100_000_000], np.int64)
    buf = (Discord * 8)()
    cnt = C.c_int64(0)
    st = Stats()
    o = Opts(1, 64, 0, 0, None)
    assert L.paddi_discords_batch(ctx._c, ptrs, ns.ctypes.data, 2, 64, 64, 1, 1, C.byref(o),
                                  buf, 8, C.byref(cnt), C.byref(st)) == UNSUPPORTED


def test_wait_limit_returns_a_device_error(monkeypatch):
    """PADDI_WAIT_LIMIT_S bounds every wait for queued device work: with a limit of zero the
    first wait that finds its work unfinished returns PADDI_ERR_DEVICE.  The kernels run to
    completion regardless (destroying the context waits for them); a new context works."""
    x = synth.random_walk(1 << 23, 5)[0].astype(np.float32)
    monkeypatch.setenv("PADDI_WAIT_LIMIT_S", "0")
    c = Context(0)
    monkeypatch.delenv("PADDI_WAIT_LIMIT_S")
    try:
        with pytest.raises(PaddiError) as e:
            c.discords(x, 128, 130, 1, 1, mode="batched")
        assert e.value.status == DEVICE
    finally:
        c.close()
    with Context(0) as d:
        y = synth.rw_planted(20000, 3, 1, 64, 12.0)
        ref = O.palmad(y, 64, 65, 1, 1)
        got = d.discords(y, 64, 65, 1, 1, mode="batched")
        assert [(g.index, g.length) for g in got] == \
            [(int(i), r["m"]) for r in ref for i in r["idx"]]


@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_lengths_near_the_largest_supported(ctx, mode):
    """m up to MAXM = 2048 (the scans' LDS ring and the exact kernel's staged windows are
    sized for it): the oracle's discords, and m = 2049 refused."""
    x = synth.rw_planted(8000, 21, 2, 2048, 300.0)
    ref = O.palmad(x, 2040, 2048, 4, 2)
    got = ctx.discords(x, 2040, 2048, 4, 2, mode=mode)
    exp = [(int(i), r["m"], float(d)) for r in ref for i, d in zip(r["idx"], r["d2"])]
    assert [(g.index, g.length) for g in got] == [(i, m) for i, m, _ in exp]
    for g, (_, _, d) in zip(got, exp):
        assert abs(g.nn_dist2 - d) <= 1e-4 * d
    with pytest.raises(PaddiError) as e:
        ctx.discords(x, 2049, 2049, 1, 1)
    assert e.value.status == UNSUPPORTED


def test_range_discords_at_the_largest_length(ctx):
    x = synth.ar2_mix(6000, 23)
    m = 2048
    nn_ref = O.profile(x, m)   # rows in the middle have no valid pair (DBL_MAX)
    fin = nn_ref[nn_ref < 1e300]
    for q in (0.5, 0.99):
        r2 = float(np.quantile(fin, q))
        flag, nn = ctx.range_discords(x, m, r2)
        cand, nn_o = O.range_discords(x, m, r2)
        near = np.abs(nn_ref - r2) <= 1e-7 * r2
        assert np.array_equal((flag == 1)[~near], (cand == 1)[~near]), q
        sel = (flag == 1) & (cand == 1) & (nn_o < 1e300)
        assert sel.any()
        assert (np.abs(nn[sel] - nn_o[sel]) / nn_o[sel]).max() <= 1e-4
