"""The learned range proposal: the double-precision restatement (oracle/mlp.py) pinned by
central finite differences, then the HIP forward/backward/Adam kernels and the feature
kernel checked against it on the GPU."""
import os

import numpy as np
import pytest

from oracle import mlp as R
from paddi_amd import model as Mdl
from paddi_amd import synth


def small_problem(B=37, seed=3):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, R.NFEAT)).astype(np.float32)
    T = (X[:, 0] * 0.5 - X[:, 3] ** 2 * 0.1 + 0.2).astype(np.float32)
    fmu = X.mean(0)
    fsd = X.std(0) + 1e-3
    p = Mdl.init_params(R.NFEAT, R.HID, seed=seed)
    return p, fmu, fsd, X, T


def test_restatement_gradient_matches_finite_differences():
    p, fmu, fsd, X, T = small_problem()
    p = p.astype(np.float64)
    loss, g = R.loss_grad(p, fmu, fsd, X, T)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(R.NPARAM - 1, 60, replace=False), [R.NPARAM - 1]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (R.loss_grad(p + e, fmu, fsd, X, T)[0] - R.loss_grad(p - e, fmu, fsd, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_planted_gradient_mistake_is_caught():
    """Dropping the ReLU mask in the backward pass is visible to the finite-difference check."""
    p, fmu, fsd, X, T = small_problem()
    p = p.astype(np.float64)
    W1, b1, W2, b2, W3, b3 = R.unpack(p)
    Y, (Z, A1, H1, A2, H2) = R.forward(p, fmu, fsd, X)
    dY = 2.0 * (Y - T) / len(Y)
    dA2 = np.outer(dY, W3)  # planted: mask (A2 > 0) dropped
    wrong_gb2 = dA2.sum(0)
    _, g = R.loss_grad(p, fmu, fsd, X, T)
    o = R.NFEAT * R.HID + R.HID + R.HID * R.HID
    assert not np.allclose(wrong_gb2, g[o:o + R.HID], rtol=1e-6, atol=1e-9)


def test_model_file_round_trip(tmp_path):
    p, fmu, fsd, X, T = small_problem()
    f = tmp_path / "m.bin"
    Mdl.save(str(f), fmu, fsd, p, 0.25)
    d = Mdl.load(str(f))
    assert np.array_equal(d["params"], p) and d["margin"] == pytest.approx(0.25)
    assert os.path.getsize(f) == 8 + 16 + 4 * (2 * R.NFEAT + R.NPARAM + 1)


def test_committed_model_loads():
    if not os.path.exists(Mdl.DEFAULT_PATH):
        pytest.skip("no trained model committed yet")
    d = Mdl.load()
    assert len(d["params"]) == R.NPARAM and np.all(np.isfinite(d["params"]))


@pytest.mark.gpu
def test_gpu_forward_matches_restatement(ctx):
    p, fmu, fsd, X, T = small_problem()
    y = ctx.mlp_forward(p, fmu, fsd, X)
    yr, _ = R.forward(p, fmu, fsd, X)
    np.testing.assert_allclose(y, yr, rtol=1e-4, atol=1e-5)


@pytest.mark.gpu
def test_gpu_gradient_matches_restatement(ctx):
    p, fmu, fsd, X, T = small_problem()
    g, loss = ctx.mlp_gradient(p, fmu, fsd, X, T)
    lr, gr = R.loss_grad(p, fmu, fsd, X, T)
    assert loss == pytest.approx(lr, rel=1e-5)
    err = np.abs(g - gr).max() / np.abs(gr).max()
    print(f"gradient: max |gpu - fp64| / max |fp64| = {err:.2e}")
    assert err < 1e-4


@pytest.mark.gpu
def test_gpu_training_reduces_loss(ctx):
    p, fmu, fsd, X, T = small_problem(B=200)
    p2, losses = ctx.mlp_train(p, fmu, fsd, X, T, 300, lr=3e-3)
    assert losses[-1] < 0.1 * losses[0]
    l2, _ = R.loss_grad(p2, fmu, fsd, X, T)
    assert l2 < 0.1 * losses[0]


@pytest.mark.gpu
def test_gpu_features_match_restatement(ctx):
    xs = [synth.ar2_mix(20000, 5), synth.rw_planted(30000, 6, 1, 64, 16.0)]
    hs = [ctx.upload(x) for x in xs]
    for m in (64, 128):
        F = ctx.features(hs, m, top_k=2)
        for q, x in enumerate(xs):
            f = R.features(x, m, sid=q, k=2)
            assert F[q, 0] == pytest.approx(f[0])
            assert F[q, 1] == pytest.approx(f[1], abs=1e-6)
            assert F[q, 2:34].sum() == pytest.approx(1.0, abs=1e-5)
            # bins agree except for pairs within rounding of a bin edge
            assert np.abs(F[q, 2:34] - f[2:34]).sum() <= 4.0 / R.NPAIR
            assert F[q, 34] == pytest.approx(f[34], abs=1e-6)  # log10(n) / 6
            assert F[q, 35] == pytest.approx(f[35])             # k / 4


def test_pair_generator_is_splitmix64():
    """The feature pairs come from splitmix64 (Steele, Lea, Flood 2014): the restatement's
    mixer reproduces the published output stream for state 0."""
    gamma = 0x9E3779B97F4A7C15
    outs = [R._mix64((q * gamma) & R.M64) for q in range(3)]
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
