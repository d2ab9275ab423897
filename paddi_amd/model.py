"""File format of the learned range-proposal model (read by the HIP library at
paddi_amd/model/r_mlp.bin unless paddi_opts.model_path names another file).

Layout, little endian: magic b"PDRMLP02"; int32 nfeat, hidden, nparam, 0; float32
fmu[nfeat], fsd[nfeat], params[nparam], margin.  Features: m/512, log10(sigma)/4, the
32-bin pair-distance histogram, log10(n)/6, k/4.  The network predicts t = log(d_k / 2m) of
the k-th discord; the library proposes r2 = 2m * exp(t - margin).
"""
from __future__ import annotations

import os

import numpy as np

MAGIC = b"PDRMLP02"
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "model", "r_mlp.bin")


def init_params(nfeat=36, hidden=64, seed=0):
    """He-initialised flat parameter vector W1 b1 W2 b2 W3 b3 (biases zero)."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((nfeat, hidden)) * np.sqrt(2.0 / nfeat)
    W2 = rng.standard_normal((hidden, hidden)) * np.sqrt(2.0 / hidden)
    W3 = rng.standard_normal(hidden) * np.sqrt(1.0 / hidden)
    return np.concatenate([W1.ravel(), np.zeros(hidden), W2.ravel(), np.zeros(hidden), W3,
                           [0.0]]).astype(np.float32)


def save(path, fmu, fsd, params, margin, nfeat=36, hidden=64):
    params = np.asarray(params, np.float32)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "wb") as f:
        f.write(MAGIC)
        np.array([nfeat, hidden, len(params), 0], "<i4").tofile(f)
        np.asarray(fmu, "<f4").tofile(f)
        np.asarray(fsd, "<f4").tofile(f)
        params.astype("<f4").tofile(f)
        np.array([margin], "<f4").tofile(f)


def load(path=DEFAULT_PATH):
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError("not a range-proposal model file")
        nfeat, hidden, nparam, _ = np.fromfile(f, "<i4", 4)
        fmu = np.fromfile(f, "<f4", nfeat)
        fsd = np.fromfile(f, "<f4", nfeat)
        params = np.fromfile(f, "<f4", nparam)
        margin = float(np.fromfile(f, "<f4", 1)[0])
    return dict(fmu=fmu, fsd=fsd, params=params, margin=margin, nfeat=int(nfeat),
                hidden=int(hidden))
