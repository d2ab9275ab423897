"""Double-precision restatement of the range-proposal MLP -- TEST INFRASTRUCTURE ONLY.

Checks the HIP kernels of paddi_amd/csrc/mlp.hip (forward, backward) and the feature
histogram.  No reference counterpart: the reference's range schedule is the blind r-descent
of src/main.cu:136-233.  Pinned by central finite differences (tests/test_mlp.py).
"""
from __future__ import annotations

import numpy as np

NFEAT, HID, NPAIR = 36, 64, 1024
NPARAM = NFEAT * HID + HID + HID * HID + HID + HID + 1
FEAT_SEED = 0x5eed5eed1234


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + NFEAT * HID].reshape(NFEAT, HID); o += NFEAT * HID
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    W3 = p[o:o + HID]; o += HID
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, fmu, fsd, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Z = (np.asarray(X, np.float64) - fmu) / fsd
    A1 = Z @ W1 + b1
    H1 = np.maximum(A1, 0.0)
    A2 = H1 @ W2 + b2
    H2 = np.maximum(A2, 0.0)
    Y = H2 @ W3 + b3
    return Y, (Z, A1, H1, A2, H2)


def loss_grad(p, fmu, fsd, X, T):
    """Mean squared error and its gradient with respect to the flat parameter vector."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Y, (Z, A1, H1, A2, H2) = forward(p, fmu, fsd, X)
    B = len(Y)
    E = Y - np.asarray(T, np.float64)
    loss = float(np.mean(E * E))
    dY = 2.0 * E / B
    gW3 = H2.T @ dY
    gb3 = dY.sum()
    dA2 = np.outer(dY, W3) * (A2 > 0)
    gW2 = H1.T @ dA2
    gb2 = dA2.sum(0)
    dA1 = (dA2 @ W2.T) * (A1 > 0)
    gW1 = Z.T @ dA1
    gb1 = dA1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g


# --- features (restatement of k_feat_pairs / k_feat_std) -------------------------------

M64 = (1 << 64) - 1


def _mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def pair_indices(sid, N, m, seed=FEAT_SEED):
    out = []
    for p in range(NPAIR):
        h0 = _mix64(seed ^ (sid << 32) ^ p)
        h1 = _mix64(h0)
        i = h0 % N
        d = m + (h1 % (N - 2 * m))
        out.append((i, (i + d) % N))
    return out


def features(x, m, sid=0, k=1):
    """Features of one series at length m in double precision (bins may differ from the
    single-precision kernel only for pair distances within rounding of a bin edge)."""
    x64 = np.asarray(x, np.float64)
    N = len(x) - m + 1
    W = np.lib.stride_tricks.sliding_window_view(x64, m)
    hist = np.zeros(32)
    for i, j in pair_indices(sid, N, m):
        a = W[i] - W[i].mean()
        b = W[j] - W[j].mean()
        na, nb = np.sqrt((a * a).sum()), np.sqrt((b * b).sum())
        q = 1.0 if (na == 0 or nb == 0) else 1.0 - (a @ b) / (na * nb)
        hist[min(max(int(q * 16.0), 0), 31)] += 1.0 / NPAIR
    f = np.empty(NFEAT)
    f[0] = m / 512.0
    f[1] = np.log10(max(x64.std(), 1e-30)) / 4.0
    f[2:34] = hist
    f[34] = np.log10(len(x)) / 6.0
    f[35] = k / 4.0
    return f
