"""numpy restatement of the initial-density network (sqminiapp_amd/csrc/sq_mlp.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
kernels.  The reference has no counterpart (its SCF starts from the
superposition of atomic densities, lsdft_readfiles.cc:1026-1040, when it
starts at all); this restates the network the product defines:

features  f_k(x) = sum over atom images J with r = |x - X_J| < rf of
                   exp(-((r - mu_k)/s)^2) * (cos(pi r / rf) + 1) / 2,
                   mu_k = k rf / (K - 1), s = rf / (K - 1), at the cell points
                   x = h i - R (the fields' grid);
network   y = w3 . tanh(W2 tanh(W1 xhat + b1) + b2) + b3, xhat = (f - mean)/std;
loss      L = (1/2N) sum (y - t)^2;
parameters packed as W1[H][K], b1[H], W2[H][H], b2[H], w3[H], b3.
"""
from __future__ import annotations

import numpy as np


def n_params(K, H):
    return H * K + H + H * H + H + H + 1


def unpack(p, K, H):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * K].reshape(H, K); o += H * K
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [b3]])


def atom_images(pos, R, rf):
    """Every periodic image (shifts 2R) of the atoms that can lie within rf of
    the cell, the positions first wrapped into [-R, R)."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    R = np.asarray(R, np.float64)
    pos = pos - 2 * R * np.floor((pos + R) / (2 * R))
    n = np.ceil(rf / (2 * R)).astype(int) + 1
    out = []
    for a in pos:
        for i in range(-n[0], n[0] + 1):
            for j in range(-n[1], n[1] + 1):
                for k in range(-n[2], n[2] + 1):
                    out.append(a + 2 * R * np.array([i, j, k]))
    return np.array(out)


def features(pos, h, R, G, K=16, rf=6.0):
    """[G2*G1*G0, K] descriptors at the cell points (i fastest)."""
    imgs = atom_images(pos, R, rf)
    h = np.asarray(h, np.float64)
    R = np.asarray(R, np.float64)
    zz, yy, xx = np.meshgrid(*[h[d] * np.arange(G[d]) - R[d] for d in (2, 1, 0)], indexing="ij")
    pts = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1)
    mu = np.arange(K) * rf / (K - 1)
    s = rf / (K - 1)
    f = np.zeros((pts.shape[0], K))
    for X in imgs:
        r = np.sqrt(np.sum((pts - X) ** 2, 1))
        m = r < rf
        if not np.any(m):
            continue
        rc = r[m]
        fc = 0.5 * (np.cos(np.pi * rc / rf) + 1.0)
        f[m] += np.exp(-(((rc[:, None] - mu[None, :]) / s) ** 2)) * fc[:, None]
    return f


def forward(p, xhat, K, H):
    W1, b1, W2, b2, w3, b3 = unpack(p, K, H)
    h1 = np.tanh(xhat @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ w3 + b3, h1, h2


def loss_grad(p, xhat, t, K, H):
    """(L, dL/dp) for L = (1/2N) sum (y - t)^2."""
    W1, b1, W2, b2, w3, b3 = unpack(p, K, H)
    y, h1, h2 = forward(p, xhat, K, H)
    N = len(t)
    e = y - t
    L = 0.5 * np.sum(e * e) / N
    d3 = e / N
    d2 = (d3[:, None] * w3[None, :]) * (1.0 - h2 * h2)
    d1 = (d2 @ W2) * (1.0 - h1 * h1)
    g = pack(d1.T @ xhat, d1.sum(0), d2.T @ h1, d2.sum(0), h2.T @ d3, d3.sum())
    return L, g


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)
