"""numpy restatement of the initial-density network (sqminiapp_amd/csrc/sq_mlp.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
kernels.  The reference has no counterpart (its SCF starts from the
superposition of atomic densities, lsdft_readfiles.cc:1026-1040, when it
starts at all); this restates the network the product defines.

features  at a cell point x = h i - R, over the atom images J with
          r_J = |X_J - x| < rf (rf = 8 bohr), fc(r) = (cos(pi r/rf) + 1)/2:
          radial   f_k = sum_J exp(-((r_J - mu_k)/s)^2) fc(r_J),
                   mu_k = k rf/15, s = rf/15, k = 0..15;
          angular  g_a = 2^(1-zeta) sum_{J<L} (1 + lam cos t_JL)^zeta
                         exp(-eta (r_J^2 + r_L^2)) fc(r_J) fc(r_L),
                   t_JL the angle at x between X_J and X_L, (eta, zeta, lam)
                   over eta in ANG_ETA, zeta in ANG_ZETA, lam in (+1, -1),
                   a = (ie * 4 + iz) * 2 + il -- Behler's angular symmetry
                   functions with the grid point as the centre;
network   y = w3 . silu(W2 silu(W1 xhat + b1) + b2) + b3, xhat = (f - mean)/std,
          silu(z) = z / (1 + exp(-z));
loss      L = (1/2N) sum (y - t)^2;
parameters packed as W1[H][K], b1[H], W2[H][H], b2[H], w3[H], b3 (K = 32, H = 64).
"""
from __future__ import annotations

import numpy as np

K_RAD, K_ANG = 16, 16
K = K_RAD + K_ANG
RF = 8.0
ANG_ETA = (0.005, 0.02)
ANG_ZETA = (1.0, 2.0, 4.0, 16.0)


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


def atom_images(pos, R, rf=RF):
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


def point_features(x, imgs, rf=RF):
    """The K descriptors at one point x from the image positions imgs."""
    d = imgs - x
    r = np.sqrt(np.sum(d * d, 1))
    m = r < rf
    d, r = d[m], r[m]
    fc = 0.5 * (np.cos(np.pi * r / rf) + 1.0)
    mu = np.arange(K_RAD) * rf / (K_RAD - 1)
    s = rf / (K_RAD - 1)
    f = np.zeros(K)
    f[:K_RAD] = np.sum(np.exp(-(((r[:, None] - mu[None, :]) / s) ** 2)) * fc[:, None], 0)
    iu, ju = np.triu_indices(len(r), 1)
    if len(iu):
        c = np.sum(d[iu] * d[ju], 1) / (r[iu] * r[ju])
        rr = r[iu] * r[iu] + r[ju] * r[ju]
        w = fc[iu] * fc[ju]
        q = K_RAD
        for eta in ANG_ETA:
            e = np.exp(-eta * rr) * w
            for zeta in ANG_ZETA:
                for lam in (1.0, -1.0):
                    f[q] = np.sum(2.0 ** (1.0 - zeta) * (1.0 + lam * c) ** zeta * e)
                    q += 1
    return f


def features(pos, h, R, G, rf=RF):
    """[G2*G1*G0, K] descriptors at the cell points (i fastest)."""
    imgs = atom_images(pos, R, rf)
    h = np.asarray(h, np.float64)
    R = np.asarray(R, np.float64)
    out = np.zeros((G[0] * G[1] * G[2], K))
    q = 0
    for k in range(G[2]):
        for j in range(G[1]):
            for i in range(G[0]):
                out[q] = point_features(np.array([h[0] * i - R[0], h[1] * j - R[1], h[2] * k - R[2]]), imgs, rf)
                q += 1
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, xhat, K, H):
    W1, b1, W2, b2, w3, b3 = unpack(p, K, H)
    z1 = xhat @ W1.T + b1
    h1 = silu(z1)
    z2 = h1 @ W2.T + b2
    h2 = silu(z2)
    return h2 @ w3 + b3, h1, h2, z1, z2


def loss_grad(p, xhat, t, K, H):
    """(L, dL/dp) for L = (1/2N) sum (y - t)^2."""
    W1, b1, W2, b2, w3, b3 = unpack(p, K, H)
    y, h1, h2, z1, z2 = forward(p, xhat, K, H)
    N = len(t)
    e = y - t
    L = 0.5 * np.sum(e * e) / N
    d3 = e / N
    d2 = (d3[:, None] * w3[None, :]) * dsilu(z2)
    d1 = (d2 @ W2) * dsilu(z1)
    g = pack(d1.T @ xhat, d1.sum(0), d2.T @ h1, d2.sum(0), h2.T @ d3, d3.sum())
    return L, g


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)
