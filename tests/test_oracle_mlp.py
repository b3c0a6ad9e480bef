"""Pins the numpy restatement of the initial-density network (oracle/mlp.py),
which the device kernels are checked against:

- the gradient of the loss by central differences in every parameter block;
- bfloat16 rounding against hand-worked bit patterns (ties to even);
- descriptors: invariant under a shift of every atom by a cell vector 2R and
  under a permutation of the atoms, zero far from every atom, and equal to the
  closed form for one atom.
"""
import numpy as np

from oracle import mlp as M

K, H = 16, 64


def test_gradient_against_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, M.n_params(K, H))
    x = rng.normal(0, 1, (40, K))
    t = rng.normal(0, 1, 40)
    L, g = M.loss_grad(p, x, t, K, H)
    offs = [0, H * K, H * K + H, H * K + H + H * H, H * K + 2 * H + H * H, H * K + 3 * H + H * H]
    idx = []
    for o in offs:
        idx += [o, o + 1, o + 7] if o < M.n_params(K, H) - 1 else [o]
    h = 1e-6
    for q in idx:
        e = np.zeros_like(p)
        e[q] = h
        fd = (M.loss_grad(p + e, x, t, K, H)[0] - M.loss_grad(p - e, x, t, K, H)[0]) / (2 * h)
        assert abs(fd - g[q]) < 1e-7 * max(1.0, abs(g[q])), (q, fd, g[q])
    assert L > 0


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -3.140625, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -3.140625, 0.0], np.float32)
    assert np.array_equal(M.bf16(vals), want)


def test_descriptor_invariances():
    rng = np.random.default_rng(3)
    R = np.array([3.0, 5.2, 4.9])
    h = 2 * R / np.array([10, 17, 16])
    G = (10, 17, 16)
    pos = rng.uniform(-R, R, (4, 3))
    f = M.features(pos, h, R, G)
    shifted = pos + 2 * R * np.array([[1, 0, 0], [0, -1, 0], [0, 0, 2], [1, 1, 1]])
    assert np.allclose(M.features(shifted, h, R, G), f, atol=1e-12)
    assert np.allclose(M.features(pos[::-1], h, R, G), f, atol=1e-12)
    # one atom in a large cell: the closed form, and zero past the cutoff
    Rb = np.array([10.0, 10.0, 10.0])
    hb = 2 * Rb / 20
    one = M.features(np.zeros((1, 3)), hb, Rb, (20, 20, 20))
    pts = np.stack(np.meshgrid(*[hb[d] * np.arange(20) - Rb[d] for d in (2, 1, 0)], indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    r = np.linalg.norm(pts, axis=1)
    mu = np.arange(K) * 6.0 / (K - 1)
    s = 6.0 / (K - 1)
    ref = np.where(r[:, None] < 6.0, np.exp(-((r[:, None] - mu) / s) ** 2) * 0.5 * (np.cos(np.pi * r[:, None] / 6.0) + 1), 0)
    assert np.allclose(one, ref, atol=1e-12)
    assert np.all(one[r >= 6.0] == 0)
