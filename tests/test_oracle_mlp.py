"""Pins the numpy restatement of the initial-density network (oracle/mlp.py),
which the device kernels are checked against:

- the gradient of the loss by central differences in every parameter block;
- bfloat16 rounding against hand-worked bit patterns (ties to even);
- descriptors: invariant under a shift of every atom by a cell vector 2R and
  under a permutation of the atoms, zero far from every atom; for one atom the
  radial closed form and no angular part; for two atoms the angular closed
  form (the angle at the point, the products of cutoffs and Gaussians);
- SiLU's derivative against differences.
"""
import numpy as np

from oracle import mlp as M

K, H = M.K, 64


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
    # one atom in a large cell: the radial closed form, no angular part, zero past the cutoff
    Rb = np.array([12.0, 12.0, 12.0])
    hb = 2 * Rb / 24
    one = M.features(np.zeros((1, 3)), hb, Rb, (24, 24, 24))
    pts = np.stack(np.meshgrid(*[hb[d] * np.arange(24) - Rb[d] for d in (2, 1, 0)], indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    r = np.linalg.norm(pts, axis=1)
    rf = M.RF
    mu = np.arange(M.K_RAD) * rf / (M.K_RAD - 1)
    s = rf / (M.K_RAD - 1)
    ref = np.where(r[:, None] < rf, np.exp(-((r[:, None] - mu) / s) ** 2) * 0.5 * (np.cos(np.pi * r[:, None] / rf) + 1), 0)
    assert np.allclose(one[:, :M.K_RAD], ref, atol=1e-12)
    assert np.all(one[:, M.K_RAD:] == 0) and np.all(one[r >= rf] == 0)


def test_angular_features_closed_form():
    """Two atoms seen from a point: g = 2^(1-zeta) (1 + lam cos t)^zeta
    exp(-eta (r1^2 + r2^2)) fc(r1) fc(r2) with t the angle at the point."""
    x = np.array([0.3, -0.2, 0.1])
    a, b = np.array([2.0, 0.5, -1.0]), np.array([-1.5, 2.5, 0.7])
    f = M.point_features(x, np.stack([a, b]))
    u, v = a - x, b - x
    r1, r2 = np.linalg.norm(u), np.linalg.norm(v)
    ct = np.dot(u, v) / (r1 * r2)
    fc = lambda r: 0.5 * (np.cos(np.pi * r / M.RF) + 1)
    q = M.K_RAD
    for eta in M.ANG_ETA:
        for zeta in M.ANG_ZETA:
            for lam in (1.0, -1.0):
                want = 2 ** (1 - zeta) * (1 + lam * ct) ** zeta * np.exp(-eta * (r1 * r1 + r2 * r2)) * fc(r1) * fc(r2)
                assert f[q] == np.float64(want) or abs(f[q] - want) < 1e-15 * max(1.0, abs(want)), (q, f[q], want)
                q += 1
    # a third atom along the first one's direction doubles nothing: the pair sums add up
    c = x + 0.5 * u
    f3 = M.point_features(x, np.stack([a, b, c]))
    assert np.all(f3[M.K_RAD:] > f[M.K_RAD:])


def test_silu_derivative():
    z = np.linspace(-6, 6, 41)
    h = 1e-6
    assert np.allclose(M.dsilu(z), (M.silu(z + h) - M.silu(z - h)) / (2 * h), atol=1e-8)
