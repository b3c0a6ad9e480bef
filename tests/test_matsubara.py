"""The stopping test of the per-node early exit, restated on numpy and pinned
against closed forms: the continued fraction the kernel evaluates is the
resolvent e1^T (z - T)^{-1} e1 = sum_i w_i / (z - lambda_i) of the Gauss rule
(eigh_tridiagonal), and the Matsubara series 1/2 + (2/beta) sum_m Re g(mu +
i w_m) of that rule tends to its Fermi quadrature sum_i w_i f(lambda_i)."""
import numpy as np
from scipy.linalg import eigh_tridiagonal
from scipy.special import polygamma


def resolvent_cf(a, b, z):
    """1/(z - a0 - b0^2/(z - a1 - ...)) evaluated from the bottom, as the kernel does."""
    t = z - a[-1]
    for j in range(len(a) - 2, -1, -1):
        t = z - a[j] - b[j] ** 2 / t
    return 1.0 / t


def test_continued_fraction_is_the_gauss_rule_resolvent():
    rng = np.random.default_rng(3)
    a = rng.uniform(-2, 20, 60)
    b = rng.uniform(0.5, 6, 59)
    lam, V = eigh_tridiagonal(a, b)
    w = V[0] ** 2
    for z in (0.3 + 0.0314j, 5.0 + 1.0j, -3.0 + 0.2j):
        ref = np.sum(w / (z - lam))
        assert abs(resolvent_cf(a, b, z) - ref) < 1e-11 * abs(ref)


def test_matsubara_series_tends_to_the_fermi_quadrature():
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 4, 30)
    b = rng.uniform(0.2, 2, 29)
    lam, V = eigh_tridiagonal(a, b)
    w = V[0] ** 2
    beta, mu = 20.0, 1.3
    exact = np.sum(w / (1.0 + np.exp(beta * (lam - mu))))
    K = 2000
    wm = (2 * np.arange(K) + 1) * np.pi / beta
    g = (w[None, :] / (mu + 1j * wm[:, None] - lam[None, :])).sum(1)
    # the terms past K from the asymptote Re g ~ -sum w (lam - mu) / w_m^2, summed in
    # closed form: sum_{m>=K} 1/(2m+1)^2 = psi'(K + 1/2)/4
    mom = np.sum(w * (lam - mu))
    tail = -mom * (beta / np.pi) ** 2 * polygamma(1, K + 0.5) / 4.0
    series = 0.5 + (2.0 / beta) * (np.sum(g.real) + tail)
    assert abs(series - exact) < 1e-8
    # truncated after 8 terms, as the kernel's test, the series is far from converged:
    # the test watches the change of these slowest terms, not their sum
    assert abs(0.5 + (2.0 / beta) * np.sum(g.real[:8]) - exact) > 1e-3
    for x in wm[:8]:
        assert abs(resolvent_cf(a, b, mu + 1j * x) - np.sum(w / (mu + 1j * x - lam))) < 1e-12
