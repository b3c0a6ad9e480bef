"""Pins the oracle's density / Fermi level / band and entropy energy.

The reference only declares these (lsdft.h:98-102,209); they are written from
SQ theory: rho_p = (2/dV) sum_i w_pi f(l_pi), 2 sum w f = N_e,
E_band = 2 sum w l f, -TS = (2/beta) sum w [f ln f + (1-f) ln(1-f)].
"""
import numpy as np
import pytest
from scipy.optimize import brentq


def _naive(lam, w, beta, lf):
    f = 1.0 / (1.0 + np.exp(beta * (lam - lf)))
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where((f > 0) & (f < 1), f * np.log(f) + (1 - f) * np.log1p(-f), 0.0)
    return f, s


@pytest.mark.parametrize("beta", [30.0, 1000.0])
def test_fermi_level_density_energy(oracle_mod, beta):
    rng = np.random.default_rng(int(beta))
    n, q = 200, 10
    lam = np.sort(rng.uniform(-0.5, 3.0, (n, q)), axis=1)
    w = rng.uniform(0.1, 1.0, (n, q))
    w /= w.sum(1, keepdims=True)
    nelec = 0.37 * 2 * n
    dV = 0.25
    lf, rho, eb, mts = oracle_mod.density_fermi_energy(lam, w, beta, nelec, dV)
    # independent root of the electron-count constraint
    g = lambda x: 2 * np.sum(w / (1 + np.exp(np.clip(beta * (lam - x), -700, 700)))) - nelec
    ref = brentq(g, lam.min() - 1, lam.max() + 1, xtol=1e-15, rtol=1e-15)
    assert lf == pytest.approx(ref, abs=1e-10)
    f, s = _naive(lam, w, beta, lf)
    assert np.allclose(rho, 2 * np.sum(w * f, 1) / dV, rtol=1e-12)
    assert rho.sum() * dV == pytest.approx(nelec, rel=1e-11)
    assert eb == pytest.approx(2 * np.sum(w * lam * f), rel=1e-11)
    assert mts == pytest.approx(2 / beta * np.sum(w * s), rel=1e-9, abs=1e-14)
    assert mts <= 0.0


def test_electron_count_monotone_and_limits(oracle_mod):
    lam = np.array([[0.0, 1.0], [0.5, 2.0]])
    w = np.array([[0.5, 0.5], [0.25, 0.75]])
    assert oracle_mod.electron_count(lam, w, 50.0, -10.0) == pytest.approx(0.0, abs=1e-12)
    assert oracle_mod.electron_count(lam, w, 50.0, 10.0) == pytest.approx(4.0, abs=1e-12)
    vals = [oracle_mod.electron_count(lam, w, 50.0, x) for x in np.linspace(-1, 3, 30)]
    assert np.all(np.diff(vals) >= 0)


def test_unbracketed_fermi_level_is_an_error(oracle_mod):
    lam = np.zeros((2, 2))
    w = np.full((2, 2), 0.5)
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.density_fermi_energy(lam, w, 30.0, 100.0, 1.0)
