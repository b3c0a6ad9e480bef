"""Pins the oracle's numerical building blocks against closed forms and libraries.

FD weights (lsdft_readfiles.cc:453-476), fract (lsdft_tools.cc:519-531), the
not-a-knot spline (lsdft_tools.cc:15-151), real Y_lm (lsdft_tools.cc:156-367)
and the QL nodes/weights (lsdft_tools.cc:912-1026).
"""
import math
from fractions import Fraction

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.linalg import eigh_tridiagonal
from scipy.special import sph_harm_y

from conftest import MG1, rel_err


def test_fract_closed_form(oracle_mod):
    for n in range(1, 8):
        for k in range(0, n + 1):
            exact = Fraction(math.factorial(n) ** 2, math.factorial(n - k) * math.factorial(n + k))
            assert oracle_mod.fract(n, k) == pytest.approx(float(exact), rel=1e-15)


def test_fd_weights_order12_known_answers(oracle_mod):
    S = oracle_mod.System(MG1, "Mg1")
    I = S.info()
    # -(1/2) Laplacian, 12th order: c0 = sum 1/a^2 = 5369/3600, c_a = -(-1)^(a+1) fract(6,a)/a^2
    known = [5369 / 3600, -6 / 7, 15 / 112, -10 / 378, 1 / 224, -1 / 1925, 1 / 33264]
    for d in range(3):
        h2 = I.delta[d] ** 2
        for p in range(7):
            assert I.coeffs[d][p] * h2 == pytest.approx(known[p], rel=1e-14)
    # the standard central weights of d^2/dx^2 are -2 * these, summing to zero with the 2*c_a pairs
    s = known[0] + 2 * sum(known[1:])
    assert abs(s) < 1e-14


def test_stencil_exact_on_quadratics(oracle_mod, tiny_deck):
    """-(1/2) Lap_h of a quadratic equals the exact value at interior points."""
    S = oracle_mod.System(*tiny_deck)
    I = S.info()
    rp, col, val = S.cube_csr()
    nx, ny, nz = I.nloc
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    h = I.delta
    x, y, z = i * h[0], j * h[1], k * h[2]
    f = (0.3 * x * x - 1.1 * y * y + 0.7 * z * z + x * y - 2 * z).ravel()
    g = np.zeros_like(f)
    for r in range(len(f)):
        g[r] = np.dot(val[rp[r]:rp[r + 1]], f[col[rp[r]:rp[r + 1]]])
    exact = -0.5 * 2 * (0.3 - 1.1 + 0.7)
    o = I.order
    inner = ((i >= o) & (i < nx - o) & (j >= o) & (j < ny - o) & (k >= o) & (k < nz - o)).ravel()
    if inner.any():
        assert np.allclose(g[inner], exact, rtol=1e-10, atol=1e-10)


def test_cube_laplacian_is_kronecker_sum(oracle_mod, tiny_deck):
    """The AIJ cube operator equals the Kronecker sum of 1-D Toeplitz stencils
    truncated at the cube faces (zero Dirichlet, lsdft_initObjs.cc:393-475)."""
    S = oracle_mod.System(*tiny_deck)
    I = S.info()
    rp, col, val = S.cube_csr()
    n = I.nc
    Acsr = np.zeros((n, n))
    for r in range(n):
        Acsr[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    nx, ny, nz = I.nloc

    def T(m, d):
        t = np.zeros((m, m))
        for a in range(m):
            t[a, a] = I.coeffs[d][0]
            for p in range(1, I.order + 1):
                if a + p < m:
                    t[a, a + p] = t[a + p, a] = I.coeffs[d][p]
        return t
    K = (np.kron(np.eye(nz), np.kron(np.eye(ny), T(nx, 0))) + np.kron(np.eye(nz), np.kron(T(ny, 1), np.eye(nx)))
         + np.kron(T(nz, 2), np.eye(ny * nx)))
    assert np.allclose(Acsr, K, rtol=1e-13, atol=1e-12)
    assert np.array_equal(Acsr, Acsr.T)
    # columns ascending in every row (the AIJ row order the SpMV sum follows)
    for r in range(n):
        assert np.all(np.diff(col[rp[r]:rp[r + 1]]) > 0)


def test_spline_matches_scipy_notaknot(oracle_mod, mg1_oracle):
    R = mg1_oracle.radial(0, 0)
    r = R["r"]
    for key in ("u", "udv", "vloc"):
        y = R[key]
        yd = oracle_mod.spline_yd(r, y)
        cs = CubicSpline(r, y, bc_type="not-a-knot")
        assert rel_err(yd, cs(r, 1)) < 1e-8
        pts = np.linspace(0.0, 5.0, 97)
        mine = np.array([oracle_mod.spline_eval(r, y, yd, p) for p in pts])
        assert rel_err(mine, cs(pts)) < 1e-9


def test_spline_exact_on_cubics(oracle_mod):
    x = np.cumsum(np.linspace(0.05, 0.3, 40))
    y = 0.5 * x ** 3 - 2 * x ** 2 + x - 3
    yd = oracle_mod.spline_yd(x, y)
    assert np.allclose(yd, 1.5 * x ** 2 - 4 * x + 1, rtol=1e-10, atol=1e-10)
    for p in np.linspace(x[0], x[-1], 53):
        assert oracle_mod.spline_eval(x, y, yd, p) == pytest.approx(0.5 * p ** 3 - 2 * p ** 2 + p - 3, rel=1e-11,
                                                                    abs=1e-11)


def test_spline_out_of_range_is_an_error(oracle_mod):
    x = np.linspace(0, 1, 10)
    y = x ** 2
    yd = oracle_mod.spline_yd(x, y)
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.spline_eval(x, y, yd, 1.5)


def _real_ylm_scipy(l, m, x, y, z):
    r = np.sqrt(x * x + y * y + z * z)
    theta = np.arccos(z / r)
    phi = np.arctan2(y, x)
    Y = sph_harm_y(l, abs(m), theta, phi)
    if m > 0:
        return math.sqrt(2) * (-1) ** m * Y.real
    if m < 0:
        return math.sqrt(2) * (-1) ** m * Y.imag
    return Y.real


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_ylm_against_scipy_and_orthonormal(oracle_mod, l):
    rng = np.random.default_rng(l)
    pts = rng.standard_normal((40, 3))
    for m in range(-l, l + 1):
        mine = np.array([oracle_mod.ylm(*p, l, m) for p in pts])
        ref = np.array([_real_ylm_scipy(l, m, *p) for p in pts])
        # the reference's 15-digit truncated constants bound the agreement
        assert np.allclose(np.abs(mine), np.abs(ref), rtol=1e-13, atol=1e-14), (l, m)
    # orthonormality by Gauss-Legendre x trapezoid quadrature on the sphere
    nt, nphi = 24, 48
    ct, wt = np.polynomial.legendre.leggauss(nt)
    ph = np.arange(nphi) * 2 * np.pi / nphi
    rows = []
    for m in range(-l, l + 1):
        vals = np.array([[oracle_mod.ylm(math.sqrt(1 - c * c) * math.cos(p), math.sqrt(1 - c * c) * math.sin(p), c, l, m)
                          for p in ph] for c in ct])
        rows.append(vals)
    for a in range(len(rows)):
        for b in range(len(rows)):
            integ = np.sum(wt[:, None] * rows[a] * rows[b]) * 2 * np.pi / nphi
            assert integ == pytest.approx(1.0 if a == b else 0.0, abs=1e-12)


def test_ylm_origin_and_bad_l(oracle_mod):
    assert oracle_mod.ylm(0, 0, 0, 0, 0) == 0.282094791773878
    assert oracle_mod.ylm(0, 0, 0, 2, 1) == 0.0
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.ylm(1, 0, 0, 2, 3)


def test_ql_matches_eigh_tridiagonal(oracle_mod):
    rng = np.random.default_rng(7)
    for n in (1, 2, 5, 10, 40, 80):
        a = rng.uniform(-3, 10, n)
        b = rng.uniform(0.5, 4, n)
        lam, w, lmin, lmax = oracle_mod.tridiag_nodes_weights(a, b, n)
        ev, V = eigh_tridiagonal(a, b[: n - 1])
        order = np.argsort(lam)
        assert np.allclose(lam[order], ev, rtol=1e-12, atol=1e-12)
        assert np.allclose(w[order], V[0] ** 2, rtol=1e-9, atol=1e-13)
        assert lmin == pytest.approx(ev[0], abs=1e-12) and lmax == pytest.approx(ev[-1], abs=1e-12)
        assert w.sum() == pytest.approx(1.0, abs=1e-13)


def test_ql_too_many_sweeps_is_an_error(oracle_mod):
    # a NaN never deflates: the 50-sweep limit (lsdft_tools.cc:950-953) trips
    a = np.array([1.0, np.nan, 2.0])
    b = np.array([1.0, 1.0, 0.0])
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.tridiag_nodes_weights(a, b, 3)
