"""Pins the oracle's force restatement (oracle/sq_oracle.c, forces section).
The reference declares its force routines only (src/lsdft.h:188-191,223-227),
so each piece is pinned by mathematics:

- grad Y_lm (l <= 3) and grad chi_lm = grad(U_l dV_l(r) Y_lm) against central
  differences of the oracle's own Y_lm and chi;
- the Fermi column f(T) e_1 of the full-eigenvector QL against
  scipy.linalg.eigh_tridiagonal;
- the local (pseudocharge) force against the central difference of the
  electrostatic energy (1/2) int (rho + b) phi - E_self at fixed rho, with phi
  from the FFT Poisson solve and b, E_self rebuilt at displaced atoms;
- the nonlocal force of single nodes against the same contraction with the
  density-matrix column f(H_c) e_c from a dense eigendecomposition of the
  node's cube Hamiltonian (small cube, hot smearing, enough Lanczos steps for
  the Gauss rule to resolve f).
"""
import os

import numpy as np
import pytest
from scipy.linalg import eigh_tridiagonal

from conftest import MG1, write_variant
from oracle import oracle as O
from oracle import scf as OS


def test_ylm_gradient_against_differences():
    rng = np.random.default_rng(11)
    h = 1e-6
    for l in range(4):
        for m in range(-l, l + 1):
            for _ in range(8):
                p = rng.uniform(-2.0, 2.0, 3)
                g = O.ylm_grad(*p, l, m)
                fd = np.array([(O.ylm(*(p + h * e), l, m) - O.ylm(*(p - h * e), l, m)) / (2 * h) for e in np.eye(3)])
                assert np.max(np.abs(g - fd)) < 1e-7 * max(1.0, np.max(np.abs(g))), (l, m, p)


def test_projector_gradient_against_differences(mg1_oracle):
    rng = np.random.default_rng(5)
    h = 1e-5
    worst = 0.0
    for l, m in [(0, 0), (2, -2), (2, -1), (2, 0), (2, 1), (2, 2)]:
        for _ in range(10):
            p = rng.uniform(-2.5, 2.5, 3)
            _, g = mg1_oracle.chi_grad(0, l, m, *p)
            fd = np.array([(mg1_oracle.chi_grad(0, l, m, *(p + h * e))[0] - mg1_oracle.chi_grad(0, l, m, *(p - h * e))[0])
                           / (2 * h) for e in np.eye(3)])
            worst = max(worst, np.max(np.abs(g - fd)) / (np.max(np.abs(g)) + 1e-3))
    assert worst < 1e-7


def test_fermi_column_against_scipy():
    rng = np.random.default_rng(2)
    for n in (1, 7, 40):
        a = rng.standard_normal(n)
        b = np.abs(rng.standard_normal(n)) + 0.05
        for beta, lf in ((5.0, 0.2), (100.0, -0.3)):
            c = O.fermi_column(a, b, n, beta, lf)
            if n == 1:
                lam, Z = a, np.ones((1, 1))
            else:
                lam, Z = eigh_tridiagonal(a, b[:n - 1])
            f = 1.0 / (1.0 + np.exp(beta * (lam - lf)))
            assert np.max(np.abs(c - Z @ (f * Z[0]))) < 1e-13
            assert c[0] == pytest.approx(np.sum(Z[0] ** 2 * f), abs=1e-14)   # the node's density sum w f


def _e_es(S, rho, coeffs, order, G, dV):
    b, es, sc = S.pseudocharge()
    b = b.reshape(G)
    phi = OS.poisson_fft(coeffs, order, rho + b)
    return 0.5 * np.sum((rho + b) * phi) * dV - es, phi, sc


def test_local_force_is_the_electrostatic_derivative(tmp_path):
    S = O.System(MG1, "Mg1")
    S.set_phantom(False)
    I = S.info()
    G = (I.npts[2] - 1, I.npts[1] - 1, I.npts[0] - 1)
    coeffs = [list(I.coeffs[d]) for d in range(3)]
    rho, _ = S.atomic_density()
    rho = rho.reshape(G)
    zz, yy, xx = np.meshgrid(*[np.arange(n) for n in G], indexing="ij")
    pert = 0.002 * np.sin(2 * np.pi * xx / G[2]) * np.cos(2 * np.pi * (yy / G[1] + zz / G[0]))
    rho = rho + pert - pert.mean()                      # a density the atoms are not centred in
    _, phi, sc = _e_es(S, rho, coeffs, I.order, G, I.dV)
    F = S.local_forces(phi.ravel(), sc)
    pos = S.atom_positions()
    dl = 1e-3
    big = 0.0
    for atom, d in ((0, 0), (3, 0), (3, 1), (1, 2)):
        E = []
        for sgn in (1, -1):
            p2 = pos.copy()
            p2[atom, d] += sgn * dl
            dd = os.path.join(str(tmp_path), f"a{atom}{d}{sgn}")
            write_variant(dd, "V", atoms=p2)
            S2 = O.System(dd, "V")
            S2.set_phantom(False)
            E.append(_e_es(S2, rho, coeffs, I.order, G, I.dV)[0])
        fd = -(E[0] - E[1]) / (2 * dl)
        assert abs(F[atom, d] - fd) < 1e-7, (atom, d, F[atom, d], fd)
        big = max(big, abs(fd))
    assert big > 1e-3                                    # the check is not on vanishing forces


def test_nonlocal_force_of_nodes_against_dense_density_matrix(tmp_path):
    write_variant(str(tmp_path), "T", rcut=1.3, nqp=60, beta=2.0)
    S = O.System(str(tmp_path), "T")
    S.set_phantom(False)
    S.build_nonlocal()
    I = S.info()
    nx, ny, _ = I.nloc
    hw = [(n - 1) // 2 for n in I.nloc]
    lx, ly = S.box[3], S.box[4]
    imgs = S.images(0)
    chis = [S.chi(0, c, im["ndim"]) for c, im in enumerate(imgs)]
    chans = [(l, m) for l in range(I.lmax[0] + 1) if l != I.lloc[0] for m in range(-l, l + 1)]
    lf = 0.1
    ids = []
    for im in imgs[:2]:
        b = im["box"]
        ids.append(((b[2] + b[5]) // 2) * lx * ly + ((b[1] + b[4]) // 2) * lx + (b[0] + b[3]) // 2)
    ids.append(ids[0] + 1)
    _, Fn = S.nonlocal_forces(I.beta, lf, veff=0.333, node_ids=ids, nthreads=4, per_node=True)
    for r, nid in enumerate(ids):
        Ii, Jj, Kk = nid % lx, (nid // lx) % ly, nid // (lx * ly)
        H = np.stack([S.apply_h((Kk, Jj, Ii), e, mode=O.MODE_KB, veff=0.333) for e in np.eye(I.nc)], 1)
        assert np.max(np.abs(H - H.T)) < 1e-12
        lam, V = np.linalg.eigh(H)
        d = V @ (V[I.center] / (1.0 + np.exp(I.beta * (lam - lf))))
        alpha = {}
        for c, im in enumerate(imgs):
            chi, idd = chis[c]
            for t in range(len(chi)):
                k, j, i = idd["k"][t], idd["j"][t], idd["i"][t]
                if abs(k - Kk) <= hw[2] and abs(j - Jj) <= hw[1] and abs(i - Ii) <= hw[0]:
                    q = ((k - Kk + hw[2]) * ny + (j - Jj + hw[1])) * nx + (i - Ii + hw[0])
                    ch = chans.index((idd["l"][t], idd["m"][t]))
                    key = (im["atom"], ch)
                    alpha[key] = alpha.get(key, 0.0) + I.gamma[0][idd["l"][t]] * chi[t] * d[q]
        Fd = np.zeros((I.nAtoms, 3))
        for c, im in enumerate(imgs):
            b = im["box"]
            if not (b[0] <= Ii <= b[3] and b[1] <= Jj <= b[4] and b[2] <= Kk <= b[5]):
                continue
            x = I.delta[0] * Ii - I.range[0] - im["center"][0]
            y = I.delta[1] * Jj - I.range[1] - im["center"][1]
            z = I.delta[2] * Kk - I.range[2] - im["center"][2]
            for ch, (l, m) in enumerate(chans):
                _, g = S.chi_grad(0, l, m, x, y, z)
                Fd[im["atom"]] += 4.0 * g * alpha.get((im["atom"], ch), 0.0)
        assert np.max(np.abs(Fd)) > 1e-5
        assert np.max(np.abs(Fn[r] - Fd)) < 1e-12 * max(1.0, np.max(np.abs(Fd))), (nid, Fn[r], Fd)
