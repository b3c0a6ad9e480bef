"""Pins the oracle's Hamiltonian and Lanczos quadrature by brute force.

- the nonlocal operators of Mult_NonlocalVector (lsdft_matvec.cc:25-149) and
  Mult_HamiltonianVector (155-293) against dense projector matrices assembled
  from the chi tables (images of one atom share the atom's alpha slot);
- Lanczos + QL (lsdft_scf.cc:87-164, lsdft_tools.cc:912-1026) against Gauss
  quadrature exactness, eigh of the dense H and the V_eff shift identity.
"""
import numpy as np
import pytest
from scipy.linalg import eigh_tridiagonal

from conftest import rel_err, write_variant


def _dense_parts(O, S, node):
    """Dense L+V_eff and the projector vectors of every (atom, channel) on the cube of node."""
    I = S.info()
    rp, col, val = S.cube_csr()
    n = I.nc
    A = np.zeros((n, n))
    for r in range(n):
        A[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    K, J, Ii = node
    hw = [(I.nloc[d] - 1) // 2 for d in range(3)]
    lo = [Ii - hw[0], J - hw[1], K - hw[2]]
    nlm = I.lmax[0] ** 2 + 2 * (I.lmax[0] - I.lloc[0])
    natoms = I.nIP
    chans = np.zeros((natoms, nlm, n))
    for c, im in enumerate(S.images(0)):
        chi, ids = S.chi(0, c, im["ndim"])
        inside = ((ids["i"] >= lo[0]) & (ids["i"] < lo[0] + I.nloc[0]) & (ids["j"] >= lo[1]) &
                  (ids["j"] < lo[1] + I.nloc[1]) & (ids["k"] >= lo[2]) & (ids["k"] < lo[2] + I.nloc[2]))
        q = ((ids["k"] - lo[2]) * I.nloc[1] + (ids["j"] - lo[1])) * I.nloc[0] + (ids["i"] - lo[0])
        ch = np.arange(im["ndim"]) % nlm
        np.add.at(chans[im["atom"]], (ch[inside], q[inside]), chi[inside])
    lch = [l for l in range(I.lmax[0] + 1) if l != I.lloc[0] for _ in range(2 * l + 1)]
    gam = np.array([I.gamma[0][l] for l in lch])
    return A, chans, gam


@pytest.mark.parametrize("mode", [0, 1])
def test_apply_h_against_dense_projectors(oracle_mod, tiny_deck, mode):
    S = oracle_mod.System(*tiny_deck)
    S.set_phantom(True)
    S.build_nonlocal()
    node = (8, 9, 5)
    A, chans, gam = _dense_parts(oracle_mod, S, node)
    veff = 0.333
    H = A + veff * np.eye(A.shape[0])
    if mode == 1:   # sum_p sum_c Gamma_c chi_pc chi_pc^T
        H += np.einsum("c,pci,pcj->ij", gam, chans, chans)
    else:           # sum_p u_p w_p^T, u = sum_c chi, w = sum_c Gamma_c chi
        H += np.einsum("pi,pj->ij", chans.sum(1), np.einsum("c,pci->pi", gam, chans))
    rng = np.random.default_rng(3)
    for _ in range(3):
        x = rng.standard_normal(A.shape[0])
        y = S.apply_h(node, x, mode=mode, veff=veff)
        assert rel_err(y, H @ x) < 1e-12
    if mode == 1:
        assert np.allclose(H, H.T, rtol=0, atol=1e-9 * np.abs(H).max())


@pytest.mark.parametrize("mode", [0, 1])
def test_nonlocal_part_accumulates_onto_v2(oracle_mod, tiny_deck, mode):
    """Mult_NonlocalVector / Mult_HamiltonianVector add V_nl x onto what V2 holds."""
    S = oracle_mod.System(*tiny_deck)
    S.set_phantom(True)
    S.build_nonlocal()
    node = (8, 9, 5)
    A, chans, gam = _dense_parts(oracle_mod, S, node)
    if mode == 1:
        V = np.einsum("c,pci,pcj->ij", gam, chans, chans)
    else:
        V = np.einsum("pi,pj->ij", chans.sum(1), np.einsum("c,pci->pi", gam, chans))
    rng = np.random.default_rng(8)
    x, y0 = rng.standard_normal((2, A.shape[0]))
    y = S.apply_nonlocal(node, x, y0, mode=mode)
    assert rel_err(y - y0, V @ x) < 1e-12
    # with V2 = (L + V_eff) x first, the sum is the full H application
    assert rel_err(S.apply_nonlocal(node, x, A @ x + 0.333 * x, mode=mode), S.apply_h(node, x, mode=mode)) < 1e-14


def test_lanczos_full_spectrum_on_tiny_cube(oracle_mod, tmp_path):
    """Against the dense eigendecomposition H = V diag(ev) V^T of a 5^3 cube:
    the k-node Gauss rule integrates ev^m against the weights V[c]^2 exactly
    for m <= 2k-1 (KB mode, symmetric H)."""
    d, name = write_variant(str(tmp_path), "Full", rcut=0.7, nqp=20)
    S = oracle_mod.System(d, name)
    S.set_phantom(True)
    S.build_nonlocal()
    I = S.info()
    assert list(I.nloc) == [5, 5, 5]
    node = (3, 4, 2)
    A, chans, gam = _dense_parts(oracle_mod, S, node)
    H = A + 0.333 * np.eye(I.nc) + np.einsum("c,pci,pcj->ij", gam, chans, chans)
    ev, V = np.linalg.eigh(H)
    a, b = S.lanczos(node, mode=1)
    # the Krylov space of e_c is at most nc-dimensional: the first k steps match
    k = 8
    lam, w, _, _ = oracle_mod.tridiag_nodes_weights(a, b, k)
    wc = V[I.center] ** 2
    # Gauss quadrature with k nodes integrates polynomials of degree 2k-1 exactly
    for m in range(2 * k):
        assert np.sum(w * lam ** m) == pytest.approx(np.sum(wc * ev ** m), rel=1e-8)


def test_gauss_quadrature_moments_mg1(oracle_mod, mg1_oracle):
    """sum_i w_i lambda_i^m = (H^m)_cc for m <= 2 N_qp - 1 (KB mode)."""
    S = mg1_oracle
    I = S.info()
    node = (4, 11, 7)
    lam, w, lmin, lmax = S.nodes_and_weights(mode=1, node_ids=[node[0] * 11 * 18 + node[1] * 11 + node[2]])
    e = np.zeros(I.nc)
    e[I.center] = 1.0
    v = e.copy()
    mom = [1.0]
    for m in range(1, 2 * I.nqp):
        v = S.apply_h(node, v, mode=1)
        mom.append(float(np.dot(e, v)))
    scale = max(abs(lmin[0]), abs(lmax[0]))
    for m in range(2 * I.nqp):
        q = float(np.sum(w[0] * lam[0] ** m))
        assert q == pytest.approx(mom[m], rel=1e-8 * (2 ** (m // 4)), abs=1e-10 * scale ** m)
    assert w[0].sum() == pytest.approx(1.0, abs=1e-13)


def test_lanczos_tridiagonal_matches_eigh(oracle_mod, mg1_oracle):
    S = mg1_oracle
    a, b = S.lanczos((8, 9, 5), mode=1)
    n = S.info().nqp
    lam, w, _, _ = oracle_mod.tridiag_nodes_weights(a, b, n)
    ev, V = eigh_tridiagonal(a[:n], b[: n - 1])
    o = np.argsort(lam)
    assert np.allclose(lam[o], ev, rtol=1e-12)
    assert np.allclose(w[o], V[0] ** 2, rtol=1e-8, atol=1e-14)


def test_veff_constant_shifts_nodes(oracle_mod, mg1_oracle):
    S = mg1_oracle
    ids = [5, 700, 2000]
    l1, w1, _, _ = S.nodes_and_weights(mode=1, veff=0.333, node_ids=ids)
    l2, w2, _, _ = S.nodes_and_weights(mode=1, veff=1.333, node_ids=ids)
    assert np.allclose(np.sort(l2, 1), np.sort(l1, 1) + 1.0, rtol=0, atol=1e-11)
    assert np.allclose(np.sort(w2, 1), np.sort(w1, 1), atol=1e-10)


def test_single_channel_modes_coincide(oracle_mod, nlm1_deck):
    """With one nonlocal channel (l_max=1, l_loc=1) the collapsed and KB operators are the same."""
    S = oracle_mod.System(*nlm1_deck)
    S.set_phantom(True)
    S.build_nonlocal()
    assert S.info().IP_length == 5
    x = np.random.default_rng(0).standard_normal(S.info().nc)
    y0 = S.apply_h((8, 9, 5), x, mode=0)
    y1 = S.apply_h((8, 9, 5), x, mode=1)
    assert np.array_equal(y0, y1)


def test_breakdown_is_reported(oracle_mod, tmp_path):
    """A 1-point cube (RADIUS_TRUNC 0) makes b_0 = 0 exactly; the reference
    divides by it (lsdft_scf.cc:131), the oracle reports it."""
    d, name = write_variant(str(tmp_path), "Zero", rcut=0, nqp=3)
    S = oracle_mod.System(d, name)
    S.build_nonlocal()
    assert S.info().nc == 1
    with pytest.raises(oracle_mod.OracleError, match="code 7"):
        S.lanczos((3, 3, 3), mode=1)


def test_nodes_and_weights_threads_agree(mg1_oracle):
    ids = list(range(0, 3366, 421))
    a = mg1_oracle.nodes_and_weights(mode=0, node_ids=ids, nthreads=1)
    b = mg1_oracle.nodes_and_weights(mode=0, node_ids=ids, nthreads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
