"""The optimised CPU baseline (oracle/sq_cpu_fast.c) against the reference
restatement (oracle/sq_oracle.c) on the same nodes.

Both compute NodesAndWeights_LSSGQ (lsdft_scf.cc:32-82) with the
Kleinman-Bylander operator (lsdft_matvec.cc:155-293).  The fast form applies
the stencil matrix-free, visits only the projector images that meet each
cube and contracts with FMA, so its sums round differently: the checks are
the spectrum bounds and smooth-function quadratures (what the density and
energy consume) to 1e-10, not the individual Ritz values.
"""
import numpy as np
import pytest

from conftest import MG1


def _compare(a, b):
    lo, wo, mno, mxo = a
    lf, wf, mnf, mxf = b
    assert np.max(np.abs(mnf - mno)) < 1e-9 and np.max(np.abs(mxf - mxo)) < 1e-9
    assert np.max(np.abs(wf.sum(1) - 1.0)) < 1e-12
    for beta, mu in ((30.0, 0.2), (30.0, 0.6), (5.0, 1.0)):
        f = lambda l: 1.0 / (1.0 + np.exp(np.clip(beta * (l - mu), -700.0, 700.0)))
        qo = np.sum(wo * f(lo), 1)
        qf = np.sum(wf * f(lf), 1)
        assert np.max(np.abs(qf - qo)) < 1e-10, (beta, mu)


def test_fast_matches_restatement_on_mg1(mg1_oracle):
    S = mg1_oracle
    total = int(np.prod(S.box[3:]))
    rng = np.random.default_rng(11)
    ids = np.unique(np.concatenate([[0, total - 1, total // 2], rng.integers(0, total, 40)])).astype(np.int32)
    a = S.nodes_and_weights(mode=1, node_ids=ids, nthreads=8)
    b = S.fast_nodes_and_weights(node_ids=ids, nthreads=8)
    _compare(a, b)


def test_fast_matches_restatement_with_a_field(oracle_mod):
    """A V_eff field on the periodic cell, the form the SCF passes."""
    S = oracle_mod.System(MG1, "Mg1")
    S.set_phantom(False)
    S.set_cell_period(True)
    S.build_nonlocal(S.cell())
    I = S.info()
    G = [I.npts[d] - 1 for d in range(3)]
    z, y, x = np.meshgrid(*[np.arange(G[d]) / G[d] for d in (2, 1, 0)], indexing="ij")
    veff = 0.2 * np.cos(2 * np.pi * x) * np.sin(2 * np.pi * (y + z)) - 0.1
    total = int(np.prod(S.box[3:]))
    ids = np.random.default_rng(5).integers(0, total, 24).astype(np.int32)
    a = S.nodes_and_weights(mode=1, veff=veff.ravel(), node_ids=ids, nthreads=8)
    b = S.fast_nodes_and_weights(veff=veff.ravel(), node_ids=ids, nthreads=8)
    _compare(a, b)


def test_fast_rejects_bad_node_ids(mg1_oracle):
    total = int(np.prod(mg1_oracle.box[3:]))
    with pytest.raises(RuntimeError):
        mg1_oracle.fast_nodes_and_weights(node_ids=[0, total], nthreads=2)
