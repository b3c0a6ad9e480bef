"""The 16-atom hcp Mg cell (29^3 cubes, 80 Lanczos steps, ~21k nodes) on the GPU.

Oracle parity on sampled nodes, and over all nodes the size-independent
properties of a Gauss rule: weights positive and summing to one, nodes inside
[lambda_min, lambda_max] and inside the Gershgorin interval of H; 80 FP64
Lanczos steps amplify the reduction-order rounding of each step, so the
per-node comparison is on the quadrature of smooth functions and on the
extreme nodes, which the recurrence converges first.
"""
import numpy as np
import pytest

from conftest import rel_err

import sqminiapp_amd as sq
from sqminiapp_amd import decks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c16(tmp_path_factory):
    d, name = decks.deck("C16", str(tmp_path_factory.mktemp("c16")))
    c = sq.Context(d, name)
    yield c, d, name
    c.close()


def test_c16_sampled_nodes_match_oracle(c16, oracle_mod):
    ctx, d, name = c16
    S = oracle_mod.System(d, name)
    S.set_phantom(False)
    S.build_nonlocal()
    ids = [0, 5000, 10499, 20999]
    lo, wo, mno, mxo = S.nodes_and_weights(mode=1, node_ids=ids, nthreads=4)
    for nid, r in zip(ids, range(len(ids))):
        ctx.run(sq.SQ_NL_KB, sq.SQ_FP64, begin=nid, count=1)
        lam, w, mn, mx = ctx.results()
        assert mn[0] == pytest.approx(mno[r], abs=1e-9) and mx[0] == pytest.approx(mxo[r], abs=1e-9)
        for beta, lf in ((30.0, 0.2), (30.0, 0.6), (5.0, 1.0)):
            f = lambda l: 1.0 / (1.0 + np.exp(beta * (l - lf)))
            assert np.sum(w[0] * f(lam[0])) == pytest.approx(np.sum(wo[r] * f(lo[r])), abs=1e-10)
        for m in range(1, 7):
            assert np.sum(w[0] * lam[0] ** m) == pytest.approx(np.sum(wo[r] * lo[r] ** m), rel=1e-10)


def test_c16_full_pass_invariants(c16):
    ctx, _, _ = c16
    I = ctx.info
    ctx.run(sq.SQ_NL_KB, sq.SQ_FP64)
    lam, w, lmin, lmax = ctx.results()
    assert lam.shape == (21000, 80)
    assert np.all(np.isfinite(lam)) and np.all(w >= 0)
    assert np.max(np.abs(w.sum(1) - 1.0)) < 1e-10
    assert np.all(lam >= lmin[:, None] - 1e-12) and np.all(lam <= lmax[:, None] + 1e-12)
    # Gershgorin: |L| row sums + V_eff; the KB term is positive semidefinite or not, bound it generously
    lap = sum(abs(I.coeffs[d][0]) + 2 * sum(abs(I.coeffs[d][p]) for p in range(1, I.order + 1)) for d in range(3))
    assert lmin.min() > -50 and lmax.max() < lap + 0.333 + 50
    lf, rho, eb, mts = ctx.density_fermi_energy()
    assert rho.sum() * I.dV == pytest.approx(I.noetot, rel=1e-10)
    assert np.all(rho > -1e-8)


def test_c16_fp32_error_against_fp64(c16):
    """FP32 recurrence, FP64 QL/Fermi/sums, against the FP64 form on every node.
    At BETA 100 (kT = 10 mHa) the occupation is still steep, so a node whose
    Gauss nodes straddle lambda_f moves its density by a few percent; the
    integrated quantities stay within the bounds below."""
    ctx, _, _ = c16
    ctx.run(sq.SQ_NL_KB, sq.SQ_FP64)
    l64, w64, mn64, mx64 = ctx.results()
    lf64, rho64, eb64, ts64 = ctx.density_fermi_energy()
    ctx.run(sq.SQ_NL_KB, sq.SQ_FP32)
    l32, w32, mn32, mx32 = ctx.results()
    lf32, rho32, eb32, ts32 = ctx.density_fermi_energy()
    natoms = ctx.info.nAtoms
    stats = dict(d_lambda_f=abs(lf32 - lf64), d_energy_per_atom=abs((eb32 + ts32) - (eb64 + ts64)) / natoms,
                 rho_rel_l2=float(np.linalg.norm(rho32 - rho64) / np.linalg.norm(rho64)),
                 rho_rel_max=rel_err(rho32, rho64), d_lmin=float(np.max(np.abs(mn32 - mn64))),
                 d_lmax=float(np.max(np.abs(mx32 - mx64))))
    print("C16 FP32 vs FP64:", stats)
    assert stats["d_lambda_f"] < 1e-4
    assert stats["d_energy_per_atom"] < 1e-4
    assert stats["rho_rel_l2"] < 1e-2 and stats["rho_rel_max"] < 0.2
    assert stats["d_lmin"] < 1e-3 and stats["d_lmax"] < 1e-3
