"""The 16-atom hcp Mg cell (29^3 cubes, 80 Lanczos steps, ~21k nodes) on the GPU.

Oracle parity on sampled nodes, and over all nodes the size-independent
properties of a Gauss rule: weights positive and summing to one, nodes inside
[lambda_min, lambda_max] and inside the Gershgorin interval of H; 80 FP64
Lanczos steps amplify the reduction-order rounding of each step, so the
per-node comparison is on the quadrature of smooth functions and on the
extreme nodes, which the recurrence converges first.  Each kernel
configuration (the one-node kernel in its 512- and 1024-thread shapes, the
batched kernel) is pinned explicitly, so every variant is tested whichever
the autotuner would pick.
"""
import numpy as np
import pytest

from conftest import rel_err

import sqminiapp_amd as sq
from sqminiapp_amd import decks

pytestmark = pytest.mark.gpu

# (SQ_BATCH, SQ_ZMARCH_WG): one-node kernel in both shapes, the batched kernel
SHAPES = [("1", "512"), ("1", "1024"), ("auto", None)]


@pytest.fixture(scope="module")
def c16_deck(tmp_path_factory):
    return decks.deck("C16", str(tmp_path_factory.mktemp("c16")))


@pytest.fixture(scope="module")
def c16_oracle(c16_deck, oracle_mod):
    d, name = c16_deck
    S = oracle_mod.System(d, name)
    S.set_phantom(False)
    S.build_nonlocal()
    ids = [0, 5000, 10499, 20999]
    return ids, S.nodes_and_weights(mode=1, node_ids=ids, nthreads=8)


def _ctx(monkeypatch, deck, batch, wg):
    for k in ("SQ_BATCH", "SQ_ZMARCH_WG", "SQ_BATCH_TRIP"):
        monkeypatch.delenv(k, raising=False)
    if batch != "auto":
        monkeypatch.setenv("SQ_BATCH", batch)
    if wg:
        monkeypatch.setenv("SQ_ZMARCH_WG", wg)
    c = sq.Context(*deck)
    for k in ("SQ_BATCH", "SQ_ZMARCH_WG"):
        monkeypatch.delenv(k, raising=False)
    if wg:
        assert c.info.workgroup[0] == int(wg) and c.info.workgroup[1] == int(wg) and c.info.batch[1] == 1
    return c


def _quadratures(lam, w):
    out = []
    for beta, lf in ((30.0, 0.2), (30.0, 0.6), (5.0, 1.0)):
        out.append(np.sum(w * 0.5 * (1.0 - np.tanh(0.5 * beta * (lam - lf)))))
    return np.array(out)


@pytest.mark.parametrize("batch,wg", SHAPES)
def test_c16_sampled_nodes_match_oracle(monkeypatch, c16_deck, c16_oracle, batch, wg):
    ids, (lo, wo, mno, mxo) = c16_oracle
    with _ctx(monkeypatch, c16_deck, batch, wg) as ctx:
        for r, nid in enumerate(ids):
            ctx.run(sq.SQ_NL_KB, sq.SQ_FP64, begin=nid, count=1)
            lam, w, mn, mx = ctx.results()
            assert mn[0] == pytest.approx(mno[r], abs=1e-9) and mx[0] == pytest.approx(mxo[r], abs=1e-9)
            assert np.max(np.abs(_quadratures(lam[0], w[0]) - _quadratures(lo[r], wo[r]))) < 1e-10
            for m in range(1, 7):
                assert np.sum(w[0] * lam[0] ** m) == pytest.approx(np.sum(wo[r] * lo[r] ** m), rel=1e-10)
            # the FP32 form of the same node against the FP64 oracle
            ctx.run(sq.SQ_NL_KB, sq.SQ_FP32, begin=nid, count=1)
            l32, w32, mn32, mx32 = ctx.results()
            dq = np.max(np.abs(_quadratures(l32[0], w32[0]) - _quadratures(lo[r], wo[r])))
            print(f"C16 node {nid} FP32 ({batch}, {wg}) smooth quadratures vs the FP64 oracle: {dq:.2e}", flush=True)
            assert dq < 1e-7, (nid, dq)
            assert abs(mn32[0] - mno[r]) < 1e-4 and abs(mx32[0] - mxo[r]) < 1e-4


def test_c16_full_pass_invariants(monkeypatch, c16_deck):
    with _ctx(monkeypatch, c16_deck, "auto", None) as ctx:
        I = ctx.info
        ctx.run(sq.SQ_NL_KB, sq.SQ_FP64)
        lam, w, lmin, lmax = ctx.results()
        lf, rho, eb, mts = ctx.density_fermi_energy()
    assert lam.shape == (21000, 80)
    assert np.all(np.isfinite(lam)) and np.all(w >= 0)
    assert np.max(np.abs(w.sum(1) - 1.0)) < 1e-10
    assert np.all(lam >= lmin[:, None] - 1e-12) and np.all(lam <= lmax[:, None] + 1e-12)
    # Gershgorin: |L| row sums + V_eff; the KB term is positive semidefinite or not, bound it generously
    lap = sum(abs(I.coeffs[d][0]) + 2 * sum(abs(I.coeffs[d][p]) for p in range(1, I.order + 1)) for d in range(3))
    assert lmin.min() > -50 and lmax.max() < lap + 0.333 + 50
    assert rho.sum() * I.dV == pytest.approx(I.noetot, rel=1e-10)
    assert np.all(rho > -1e-8)


@pytest.mark.parametrize("batch,wg", SHAPES)
def test_c16_fp32_error_against_fp64(monkeypatch, c16_deck, batch, wg):
    """FP32 recurrence (Dot2-compensated dots), FP64 QL/Fermi/sums, against the
    FP64 form on every node.  Bounds about 3x the measured errors (FP32:
    lambda_f 3.2e-5, energy 8.8e-6 Ha/atom, density 7.5e-4 relative L2)."""
    with _ctx(monkeypatch, c16_deck, batch, wg) as ctx:
        ctx.run(sq.SQ_NL_KB, sq.SQ_FP64)
        l64, w64, mn64, mx64 = ctx.results()
        lf64, rho64, eb64, ts64 = ctx.density_fermi_energy()
        ctx.run(sq.SQ_NL_KB, sq.SQ_FP32)
        l32, w32, mn32, mx32 = ctx.results()
        lf32, rho32, eb32, ts32 = ctx.density_fermi_energy()
        natoms = ctx.info.nAtoms
        shape = (ctx.info.batch[1], ctx.info.workgroup[1], ctx.info.gather_trip[1])
    stats = dict(d_lambda_f=abs(lf32 - lf64), d_energy_per_atom=abs((eb32 + ts32) - (eb64 + ts64)) / natoms,
                 rho_rel_l2=float(np.linalg.norm(rho32 - rho64) / np.linalg.norm(rho64)),
                 rho_rel_max=rel_err(rho32, rho64), d_lmin=float(np.max(np.abs(mn32 - mn64))),
                 d_lmax=float(np.max(np.abs(mx32 - mx64))))
    print("C16 FP32 vs FP64 (batch, workgroup, trip)", shape, stats, flush=True)
    assert stats["d_lambda_f"] < 1e-4
    assert stats["d_energy_per_atom"] < 3e-5
    assert stats["rho_rel_l2"] < 2.5e-3 and stats["rho_rel_max"] < 0.2
    assert stats["d_lmin"] < 1e-4 and stats["d_lmax"] < 1e-4
