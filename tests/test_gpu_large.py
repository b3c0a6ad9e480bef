"""The 128- and 1024-atom cells on the GPU.

C128 (hcp Mg 4x2x4, 4-site vacancy cluster, 162582 nodes):
- oracle parity on 8 sampled nodes (smooth-function quadratures and extreme
  nodes, as for the 16-atom cell) and a whole FP64 pass with the Gauss-rule
  invariants and the electron count;
- the FP32 form (the autotuned kernel) against FP64 on every node: Fermi
  level, band + entropy energy per atom, density;
- per-node early exit (30-80 steps): a stopped node's rule equals, bit for
  bit, a run of fixed shorter length in the same kernel configuration.
C1024 (1287384 nodes): one whole pass in FP64 and in FP32 with the Gauss-rule
invariants, the electron count, FP32 against FP64 on every node, and the
oracle on 4 sampled nodes.
"""
import os

import numpy as np
import pytest

from conftest import kernel_like

import sqminiapp_amd as sq
from sqminiapp_amd import decks

pytestmark = pytest.mark.gpu


def _rule_invariants(lam, w, lmin, lmax):
    assert np.all(np.isfinite(lam)) and np.all(w >= 0)
    assert np.max(np.abs(w.sum(1) - 1.0)) < 1e-10
    assert np.all(lam >= lmin[:, None] - 1e-12) and np.all(lam <= lmax[:, None] + 1e-12)


def _against_oracle(S, ids, lam, w, mn, mx, oracle_mod, tag):
    lo, wo, mno, mxo = S.nodes_and_weights(mode=1, node_ids=ids, nthreads=min(16, os.cpu_count() or 1))
    worst = 0.0
    for r, nid in enumerate(ids):
        assert mn[r] == pytest.approx(mno[r], abs=1e-9) and mx[r] == pytest.approx(mxo[r], abs=1e-9)
        for beta, lf in ((30.0, 0.2), (30.0, 0.6), (5.0, 1.0)):
            f = lambda l: 0.5 * (1.0 - np.tanh(0.5 * beta * (l - lf)))
            d = abs(np.sum(w[r] * f(lam[r])) - np.sum(wo[r] * f(lo[r])))
            worst = max(worst, d)
            assert d < 1e-10, (tag, nid)
    print(f"{tag}: {len(ids)} nodes against the oracle, worst smooth quadrature difference {worst:.2e}", flush=True)


def _fp32_vs_fp64(r64, r32, natoms):
    (l64, w64, mn64, mx64, (lf64, rho64, eb64, ts64)) = r64
    (l32, w32, mn32, mx32, (lf32, rho32, eb32, ts32)) = r32
    return dict(d_lambda_f=abs(lf32 - lf64), d_energy_per_atom=abs((eb32 + ts32) - (eb64 + ts64)) / natoms,
                rho_rel_l2=float(np.linalg.norm(rho32 - rho64) / np.linalg.norm(rho64)),
                d_lmin=float(np.max(np.abs(mn32 - mn64))), d_lmax=float(np.max(np.abs(mx32 - mx64))))


@pytest.fixture(scope="module")
def c128(tmp_path_factory):
    return decks.deck("C128", str(tmp_path_factory.mktemp("c128")))


def test_c128_sampled_nodes_and_full_pass(c128, oracle_mod):
    d, name = c128
    with sq.Context(d, name) as ctx:
        I = ctx.info
        n = I.n_nodes
        ids = np.linspace(0, n - 1, 8).astype(int).tolist()
        ctx.run(sq.SQ_NL_KB, sq.SQ_FP64)
        lam, w, lmin, lmax = ctx.results()
        assert lam.shape == (n, I.nqp)
        _rule_invariants(lam, w, lmin, lmax)
        lf, rho, eb, mts = ctx.density_fermi_energy()
        assert rho.sum() * I.dV == pytest.approx(I.noetot, rel=1e-10)
        assert I.noetot == 2 * (128 - 4)          # the vacancy cluster removed four atoms
    S = oracle_mod.System(d, name)
    S.set_phantom(False)
    S.build_nonlocal()
    _against_oracle(S, ids, lam[ids], w[ids], lmin[ids], lmax[ids], oracle_mod, "C128 FP64")


def test_c128_fp32_against_fp64_on_every_node(c128):
    d, name = c128
    with sq.Context(d, name) as ctx:
        res = {}
        for prec in (sq.SQ_FP64, sq.SQ_FP32):
            ctx.run(sq.SQ_NL_KB, prec)
            res[prec] = ctx.results() + (ctx.density_fermi_energy(),)
        stats = _fp32_vs_fp64(res[sq.SQ_FP64], res[sq.SQ_FP32], ctx.info.nAtoms)
        print(f"C128 FP32 (batch {ctx.info.batch[1]}) vs FP64 (batch {ctx.info.batch[0]}):", stats, flush=True)
    _rule_invariants(*res[sq.SQ_FP32][:4])
    assert stats["d_lambda_f"] < 1e-4 and stats["d_energy_per_atom"] < 3e-5
    assert stats["rho_rel_l2"] < 2e-3 and stats["d_lmin"] < 1e-5 and stats["d_lmax"] < 1e-5


@pytest.mark.parametrize("tol", [1e-4, 1e-6])
def test_c128_early_exit_equals_fixed_shorter_runs(c128, monkeypatch, tol):
    """Per-node retirement in the batched FP32 kernel: steps 30..80 checked
    every 10 at a tolerance on the node's Fermi quadrature (Matsubara terms);
    every stopped node equals a fixed run of its length.  The printed
    histogram is the record of how many nodes retire at each tolerance."""
    d, name = c128
    monkeypatch.setenv("SQ_BATCH", "4")
    with sq.Context(d, name) as ctx:
        assert ctx.info.batch[1] == 4
        info = ctx.info
        n = ctx.info.n_nodes
        begin, count = n // 2 - 8000, 16000
        ctx.run(sq.SQ_NL_KB, sq.SQ_FP32)
        lf, _, _, _ = ctx.density_fermi_energy()
        ctx.set_early_exit(tol, lf, 0.0, 30, 10)
        ctx.run(sq.SQ_NL_KB, sq.SQ_FP32, begin=begin, count=count)
        lam, w, lmin, lmax = ctx.results()
        sizes = ctx.rule_sizes()
    hist = np.bincount(sizes, minlength=81)[30::10].tolist()
    print(f"C128 early exit at {tol:g}: rule sizes 30, 40, ..., 80 ->", hist, "mean", sizes.mean(), flush=True)
    assert sizes.min() >= 30 and sizes.max() <= 80 and np.all((sizes - 30) % 10 == 0)
    checked = 0
    for s in sorted(set(sizes.tolist())):
        if s == 80:
            continue
        ids = np.nonzero(sizes == s)[0][:3]
        with kernel_like(info, sq.SQ_FP32):
            cs = sq.Context(d, name, nqp=int(s))
        with cs:
            assert cs.info.batch[1] == 4 and cs.info.gather_trip[1] == info.gather_trip[1]
            for i in ids:
                cs.run(sq.SQ_NL_KB, sq.SQ_FP32, begin=begin + int(i), count=1)
                ls, ws, mn, mx = cs.results()
                assert np.array_equal(lam[i, :s], ls[0]) and np.array_equal(w[i, :s], ws[0]), (i, s)
                assert lmin[i] == mn[0] and lmax[i] == mx[0] and np.all(w[i, s:] == 0.0)
                checked += 1
    print(f"C128 early exit: {checked} stopped nodes equal their fixed shorter runs", flush=True)


def test_c1024_full_pass_fp64_and_fp32(tmp_path, oracle_mod):
    d, name = decks.deck("C1024", str(tmp_path))
    with sq.Context(d, name) as ctx:
        I = ctx.info
        n = I.n_nodes
        ids = [0, n // 3 + 11, (2 * n) // 3 + 5, n - 1]
        res = {}
        for prec in (sq.SQ_FP64, sq.SQ_FP32):
            ctx.run(sq.SQ_NL_KB, prec)
            lam, w, lmin, lmax = ctx.results()
            assert lam.shape == (n, I.nqp)
            _rule_invariants(lam, w, lmin, lmax)
            lf, rho, eb, mts = ctx.density_fermi_energy()
            assert rho.sum() * I.dV == pytest.approx(I.noetot, rel=1e-10)
            print(f"C1024 pass prec={prec}: {ctx.timing()[0] / 1e3:.1f} s Lanczos, lambda_f {lf:.8f}", flush=True)
            res[prec] = (lam[ids].copy(), w[ids].copy(), lmin[ids].copy(), lmax[ids].copy(),
                         (lf, rho, eb, mts), (lam, w, lmin, lmax) if prec == sq.SQ_FP64 else None)
            del lam, w
        full64 = res[sq.SQ_FP64][5]
        lam32, w32, mn32, mx32 = ctx.results()
        stats = _fp32_vs_fp64(full64 + (res[sq.SQ_FP64][4],), (lam32, w32, mn32, mx32, res[sq.SQ_FP32][4]), I.nAtoms)
        print("C1024 FP32 vs FP64:", stats, flush=True)
        assert stats["d_lambda_f"] < 1e-4 and stats["d_energy_per_atom"] < 3e-5 and stats["rho_rel_l2"] < 2e-3
    S = oracle_mod.System(d, name)
    S.set_phantom(False)
    S.build_nonlocal()
    l64, w64, mn64, mx64 = res[sq.SQ_FP64][:4]
    _against_oracle(S, ids, l64, w64, mn64, mx64, oracle_mod, "C1024 FP64")
