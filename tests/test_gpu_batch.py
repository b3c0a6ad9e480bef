"""The batched Lanczos kernel (B nodes stacked in z per workgroup, projector
gather on the matrix cores) against the oracle and against the one-node
kernel.  SQ_BATCH fixes the batch size at context creation; info.batch
reports it.

Bars: FP64 as the one-node kernel's (Gauss nodes 1e-12 relative, weights
1e-11, Fermi level 1e-11 on Mg1 against the oracle golden); FP32 as its FP32
form's; and bit-for-bit independence of the batch a node runs in (rank-box
partitions, sub-ranges), which the kernel gets by forming every per-node sum
in an order fixed by the node and its images.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, MG1, rel_err, write_variant

import sqminiapp_amd as sq
from sqminiapp_amd import decks

pytestmark = pytest.mark.gpu

CONFIGS = [(sq.SQ_FP64, 2, 1), (sq.SQ_FP64, 2, 2), (sq.SQ_FP32, 2, 4), (sq.SQ_FP32, 2, 2), (sq.SQ_FP32, 4, 4),
           (sq.SQ_FP32, 4, 2)]


def _ctx(monkeypatch, B, trip=None, deck=(MG1, "Mg1"), **kw):
    monkeypatch.setenv("SQ_BATCH", str(B))
    if trip:
        monkeypatch.setenv("SQ_BATCH_TRIP", str(trip))
    else:
        monkeypatch.delenv("SQ_BATCH_TRIP", raising=False)
    c = sq.Context(deck[0], deck[1], **kw)
    monkeypatch.delenv("SQ_BATCH")
    monkeypatch.delenv("SQ_BATCH_TRIP", raising=False)
    return c


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(os.path.join(GOLDEN, "mg1_oracle.npz")))


@pytest.mark.parametrize("prec,B,trip", CONFIGS)
def test_mg1_batched_against_golden(monkeypatch, golden, prec, B, trip):
    with _ctx(monkeypatch, B, trip, phantom_atom=True) as c:
        assert c.info.batch[prec] == B and c.info.gather_trip[prec] == trip
        c.run(sq.SQ_NL_KB, prec)
        lam, w, lmin, lmax = c.results()
        lf, rho, eb, mts = c.density_fermi_energy()
    o = np.argsort(lam, 1)
    oo = np.argsort(golden["kb_lam"], 1)
    ls, lo = np.take_along_axis(lam, o, 1), np.take_along_axis(golden["kb_lam"], oo, 1)
    ws, wo = np.take_along_axis(w, o, 1), np.take_along_axis(golden["kb_w"], oo, 1)
    stats = dict(lam=rel_err(ls, lo), w=float(np.max(np.abs(ws - wo))), lf=abs(lf - float(golden["kb_lf"])),
                 rho=rel_err(rho, golden["kb_rho"]))
    print(f"Mg1 batched prec={prec} B={B} trip={trip}:", stats)
    assert np.max(np.abs(w.sum(1) - 1.0)) < (1e-12 if prec == sq.SQ_FP64 else 1e-6)
    if prec == sq.SQ_FP64:
        assert stats["lam"] < 1e-12 and stats["w"] < 1e-11
        assert np.max(np.abs(lmin - golden["kb_lmin"])) < 1e-11 and np.max(np.abs(lmax - golden["kb_lmax"])) < 1e-11
        assert stats["lf"] < 1e-11 and stats["rho"] < 1e-10
        assert eb == pytest.approx(float(golden["kb_eband"]), rel=1e-11)
    else:
        # the one-node FP32 form's bars against the FP64 oracle
        assert stats["lam"] < 2e-6 and stats["lf"] < 1e-6 and stats["rho"] < 5e-5
        assert abs(eb - float(golden["kb_eband"])) / 5 < 1e-5


def test_mg1_batched_collapsed_mode(monkeypatch, golden):
    with _ctx(monkeypatch, 2, phantom_atom=True) as c:
        c.run(sq.SQ_NL_REF_COLLAPSED, sq.SQ_FP64)
        lam, w, lmin, lmax = c.results()
        a, b = c.lanczos()
    s = golden["sample"]
    assert rel_err(np.sort(lam[s], 1), np.sort(golden["ref_lam"], 1)) < 1e-10
    assert rel_err(a[s], golden["ref_a"]) < 1e-10 and rel_err(b[s], golden["ref_b"]) < 1e-10


@pytest.mark.parametrize("prec,B,trip", [(sq.SQ_FP64, 2, 2), (sq.SQ_FP32, 4, 4), (sq.SQ_FP32, 4, 2)])
def test_batched_results_do_not_depend_on_the_batch(monkeypatch, prec, B, trip):
    """Per kernel build (batch size and gather engine fixed in both contexts)."""
    with _ctx(monkeypatch, B, trip, phantom_atom=True) as c:
        assert c.info.batch[prec] == B and c.info.gather_trip[prec] == trip
        c.run(sq.SQ_NL_KB, prec)
        lf_, wf, _, _ = c.results()
        # a range starting mid-plane and mid-batch: its nodes run in other batches
        c.run(sq.SQ_NL_KB, prec, begin=5 * 198 + 17, count=891)
        l1, w1, _, _ = c.results()
        r = slice(5 * 198 + 17, 5 * 198 + 17 + 891)
        assert np.array_equal(l1, lf_[r]) and np.array_equal(w1, wf[r])
    # a rank box of a 4-rank split
    with _ctx(monkeypatch, B, trip, phantom_atom=True, rank_box=(0, 9, 9, 11, 9, 8)) as rc:
        assert rc.info.batch[prec] == B and rc.info.gather_trip[prec] == trip
        rc.run(sq.SQ_NL_KB, prec)
        lr, wr, _, _ = rc.results()
    ids = [(k * 18 + j) * 11 + i for k in range(9, 17) for j in range(9, 18) for i in range(11)]
    assert np.array_equal(lr, lf_[ids]) and np.array_equal(wr, wf[ids])


@pytest.fixture(scope="module")
def c16_deck(tmp_path_factory):
    return decks.deck("C16", str(tmp_path_factory.mktemp("c16b")))


def test_c16_batched_sampled_nodes_match_oracle(monkeypatch, c16_deck, oracle_mod):
    """32 nodes spread over the cell, oracle in FP64, through a range run of
    the batched kernel (each node in its own batch of the full pass)."""
    d, name = c16_deck
    S = oracle_mod.System(d, name)
    S.set_phantom(False)
    S.build_nonlocal()
    ids = np.linspace(0, 20999, 32).astype(int).tolist()
    lo, wo, mno, mxo = S.nodes_and_weights(mode=1, node_ids=ids, nthreads=8)
    with _ctx(monkeypatch, 2, deck=c16_deck) as c:
        assert c.info.batch[0] == 2
        c.run(sq.SQ_NL_KB, sq.SQ_FP64)
        lam, w, mn, mx = c.results()
    worst = 0.0
    for r, nid in enumerate(ids):
        assert mn[nid] == pytest.approx(mno[r], abs=1e-9) and mx[nid] == pytest.approx(mxo[r], abs=1e-9)
        for beta, lf in ((30.0, 0.2), (30.0, 0.6), (5.0, 1.0)):
            f = lambda l: 0.5 * (1.0 - np.tanh(0.5 * beta * (l - lf)))
            d_ = abs(np.sum(w[nid] * f(lam[nid])) - np.sum(wo[r] * f(lo[r])))
            worst = max(worst, d_)
            assert d_ < 1e-10
        for m in range(1, 7):
            assert np.sum(w[nid] * lam[nid] ** m) == pytest.approx(np.sum(wo[r] * lo[r] ** m), rel=1e-10)
    print("C16 batched FP64, 32 nodes against the oracle: worst smooth-quadrature difference", worst)


def test_c16_batched_fp32_against_fp64(monkeypatch, c16_deck):
    """The chosen FP32 batched form against the FP64 batched form, every node."""
    with _ctx(monkeypatch, 4, deck=c16_deck) as c:
        assert c.info.batch[1] == 4
        c.run(sq.SQ_NL_KB, sq.SQ_FP32)
        l32, w32, mn32, mx32 = c.results()
        lf32, rho32, eb32, ts32 = c.density_fermi_energy()
    with _ctx(monkeypatch, 2, deck=c16_deck) as c:
        c.run(sq.SQ_NL_KB, sq.SQ_FP64)
        l64, w64, mn64, mx64 = c.results()
        lf64, rho64, eb64, ts64 = c.density_fermi_energy()
        natoms = c.info.nAtoms
    stats = dict(d_lambda_f=abs(lf32 - lf64), d_energy_per_atom=abs((eb32 + ts32) - (eb64 + ts64)) / natoms,
                 rho_rel_l2=float(np.linalg.norm(rho32 - rho64) / np.linalg.norm(rho64)),
                 d_lmin=float(np.max(np.abs(mn32 - mn64))), d_lmax=float(np.max(np.abs(mx32 - mx64))))
    print("C16 batched FP32 vs FP64:", stats)
    assert stats["d_lambda_f"] < 1e-4 and stats["d_energy_per_atom"] < 3e-5
    assert stats["rho_rel_l2"] < 2e-3
    assert stats["d_lmin"] < 1e-3 and stats["d_lmax"] < 1e-3


def test_batched_matches_one_node_kernel_on_c16(monkeypatch, c16_deck):
    """Both kernels on the whole 16-atom pass: every node's rule, by its
    smooth quadratures and extreme nodes (80 steps amplify the different
    summation orders)."""
    res = {}
    for B in (1, 2):
        with _ctx(monkeypatch, B, deck=c16_deck) as c:
            assert c.info.batch[0] == B
            c.run(sq.SQ_NL_KB, sq.SQ_FP64)
            res[B] = c.results() + (c.density_fermi_energy(),)
    (l1, w1, mn1, mx1, e1), (l2, w2, mn2, mx2, e2) = res[1], res[2]
    assert np.max(np.abs(mn1 - mn2)) < 1e-9 and np.max(np.abs(mx1 - mx2)) < 1e-9
    f = lambda l: 0.5 * (1.0 - np.tanh(15.0 * (l - 0.6)))
    assert np.max(np.abs(np.sum(w1 * f(l1), 1) - np.sum(w2 * f(l2), 1))) < 1e-10
    assert abs(e1[0] - e2[0]) < 1e-10 and abs(e1[2] - e2[2]) < 1e-9


def test_batched_early_exit_equals_shorter_runs(monkeypatch, tmp_path):
    d, n = write_variant(str(tmp_path), "L", nqp=40)
    with _ctx(monkeypatch, 2, deck=(d, n)) as c:
        c.run(sq.SQ_NL_KB, sq.SQ_FP64)
        lf, _, _, _ = c.density_fermi_energy()
        c.set_early_exit(1e-3, lf, 0.0, 12, 4)
        c.run(sq.SQ_NL_KB, sq.SQ_FP64)
        lam, w, lmin, lmax = c.results()
        sizes = c.rule_sizes()
    assert sizes.min() >= 12 and np.any(sizes < 40)
    checked = 0
    for s in sorted(set(sizes.tolist()))[:3]:
        ids = np.nonzero(sizes == s)[0][:4]
        with _ctx(monkeypatch, 2, c.info.gather_trip[0], deck=(d, n), nqp=int(s)) as cs:
            for i in ids:
                cs.run(sq.SQ_NL_KB, sq.SQ_FP64, begin=int(i), count=1)
                ls, ws, mn, mx = cs.results()
                assert np.array_equal(lam[i, :s], ls[0]) and np.array_equal(w[i, :s], ws[0]), (i, s)
                assert np.all(w[i, s:] == 0.0)
                checked += 1
    assert checked > 0


def test_autotuner_reports_its_candidates():
    with sq.Context(MG1, "Mg1", phantom_atom=True) as c:
        I = c.info
        for prec in (0, 1):
            t = list(I.tune_candidates_ms[prec])
            print("autotuner", "FP64" if prec == 0 else "FP32", "candidates ms:", [round(x, 3) for x in t],
                  "batch", I.batch[prec], "trip", I.gather_trip[prec], "workgroup", I.workgroup[prec])
            timed = [x for x in t if x > 0]
            assert len(timed) >= 3
            # the chosen configuration is the fastest timed one
            k = (0 if I.workgroup[prec] == 512 else 1) if I.batch[prec] == 1 else \
                (2 if I.batch[prec] == 2 else 4) + (0 if I.gather_trip[prec] in (1, 4) else 1)
            assert t[k] == min(timed)
