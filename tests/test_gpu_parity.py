"""HIP path (libsqamd.so through its C ABI) against the oracle.

Tolerances (FP64 form): per-point H application 1e-13 relative (only the
projector coefficient alpha is a parallel reduction); Lanczos coefficients and
Gauss nodes 1e-12 relative in KB mode, 1e-10 in the collapsed mode whose
spectrum spans +-6e3 Ha; QL bit-identical for identical tridiagonals; Fermi
level 1e-11 absolute.  FP32 form: stated against the FP64 oracle below.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, MG1, kernel_like, rel_err, write_variant

import sqminiapp_amd as sq

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sq.Context(MG1, "Mg1", phantom_atom=True, device=0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(os.path.join(GOLDEN, "mg1_oracle.npz")))


def _node(nid):
    return (nid // (11 * 18), (nid // 11) % 18, nid % 11)


@pytest.mark.parametrize("mode", [sq.SQ_NL_KB, sq.SQ_NL_REF_COLLAPSED])
def test_apply_h_matches_oracle(ctx, mg1_oracle, mode):
    rng = np.random.default_rng(11 + mode)
    for nid in (0, 197, 1700, 3365):
        x = rng.standard_normal(ctx.info.nc)
        y = ctx.apply_h(_node(nid), x, mode=mode)
        yo = mg1_oracle.apply_h(_node(nid), x, mode=mode)
        assert rel_err(y, yo) < 1e-13, nid


def test_apply_h_with_veff_field(mg1_oracle):
    I = mg1_oracle.info()
    g = np.indices((17, 18, 11)).astype(float)
    field = 0.3 + 0.1 * np.sin(0.7 * g[0]) * np.cos(0.4 * g[1]) + 0.05 * g[2] / 11
    with sq.Context(MG1, "Mg1", phantom_atom=True) as c:
        c.set_veff(field)
        x = np.random.default_rng(5).standard_normal(I.nc)
        for nid in (0, 1234):
            y = c.apply_h(_node(nid), x, mode=sq.SQ_NL_KB)
            yo = mg1_oracle.apply_h(_node(nid), x, mode=1, veff=field.ravel())
            assert rel_err(y, yo) < 1e-13
        c.run(sq.SQ_NL_KB, sq.SQ_FP64, begin=1234, count=1)
        lam, w, lo, hi = c.results()
        lo_o, wo, _, _ = mg1_oracle.nodes_and_weights(mode=1, veff=field.ravel(), node_ids=[1234])
        assert rel_err(np.sort(lam), np.sort(lo_o)) < 1e-12


def test_ql_bit_identical(ctx, golden, oracle_mod):
    rng = np.random.default_rng(2)
    n = 10
    diag = np.concatenate([golden["kb_a"][:, :n], golden["ref_a"][:, :n], rng.uniform(-5, 20, (50, n))])
    sub = np.concatenate([golden["kb_b"][:, :n], golden["ref_b"][:, :n], rng.uniform(0.1, 8, (50, n))])
    lam, w, lmin, lmax = ctx.tridiag_nodes_weights(diag, sub)
    for r in range(len(diag)):
        lo, wo, mn, mx = oracle_mod.tridiag_nodes_weights(diag[r], sub[r], n)
        assert np.array_equal(lam[r], lo) and np.array_equal(w[r], wo) and lmin[r] == mn and lmax[r] == mx, r
    # a longer recurrence (N_qp = 80)
    d80 = rng.uniform(-1, 25, (64, 80))
    s80 = rng.uniform(0.5, 9, (64, 80))
    lam, w, _, _ = ctx.tridiag_nodes_weights(d80, s80)
    for r in range(0, 64, 9):
        lo, wo, _, _ = oracle_mod.tridiag_nodes_weights(d80[r], s80[r], 80)
        assert np.array_equal(lam[r], lo) and np.array_equal(w[r], wo)


@pytest.mark.parametrize("mode,tag,tol", [(sq.SQ_NL_KB, "kb", 1e-12), (sq.SQ_NL_REF_COLLAPSED, "ref", 1e-10)])
def test_lanczos_coefficients_match_oracle(ctx, golden, mode, tag, tol):
    ctx.run(mode, sq.SQ_FP64)
    a, b = ctx.lanczos()
    s = golden["sample"]
    assert rel_err(a[s], golden[f"{tag}_a"]) < tol
    assert rel_err(b[s], golden[f"{tag}_b"]) < tol


def test_mg1_nodes_weights_density_kb(ctx, golden, oracle_mod):
    ctx.run(sq.SQ_NL_KB, sq.SQ_FP64)
    lam, w, lmin, lmax = ctx.results()
    o = np.argsort(lam, 1)
    oo = np.argsort(golden["kb_lam"], 1)
    assert rel_err(np.take_along_axis(lam, o, 1), np.take_along_axis(golden["kb_lam"], oo, 1)) < 1e-12
    assert np.max(np.abs(np.take_along_axis(w, o, 1) - np.take_along_axis(golden["kb_w"], oo, 1))) < 1e-11
    assert np.max(np.abs(lmin - golden["kb_lmin"])) < 1e-11 and np.max(np.abs(lmax - golden["kb_lmax"])) < 1e-11
    assert np.max(np.abs(w.sum(1) - 1.0)) < 1e-12
    lf, rho, eb, mts = ctx.density_fermi_energy()
    assert lf == pytest.approx(float(golden["kb_lf"]), abs=1e-11)
    assert eb == pytest.approx(float(golden["kb_eband"]), rel=1e-11)
    assert mts == pytest.approx(float(golden["kb_mts"]), rel=1e-10)
    assert rel_err(rho, golden["kb_rho"]) < 1e-10
    assert rho.sum() * ctx.info.dV == pytest.approx(8.0, rel=1e-11)
    # the device reduction on the oracle's own nodes/weights gives the same level as the oracle
    lf2, rho2, eb2, mts2 = oracle_mod.density_fermi_energy(lam, w, ctx.info.beta, 8.0, ctx.info.dV)
    assert lf == pytest.approx(lf2, abs=1e-12)


@pytest.mark.parametrize("wg", ["512", "1024"])
@pytest.mark.parametrize("prec", [sq.SQ_FP64, sq.SQ_FP32])
def test_mg1_each_one_node_shape_against_golden(monkeypatch, golden, wg, prec):
    """The one-node kernel pinned to each workgroup shape (SQ_BATCH=1,
    SQ_ZMARCH_WG), whichever the autotuner would pick, on all 3366 nodes."""
    monkeypatch.setenv("SQ_BATCH", "1")
    monkeypatch.setenv("SQ_ZMARCH_WG", wg)
    with sq.Context(MG1, "Mg1", phantom_atom=True) as c:
        monkeypatch.delenv("SQ_BATCH")
        monkeypatch.delenv("SQ_ZMARCH_WG")
        assert c.info.workgroup[prec] == int(wg) and c.info.batch[prec] == 1
        c.run(sq.SQ_NL_KB, prec)
        lam, w, lmin, lmax = c.results()
        lf, rho, eb, mts = c.density_fermi_energy()
    ref = np.sort(golden["kb_lam"], 1)
    if prec == sq.SQ_FP64:
        assert rel_err(np.sort(lam, 1), ref) < 1e-12
        assert np.max(np.abs(lmin - golden["kb_lmin"])) < 1e-11 and np.max(np.abs(lmax - golden["kb_lmax"])) < 1e-11
        assert lf == pytest.approx(float(golden["kb_lf"]), abs=1e-11)
        assert rel_err(rho, golden["kb_rho"]) < 1e-10
    else:
        assert rel_err(np.sort(lam, 1), ref) < 2e-6
        assert abs(lf - float(golden["kb_lf"])) < 1e-6
        assert abs(eb - float(golden["kb_eband"])) / 5 < 1e-5
        assert rel_err(rho, golden["kb_rho"]) < 5e-5


def test_mg1_collapsed_nodes_match_oracle(ctx, golden):
    ctx.run(sq.SQ_NL_REF_COLLAPSED, sq.SQ_FP64)
    lam, w, lmin, lmax = ctx.results()
    s = golden["sample"]
    assert rel_err(np.sort(lam[s], 1), np.sort(golden["ref_lam"], 1)) < 1e-10
    assert rel_err(lmin[s], golden["ref_lmin"]) < 1e-10 and rel_err(lmax[s], golden["ref_lmax"]) < 1e-10


def test_fp32_form_error_against_fp64_oracle(ctx, golden):
    """FP32 vectors + compensated dots; FP64 QL, Fermi level and sums."""
    ctx.run(sq.SQ_NL_KB, sq.SQ_FP32)
    lam, w, lmin, lmax = ctx.results()
    ref = np.sort(golden["kb_lam"], 1)
    assert rel_err(np.sort(lam, 1), ref) < 2e-6
    lf, rho, eb, mts = ctx.density_fermi_energy()
    assert abs(lf - float(golden["kb_lf"])) < 1e-6
    # band energy per atom (5 atoms incl. the phantom) within 1e-5 Ha
    assert abs(eb - float(golden["kb_eband"])) / 5 < 1e-5
    assert rel_err(rho, golden["kb_rho"]) < 5e-5   # measured 1.2e-5


def test_runs_are_deterministic_and_partition_independent(ctx):
    ctx.run(sq.SQ_NL_KB, sq.SQ_FP64, begin=1683, count=891)
    l1, w1, _, _ = ctx.results()
    ctx.run(sq.SQ_NL_KB, sq.SQ_FP64, begin=1683, count=891)
    l2, w2, _, _ = ctx.results()
    assert np.array_equal(l1, l2) and np.array_equal(w1, w2)
    # the same nodes owned by a rank box (0,9,9 | 11,9,8) of a 4-rank split, same kernel
    with kernel_like(ctx.info, sq.SQ_FP64):
        rc = sq.Context(MG1, "Mg1", phantom_atom=True, rank_box=(0, 9, 9, 11, 9, 8))
    with rc:
        rc.run(sq.SQ_NL_KB, sq.SQ_FP64)
        lr, wr, _, _ = rc.results()
    ctx.run(sq.SQ_NL_KB, sq.SQ_FP64)
    lf, wf, _, _ = ctx.results()
    ids = [(k * 18 + j) * 11 + i for k in range(9, 17) for j in range(9, 18) for i in range(11)]
    assert np.array_equal(lr, lf[ids]) and np.array_equal(wr, wf[ids])


def test_breakdown_reported_by_device(tmp_path):
    d, name = write_variant(str(tmp_path), "Zero", rcut=0, nqp=3)
    with sq.Context(d, name) as c:
        assert c.info.nc == 1
        with pytest.raises(sq.SQError) as e:
            c.run(sq.SQ_NL_KB, sq.SQ_FP64, begin=0, count=4)
        assert e.value.code == 7


def test_bad_node_range_rejected(ctx):
    with pytest.raises(sq.SQError) as e:
        ctx.run(sq.SQ_NL_KB, sq.SQ_FP64, begin=3000, count=400)
    assert e.value.code == 4
