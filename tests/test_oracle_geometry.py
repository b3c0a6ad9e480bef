"""Pins the oracle's setup geometry against the numbers the reference's own logs hold.

test/a1.out:12, test/a2.out:20,29, test/a4.out:36-63 print
AtomOverlap_local[0].Natoms = 115 (1 rank), 88/88 (2 ranks), 68 x4 (4 ranks);
test/a4a.out:70-76 gives 891/891/792/792 owned points, a 1x2x2 split of the
11x18x17 grid.  These counts depend on the pseudocharge cutoff
(lsdft_initObjs.cc:628-817), the periodic image range, the single-precision
roundf of the image position (lsdft_initObjs.cc:1534-1536), Check_Overlap
(2612-2663) and the atom list -- which, per endPosSC = 3*nAtoms
(lsdft_initObjs.cc:3038), holds a fifth atom read past the end of the
position array.  With that phantom atom at the origin every count matches.
"""
import numpy as np
import pytest

from conftest import MG1

ONE = [(0, 0, 0, 11, 18, 17)]
TWO = [(0, 0, 0, 11, 9, 17), (0, 9, 0, 11, 9, 17)]                      # y split 9/9
FOUR = [(0, 0, 0, 11, 9, 9), (0, 9, 0, 11, 9, 9), (0, 0, 9, 11, 9, 8), (0, 9, 9, 11, 9, 8)]


def test_partition_matches_logged_point_counts():
    # 891/891/792/792 points per rank (test/a4a.out:70-76)
    assert sorted(b[3] * b[4] * b[5] for b in FOUR) == [792, 792, 891, 891]


def test_overlap_counts_match_reference_logs_with_phantom(oracle_mod):
    S = oracle_mod.System(MG1, "Mg1")
    S.set_phantom(True, (0.0, 0.0, 0.0))
    assert [S.count_local_overlaps(b)[0] for b in ONE] == [115]
    assert [S.count_local_overlaps(b)[0] for b in TWO] == [88, 88]
    assert [S.count_local_overlaps(b)[0] for b in FOUR] == [68, 68, 68, 68]


def test_overlap_counts_without_phantom_do_not_match(oracle_mod):
    S = oracle_mod.System(MG1, "Mg1")
    S.set_phantom(False)
    assert [S.count_local_overlaps(b)[0] for b in ONE] == [88]
    assert [S.count_local_overlaps(b)[0] for b in FOUR] == [56] * 4


def test_mg1_deck_quantities(oracle_mod):
    S = oracle_mod.System(MG1, "Mg1")
    I = S.info()
    assert list(I.npts) == [11, 18, 17]
    assert list(I.nloc) == [21, 21, 21]           # 2*ceil(6/h)+1
    assert I.nqp == 10 and I.order == 6 and I.noetot == 8
    assert I.beta == 30.0
    # h = 2R/(N-1)
    for d, (R, N) in enumerate(zip([3.0214716611851919, 5.2333424308022884, 4.9240923662335074], [11, 18, 17])):
        assert I.delta[d] == pytest.approx(2 * R / (N - 1), rel=1e-15)
    assert I.lmax[0] == 2 and I.lloc[0] == 1 and I.psd_size[0] == 1062
    assert I.rc[0][0] == pytest.approx(3.04964630763)
    assert I.gamma[0][1] == 0.0


def test_gamma_against_scipy_quadrature(oracle_mod):
    """Gamma_l = dV / int_0^{rc+max h} U dV U r^2 dr (rectangle rule, dr=1e-6)."""
    from scipy.integrate import quad
    from scipy.interpolate import CubicSpline
    S = oracle_mod.System(MG1, "Mg1")
    I = S.info()
    hmax = max(I.delta)
    for l in (0, 2):
        R = S.radial(0, l)
        f = CubicSpline(R["r"], R["udv"], bc_type="not-a-knot")
        u = CubicSpline(R["r"], R["u"], bc_type="not-a-knot")
        rc = I.rc[0][l]
        val, _ = quad(lambda r: f(r) * u(r) * r * r, 0.0, rc + hmax, limit=400, epsabs=0, epsrel=1e-12)
        assert I.gamma[0][l] == pytest.approx(I.dV / val, rel=2e-5)


def test_pseudocharge_cutoff_is_grid_multiple(oracle_mod):
    S = oracle_mod.System(MG1, "Mg1")
    I = S.info()
    for d in range(3):
        n = I.cutoff[0][d] / I.delta[d]
        assert abs(n - round(n)) < 1e-9 and 3.0 < I.cutoff[0][d] <= 15.0 + I.delta[d]


def test_ip_index_slots(mg1_oracle):
    # Nlm = lmax^2 + 2(lmax - lloc) = 6 slots per supercell atom, phantom included
    assert list(mg1_oracle.ip_index()) == [0, 6, 12, 18, 24]
    assert mg1_oracle.info().IP_length == 30


def test_projector_flattening_order(mg1_oracle):
    """k,j,i outer and (l,m) inner with m-l stored (lsdft_nl.cc:81-123)."""
    im = mg1_oracle.images(0)[3]
    chi, ids = mg1_oracle.chi(0, 3, im["ndim"])
    box = im["box"]
    nx, ny, nz = box[3] - box[0] + 1, box[4] - box[1] + 1, box[5] - box[2] + 1
    assert im["ndim"] == 6 * nx * ny * nz
    assert list(ids["l"][:6]) == [0, 2, 2, 2, 2, 2]
    assert list(ids["m"][:6]) == [0, -2, -1, 0, 1, 2]
    assert np.all(ids["i"][:6] == box[0]) and np.all(ids["i"][6:12] == box[0] + 1)
    assert ids["k"][-1] == box[5] and ids["j"][-1] == box[4] and ids["i"][-1] == box[3]
