"""The product's host setup (libsqamd.so, host-only context) against the oracle, on the CPU.

Projector tables must be bit-identical: chi of every image (lsdft_nl.cc:141-208
flattened as lsdft_nl.cc:81-123), image boxes/centers/atoms
(lsdft_initObjs.cc:1671-1946), Gamma_l (822-869), the pseudocharge cutoff
(628-817) and the FD weights (lsdft_readfiles.cc:453-476).
"""
import os
import re

import numpy as np
import pytest

from conftest import MG1, REPO

import sqminiapp_amd as sq
from sqminiapp_amd import decks
from sqminiapp_amd._lib import EXPORTS, SQError


def _compare(ctx, S):
    I, OI = ctx.info, S.info()
    assert I.n_images == OI.nimages[0]
    assert list(I.nloc) == list(OI.nloc) and I.nqp == OI.nqp
    assert I.ip_length == OI.IP_length
    for d in range(3):
        assert I.delta[d] == OI.delta[d]
        assert I.cutoff[0][d] == OI.cutoff[0][d]
        for p in range(OI.order + 1):
            assert I.coeffs[d][p] == OI.coeffs[d][p]
    for l in range(4):
        assert I.gamma[0][l] == OI.gamma[0][l]
    for c, im in enumerate(S.images(0)):
        mine = ctx.image(c)
        chi, _ = S.chi(0, c, im["ndim"])
        assert mine["atom"] == im["atom"]
        assert np.array_equal(mine["box"], im["box"])
        assert np.array_equal(mine["center"], im["center"])
        assert np.array_equal(mine["chi"], chi), f"image {c}"


@pytest.mark.parametrize("phantom", [True, False])
def test_mg1_setup_bit_identical(oracle_mod, phantom):
    ctx = sq.Context(MG1, "Mg1", phantom_atom=phantom, device=-1)
    S = oracle_mod.System(MG1, "Mg1")
    S.set_phantom(phantom)
    S.build_nonlocal()
    _compare(ctx, S)
    assert ctx.info.n_sc_atoms == (5 if phantom else 4)


def test_rank_box_setup_bit_identical(oracle_mod):
    box = (0, 9, 9, 11, 9, 8)
    ctx = sq.Context(MG1, "Mg1", phantom_atom=True, rank_box=box, device=-1)
    S = oracle_mod.System(MG1, "Mg1")
    S.set_phantom(True)
    S.build_nonlocal(box)
    _compare(ctx, S)
    assert ctx.info.n_nodes == 792


def test_c16_deck_setup_bit_identical(oracle_mod, tmp_path):
    d, name = decks.deck("C16", str(tmp_path))
    ctx = sq.Context(d, name, device=-1)
    assert list(ctx.info.nloc) == [29, 29, 29] and ctx.info.n_nodes == 21000 and ctx.info.nqp == 80
    S = oracle_mod.System(d, name)
    S.set_phantom(False)
    S.build_nonlocal()
    _compare(ctx, S)


def test_deck_generator_is_deterministic(tmp_path):
    a = decks.CONFIGS["C128"].positions()
    b = decks.CONFIGS["C128"].positions()
    assert np.array_equal(a, b) and len(a) == 124
    assert decks.CONFIGS["C128"].grid() == [49, 42, 79]
    assert decks.CONFIGS["C1024"].grid() == [97, 84, 158]
    L = decks.CONFIGS["C16"].lengths()
    for Ld, N in zip(L, decks.CONFIGS["C16"].grid()):
        assert Ld / (N - 1) >= 0.5           # keeps n_loc = 29 at Rcut = 7


def test_abi_exports_every_declared_symbol():
    lib = sq.lib()
    header = open(os.path.join(REPO, "include", "sq_amd.h")).read()
    declared = sorted(set(re.findall(r"\b(sq_[a-z_]+)\s*\(", header)))
    assert declared == sorted(EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_abi_errors():
    with pytest.raises(SQError) as e:
        sq.Context("/nonexistent", "Mg1", device=-1)
    assert e.value.code == 1
    ctx = sq.Context(MG1, "Mg1", device=-1)
    with pytest.raises(SQError) as e:
        ctx.run()
    assert e.value.code == 9  # host-only context
    with pytest.raises(SQError) as e:
        sq.Context(MG1, "Mg1", rank_box=(0, 0, 0, 12, 18, 17), device=-1)
    assert e.value.code == 4


def test_malformed_deck_is_an_error(tmp_path):
    from conftest import write_variant
    d, name = write_variant(str(tmp_path), "Bad")
    with open(os.path.join(d, name + ".inpt"), "a") as f:
        f.write("FD_GRID: 11 eighteen 17\n")
    with pytest.raises(SQError) as e:
        sq.Context(d, name, device=-1)
    assert e.value.code == 2
