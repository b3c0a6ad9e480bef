"""Projector tables built on the device (sq_setup.hip) against the host fill
(sq_setup.cpp, itself bit-identical to the oracle's chi, tests/test_host_setup.py):
every chi entry of every image equal bit for bit, on Mg1 and on the 16-atom
cell, and the Lanczos results of the two contexts equal."""
import os
import time

import numpy as np
import pytest

from conftest import MG1, kernel_like

import sqminiapp_amd as sq
from sqminiapp_amd import decks

pytestmark = pytest.mark.gpu


def _contexts(d, name, **kw):
    t0 = time.perf_counter()
    dev = sq.Context(d, name, **kw)
    t1 = time.perf_counter()
    os.environ["SQ_HOST_TABLES"] = "1"
    try:
        with kernel_like(dev.info, sq.SQ_FP64):   # the same kernel, so results compare bit for bit
            host = sq.Context(d, name, **kw)
    finally:
        del os.environ["SQ_HOST_TABLES"]
    t2 = time.perf_counter()
    print(f"{name}: context with device tables {t1 - t0:.2f} s, with host tables {t2 - t1:.2f} s", flush=True)
    return dev, host


@pytest.mark.parametrize("which", ["Mg1", "C16"])
def test_device_tables_bit_identical(which, tmp_path):
    if which == "Mg1":
        dev, host = _contexts(MG1, "Mg1", phantom_atom=True)
    else:
        d, n = decks.deck("C16", str(tmp_path))
        dev, host = _contexts(d, n)
    with dev, host:
        assert dev.info.chi_entries == host.info.chi_entries and dev.info.n_images == host.info.n_images
        for g in range(dev.info.n_images):
            a, b = dev.image(g)["chi"], host.image(g)["chi"]
            assert np.array_equal(a, b), g
        ids = (0, 7, dev.info.n_nodes - 1)
        for nid in ids:
            dev.run(sq.SQ_NL_KB, sq.SQ_FP64, begin=nid, count=1)
            host.run(sq.SQ_NL_KB, sq.SQ_FP64, begin=nid, count=1)
            la, wa, _, _ = dev.results()
            lb, wb, _, _ = host.results()
            assert np.array_equal(la, lb) and np.array_equal(wa, wb)
