"""AddressSanitizer + UndefinedBehaviorSanitizer builds (g++/gcc, CPU only) of
the product's host setup (sqminiapp_amd/csrc/sq_setup.cpp: deck parsing, image
lists, host projector tables, pseudocharge, atomic density, local forces) and
of the CPU oracle (oracle/sq_oracle.c: load, overlaps, Lanczos + QL in both
nonlocal modes, pseudocharge, atomic density, local and nonlocal forces), run
over decks by the drivers in tests/sanitize/.  Any sanitizer report fails."""
import os
import shutil
import subprocess

import pytest

from conftest import MG1

from sqminiapp_amd import decks

HERE = os.path.dirname(os.path.abspath(__file__))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                                reason="needs g++ and make")


@pytest.fixture(scope="module")
def san_build(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("sanbuild"))
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize"), f"OUT={out}"], check=True,
                   capture_output=True, text=True)
    return out


def _run(exe, deck):
    r = subprocess.run([exe, *deck], env=ENV, capture_output=True, text=True, timeout=600)
    text = r.stdout + r.stderr
    assert r.returncode == 0, text[-3000:]
    for bad in ("AddressSanitizer", "LeakSanitizer", "runtime error"):
        assert bad not in text, text[-3000:]
    return r.stdout


def test_host_setup_clean_on_mg1_and_c16(san_build, tmp_path):
    out = _run(os.path.join(san_build, "host_setup_asan"), MG1_DECK)
    assert "host setup driver done" in out and "int b dV -8.000000000000" in out
    out = _run(os.path.join(san_build, "host_setup_asan"), decks.deck("C16", str(tmp_path)))
    assert "host setup driver done" in out and "int b dV -32.000000000000" in out


def test_oracle_clean_on_mg1(san_build):
    out = _run(os.path.join(san_build, "oracle_asan"), MG1_DECK)
    assert "oracle driver done" in out and out.count("int rho dV 8.0000000000") == 2


MG1_DECK = (MG1, "Mg1")
