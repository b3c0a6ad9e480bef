import contextlib
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")
MG1 = os.path.join(GOLDEN, "Mg1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) -- run with -m gpu")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def mg1_oracle(oracle_mod):
    """Mg1 as the reference ran it: phantom atom at the origin, one rank."""
    S = oracle_mod.System(MG1, "Mg1")
    S.set_phantom(True, (0.0, 0.0, 0.0))
    S.build_nonlocal()
    return S


def write_variant(tmp_dir, name, *, rcut=None, nqp=None, grid=None, lmax_keep=None, beta=None, atoms=None,
                  src_dir=MG1, src_name="Mg1"):
    """A copy of a deck with some keywords replaced; lmax_keep drops the
    pseudopotential and wave-function blocks with l > lmax_keep."""
    os.makedirs(os.path.join(tmp_dir, "pseudopotential"), exist_ok=True)
    lines = open(os.path.join(src_dir, src_name + ".inpt")).read().splitlines()
    out = []
    for ln in lines:
        key = ln.split(":")[0].strip() + ":" if ":" in ln else ""
        if key == "RADIUS_TRUNC:" and rcut is not None:
            ln = f"RADIUS_TRUNC: {rcut}"
        elif key == "LANCZOS_DEGREE:" and nqp is not None:
            ln = f"LANCZOS_DEGREE: {nqp}"
        elif key == "FD_GRID:" and grid is not None:
            ln = "FD_GRID: %d %d %d" % tuple(grid)
        elif key == "BETA:" and beta is not None:
            ln = f"BETA: {beta}"
        out.append(ln)
    open(os.path.join(tmp_dir, name + ".inpt"), "w").write("\n".join(out) + "\n")
    if atoms is None:
        ion = open(os.path.join(src_dir, src_name + ".ion")).read()
    else:
        ion = "%d\n\nMg 2 %d\n" % (len(atoms), len(atoms)) + "".join("%.16f %.16f %.16f\n" % tuple(a) for a in atoms)
    open(os.path.join(tmp_dir, name + ".ion"), "w").write(ion)
    pot_src = os.path.join(src_dir, "pseudopotential", "psd_Mg.pot")
    pot = open(pot_src).read().split("\n")
    if lmax_keep is not None:
        kept, skip = [], False
        i = 0
        while i < len(pot):
            ln = pot[i]
            s = ln.strip()
            if s.startswith("Pseudopotential follows") or s.startswith("Pseudo-wave-function follows"):
                lval = int(pot[i + 1].split()[0])
                skip = lval > lmax_keep
            elif s.endswith("follows") or s.startswith("Core charge") or s.startswith("Valence charge"):
                skip = False
            if not skip:
                kept.append(ln)
            i += 1
        pot = kept
    text = "\n".join(pot)
    if not text.endswith("\n"):
        text += "\n"  # the .pot reader stops a block list on the line after the last value
    open(os.path.join(tmp_dir, "pseudopotential", "psd_Mg.pot"), "w").write(text)
    return tmp_dir, name


@pytest.fixture(scope="session")
def tiny_deck(tmp_path_factory):
    """Mg1 geometry with a 7^3 cube (RADIUS_TRUNC 1.3) for brute-force checks."""
    d = str(tmp_path_factory.mktemp("tiny"))
    return write_variant(d, "Tiny", rcut=1.3, nqp=8)


@pytest.fixture(scope="session")
def nlm1_deck(tmp_path_factory):
    """Mg pseudopotential cut to l<=1 with l_loc=1: one nonlocal channel."""
    d = str(tmp_path_factory.mktemp("nlm1"))
    return write_variant(d, "One", rcut=2.5, nqp=8, lmax_keep=1)


def rel_err(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


KERNEL_ENV = ("SQ_BATCH", "SQ_ZMARCH_WG", "SQ_BATCH_TRIP")


@contextlib.contextmanager
def kernel_like(info, prec):
    """Contexts created inside the block run, for precision `prec`, the kernel
    that `info` reports.  Every context creation times the kernel candidates
    again, and where two run within noise of each other a second context may
    pick the other one: equally valid, but its sums round differently, so a
    bit-for-bit comparison across contexts pins the kernel first."""
    b = info.batch[prec]
    env = {"SQ_BATCH": str(b)}
    if b == 1:
        env["SQ_ZMARCH_WG"] = str(info.workgroup[prec])
    else:
        env["SQ_BATCH_TRIP"] = str(info.gather_trip[prec])
    old = {k: os.environ.get(k) for k in KERNEL_ENV}
    for k in KERNEL_ENV:
        os.environ.pop(k, None)
    os.environ.update(env)
    try:
        yield
    finally:
        for k in KERNEL_ENV:
            os.environ.pop(k, None)
            if old[k] is not None:
                os.environ[k] = old[k]
