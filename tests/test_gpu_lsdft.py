"""The reference-named layer on the GPU: Mult_NonlocalVector /
Mult_HamiltonianVector semantics (V2 += V_nl V1) against the oracle, and the
sq_main driver (src/main.cc's steps) on the Mg1 deck against the golden
oracle results, through its .lambdas / .den files."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, MG1, REPO, rel_err

import sqminiapp_amd as sq

pytestmark = pytest.mark.gpu
EXE = os.path.join(REPO, "sqminiapp_amd", "sq_main")


@pytest.mark.parametrize("mode", [sq.SQ_NL_KB, sq.SQ_NL_REF_COLLAPSED])
def test_apply_nonlocal_accumulates_like_the_oracle(mg1_oracle, mode):
    rng = np.random.default_rng(21 + mode)
    with sq.Context(MG1, "Mg1", phantom_atom=True) as c:
        for node in ((0, 0, 0), (9, 4, 7), (16, 17, 10)):
            x, y0 = rng.standard_normal((2, c.info.nc))
            y = c.apply_nonlocal(node, x, y0, mode=mode)
            yo = mg1_oracle.apply_nonlocal(node, x, y0, mode=mode)
            assert rel_err(y - y0, yo - y0) < 1e-12, node
            # local part + nonlocal part = the full H application
            full = c.apply_h(node, x, mode=mode)
            local = full - c.apply_nonlocal(node, x, np.zeros_like(x), mode=mode)
            assert rel_err(c.apply_nonlocal(node, x, local, mode=mode), full) < 1e-14


def test_driver_mg1_kb_against_golden(tmp_path):
    g = dict(np.load(os.path.join(GOLDEN, "mg1_oracle.npz")))
    print("sq_main Mg1 kb -density", flush=True)
    r = subprocess.run([EXE, "-name", os.path.join(MG1, "Mg1"), "-phantom", "-mode", "kb", "-density",
                        "-verify_files", "-out", str(tmp_path / "Mg1")],
                       capture_output=True, text=True, cwd=str(tmp_path), timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Done LSSGQ" in r.stdout and "files read back" in r.stdout
    lf = float(re.search(r"Fermi level\s+=\s+(\S+)", r.stdout).group(1))
    assert lf == pytest.approx(float(g["kb_lf"]), abs=1e-11)
    rows = np.loadtxt(tmp_path / "Mg1_0.lambdas", skiprows=1)
    assert open(tmp_path / "Mg1_0.lambdas").readline().split() == ["0", "0", "0", "11", "18", "17"]
    assert np.array_equal(rows[:, 0], np.arange(3366))
    assert np.max(np.abs(rows[:, 1] - g["kb_lmin"])) < 1e-11
    assert np.max(np.abs(rows[:, 2] - g["kb_lmax"])) < 1e-11
    den = np.loadtxt(tmp_path / "Mg1_0.den", skiprows=1)
    k, j, i = den[:, 0].astype(int), den[:, 1].astype(int), den[:, 2].astype(int)
    assert np.array_equal((k * 18 + j) * 11 + i, np.arange(3366))
    assert rel_err(den[:, 3], g["kb_rho"]) < 1e-10


def test_driver_scf_matches_the_api(tmp_path):
    """sq_main -scf -forces: SCF_LSSGQ on the Mg1 cell with the deck's SCF
    keywords (MIXING_PARAMETER 0.05, history 7, Pulay every 2nd step, TOL_SCF
    1e-5), then Calculate_localforce / Calculate_nonlocalforce / Display_force."""
    print("sq_main Mg1 -scf -forces", flush=True)
    r = subprocess.run([EXE, "-name", os.path.join(MG1, "Mg1"), "-mode", "kb", "-scf", "-forces", "-out",
                        str(tmp_path / "Mg1")], capture_output=True, text=True, cwd=str(tmp_path), timeout=300)
    assert r.returncode == 0, r.stderr
    e = float(re.search(r"Total energy\s+=\s+(\S+)", r.stdout).group(1))
    its = int(re.search(r"SCF converged in (\d+) iterations", r.stdout).group(1))
    with sq.Context(MG1, "Mg1", cell=True) as c:
        res = c.scf(sq.SQ_NL_KB, sq.SQ_FP64)
        F, _, _, _ = c.forces()
    assert res["converged"] and res["iterations"] == its
    Ff = np.loadtxt(tmp_path / "Mg1.force")
    assert Ff.shape == (4, 3) and np.allclose(Ff, F, rtol=0, atol=1e-15)
    assert "Maximum force component" in r.stdout
    assert e == pytest.approx(res["e_total"], abs=1e-11)
    den = np.loadtxt(tmp_path / "Mg1_0.den", skiprows=1)
    assert den.shape[0] == 10 * 17 * 16
    assert den[:, 3].sum() * c.info.dV == pytest.approx(8.0, rel=1e-9)
