"""Forces on the GPU (sq_forces) against the oracle and against the energy.

- Parity: at the device's converged SCF state (its phi and V_eff fields, its
  Fermi level) the oracle's local and nonlocal forces (oracle/sq_oracle.c:
  explicit density-matrix columns Q f(T) e_1, a different route from the
  device's recorded projections) agree with sq_forces.
- Consistency: -dE/dR by central differences of three SCF energies (the atom
  moved by +-delta) against the force; SQ's truncated cubes and finite
  Gauss rules make the force the derivative only up to their own error,
  so the tolerance is the measured gap with margin (see the assertion).
- Precision: the FP32 form's forces against the plain FP64 form's.
"""
import numpy as np
import pytest

from conftest import MG1, write_variant

import sqminiapp_amd as sq

pytestmark = pytest.mark.gpu


def _displaced(tmp, name, shift=None, seed=3, sigma=0.1, **kw):
    from oracle import oracle as O
    S = O.System(MG1, "Mg1")
    pos = S.atom_positions() + np.random.default_rng(seed).normal(0.0, sigma, (4, 3))
    if shift is not None:
        atom, d, dl = shift
        pos[atom, d] += dl
    write_variant(str(tmp), name, atoms=pos, **kw)
    return str(tmp), name


def _scf_forces(d, name, precision=sq.SQ_FP64, tol=1e-9, tol_poisson=1e-12):
    with sq.Context(d, name, cell=True) as c:
        r = c.scf(precision=precision, tol=tol, max_iter=300, mixing=0.3, tol_poisson=tol_poisson)
        assert r["converged"], (r["iterations"], r["residual"])
        F, Fl, Fn, info = c.forces()
        phi, veff = c.cell_field("phi"), c.cell_field("veff")
    return r, F, Fl, Fn, info, phi, veff


def test_forces_match_oracle(tmp_path):
    from oracle import oracle as O
    d, n = _displaced(tmp_path, "D")
    r, F, Fl, Fn, info, phi, veff = _scf_forces(d, n, tol=1e-8, tol_poisson=0.0)
    print(f"SCF {r['iterations']} it, forces {info['wall_s']:.3f} s ({info['nodes']} nodes, {info['slots']} slots)", flush=True)
    S = O.System(d, n)
    S.set_phantom(False)
    S.set_cell_period(True)
    S.build_nonlocal(S.cell())
    I = S.info()
    _, _, sc = S.pseudocharge()
    Fl_o = S.local_forces(phi.ravel(), sc)
    Fn_o = S.nonlocal_forces(I.beta, info["lambda_f"], veff=veff.ravel(), nthreads=16)
    print("local  dev", Fl.tolist(), "\nlocal  ora", Fl_o.tolist(), flush=True)
    print("nonloc dev", Fn.tolist(), "\nnonloc ora", Fn_o.tolist(), flush=True)
    assert np.max(np.abs(Fl)) > 1e-3 and np.max(np.abs(Fn)) > 1e-3
    assert np.max(np.abs(Fl - Fl_o)) < 1e-11 * max(1.0, np.max(np.abs(Fl_o)))
    assert np.max(np.abs(Fn - Fn_o)) < 1e-10 * max(1.0, np.max(np.abs(Fn_o)))
    assert np.array_equal(F, Fl + Fn)


def _c16_deck(tmp, shift=None):
    """The 16-atom cell with a 5-bohr cube (narrower than the cell, so a cube
    holds one image of an atom), kT = 0.1 Ha and 40 Lanczos steps: a regime
    where the truncated cube resolves the density matrix."""
    import dataclasses
    from sqminiapp_amd import decks
    cfg = dataclasses.replace(decks.CONFIGS["C16"], rcut=5.0, beta=10.0, nqp=40)
    decks.write_deck(cfg, str(tmp))
    if shift is not None:
        atom, k, dl = shift
        ion = tmp / (cfg.name + ".ion")
        lines = ion.read_text().splitlines()
        v = [float(x) for x in lines[3 + atom].split()]
        v[k] += dl
        lines[3 + atom] = "%.16f %.16f %.16f" % tuple(v)
        ion.write_text("\n".join(lines) + "\n")
    return str(tmp), cfg.name


def test_forces_are_the_energy_derivative(tmp_path):
    """-dE/dR by central differences (delta = 0.01 bohr) against the force on
    one atom.  SQ's energy is not variational in its truncated cubes: the
    Hellmann-Feynman force is its derivative up to the cubes' truncation
    error, which grows as kT falls at a fixed cube (measured on this cell:
    gap ~2e-4 at kT = 0.1 Ha with a 5-bohr cube, ~3e-3 at kT = 0.033 Ha).
    The bound is twice the measured gap here."""
    atom, dl = 0, 0.01
    d, n = _c16_deck(tmp_path / "c")
    r0, F, Fl, Fn, info, _, _ = _scf_forces(d, n)
    gaps = []
    for dd in range(3):
        E = []
        for sgn in (1, -1):
            dx, nx = _c16_deck(tmp_path / f"s{dd}{sgn}", shift=(atom, dd, sgn * dl))
            E.append(_scf_forces(dx, nx)[0]["e_total"])
        fd = -(E[0] - E[1]) / (2 * dl)
        gaps.append(F[atom, dd] - fd)
        print(f"dir {dd}: force {F[atom, dd]:+.6e} (local {Fl[atom, dd]:+.3e}, nonlocal {Fn[atom, dd]:+.3e})"
              f"  -dE/dR {fd:+.6e}  gap {F[atom, dd] - fd:+.3e}", flush=True)
    print("sum of forces", F.sum(0).tolist(), flush=True)
    assert np.max(np.abs(F[atom])) > 2e-3
    assert np.max(np.abs(gaps)) < 5e-4


def test_fp32_forces_against_fp64(tmp_path):
    d, n = _displaced(tmp_path, "D")
    _, F64, *_ = _scf_forces(d, n, tol=1e-8, tol_poisson=0.0)
    r32, F32, *_ = _scf_forces(d, n, precision=sq.SQ_FP32, tol=1e-5, tol_poisson=0.0)
    diff = np.max(np.abs(F32 - F64))
    print(f"max |F32 - F64| = {diff:.3e} Ha/bohr (FP32 SCF {r32['iterations']} it)", flush=True)
    assert diff < 1e-4


def test_force_and_network_error_paths(tmp_path):
    """The new entry points fail loudly, with the documented status."""
    with sq.Context(MG1, "Mg1", cell=True) as c:
        with pytest.raises(sq.SQError) as e:
            c.forces()                                   # no SCF yet
        assert e.value.code == 9
        r = c.scf(sq.SQ_NL_REF_COLLAPSED, sq.SQ_FP64, tol=1e-4, max_iter=100, mixing=0.3)
        assert r["iterations"] > 0
        with pytest.raises(sq.SQError) as e:
            c.forces()                                   # the collapsed operator has no forces here
        assert e.value.code == 4
        with pytest.raises(sq.SQError) as e:
            c.kerker(np.zeros(c.cell_shape()), 0.0)
        assert e.value.code == 4
    with sq.Mlp() as m:
        with pytest.raises(sq.SQError) as e:
            m.train(10)                                  # no samples
        assert e.value.code == 9
    with sq.Context(MG1, "Mg1") as c:                    # node box is not the cell
        with pytest.raises(sq.SQError) as e:
            c.descriptors()
        assert e.value.code == 9
