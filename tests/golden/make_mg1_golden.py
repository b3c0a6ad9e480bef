"""Writes tests/golden/mg1_oracle.npz: oracle results on the shipped Mg1 deck
as the reference ran it (phantom atom at the origin, one rank, V_eff = 0.333).

Contents: for every one of the 3366 nodes in KB mode the nodes/weights and
min/max; the density, Fermi level, band and -TS energy from them; for a sample
of nodes the Lanczos coefficients a,b in both nonlocal modes.  Run:
    python tests/golden/make_mg1_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import oracle as O  # noqa: E402

SAMPLE = np.array([0, 1, 10, 11, 197, 198, 1683, 1700, 2000, 3000, 3365], dtype=np.int32)


def main():
    S = O.System(os.path.join(HERE, "Mg1"), "Mg1")
    S.set_phantom(True, (0.0, 0.0, 0.0))
    S.build_nonlocal()
    I = S.info()
    nth = os.cpu_count() or 1
    lam, w, lmin, lmax = S.nodes_and_weights(mode=O.MODE_KB, nthreads=nth)
    lf, rho, eb, mts = O.density_fermi_energy(lam, w, I.beta, float(I.noetot), I.dV)
    out = dict(kb_lam=lam, kb_w=w, kb_lmin=lmin, kb_lmax=lmax, kb_rho=rho, kb_lf=lf, kb_eband=eb, kb_mts=mts,
               sample=SAMPLE)
    for mode, tag in ((O.MODE_REF_COLLAPSED, "ref"), (O.MODE_KB, "kb")):
        A, B = [], []
        for nid in SAMPLE:
            i, j, k = nid % 11, (nid // 11) % 18, nid // (11 * 18)
            a, b = S.lanczos((k, j, i), mode=mode)
            A.append(a)
            B.append(b)
        out[f"{tag}_a"] = np.array(A)
        out[f"{tag}_b"] = np.array(B)
    lam0, w0, lmin0, lmax0 = S.nodes_and_weights(mode=O.MODE_REF_COLLAPSED, node_ids=SAMPLE, nthreads=nth)
    out.update(ref_lam=lam0, ref_w=w0, ref_lmin=lmin0, ref_lmax=lmax0)
    np.savez_compressed(os.path.join(HERE, "mg1_oracle.npz"), **out)
    print("lambda_f", lf, "E_band", eb, "-TS", mts, "N", rho.sum() * I.dV)


if __name__ == "__main__":
    main()
