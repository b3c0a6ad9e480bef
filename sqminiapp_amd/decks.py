"""Synthetic hcp-Mg supercell decks in SQMiniApp's input format.

A deck is <dir>/<name>.inpt + <dir>/<name>.ion + <dir>/pseudopotential/<pot>,
the layout Read_Parameters / Read_Ion / Read_Pseudopotential read
(lsdft_readfiles.cc:21-1101).  Grid convention of the reference (BC=1):
CELL holds half-lengths R, the N grid points span [-R, R] with h = 2R/(N-1),
and periodic images sit at +-2R (lsdft_readfiles.cc:356-360,
lsdft_initObjs.cc:1529-1531).  N_d = floor(L_d/h_target)+1 keeps h_d >= 0.5,
so a 7 bohr truncation radius gives the 29-point cube edge.

The 4-atom orthorhombic hcp cell (a, a*sqrt(3), c) holds atoms at fractions
(0,0,0), (1/2,1/2,0), (1/2,1/6,1/2), (0,2/3,1/2) -- the cell of the shipped
test/Mg1 deck.  Displacements are Gaussian (numpy default_rng(seed)).
"""
from __future__ import annotations

import math
import os
import shutil
from dataclasses import dataclass, field

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MG_POT = os.path.join(REPO, "tests", "golden", "Mg1", "pseudopotential", "psd_Mg.pot")

HCP_FRAC = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 1.0 / 6.0, 0.5], [0.0, 2.0 / 3.0, 0.5]])


@dataclass
class Supercell:
    name: str
    reps: tuple
    a: float = 6.06
    c: float = 9.83
    sigma: float = 0.1
    seed: int = 0
    h_target: float = 0.5
    rcut: float = 7.0
    nqp: int = 80
    # kT = 10 mHa: with 80 Gauss nodes per node the quadrature resolves the
    # Fermi function only to ~ (spectral width)/80, so at kT = 1 mHa the SQ
    # density jumps as the Fermi level moves and the SCF fixed point is not
    # reached (measured on the 16-atom cell); kT = 10 mHa converges in ~25 steps
    beta: float = 100.0
    vacancy_cluster: int = 0          # remove the atom nearest the centre and its n-1 nearest neighbours
    extra: dict = field(default_factory=dict)

    def lengths(self):
        return np.array([self.a * self.reps[0], self.a * math.sqrt(3.0) * self.reps[1], self.c * self.reps[2]])

    def grid(self):
        L = self.lengths()
        return [int(math.floor(L[d] / self.h_target + 1e-12)) + 1 for d in range(3)]

    def positions(self):
        L = self.lengths()
        cell = np.array([self.a, self.a * math.sqrt(3.0), self.c])
        pos = []
        for i in range(self.reps[0]):
            for j in range(self.reps[1]):
                for k in range(self.reps[2]):
                    for f in HCP_FRAC:
                        pos.append((f + np.array([i, j, k])) * cell - L / 2.0)
        pos = np.array(pos)
        if self.vacancy_cluster > 0:
            d0 = np.linalg.norm(pos, axis=1)
            c0 = int(np.argmin(d0))
            d = np.linalg.norm(pos - pos[c0], axis=1)
            drop = np.argsort(d, kind="stable")[: self.vacancy_cluster]
            pos = np.delete(pos, drop, axis=0)
        rng = np.random.default_rng(self.seed)
        pos = pos + rng.normal(0.0, self.sigma, size=pos.shape)
        return pos


CONFIGS = {
    # 16 atoms, ~21k nodes, 29^3 cubes, 80 Lanczos steps
    "C16": Supercell("C16", (2, 1, 2), sigma=0.1, seed=16),
    # 128 atoms minus a 4-site vacancy cluster, ~163k nodes
    "C128": Supercell("C128", (4, 2, 4), sigma=0.15, seed=128, vacancy_cluster=4),
    # 1024 atoms, ~1.29M nodes
    "C1024": Supercell("C1024", (8, 4, 8), sigma=0.1, seed=1024),
}


def write_deck(cfg: Supercell, directory: str, pot: str = MG_POT, nqp: int | None = None) -> str:
    """Write <directory>/<cfg.name>.{inpt,ion} and the pseudopotential; returns directory."""
    os.makedirs(os.path.join(directory, "pseudopotential"), exist_ok=True)
    L = cfg.lengths()
    N = cfg.grid()
    pos = cfg.positions()
    with open(os.path.join(directory, cfg.name + ".inpt"), "w") as f:
        f.write("FD_ORDER: 12\n")
        f.write("CELL: %.16f %.16f %.16f\n" % tuple(L / 2.0))
        f.write("FD_GRID: %d %d %d\n" % tuple(N))
        f.write("BOUNDARY_CONDITION: 1\n")
        f.write("BETA: %g\n" % cfg.beta)
        f.write("LANCZOS_DEGREE: %d\n" % (cfg.nqp if nqp is None else nqp))
        f.write("NTYPE: 1\n")
        f.write("RADIUS_TRUNC: %g\n" % cfg.rcut)
        f.write("PSEUDOPOTENTIAL_LOCAL: 1\n")
        f.write("PSEUDOPOTENTIAL_FILE: psd_Mg.pot\n")
    with open(os.path.join(directory, cfg.name + ".ion"), "w") as f:
        f.write("%d\n\n" % len(pos))
        f.write("Mg 2 %d\n" % len(pos))
        for p in pos:
            f.write("%.16f %.16f %.16f\n" % tuple(p))
    dst = os.path.join(directory, "pseudopotential", "psd_Mg.pot")
    if not os.path.exists(dst):
        shutil.copyfile(pot, dst)
    return directory


def deck(name: str, root: str, nqp: int | None = None) -> tuple:
    """(directory, name) of config `name` written under root."""
    cfg = CONFIGS[name]
    d = os.path.join(root, name if nqp is None else f"{name}_q{nqp}")
    write_deck(cfg, d, nqp=nqp)
    return d, cfg.name
