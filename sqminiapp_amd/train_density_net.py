"""Trains the initial-density network on converged SQ densities of 16- and
128-atom Mg cells and writes its weights.

    python -m sqminiapp_amd.train_density_net [--c16 24] [--c128 4] [--epochs 2000] [--out FILE]
    python -m sqminiapp_amd.train_density_net --evaluate C16,C128 [--out FILE]

Training cells: the 16-atom hcp Mg cell and the 128-atom cell with its 4-site
vacancy cluster (sqminiapp_amd.decks: h = 0.5, 29^3 cubes, 80 Lanczos steps,
kT = 10 mHa), each with its own seeded displacements, sigma spread over
0.1-0.2 bohr: seeds 101, 102, ... (16 atoms) and 201, 202, ... (128 atoms),
none of them the cells the bench and the tests hold out (16-atom seed 16,
128-atom seed 128).  Each is converged on the GPU by the chosen SCF form
(FP32 Lanczos, Kerker mixing) to a density residual of 1e-5; every stride-th
cell point's descriptors and rho - rho_atomic join the training set
(sq_mlp_add_sample_every); the network is then trained by the library's own
forward, backward and Adam kernels (sq_mlp_train) and saved as float32 arrays
"params" and "norm" plus a record of the training set (numpy .npz, no
pickles).
"""
from __future__ import annotations

import argparse
import dataclasses
import os
import tempfile
import time

import numpy as np

from . import decks
from ._lib import SQ_FP32, SQ_NL_KB, Context, Mlp

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "density_net.npz")


def _cells(n, base, seed0):
    for i in range(n):
        yield dataclasses.replace(decks.CONFIGS[base], seed=seed0 + i, sigma=0.1 + 0.1 * i / max(1, n - 1))


def train(c16=24, c128=4, epochs=2000, lr=3e-3, out=DEFAULT_OUT, seed=0, stride16=2, stride128=4, verbose=True):
    root = tempfile.mkdtemp(prefix="sq_train_")
    t_all = time.time()
    iters = []
    with Mlp(seed=seed) as m:
        for cfg, stride in [(c, stride16) for c in _cells(c16, "C16", 101)] + \
                           [(c, stride128) for c in _cells(c128, "C128", 201)]:
            d = os.path.join(root, f"{cfg.name}_{cfg.seed}")
            decks.write_deck(cfg, d)
            t = time.time()
            with Context(d, cfg.name, cell=True) as c:
                r = c.scf(SQ_NL_KB, SQ_FP32, tol=1e-5, max_iter=80, kerker=1.0, mixing=0.7)
                if not r["converged"]:
                    raise RuntimeError(f"training cell {cfg.name} seed {cfg.seed} did not converge: {r['residual']}")
                m.add_sample(c, stride=stride)
            iters.append(r["iterations"])
            if verbose:
                print(f"{cfg.name} seed {cfg.seed} sigma {cfg.sigma:.3f}: SCF {r['iterations']} iterations, "
                      f"{time.time() - t:.1f} s", flush=True)
        t = time.time()
        loss0 = m.train(0)
        loss = m.train(epochs, lr=lr)
        if verbose:
            print(f"trained {epochs} epochs in {time.time() - t:.1f} s: loss {loss0:.4f} -> {loss:.4f}", flush=True)
        p, norm = m.params()
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    record = np.array([c16, c128, stride16, stride128, epochs, lr, loss0, loss, time.time() - t_all], np.float64)
    np.savez(out, params=p, norm=norm, record=record, scf_iterations=np.array(iters, np.int32))
    return out, loss0, loss


def load(path=DEFAULT_OUT, device=0):
    """An Mlp with the saved weights, or None when the file is absent or was
    written for another network shape."""
    if not os.path.exists(path):
        return None
    z = np.load(path, allow_pickle=False)
    if z["params"].size != Mlp.PARAMS or z["norm"].size != Mlp.NORM:
        return None
    m = Mlp(device=device)
    m.set_params(z["params"], z["norm"])
    return m


def evaluate(cells=("C16", "C128"), path=DEFAULT_OUT, verbose=True):
    """The chosen SCF form on held-out cells from the superposed atoms and from
    the network's BF16 guess: iterations, wall time, initial errors, energy gap."""
    root = tempfile.mkdtemp(prefix="sq_eval_")
    out = []
    for name in cells:
        d, nm = decks.deck(name, root)
        net = load(path)
        if net is None:
            raise RuntimeError(f"no network weights at {path}")
        with Context(d, nm, cell=True) as c:
            kw = dict(tol=1e-5, max_iter=80, kerker=1.0, mixing=0.7)
            ra = c.scf(SQ_NL_KB, SQ_FP32, **kw)
            rho = c.cell_field("rho")
            atomic = c.cell_field("rho_atomic")
            atomic = atomic * (c.info.noetot / (atomic.sum() * c.info.dV))
            t0 = time.time()
            g = net.guess(c, bf16=True)
            tg = time.time() - t0
            rl = c.scf(SQ_NL_KB, SQ_FP32, rho0=g, **kw)
            norm = np.linalg.norm(rho)
            row = dict(cell=name, atoms_iterations=ra["iterations"], network_iterations=rl["iterations"],
                       atoms_wall_s=ra["wall_s"], network_wall_s=rl["wall_s"], guess_s=tg,
                       error_atoms=float(np.linalg.norm(atomic - rho) / norm),
                       error_network=float(np.linalg.norm(g - rho) / norm),
                       d_energy_per_atom=abs(rl["e_total"] - ra["e_total"]) / c.info.nAtoms)
        net.close()
        out.append(row)
        if verbose:
            print(row, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--c16", type=int, default=24)
    ap.add_argument("--c128", type=int, default=4)
    ap.add_argument("--stride16", type=int, default=2)
    ap.add_argument("--stride128", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--evaluate", default=None, help="held-out cells to evaluate the saved weights on (no training)")
    a = ap.parse_args()
    if a.evaluate:
        evaluate(a.evaluate.split(","), a.out)
        return
    path, l0, l1 = train(a.c16, a.c128, a.epochs, a.lr, a.out, stride16=a.stride16, stride128=a.stride128)
    print(f"wrote {path} (loss {l0:.4f} -> {l1:.4f})")


if __name__ == "__main__":
    main()
