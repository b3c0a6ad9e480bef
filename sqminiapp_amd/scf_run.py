"""Runs one SCF on a synthetic deck and prints its iterations and a JSON summary.

    python -m sqminiapp_amd.scf_run [--config C128] [--precision fp64|fp32] [--mixing deck|kerker]
                                    [--max-iter 30] [--tol 1e-5] [--start atoms|network]

--mixing deck is the deck's Periodic Pulay (MIXING_PARAMETER, MIXING_HISTORY,
PULAY_FREQUENCY; the reference's defaults 0.3 / 7 / 1); kerker is the
Kerker-preconditioned Pulay at k0 = 1/bohr and mixing 0.7.  One line per
iteration goes to stderr (the library's verbose output), the summary to stdout.
"""
from __future__ import annotations

import argparse
import json
import tempfile
import time

from . import decks
from ._lib import SQ_FP32, SQ_FP64, SQ_NL_KB, Context


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="C128")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--mixing", choices=["deck", "kerker"], default="deck")
    ap.add_argument("--max-iter", type=int, default=30)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--start", choices=["atoms", "network"], default="atoms")
    a = ap.parse_args()
    d, name = decks.deck(a.config, tempfile.mkdtemp(prefix="sq_scf_"))
    prec = SQ_FP64 if a.precision == "fp64" else SQ_FP32
    kw = dict(kerker=1.0, mixing=0.7) if a.mixing == "kerker" else {}
    with Context(d, name, cell=True) as c:
        rho0 = None
        if a.start == "network":
            from . import train_density_net
            net = train_density_net.load()
            if net is None:
                raise SystemExit("no network weights")
            rho0 = net.guess(c, bf16=True)
            net.close()
        t = time.time()
        r = c.scf(SQ_NL_KB, prec, tol=a.tol, max_iter=a.max_iter, rho0=rho0, verbose=True, **kw)
        r["wall_clock_s"] = time.time() - t
        r.update(config=a.config, precision=a.precision, mixing=a.mixing, start=a.start, nodes=c.info.n_nodes)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
