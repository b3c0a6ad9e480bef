"""SQ grid-nodes/s of the per-node Lanczos hot path on one MI355X.

One step = one SQ pass (N_qp+1 H applications + QL per node) over every grid
node of the 16-atom hcp Mg cell (25x21x40 = 21000 nodes, 29^3 cubes, 80
Lanczos steps, KB nonlocal projectors, V_eff = 0.333 as the reference),
inputs resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config C16]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "SQ grid-nodes/sec + SCF wall-time at <=1e-5 Ha/atom, chosen vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_node(nqp: int, nc: int, s: int) -> float:
    """m*3*n_c*s + 8*n_c: two vector reads and one write per Lanczos step, one
    V_eff cube read per node; projector tables are cache-resident."""
    return (nqp + 1) * 3 * nc * s + 8 * nc


def cpu_baseline(deck_dir, name, n_sample, threads):
    """The oracle (reference-faithful FP64 C restatement) on a bounded sample of nodes."""
    from oracle import oracle as O
    O.build()
    S = O.System(deck_dir, name)
    S.build_nonlocal()
    total = S.box[3] * S.box[4] * S.box[5]
    ids = np.linspace(0, total - 1, n_sample).astype(np.int32)
    t0 = time.perf_counter()
    S.nodes_and_weights(mode=O.MODE_KB, node_ids=ids, nthreads=threads)
    dt = time.perf_counter() - t0
    return dict(value=n_sample / dt, unit="grid-nodes/s", cores=threads, kind="port",
                sample=f"{n_sample} evenly spaced nodes of the same deck, KB mode, FP64, oracle/sq_oracle.c "
                       f"(per-node operator rebuild, AIJ SpMV, projector scan over all images, full-z QL), "
                       f"{dt:.1f} s on {threads} threads", host=_cpu_model())


def cpu_optimised(deck_dir, name, n_sample, threads):
    """The optimised CPU form of the same pass (oracle/sq_cpu_fast.c: matrix-free
    stencil, per-node image lists, AVX2/FMA, OpenMP), on a larger sample."""
    from oracle import oracle as O
    O.build()
    S = O.System(deck_dir, name)
    S.build_nonlocal()
    total = S.box[3] * S.box[4] * S.box[5]
    ids = np.linspace(0, total - 1, n_sample).astype(np.int32)
    t0 = time.perf_counter()
    S.fast_nodes_and_weights(node_ids=ids, nthreads=threads)
    dt = time.perf_counter() - t0
    return dict(value=n_sample / dt, unit="grid-nodes/s", cores=threads, kind="port",
                sample=f"{n_sample} evenly spaced nodes of the same deck, KB mode, FP64, oracle/sq_cpu_fast.c, "
                       f"{dt:.1f} s on {threads} threads", host=_cpu_model())


def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


KERNEL_NAMES = {0: "lanczos_kernel (generic)", 1: "lanczos_zmarch (column-marching)"}
KERKER_K0 = 1.0       # Kerker wavenumber, 1/bohr
KERKER_MIX = 0.7      # Pulay mixing parameter with the Kerker preconditioner
SCF_TOL = 1e-5        # the stopping tolerance of every timed form (the FP32 floor is about 5e-6)
SCF_TOL_REF = 1e-6    # the accuracy yardstick: plain FP64 converged further


def scf_wall_time(d, name, max_iter, mode, verbose=False):
    """SCF wall time on the periodic cell of the same deck.

    The yardstick is plain FP64 (the deck's mixing, superposed atoms) converged
    to a density residual of 1e-6; its energy and forces are what every other
    form is compared with (targets: 1e-5 Ha/atom, 1e-4 Ha/bohr).  Every timed
    form stops at the same residual, 1e-5: plain FP64, FP32, FP64 with Kerker
    mixing, the chosen form (FP32 with Kerker mixing) and the chosen form
    started from the learned initial density.  The speed-ups are wall-time
    ratios at that common tolerance."""
    import sqminiapp_amd as sq
    from sqminiapp_amd import train_density_net
    kb = mode == sq.SQ_NL_KB
    kerker = dict(kerker=KERKER_K0, mixing=KERKER_MIX)
    with sq.Context(d, name, device=0, cell=True) as c:
        I = c.info
        ref = c.scf(mode, sq.SQ_FP64, tol=SCF_TOL_REF, max_iter=max_iter, verbose=verbose)
        Fref, _, _, firef = c.forces() if kb else (None, None, None, None)
        r64 = c.scf(mode, sq.SQ_FP64, tol=SCF_TOL, max_iter=max_iter, verbose=verbose)
        r32 = c.scf(mode, sq.SQ_FP32, tol=SCF_TOL, max_iter=max_iter, verbose=verbose)
        F32, _, _, fi32 = c.forces() if kb else (None, None, None, None)
        # FP64 with Kerker mixing: the precision-only ratio of the chosen form
        rk = c.scf(mode, sq.SQ_FP64, tol=SCF_TOL, max_iter=max_iter, verbose=verbose, **kerker)
        # the chosen form: FP32 Lanczos, Kerker-preconditioned Pulay mixing, superposed atoms
        rc = c.scf(mode, sq.SQ_FP32, tol=SCF_TOL, max_iter=max_iter, verbose=verbose, **kerker)
        Fc = c.forces()[0] if kb else None
        # the same from the learned initial density (weights trained on other 16-atom cells)
        net = train_density_net.load() if kb else None
        rl = None
        if net is not None:
            t0 = time.perf_counter()
            guess = net.guess(c, bf16=True)
            t_guess = time.perf_counter() - t0
            rl = c.scf(mode, sq.SQ_FP32, tol=SCF_TOL, max_iter=max_iter, rho0=guess, verbose=verbose, **kerker)
            atomic = c.cell_field("rho_atomic")
            atomic = atomic * (I.noetot / (atomic.sum() * I.dV))
            rho = c.cell_field("rho")
            err = (float(np.linalg.norm(atomic - rho) / np.linalg.norm(rho)),
                   float(np.linalg.norm(guess - rho) / np.linalg.norm(rho)))
            net.close()
    keep = ("iterations", "converged", "residual", "wall_s", "sq_s", "poisson_s", "e_total", "lambda_f")
    de = lambda r: abs(r["e_total"] - ref["e_total"]) / I.nAtoms
    out = {"workload": f"periodic cell {I.rank_box[3]}x{I.rank_box[4]}x{I.rank_box[5]} = {I.n_nodes} nodes, "
                       f"BETA {I.beta:g}, LDA-PW, start: superposed atoms",
           "tol": SCF_TOL,
           "reference_fp64": {"tol": SCF_TOL_REF, "mixing": "deck (Periodic Pulay)", **{k: ref[k] for k in keep}},
           "plain_fp64": {"mixing": "deck (Periodic Pulay)", **{k: r64[k] for k in keep}, "d_energy_per_atom": de(r64)},
           "fp32_form": {"mixing": "deck (Periodic Pulay)", **{k: r32[k] for k in keep}, "d_energy_per_atom": de(r32)},
           "wall_speedup_fp32": r64["wall_s"] / r32["wall_s"] if r32["wall_s"] > 0 else None}
    if Fref is not None:
        out["forces"] = {"reference_fp64_s": firef["wall_s"], "fp32_s": fi32["wall_s"], "nodes": firef["nodes"],
                         "max_abs_reference": float(abs(Fref).max()),
                         "max_diff_fp32_vs_reference": float(abs(F32 - Fref).max()), "unit": "Ha/bohr"}
    kmix = f"Kerker-preconditioned Pulay (k0 = {KERKER_K0:g}/bohr, mixing {KERKER_MIX:g})"
    out["plain_fp64_kerker"] = {"mixing": kmix, **{k: rk[k] for k in keep}, "d_energy_per_atom": de(rk)}
    out["chosen_form"] = {"form": f"FP32 Lanczos, {kmix}, superposed atoms", **{k: rc[k] for k in keep},
                          "d_energy_per_atom": de(rc),
                          "max_force_diff_vs_reference": float(abs(Fc - Fref).max()) if Fc is not None else None,
                          "wall_speedup_vs_fp64": r64["wall_s"] / rc["wall_s"],
                          "wall_speedup_vs_fp64_kerker": rk["wall_s"] / rc["wall_s"]}
    if rl is not None:
        wall = rl["wall_s"] + t_guess
        out["chosen_from_learned_density"] = {
            "form": "the chosen form from the learned initial density (BF16 network inference)",
            **{k: rl[k] for k in keep}, "guess_s": t_guess, "wall_with_guess_s": wall,
            "initial_error_atoms": err[0], "initial_error_network": err[1], "d_energy_per_atom": de(rl),
            "wall_speedup_vs_fp64": r64["wall_s"] / wall}
    return out


def load_traffic(tag):
    path = os.path.join(REPO, "profiles", f"{tag}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C16")
    ap.add_argument("--mode", choices=["kb", "ref"], default="kb")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=512)
    ap.add_argument("--fp32-steps", type=int, default=None)
    ap.add_argument("--no-scf", action="store_true", help="skip the SCF wall-time measurement")
    ap.add_argument("--scf-iter", type=int, default=60)
    ap.add_argument("--scf-config", default=None, help="deck of the SCF measurement (default: --config)")
    ap.add_argument("--scf-verbose", action="store_true", help="one line per SCF iteration on stderr")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the SQ node loop would shard across GPUs by 3-D node blocks "
              "with no collective inside a pass, and this build runs and reports the one-GPU share.", file=sys.stderr)
        sys.exit(2)

    import sqminiapp_amd as sq
    from sqminiapp_amd import decks

    root = tempfile.mkdtemp(prefix="sq_bench_")
    d, name = decks.deck(args.config, root)
    mode = sq.SQ_NL_KB if args.mode == "kb" else sq.SQ_NL_REF_COLLAPSED
    ctx = sq.Context(d, name, device=0)
    I = ctx.info
    n_nodes = I.n_nodes

    def timed(precision, steps, warmup):
        for _ in range(warmup):
            ctx.run(mode, precision)
        ctx.synchronize()
        t0 = time.perf_counter()
        kernel_ms = []
        for _ in range(steps):
            ctx.run(mode, precision)
            kernel_ms.append(ctx.timing())
        ctx.synchronize()
        dt = time.perf_counter() - t0
        return dt, np.array(kernel_ms)

    dt64, k64 = timed(sq.SQ_FP64, args.steps, args.warmup)
    lf, rho, eb, mts = ctx.density_fermi_energy()
    fp32_steps = args.steps if args.fp32_steps is None else args.fp32_steps
    dt32, k32 = timed(sq.SQ_FP32, fp32_steps, 1) if fp32_steps > 0 else (None, None)

    value = n_nodes * args.steps / dt64
    lanczos_ms = float(k64[:, 0].mean())
    bytes_launch = algorithmic_bytes_per_node(I.nqp, I.nc, 8) * n_nodes
    achieved = bytes_launch / (lanczos_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(f"lanczos_{args.config.lower()}_fp64"),
                "kernel": KERNEL_NAMES[I.kernel], "kernel_ms": round(lanczos_ms, 3),
                "algorithmic_bytes_per_launch": bytes_launch}
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "grid-nodes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt64 * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.config}: {I.nAtoms}-atom hcp Mg cell, grid {I.npts[0]}x{I.npts[1]}x"
                               f"{I.npts[2]} = {n_nodes} nodes, {I.nloc[0]}^3 cubes, {I.nqp} Lanczos steps, "
                               f"{args.mode} nonlocal, V_eff 0.333",
                   "nodes": n_nodes, "nqp": I.nqp, "cube": I.nloc[0], "form": "plain FP64",
                   "workgroup": I.workgroup[0], "workgroup_fp32": I.workgroup[1],
                   "autotune_ms": [[round(x, 3) for x in row] for row in I.tune_ms]},
        "roofline": roofline,
        "ql_ms": round(float(k64[:, 1].mean()), 3),
        "fermi_level": lf, "band_energy": eb, "minus_ts": mts,
    }
    if dt32 is not None:
        v32 = n_nodes * fp32_steps / dt32
        b32 = algorithmic_bytes_per_node(I.nqp, I.nc, 4) * n_nodes
        out["fp32_form"] = {"value": round(v32, 1), "unit": "grid-nodes/s", "speedup_vs_fp64": round(v32 / value, 3),
                            "lanczos_ms": round(float(k32[:, 0].mean()), 3),
                            "achieved_GBs": round(b32 / (float(k32[:, 0].mean()) * 1e-3) / 1e9, 1)}
    if not args.no_scf:
        if args.scf_config and args.scf_config != args.config:
            ds, ns = decks.deck(args.scf_config, root)
        else:
            ds, ns = d, name
        out["scf"] = scf_wall_time(ds, ns, args.scf_iter, mode, args.scf_verbose)
    if not args.no_cpu_baseline:
        threads = max(1, min(16, os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline(d, name, args.cpu_sample, threads)
        out["cpu_optimised"] = cpu_optimised(d, name, 8 * args.cpu_sample, threads)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
