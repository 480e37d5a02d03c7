"""RK4-stage benchmarks of the hot path beyond the headline metric.

    python -m trans_bl_amd.bench_stage --config uneven32   # 32 uneven problems
    python -m trans_bl_amd.bench_stage --config large      # 2049x257x128, evolving RHS

uneven32: 32 independent problems with nx drawn from {257, 513, 1025, 2049},
ny from {65, 129, 193, 257} and mode counts from {16, 32, 64, 128} (seeded),
wavepacket right-hand sides; one stage (d/dx, d2/dx2, d/dy, d2/dy2 of v and
the v-Poisson solve of every mode, warm-started from v) of all problems is
captured as one hipGraph and replayed.  Line shapes come from the autotuner
per size class.

large: one 2049x257x128 problem (dy0=2e-4, stretch 1.012, kz = 0.05k) over
consecutive stages whose wavepackets drift downstream by 0.4 per stage; each
stage solves from the previous stage's v (the reference's warm start) or from
the learned guess.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import trans_bl_amd as tb  # noqa: E402
from trans_bl_amd import autotune, batch as B, workload as wl  # noqa: E402
from trans_bl_amd.train_guess import packet_params, rhs_torch  # noqa: E402


def uneven_grids(n=32, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(wl.Grid(int(rng.choice([257, 513, 1025, 2049])),
                           int(rng.choice([65, 129, 193, 257])),
                           int(rng.choice([16, 32, 64, 128])), 4, 0.02, 2e-4, 1.015, 0.0, 0.1))
    return out


def run_uneven(a, torch):
    grids = uneven_grids()
    bt = B.Batch(grids)
    tuned = {}
    for ctx in bt.contexts:
        tuned.update(autotune.cached(ctx, sweeps=3))
    fs, vs, ds = [], [], []
    for p, g in enumerate(grids):
        fs.append(rhs_torch(torch, g, packet_params(g, 7000 + p), 0.0))
        vs.append(torch.zeros_like(fs[-1]))
        ds.extend(torch.zeros_like(fs[-1]) for _ in range(4))
    torch.cuda.synchronize()
    fp = [t.data_ptr() for t in fs]
    vp = [t.data_ptr() for t in vs]
    dp = [t.data_ptr() for t in ds]
    # --converge: solved to the tolerance (tbl_batch_stage_converge), with
    # --form fmg the full-multigrid first correction and three pre-sweeps in
    # FP32 (the coarsest sweeps by the default rule: the uneven batch's
    # ny = 257 problems need more than three); else a.cycles fixed V-cycles
    # (tbl_batch_stage)
    form = dict(fmg=1, nu1=3) if a.precision == 1 and a.form == "fmg" else {}
    times, runs = [], []
    for it in range(a.warmup + a.steps):
        for v in vs:
            v.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if a.converge:
            ms, run = bt.stage_converge(fp, vp, dp, a.tol, max_cycles=a.cycles,
                                        precision=a.precision, **form)
            runs.append(run)
        else:
            ms = bt.stage(fp, vp, dp, a.tol, a.cycles, a.precision, 1)
        wall = time.perf_counter() - t0
        if it >= a.warmup:
            times.append((ms, wall))
    cyc = [bt.results(p)[0] for p in range(len(grids))]
    res = [bt.results(p)[1] for p in range(len(grids))]
    modes = sum(g.nmodes for g in grids)
    points = sum(g.nx * g.ny * g.nmodes for g in grids)
    ms = float(np.mean([t[0] for t in times]))
    out = {"config": "uneven32", "problems": len(grids), "modes": modes, "points": points,
           "stage_ms": round(ms, 3), "wall_ms": round(1e3 * float(np.mean([t[1] for t in times])), 3),
           "mode_solves_per_s": round(modes / (ms / 1e3), 1),
           "converged_modes": int(sum(int(np.sum(r < a.tol)) for r in res)),
           "cycles_max": int(max(c.max() for c in cyc)),
           "precision": a.precision,
           "cycles": ({"mode": "to the tolerance", "cap": a.cycles, "run_per_call": runs,
                       "form": form} if a.converge else {"mode": "fixed", "count": a.cycles}),
           "line_shapes": {k: {"fp64": v.get("fp64"), "fp32": v.get("fp32")}
                           for k, v in sorted(tuned.items())},
           "grids": [f"{g.nx}x{g.ny}x{g.nmodes}" for g in grids]}
    bt.close()
    return out


def run_large(a, torch):
    g = wl.LARGE_2049x257x128
    ctx = tb.Context(g)
    tuned = autotune.cached(ctx, sweeps=3)
    guess = None
    if a.learned:
        from trans_bl_amd import guess as G
        guess = G.load_default(ctx)
    params = packet_params(g, 9000)
    v = torch.zeros((g.nmodes, g.ny, g.nx), dtype=torch.complex128, device="cuda")
    ds = [torch.zeros_like(v) for _ in range(4)]
    stats, t_all = [], []
    for s in range(a.stages):
        f = rhs_torch(torch, g, params, 0.4 * s)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for q in range(4):
            ctx.deriv_dev(q, v.data_ptr(), ds[q].data_ptr())
        ctx.synchronize()
        kw = dict(precision=a.precision, max_cycles=40, use_graph=1)
        if a.precision == 1 and a.form == "fmg":
            kw.update(fmg=1, nu1=3)
        if guess is not None:
            cyc, res, st = ctx.solve_rhs_dev(f.data_ptr(), v.data_ptr(), a.tol, guess=2, **kw)
        else:
            cyc, res, st = ctx.solve_rhs_dev(f.data_ptr(), v.data_ptr(), a.tol, guess=0, **kw)
        t_all.append(time.perf_counter() - t0)
        stats.append((cyc.mean(), res.max(), st["fallback"]))
        if s % 20 == 0:
            print(json.dumps({"stage": s, "cycles_mean": float(cyc.mean()),
                              "residual_max": float(res.max())}), flush=True)
    ms = 1e3 * float(np.mean(t_all[1:]))
    return {"config": "large", "grid": "2049x257x128", "stages": a.stages,
            "stage_ms": round(ms, 3), "mode_solves_per_s": round(g.nmodes / (ms / 1e3), 1),
            "cycles_mean": float(np.mean([s[0] for s in stats])),
            "residual_max": float(max(s[1] for s in stats)),
            "fallback_total": int(sum(s[2] for s in stats)),
            "initial_guess": "learned" if guess is not None else "previous stage",
            "precision": a.precision, "form": a.form,
            "line_shapes": {k: {"fp64": v.get("fp64"), "fp32": v.get("fp32")}
                            for k, v in sorted(tuned.items())}}


def main():
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", choices=["uneven32", "large"], default="uneven32")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed stages (default 1; 2 with --converge, whose second "
                         "call re-captures its graph for the cycles the first needed)")
    ap.add_argument("--cycles", type=int, default=12)
    ap.add_argument("--stages", type=int, default=200)
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--learned", action="store_true")
    ap.add_argument("--converge", action="store_true",
                    help="uneven32: solve to the tolerance (tbl_batch_stage_converge)")
    ap.add_argument("--form", choices=["default", "fmg"], default="default",
                    help="the solver form (uneven32: with --converge; fmg: full-multigrid "
                         "first correction, three pre-sweeps)")
    a = ap.parse_args()
    if a.warmup is None:
        a.warmup = 2 if a.converge else 1
    out = run_uneven(a, torch) if a.config == "uneven32" else run_large(a, torch)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
