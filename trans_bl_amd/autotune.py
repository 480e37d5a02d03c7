"""Autotuner of the line smoother's shape.

For every multigrid level and both smoother precisions it times each valid
line shape -- 16, 32 or 64 lanes of a wavefront per y-line, with direct
(lane-contiguous) or LDS-staged (coalesced) global access per colour phase,
the fused-sweep strip form (whole sweep per launch in LDS), or the one-launch
out-of-place sweep (a strip of columns per workgroup, all colour phases) -- on the
context's own grid with hipEvents (tbl_time_smoother), and keeps the fastest.
Results are per size class (level shape), so contexts of the same grid reuse
them through the JSON cache.
"""
from __future__ import annotations

import json
import os

from . import TblError

# direct, LDS-staged (negative), x-terms-staged (2000 + lanes), fused-sweep
# strip (1000 + lanes) and one-launch out-of-place sweep (3000 + lanes; FP32
# only) forms; 1 = the coarsest level's sequential kernels (rejected on
# other levels; any lane form there selects the lane-scan coarsest kernel)
CANDIDATES = (1, 64, 32, 16, -64, -32, -16, 2064, 2032, 2016, 1064, 1032, 1016, 3064, 3032,
              3016)


def size_class(ctx, level):
    """Key of a level's shape: nx x ny x modes, with a trailing "c" on the
    coarsest level (whose solve takes other shapes: no strip form)."""
    nx, ny = ctx.grid.level_shape(level)
    tail = "c" if level == ctx.grid.levels else ""
    return f"{nx}x{ny}x{ctx.grid.nmodes}{tail}"


def tune(ctx, sweeps: int = 6, levels=None, precisions=(0, 1), verbose=False):
    """Pick the fastest shape per (level, precision).  Returns
    {size_class: {"fp64": lanes, "fp32": lanes, "us_per_sweep": {...}}}."""
    out = {}
    levels = levels or range(1, ctx.grid.levels + 1)
    for lvl in levels:
        entry = {"level": lvl, "us_per_sweep": {}}
        for prec in precisions:
            best, best_t = None, None
            orig = ctx.line_shape(prec, lvl)
            for lanes in CANDIDATES:
                try:
                    ctx.set_line_shape(prec, lvl, lanes)
                    ctx.time_smoother(prec, lvl, 1)
                    ms, _ = ctx.time_smoother(prec, lvl, sweeps)
                except TblError:
                    continue
                us = 1e3 * ms / sweeps
                entry["us_per_sweep"][f"{'fp32' if prec else 'fp64'}:{lanes}"] = round(us, 2)
                if best_t is None or us < best_t:
                    best, best_t = lanes, us
            ctx.set_line_shape(prec, lvl, best if best is not None else orig)
            entry["fp32" if prec else "fp64"] = best
        out[size_class(ctx, lvl)] = entry
        if verbose:
            print(json.dumps({size_class(ctx, lvl): entry}))
    return out


def apply(ctx, table):
    """Set the shapes of a previous tune() result on a context of the same grid."""
    for lvl in range(1, ctx.grid.levels + 1):
        e = table.get(size_class(ctx, lvl))
        if not e:
            continue
        for prec, key in ((0, "fp64"), (1, "fp32")):
            if e.get(key):
                ctx.set_line_shape(prec, lvl, int(e[key]))


def cached(ctx, path=None, **kw):
    """tune() with a JSON cache keyed by size class."""
    path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                                "linetune.json")
    table = {}
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
    missing = [l for l in range(1, ctx.grid.levels + 1) if size_class(ctx, l) not in table]
    if missing:
        table.update(tune(ctx, levels=missing, **kw))
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
        except OSError:
            pass
    apply(ctx, table)
    return {size_class(ctx, l): table[size_class(ctx, l)] for l in range(1, ctx.grid.levels + 1)}
