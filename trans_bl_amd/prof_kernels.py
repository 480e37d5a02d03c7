"""Kernel workload for hardware-counter runs of the hot path.

    rocprofv3 --pmc FETCH_SIZE --kernel-trace --output-format csv -d DIR -- \
        python3 -m trans_bl_amd.prof_kernels

Runs, on the 1025x193x64 v-Poisson workload:
  1. a plain device copy of 256 MiB (torch copy kernel) -- the calibration
     of the byte counters (its algorithmic traffic is exactly 256 MiB read and
     256 MiB written);
  2. two sweeps of the finest-level line smoother, FP64 then FP32;
  3. one flagship solve (FP32 smoothing under FP64 defect correction) and one
     plain FP64 solve, each from zero to 1e-9.
Prints one JSON line per stage so the counter rows can be matched to stages
by dispatch order.
"""
from __future__ import annotations

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

COPY_BYTES = 256 << 20


def main():
    import torch
    import trans_bl_amd as tb
    from trans_bl_amd import autotune, workload as wl

    grid = wl.BATCH_1025x193x64
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"stage": "copy", "bytes": COPY_BYTES}), flush=True)

    ctx = tb.Context(grid)
    autotune.cached(ctx, sweeps=3)
    for prec in (0, 1):
        ms, launches = ctx.time_smoother(prec, 1, 2)
        print(json.dumps({"stage": "smoother", "precision": prec, "launches": launches,
                          "ms": ms, "lanes": ctx.line_shape(prec, 1)}), flush=True)
    f = torch.from_numpy(wl.wavepacket_rhs(grid, tag="bench")).to("cuda")
    u = torch.zeros_like(f)
    for prec in (1, 0):
        u.zero_()
        torch.cuda.synchronize()
        cyc, res, st = ctx.solve_rhs_dev(f.data_ptr(), u.data_ptr(), 1e-9, precision=prec,
                                         max_cycles=40, guess=1, use_graph=0)
        print(json.dumps({"stage": "solve", "precision": prec, "cycles_max": int(cyc.max()),
                          "residual_max": float(res.max())}), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
