"""Kernel workload for hardware-counter runs of the hot path.

    rocprofv3 --pmc FETCH_SIZE --kernel-trace --output-format csv -d DIR -- \
        python3 -m trans_bl_amd.prof_kernels

Runs, on the 1025x193x64 v-Poisson workload:
  1. a plain device copy of 256 MiB (torch copy kernel) -- the calibration
     of the byte counters (its algorithmic traffic is exactly 256 MiB read and
     256 MiB written);
  2. two sweeps of the finest-level line smoother, FP64 then FP32;
  3. one flagship solve (FP32 smoothing under FP64 defect correction) and one
     plain FP64 solve, each from zero to 1e-9;
  4. the four derivatives of the solution (tbl_derivs_dev: one x and one y
     launch), twice;
  5. the x-line solves of loop on those fields, twice each: poi_ux
     (tbl_poi_ux_dev: the k = 0 pentadiagonal solve and the k = 1
     integration), poi_uz (tbl_poi_uz_dev) and filter_trid (tbl_filter_trid_dev,
     three fields in place).
Prints one JSON line per stage so the counter rows can be matched to stages
by dispatch order.

    python3 -m trans_bl_amd.prof_kernels --summarize FETCH_DIR WRITE_DIR OUT.json

reads the two counter runs' CSV output (rocprofv3 --output-format csv) and
writes, per kernel, the average FETCH_SIZE and WRITE_SIZE per dispatch, each
scaled so that the 256 MiB copy (the copy kernel's first dispatch) reads and
writes exactly 256 MiB, as HBM bytes per launch.  For the smoother kernels it
also writes the average over their first SWEEP_LAUNCHES dispatches alone
(`sweep_bytes_per_launch`): the plain sweeps of step 2 (with the line shapes
cached, nothing launches a smoother before them), the same launches that
`tbl_time_smoother` times; the solves' sweeps also include first sweeps from
zero (fewer reads) and sweeps that update the FP64 field (more writes).
Likewise for the derivative kernels, the average over their first
DERIV_LAUNCHES dispatches (`derivs_bytes_per_launch`): the two
tbl_derivs_dev calls of step 4, which write two derivatives per launch (the
x-line solves of step 5 launch the same kernels for one derivative).
"""
from __future__ import annotations

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

COPY_BYTES = 256 << 20
SWEEP_LAUNCHES = 8   # two sweeps of four launches (border + three colours)
DERIV_LAUNCHES = 2   # two tbl_derivs_dev calls: one x and one y launch each


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
        # the mixed solve in the benchmark's zero-start form (full-multigrid
        # first correction, three pre-sweeps, three coarsest sweeps)
        form = dict(fmg=1, nu1=3, ncoarse=3) if prec == 1 else {}
        cyc, res, st = ctx.solve_rhs_dev(f.data_ptr(), u.data_ptr(), 1e-9, precision=prec,
                                         max_cycles=40, guess=1, use_graph=0, **form)
        print(json.dumps({"stage": "solve", "precision": prec, "cycles_max": int(cyc.max()),
                          "residual_max": float(res.max())}), flush=True)
    outs = [torch.empty_like(u) for _ in range(4)]
    for _ in range(2):
        ctx.derivs_dev(u.data_ptr(), *[o.data_ptr() for o in outs])
    ctx.synchronize()
    print(json.dumps({"stage": "derivs", "launches": 4}), flush=True)
    uy, duydy, wy, d2 = outs
    ux, uz = torch.empty_like(u), torch.empty_like(u)
    for _ in range(2):
        ctx.poi_ux_dev(uy.data_ptr(), wy.data_ptr(), ux.data_ptr(), duydy.data_ptr())
        ctx.poi_uz_dev(wy.data_ptr(), duydy.data_ptr(), uz.data_ptr())
        ctx.filter_trid_dev(ux.data_ptr(), uz.data_ptr(), d2.data_ptr())
    ctx.synchronize()
    print(json.dumps({"stage": "xline", "calls": 6}), flush=True)
    ctx.close()


def library_sha256():
    """sha256 of the built HIP library (None when it is missing)."""
    import hashlib
    path = os.path.join(ROOT, "trans_bl_amd", "libtrans_bl_amd.so")
    try:
        with open(path, "rb") as f:
            return hashlib.sha256(f.read()).hexdigest()
    except OSError:
        return None


def _counter_rows(directory, counter):
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"),
                          recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return rows


def summarize(fetch_dir, write_dir, out):
    """Per-kernel HBM bytes per dispatch from the two counter runs."""
    fetch = _counter_rows(fetch_dir, "FETCH_SIZE")
    write = _counter_rows(write_dir, "WRITE_SIZE")
    copy_kib = COPY_BYTES / 1024.0

    def calib(rows):
        for _, name, v in rows:
            if "copyBuffer" in name or "copy" in name.lower() and "kernel" in name.lower():
                return copy_kib / v
        raise SystemExit("no copy dispatch found for the calibration")

    fs, ws = calib(fetch), calib(write)
    table = {}
    for rows, key, scale in ((fetch, "fetch", fs), (write, "write", ws)):
        for _, name, v in rows:
            e = table.setdefault(name, {"fetch_kib": [], "write_kib": []})
            e[key + "_kib"].append(v * scale)
    res = {"fetch_scale": fs, "write_scale": ws, "kernels": {},
           # the library the counters were read from (run --summarize on the
           # machine and tree of the counter runs): bench.py says whether its
           # own library is the same build
           "library_sha256": library_sha256()}
    for name, e in table.items():
        if not e["fetch_kib"] or not e["write_kib"]:
            continue
        f = sum(e["fetch_kib"]) / len(e["fetch_kib"])
        w = sum(e["write_kib"]) / len(e["write_kib"])
        res["kernels"][name] = {"dispatches": len(e["fetch_kib"]),
                                "fetch_bytes": round(f * 1024), "write_bytes": round(w * 1024),
                                "hbm_bytes_per_launch": round((f + w) * 1024)}
        for tag, key, n in (("smooth_", "sweep_bytes_per_launch", SWEEP_LAUNCHES),
                            ("deriv_", "derivs_bytes_per_launch", DERIV_LAUNCHES)):
            if tag in name and len(e["fetch_kib"]) >= n:
                fn = sum(e["fetch_kib"][:n]) / n
                wn = sum(e["write_kib"][:n]) / n
                res["kernels"][name][key] = round((fn + wn) * 1024)
    with open(out, "w") as fo:
        json.dump(res, fo, indent=1, sort_keys=True)
    print(json.dumps({"kernels": len(res["kernels"]), "fetch_scale": fs, "write_scale": ws}))


if __name__ == "__main__":
    if len(sys.argv) == 5 and sys.argv[1] == "--summarize":
        summarize(sys.argv[2], sys.argv[3], sys.argv[4])
    else:
        main()
