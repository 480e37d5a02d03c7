#!/usr/bin/env python3
"""Benchmark: v-Poisson mode-solves/s to a 1e-9 residual on one MI355X.

Workload (the batched configuration the metric is quoted on): nx=1025 (dx=0.02)
x ny=193 (dy0=2e-4, stretch 1.015), 64 spanwise modes kz=0.1k, 4 multigrid
levels, synthetic wavepacket right-hand sides (6 packets per mode and part).
One step = one solve of all 64 modes from the initial guess to
max|S - L v| < 1e-9 (source assembly, cycles and copy-out included), with
inputs resident in HBM.

Paths: plain FP64 FAS multigrid (the reference's cycle with the GPU's colour
line ordering), and FP32 line smoothing under FP64 defect correction started
from zero, from the learned initial guess (when a trained network is
present, trans_bl_amd/weights/guess_*.bin) and from the previous stage's
solution (packets 0.4 upstream: the reference's warm start, reported beside);
the mixed solves correct first by full multigrid and then by V-cycles, with
three pre-sweeps per level and three coarsest-level sweeps.
`value` is the whole-job mode-solves/s of the mixed path from zero, or from
the learned guess when that is faster by more than 1 %; the plain FP64 path
is reported beside it.

Prints ONE JSON line.  Also reports the smoother's HBM roofline (live hipEvent
timing of the level-1 smoother kernel), the four derivatives' roofline, and a
CPU baseline: the reference's own Fortran poi_uy (compiled from its sources
into oracle/_ref) on a bounded sample of the workload, one process per host
core, with the CPU oracle timed on the same sample for calibration (the
oracle alone when the reference driver is absent).
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the mixed solve (every start): the first correction by full multigrid,
# three pre-sweeps and one post-sweep per level, three coarsest-level sweeps
MIXED_FORM = dict(fmg=1, nu1=3, ncoarse=3)
# per-kernel HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter runs
# (rocprofv3 --pmc ... -- python3 -m trans_bl_amd.prof_kernels, summarized by
# trans_bl_amd.prof_kernels --summarize)
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "hbm_traffic_1025x193x64.json")
ROWS = (1, 2, 3, 4, 5, 6, 7, 9, 13, 17)


def smoother_kernel(prec, lanes, ny):
    """Name of the smoother kernel a line shape runs in a plain sweep (as
    rocprofv3 prints it): G direct, -G staged (x terms, v, S in LDS), 2000+G
    staged (x terms only), 1000+G fused-sweep strip.  The staged kernel's
    last template argument selects its updating form, which plain sweeps do
    not use.  A line's lanes hold its rows 1..ny-1 (row 0 is not relaxed)."""
    g = lanes - 2000 if lanes > 2000 else lanes - 1000 if lanes > 1000 else abs(lanes)
    need = -(-(ny - 1) // g)
    r = next(x for x in ROWS if x >= need)
    t = "float" if prec else "double"
    if lanes > 2000:
        return f"smooth_staged_kernel<{t}, {g}, {r}, false, false>"
    if lanes > 1000:
        return f"smooth_strip_kernel<{t}, {g}, {r}>"
    if lanes < 0:
        return f"smooth_staged_kernel<{t}, {g}, {r}, true, false>"
    return f"smooth_phase_kernel<{t}, {g}, {r}>"


def measured_traffic(kernel):
    try:
        with open(TRAFFIC_JSON) as f:
            table = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    for name, e in table.items():
        if kernel + "(" in name:
            # the bytes of the launches the roofline times (a smoother's plain
            # sweeps, a derivative kernel's two-derivative launches) when
            # recorded, else the average over all of the kernel's launches
            return e.get("sweep_bytes_per_launch",
                         e.get("derivs_bytes_per_launch", e["hbm_bytes_per_launch"]))
    return None


def traffic_source():
    """Where `traffic` comes from: the counter summary and whether it was
    read from this same library build (its sha256 recorded at profiling)."""
    import hashlib
    try:
        with open(TRAFFIC_JSON) as f:
            rec = json.load(f).get("library_sha256")
        with open(os.path.join(ROOT, "trans_bl_amd", "libtrans_bl_amd.so"), "rb") as f:
            cur = hashlib.sha256(f.read()).hexdigest()
    except (OSError, ValueError):
        return None
    return {"file": os.path.relpath(TRAFFIC_JSON, ROOT),
            "counters": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of trans_bl_amd.prof_kernels",
            "same_library_build": rec is not None and rec == cur}
METRIC = "v-Poisson mode-solves/s to 1e-9 residual; learned+MG vs plain MG on 1 GPU"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-modes", type=int, default=64, help="modes in the CPU sample")
    ap.add_argument("--no-guess", action="store_true", help="ignore a trained network")
    ap.add_argument("--no-rk4", action="store_true", help="skip the RK4-step measurement")
    return ap.parse_args()


def cpu_worker(args):
    import oracle
    g, f = args
    p = oracle.make_params(g.nx, g.ny, g.nmodes, g.levels, g.dx, g.dy0, g.stf, g.alpha, g.beta)
    z = np.zeros_like(f)
    t = time.perf_counter()
    _, cyc, res = oracle.poi_uy(p, z, -f, z, 1e-9, 11, 0)
    return time.perf_counter() - t, cyc.tolist(), res.tolist()


def cpu_baseline(grid, f, ncpu):
    """The reference's poi_uy on modes k=0..7 of the workload (its driver
    built for the same grid with 8 modes), `ncpu` processes at once; the CPU
    oracle (port, reference line order, FP64) on the same 8 modes in one
    process for the calibration ratio.  None when the driver is absent."""
    import dataclasses
    from oracle import reference as R
    name = "1025x193x8"
    if not (os.path.exists(R.exe_path(name)) and os.path.exists(R.coefs_path(name))):
        print("reference driver for the CPU baseline not built", file=sys.stderr)
        return None
    f8 = np.ascontiguousarray(f[:8])
    z = np.zeros_like(f8)
    try:
        wall, secs, cyc = R.run_parallel(name, z, -f8, z, 1e-9, ncpu)
    except Exception as e:  # noqa: BLE001 - report, fall back to the port
        print(f"reference driver failed: {e}", file=sys.stderr)
        return None
    g8 = dataclasses.replace(grid, nmodes=8)
    t_port, cyc_port, _ = cpu_worker((g8, f8))
    ref_one = float(np.median(secs))
    return {"value": round(8 * ncpu / wall, 4), "unit": "mode-solves/s", "cores": ncpu,
            "kind": "reference",
            "sample": f"modes k=0..7 of the same workload (1025x193, kz=0.1k) from zero to "
                      f"1e-9 by the reference's own poi_uy (Fortran, compiled from its sources, "
                      f"one MPI rank), {ncpu} processes at once on {ncpu} cores; {wall:.1f} s",
            "cycles": cyc.tolist(),
            "calibration": {"reference_s_per_8_modes": round(ref_one, 3),
                            "oracle_s_per_8_modes": round(t_port, 3),
                            "oracle_over_reference_speed": round(ref_one / t_port, 3),
                            "oracle_cycles": cyc_port}}


def cpu_baseline_port(grid, f, nmodes, workers):
    """CPU oracle (reference ordering, FP64) on `nmodes` modes, one process
    per mode, `workers` at a time."""
    import dataclasses
    import multiprocessing as mp
    jobs = []
    for k in range(nmodes):
        g1 = dataclasses.replace(grid, nmodes=1)
        # mode k of the batch has kz = k*beta: keep it by shifting beta
        g1 = dataclasses.replace(g1, beta=grid.beta * k if k else 0.0, nmodes=2) if k else g1
        fk = f[k:k + 1] if not k else np.stack([np.zeros_like(f[k]), f[k]])
        jobs.append((g1, fk))
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        out = pool.map(cpu_worker, jobs)
    wall = time.perf_counter() - t0
    return wall, out


def rk4_step_bench(torch, ctx, grid, steps=2):
    """The whole RK-stage update on the bench grid: one RK4 time step of the
    Goertler case (tbl_rk4_step_dev: four drv_gv with the linear terms +
    stage update + loop, whose poi_uy runs to 1e-5 / 1e-6 as the reference
    calls it), on synthetic fields (GPU wavepackets, vorticities scaled by
    0.1), an analytic boundary-layer base flow, the reference's Re, the
    buffer-zone table with the reference's proportions; FP64 and the mixed
    cycle.  Also drv_gv and loop alone.  Timed with hipEvents on the
    library's stream, fields re-copied outside the timed region."""
    import trans_bl_amd as tb
    from trans_bl_amd.train_guess import packet_params, rhs_torch
    names = ("ux", "uy", "uz", "wx", "wy", "wz")
    F0 = {k: rhs_torch(torch, grid, packet_params(grid, 500 + q), 0.0) * (0.1 if k[0] == "w" else 1.0)
          for q, k in enumerate(names)}
    x = torch.as_tensor(grid.x(), device="cuda")
    y = torch.as_tensor(grid.y(), device="cuda")
    X, Y = torch.meshgrid(x, y, indexing="xy")
    d = 0.002 * torch.sqrt(X + 1)
    base = {"uxb": torch.tanh(Y / d), "uyb": 0.01 * (1 - torch.exp(-Y / d)),
            "wzb": -(1 - torch.tanh(Y / d) ** 2) / d, "varg": 0.05 + 0.01 * X,
            "dvargdx": 0.01 + 0 * X}
    base = {k: v.contiguous() for k, v in base.items()}
    bp = {k: v.data_ptr() for k, v in base.items()}
    bd = torch.from_numpy(tb.bdfc_generate(grid.nx, 10, int(grid.nx * 0.785),
                                           int(grid.nx * 0.88))).to("cuda")
    d2 = torch.zeros((grid.nmodes, grid.nx), dtype=torch.complex128, device="cuda")
    F = {k: torch.empty_like(v) for k, v in F0.items()}
    dv = [torch.empty_like(F0["ux"]) for _ in range(3)]
    Re, dt = 33124.00001004055, 1e-4
    st = torch.cuda.Stream()
    out = {}

    def timed(fn):
        for k in names:
            F[k].copy_(F0[k])
        torch.cuda.synchronize()
        with torch.cuda.stream(st):
            ctx.set_stream(st.cuda_stream)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            r = fn()
            e1.record(st)
        torch.cuda.synchronize()
        ctx.set_stream(None)
        return e0.elapsed_time(e1), r

    fp = {k: F[k].data_ptr() for k in names}
    for prec in (0, 1):
        # the cycle's default form: at the loop's 1e-5 / 1e-6 from the previous
        # stage the full-multigrid first correction saves a cycle per stage
        # but costs about as much (38.4 ms per step either way)
        kw = dict(precision=prec, use_graph=1, max_cycles=11)
        timed(lambda: ctx.rk4_step_dev(Re, dt, 0, bp, bd.data_ptr(), d2.data_ptr(), fp, **kw))
        ms, cyc = [], None
        for _ in range(steps):
            t, (cyc, res) = timed(lambda: ctx.rk4_step_dev(Re, dt, 0, bp, bd.data_ptr(),
                                                            d2.data_ptr(), fp, **kw))
            ms.append(t)
        out["fp64" if prec == 0 else "mixed"] = {
            "ms_per_step": round(float(np.mean(ms)), 3),
            "poi_uy_cycles_mean_per_stage": [round(float(c.mean()), 2) for c in cyc]}
    t_drv, _ = timed(lambda: ctx.drv_gv_dev(Re, 0, bp, *[fp[k] for k in names],
                                             *[v.data_ptr() for v in dv]))
    t_loop, _ = timed(lambda: ctx.loop_dev(1e-5, bd.data_ptr(), d2.data_ptr(),
                                           *[fp[k] for k in names], precision=1, use_graph=1))
    out.update({"drv_gv_ms": round(t_drv, 3), "loop_ms_mixed": round(t_loop, 3),
                "grid": f"{grid.nx}x{grid.ny}x{grid.nmodes}", "Re": Re, "dt": dt,
                "terms": "linear (lterms_gv, the reference's drv_gv)",
                "data": "synthetic: GPU wavepackets, analytic base flow"})
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the path shards "
              "by mode with no communication, see the design notes", file=sys.stderr)
        sys.exit(2)
    from trans_bl_amd import workload as wl
    grid = wl.BATCH_1025x193x64
    f = wl.wavepacket_rhs(grid, tag="bench")

    # CPU baseline first: its worker processes are forked before anything in
    # this process touches the GPU
    cpu = None
    if not a.no_cpu:
        ncpu = min(len(os.sched_getaffinity(0)), 16)
        cpu = cpu_baseline(grid, f, ncpu)
        if cpu is None:
            nm = a.cpu_modes or ncpu
            wall_c, out = cpu_baseline_port(grid, f, nm, ncpu)
            cpu = {"value": round(nm / wall_c, 4), "unit": "mode-solves/s", "cores": ncpu,
                   "kind": "port",
                   "sample": f"{nm} of the 64 modes (k=0..{nm - 1}) of the same workload, CPU "
                             f"oracle FP64 with the reference's lexicographic line order, one "
                             f"process per mode, {ncpu} at a time; {wall_c:.1f} s",
                   "cycles": [o[1][-1] for o in out]}

    import torch
    import trans_bl_amd as tb

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    from trans_bl_amd import autotune
    K = grid.nmodes
    ctx = tb.Context(grid)
    tuned = autotune.cached(ctx, sweeps=3)
    fd = torch.from_numpy(f).to("cuda")
    ud = torch.zeros_like(fd)
    torch.cuda.synchronize()

    guess = None
    if not a.no_guess:
        try:
            from trans_bl_amd import guess as gmod
            guess = gmod.load_default(ctx)
        except Exception as e:  # noqa: BLE001 - report and run without it
            print(f"learned guess unavailable: {e}", file=sys.stderr)
            guess = None

    # the previous stage's solution: the same packets 0.4 upstream
    fprev = torch.from_numpy(wl.wavepacket_rhs(grid, tag="bench", shift=-0.4)).to("cuda")
    uprev = torch.zeros_like(fd)
    torch.cuda.synchronize()
    ctx.solve_rhs_dev(fprev.data_ptr(), uprev.data_ptr(), a.tol, precision=1, max_cycles=40,
                      use_graph=1, guess=1)

    def run(prec, learned, steps, warm, start="zero"):
        # no garbage-collector pass inside a timed loop
        gc.collect()
        gc.disable()
        stats = []
        for it in range(warm + steps):
            if it == warm:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            kw = dict(precision=prec, max_cycles=40, use_graph=1)
            if prec == 1:
                kw.update(MIXED_FORM)
            if learned:
                ud.zero_()
                cyc, res, st = guess.solve(ctx, fd, ud, a.tol, **kw)
            elif start == "warm":
                # warm start: v = the previous stage's solution (guess 0)
                ud.copy_(uprev)
                cyc, res, st = ctx.solve_rhs_dev(fd.data_ptr(), ud.data_ptr(), a.tol, guess=0,
                                                 **kw)
            else:
                # the zero start (guess 1) reads nothing of uy: no clearing
                cyc, res, st = ctx.solve_rhs_dev(fd.data_ptr(), ud.data_ptr(), a.tol, guess=1, **kw)
            if it >= warm:
                stats.append((cyc.copy(), res.copy(), st))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gc.enable()
        return wall, stats

    wall64, st64 = run(0, False, a.steps, a.warmup)
    v64 = ud.clone()
    wallz, stz = run(1, False, a.steps, a.warmup)
    vz = ud.clone()
    learned = None
    if guess is not None:
        walll, stl = run(1, True, a.steps, a.warmup)
        learned = (walll, stl, ud.clone())
    wallw, stw = run(1, False, a.steps, a.warmup, start="warm")
    # value: the mixed path from zero, or from the learned guess when that is
    # faster by more than 1 %
    flag_learned = learned is not None and learned[0] < wallz / 1.01
    wallm, stm, vm = (learned if flag_learned else (wallz, stz, vz))

    # residual-tested form of the learned start (guess=2): fallback rate
    fallback = None
    if guess is not None:
        ud.zero_()
        _, _, stt = guess.solve(ctx, fd, ud, a.tol, test=True, precision=1, max_cycles=40,
                                use_graph=1)
        fallback = {"modes_falling_back": int(stt["fallback"]), "of": K,
                    "rule": "prediction's max residual on its multigrid level after two line "
                            "sweeps above 10x the zero start's"}

    def summar(stats):
        cyc = np.array([s[0] for s in stats])
        res = np.array([s[1] for s in stats])
        return {"cycles_mean": float(cyc.mean()), "cycles_max": int(cyc.max()),
                "residual_max": float(res.max()),
                "converged": bool(np.all(res < a.tol)),
                "device_ms_mean": float(np.mean([s[2]["device_ms"] for s in stats]))}

    s64, sm = summar(st64), summar(stm)
    rate64 = K * a.steps / wall64
    ratem = K * a.steps / wallm
    diff = float(torch.linalg.vector_norm(vm - v64) / torch.linalg.vector_norm(v64))

    def path(wall, stats):
        return dict(summar(stats), value=round(K * a.steps / wall, 2),
                    ms_per_step=round(wall / a.steps * 1e3, 3))

    # ---- the four derivatives of one field (two launches) on the same grid,
    # timed with hipEvents on the stream the library launches them on
    dstream = torch.cuda.Stream()
    outs = [torch.empty_like(fd) for _ in range(4)]
    torch.cuda.synchronize()
    with torch.cuda.stream(dstream):
        ctx.set_stream(dstream.cuda_stream)
        ctx.derivs_dev(vz.data_ptr(), *[o.data_ptr() for o in outs])
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(dstream)
        nrep = 10
        for _ in range(nrep):
            ctx.derivs_dev(vz.data_ptr(), *[o.data_ptr() for o in outs])
        e1.record(dstream)
    torch.cuda.synchronize()
    ctx.set_stream(None)
    d_us = e0.elapsed_time(e1) / nrep * 1e3
    n1 = grid.nx * grid.ny * K
    # algorithmic: each direction reads v once and writes two derivatives
    d_bytes = 6 * 16 * n1
    tx, ty = measured_traffic("deriv_x_kernel<17>"), measured_traffic("deriv_y_kernel<32, 7>")
    roof_d = {"bound": "hbm", "achieved": round(d_bytes / d_us / 1e3, 1), "peak": HBM_PEAK_GBS,
              "unit": "GB/s", "frac": round(d_bytes / d_us / 1e3 / HBM_PEAK_GBS, 4),
              "traffic": tx + ty if tx is not None and ty is not None else None,
              "kernel": "deriv_x_kernel<17> + deriv_y_kernel<32, 7> (d/dx, d2/dx2, d/dy, d2/dy2 "
                        "of one field)",
              "us_per_call": round(d_us, 1), "bytes_per_call": d_bytes}
    del outs

    # ---- roofline of the dominant kernel: level-1 line smoother ----------
    # algorithmic bytes per sweep: read v, read S, write v over every point of
    # the level (3 passes), complex FP64 = 16 B, FP32 = 8 B
    roof = {}
    for prec, esize in ((1, 8), (0, 16)):
        name = smoother_kernel(prec, ctx.line_shape(prec, 1), grid.ny)
        sweeps = 20
        ctx.time_smoother(prec, 1, 2)            # warm
        ms, launches = ctx.time_smoother(prec, 1, sweeps)
        per_launch_s = ms / 1e3 / launches
        bytes_per_launch = 3 * esize * n1 * sweeps / launches
        gbs = bytes_per_launch / per_launch_s / 1e9
        roof[prec] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                      "traffic": measured_traffic(name),
                      "kernel": name + " (finest level)", "avg_launch_us": round(per_launch_s * 1e6, 2),
                      "bytes_per_launch": int(bytes_per_launch),
                      "lanes_per_line": ctx.line_shape(prec, 1)}

    rk4 = None
    if not a.no_rk4:
        rk4 = rk4_step_bench(torch, ctx, grid)

    line = {
        "metric": METRIC,
        "value": round(ratem, 2),
        "unit": "mode-solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wallm / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 smoothing + f64 defect correction",
        "data": "synthetic wavepacket RHS (6 packets/mode/part, seeded)",
        "config": {"workload": "v-Poisson 1025x193x64 modes to 1e-9",
                   "nx": grid.nx, "ny": grid.ny, "modes": K, "levels": grid.levels,
                   "dx": grid.dx, "dy0": grid.dy0, "stf": grid.stf, "beta": grid.beta,
                   "initial_guess": "learned" if flag_learned else "zero",
                   "mixed_cycles": "first correction by full multigrid, then V-cycles; "
                                   "3 pre-sweeps, 1 post-sweep, 3 coarsest sweeps",
                   "parallelism": "1 GPU, all modes batched"},
        "flagship": sm,
        "plain_fp64": dict(s64, value=round(rate64, 2),
                           ms_per_step=round(wall64 / a.steps * 1e3, 3)),
        "speedup_vs_plain_fp64": round(ratem / rate64, 3),
        "learned_guess_residual_test": fallback,
        "mixed_zero_start": path(wallz, stz),
        "mixed_learned_start": None if learned is None else path(learned[0], learned[1]),
        "mixed_warm_start": dict(path(wallw, stw),
                                 start="previous stage's solution (packets 0.4 upstream)"),
        "flagship_vs_fp64_rel_l2": diff,
        "line_shapes": {k: {"fp64": v.get("fp64"), "fp32": v.get("fp32")}
                        for k, v in sorted(tuned.items())},
        "roofline": roof[1],
        "roofline_fp64_smoother": roof[0],
        "roofline_derivs": roof_d,
        "traffic_source": traffic_source(),
        "rk4_step": rk4,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
