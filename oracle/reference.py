"""Build and run the reference's own Fortran hot path (test infrastructure).

The recipe compiles the reference's sources where they lie
(``/root/reference``: the ten objects of its makefile, unchanged) with
amdflang and the image's MPICH, against a ``par.for`` configured for one
workload -- the reference's own per-run configuration step (every size is a
PARAMETER of par.for) -- and links them with ``oracle/ref_harness.f``
(poi_uy and the derivatives), ``oracle/ref_loop.f`` (the other routines
of ``loop``) and ``oracle/ref_rhs.f`` (drv_gv, nlterms_gv and the RK4 step).
The
coefficient file comes from the reference's own generator
(pre_processing/coef_generation.f90 and its modules, with the constants
module written by the reference's convert.pl from the same par.for), run
once per workload.  Intermediate files (the configured par.for, links to the
sources, objects) live in a temporary directory that is removed afterwards;
only the linked driver and the generated coefs.bin are kept, in
``oracle/_ref/<workload>/``.

Only tests, ``__graft_entry__`` (build) and ``bench.py``'s CPU-baseline leg
use this module, as the checker and the baseline; the product never does.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
REF = os.environ.get("TBL_REFERENCE", "/root/reference")
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
MPI_INC = "/opt/conda/include"
MPI_LIB = "/opt/conda/lib"

# the ten objects of the reference's makefile (makefile:2)
OBJECTS = ["main_par", "derivs_par", "loop_par", "nlterms_par", "escreve_par", "penta_par",
           "poisson_par", "filter_par", "immersed_par", "les_par"]
GENERATOR = ["solvers", "derivative", "at_wall", "coef_generation"]

# workloads: (imax, jmax, kfour, msh, dx, dy, stf, alpha, beta)
CONFIGS = {
    "129x33x4": (129, 33, 4, 4, 0.03, 1e-3, 1.03, 0.2, 0.6),
    "385x97x8": (385, 97, 8, 4, 0.02, 5e-4, 1.02, 0.0, 0.2),
    "1025x193x64": (1025, 193, 64, 4, 0.02, 2e-4, 1.015, 0.0, 0.1),
    # the benchmark grid with its first 8 modes: the CPU baseline's sample
    "1025x193x8": (1025, 193, 8, 4, 0.02, 2e-4, 1.015, 0.0, 0.1),
}


def _fnum(x: float) -> str:
    """A real*8 literal."""
    s = repr(float(x))
    if "e" in s:
        return s.replace("e", "d")
    return s + "d0"


def configure_par(text: str, cfg) -> str:
    """par.for with the workload's PARAMETERs (one rank, my_form 1)."""
    imax, jmax, kfour, msh, dx, dy, stf, alpha, beta = cfg
    subs = [
        (r"parameter \( my_form = \d+ \)", f"parameter ( my_form = 1 )"),
        (r"parameter \(fac_y = [^)]*\)", "parameter (fac_y = 1.d0)"),
        (r"parameter \( alpha = [^)]*\)", f"parameter ( alpha = {_fnum(alpha)} )"),
        (r"parameter \( beta     = [^)]*\)", f"parameter ( beta     = {_fnum(beta)} )"),
        (r"parameter \( jmax = [^\n]*\)\n", f"parameter ( jmax = {jmax}, dy = {_fnum(dy)} )\n"),
        (r"parameter \( stf = [^)]*\)", f"parameter ( stf = {_fnum(stf)} )"),
        (r"parameter \( np = \d+ \)", "parameter ( np = 1 )"),
        (r"parameter \( imax = \d+,", f"parameter ( imax = {imax},"),
        (r"parameter \( dx = [0-9.d]+,", f"parameter ( dx = {_fnum(dx)},"),
        (r"parameter \( msh = \d+ \)", f"parameter ( msh = {msh} )"),
        (r"parameter \( kfour = \d+, kphys = \d+ \)",
         f"parameter ( kfour = {kfour}, kphys = {2 * kfour} )"),
    ]
    for pat, rep in subs:
        # only active (not commented) lines: they start with blanks
        new, n = re.subn(r"(?m)^(\s+)" + pat, lambda m: m.group(1) + rep, text)
        if n != 1:
            raise RuntimeError(f"par.for: pattern {pat!r} matched {n} active lines")
        text = new
    return text


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed:\n{r.stdout[-2000:]}\n{r.stderr[-3000:]}")
    return r


def exe_path(name: str) -> str:
    return os.path.join(OUT, name, "ref_harness")


def loop_exe_path(name: str) -> str:
    return os.path.join(OUT, name, "ref_loop")


def rhs_exe_path(name: str) -> str:
    return os.path.join(OUT, name, "ref_rhs")


def coefs_path(name: str) -> str:
    return os.path.join(OUT, name, "coefs.bin")


def available(name: str) -> bool:
    return all(os.path.exists(f) for f in (exe_path(name), loop_exe_path(name),
                                            rhs_exe_path(name), coefs_path(name)))


def buildable() -> bool:
    return os.path.isdir(REF) and os.path.exists(FC) and os.path.exists(
        os.path.join(MPI_INC, "mpif.h"))


def build(name: str, verbose: bool = True) -> str:
    cfg = CONFIGS[name]
    if not buildable():
        raise RuntimeError("the reference tree, amdflang or MPICH is missing")
    dest = os.path.join(OUT, name)
    os.makedirs(dest, exist_ok=True)
    tmp = tempfile.mkdtemp(prefix="tbl_ref_")
    try:
        par = configure_par(open(os.path.join(REF, "par.for")).read(), cfg)
        with open(os.path.join(tmp, "par.for"), "w") as f:
            f.write(par)
        # sources linked into the build directory, so that their
        # `include 'par.for'` finds the configured one first
        flags = ["-ffixed-form", "-O3", "-mcmodel=medium", "-I", tmp, "-I", REF, "-I", MPI_INC]
        objs = []
        for o in OBJECTS:
            os.symlink(os.path.join(REF, o + ".f"), os.path.join(tmp, o + ".f"))
            _run([FC, *flags, "-c", o + ".f", "-o", o + ".o"], tmp)
            objs.append(o + ".o")
        # the reference's program unit is not the entry point here
        _run(["objcopy", "--redefine-sym", "_QQmain=tbl_ref_unused_QQmain", "--redefine-sym",
              "main=tbl_ref_unused_main", "main_par.o"], tmp)
        _run([FC, *flags, "-c", os.path.join(HERE, "ref_harness.f"), "-o", "harness.o"], tmp)
        exe = exe_path(name)
        # GNU ld: it places the COMMON blocks (gigabytes at the larger sizes)
        # after the Fortran runtime's own data, which is built for the small
        # code model; no relaxation of the GOT loads of those blocks
        _run([FC, *flags, "-c", os.path.join(HERE, "ref_loop.f"), "-o", "loop.o"], tmp)
        _run([FC, *flags, "-c", os.path.join(HERE, "ref_rhs.f"), "-o", "rhs.o"], tmp)
        for main, out in (("harness.o", exe), ("loop.o", loop_exe_path(name)),
                          ("rhs.o", rhs_exe_path(name))):
            _run([FC, "-fuse-ld=bfd", "-mcmodel=medium", "-o", out, main, *objs, "-L", MPI_LIB,
                  "-lmpifort", "-lmpi", f"-Wl,-rpath,{MPI_LIB}", "-Wl,--no-relax"], tmp)
        # coefficient file: the reference's generator for this par.for
        gen = os.path.join(tmp, "gen")
        os.makedirs(os.path.join(gen, "pre_processing"))
        with open(os.path.join(gen, "par.for"), "w") as f:
            f.write(par)
        os.symlink(os.path.join(REF, "constants"), os.path.join(gen, "constants"))
        _run(["perl", os.path.join(REF, "convert.pl")], gen)
        srcs = [os.path.join("pre_processing", "constants.f90")]
        for s in GENERATOR:
            os.symlink(os.path.join(REF, "pre_processing", s + ".f90"), os.path.join(gen, s + ".f90"))
            srcs.append(s + ".f90")
        _run([FC, "-O2", "-J", gen, *srcs, "-o", "coef"], gen)
        _run([os.path.join(gen, "coef")], gen)
        shutil.copyfile(os.path.join(gen, "coefs.bin"), coefs_path(name))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    if verbose:
        print(f"built {exe_path(name)}")
    return exe_path(name)


def build_all(verbose: bool = True):
    for name in CONFIGS:
        build(name, verbose)


def read_coefs(name: str) -> dict:
    """Tables of the generated coefs.bin in the oracle's dict layout."""
    b = open(coefs_path(name), "rb").read()
    recs, i = [], 0
    while i < len(b):
        n = int.from_bytes(b[i:i + 4], "little")
        recs.append(np.frombuffer(b[i + 4:i + 4 + n], dtype="<f8").copy())
        i += 8 + n
    msh = CONFIGS[name][3]
    names = ["fp_fd_e", "sp_fd", "cp_fd", "pp_fd", "lp_fd", "fp_sd", "sp_sd", "cp_sd",
             "pp_sd", "lp_sd"]
    out = {n: recs[q] for q, n in enumerate(names)}
    for q, (n, w) in enumerate((("sp_poi", 7), ("cp_poi", 8), ("pp_poi", 4), ("lp_poi", 5))):
        out[n] = recs[10 + q].reshape(msh, w)
    out["w_at_w"] = recs[14]
    out["dwydy"] = recs[15]
    return out


def run(name: str, wx, dwzdx, uy, erromax: float, calls: int = 1, timeout: float = 600.0,
        checkpoint_t: int | None = None):
    """poi_uy (calls in a row) and the five derivatives of the result, by the
    reference's code.  Fields are complex (nk, ny, nx) arrays (C view of the
    Fortran (ptsx, jmax, kfour)).  Returns a dict with uy, the derivatives
    (x, xx, y, yy, yfv), per-mode cycles and residuals of the last call, and
    the wall time of the calls.  With checkpoint_t the reference's own
    checkpoint writer runs after the calls and the file's bytes are returned
    as "data_bin"."""
    if not available(name):
        raise RuntimeError(f"reference driver {name} is not built")
    nx, ny, nk = CONFIGS[name][0], CONFIGS[name][1], CONFIGS[name][2]
    shape = (nk, ny, nx)
    tmp = tempfile.mkdtemp(prefix="tbl_refrun_")
    try:
        fin = os.path.join(tmp, "in.bin")
        fout = os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            for a in (wx, dwzdx, uy):
                a = np.ascontiguousarray(a, dtype=np.complex128)
                assert a.shape == shape
                f.write(a.tobytes())
        env = dict(os.environ)
        t0 = time.perf_counter()
        # poi_uy's level planes are local arrays: give the stack room
        extra = [] if checkpoint_t is None else [str(int(checkpoint_t))]
        r = subprocess.run(["bash", "-c", 'ulimit -s unlimited && exec "$@"', "ref",
                            exe_path(name), coefs_path(name), fin, fout, repr(float(erromax)),
                            str(calls), *extra], capture_output=True, text=True,
                           timeout=timeout, env=env, cwd=tmp)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        data = np.fromfile(fout, dtype=np.complex128)
        fields = data.reshape(6, *shape)
        cycles = np.zeros(nk, np.int32)
        res = np.zeros(nk, np.float64)
        seconds = None
        for line in r.stdout.splitlines():
            t = line.split()
            if len(t) == 3 and t[0].isdigit():
                k = int(t[0])
                cycles[k - 1] = int(t[1])
                res[k - 1] = float(t[2].replace("D", "E"))
            elif line.strip().startswith("poi_uy seconds"):
                seconds = float(t[-1])
        out = {"uy": fields[0], "x": fields[1], "xx": fields[2], "y": fields[3],
               "yy": fields[4], "yfv": fields[5], "cycles": cycles, "residual": res,
               "seconds": seconds, "wall": wall}
        if checkpoint_t is not None:
            with open(os.path.join(tmp, "data_00.bin"), "rb") as f:
                out["data_bin"] = f.read()
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def run_parallel(name: str, wx, dwzdx, uy, erromax: float, copies: int,
                 timeout: float = 600.0):
    """`copies` processes of the reference driver at once on the same inputs
    (one core each): returns the wall time of the batch, each process's
    poi_uy seconds and the per-mode cycle counts of the first."""
    if not available(name):
        raise RuntimeError(f"reference driver {name} is not built")
    nx, ny, nk = CONFIGS[name][0], CONFIGS[name][1], CONFIGS[name][2]
    shape = (nk, ny, nx)
    tmp = tempfile.mkdtemp(prefix="tbl_refpar_")
    try:
        fin = os.path.join(tmp, "in.bin")
        with open(fin, "wb") as f:
            for a in (wx, dwzdx, uy):
                a = np.ascontiguousarray(a, dtype=np.complex128)
                assert a.shape == shape
                f.write(a.tobytes())
        procs = []
        t0 = time.perf_counter()
        for c in range(copies):
            fout = os.path.join(tmp, f"out{c}.bin")
            procs.append(subprocess.Popen(
                ["bash", "-c", 'ulimit -s unlimited && exec "$@"', "ref", exe_path(name),
                 coefs_path(name), fin, fout, repr(float(erromax)), "1"],
                stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, cwd=tmp))
        outs = [p.communicate(timeout=timeout) for p in procs]
        wall = time.perf_counter() - t0
        secs, cycles = [], np.zeros(nk, np.int32)
        for c, (so, se) in enumerate(outs):
            if procs[c].returncode != 0:
                raise RuntimeError(f"reference run failed:\n{so[-1000:]}\n{se[-1000:]}")
            for line in so.splitlines():
                t = line.split()
                if line.strip().startswith("poi_uy seconds"):
                    secs.append(float(t[-1]))
                elif c == 0 and len(t) == 3 and t[0].isdigit():
                    cycles[int(t[0]) - 1] = int(t[1])
        return wall, secs, cycles
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    build_all()


LOOP_FIELDS = ("ux", "uy", "uz", "wx", "wy", "wz", "th")


def run_loop(name: str, fields: dict, d2uydx2, bdfc, erro: float, what: str = "loop",
             timeout: float = 600.0):
    """One of the reference's loop routines (ref_loop.f: loop, poi, wall,
    bzone, bzonew, filter, outuy) on the given fields -- complex (nk, ny, nx)
    arrays ux .. th (missing ones zero), d2uydx2 complex (nk, nx), bdfc real
    (nx, 2).  Returns the fields after the call, the extra output ("duydy",
    "lapv" or "dwzdx") and, for loop, poi_uy's per-mode cycles and residuals."""
    if not available(name):
        raise RuntimeError(f"reference driver {name} is not built")
    nx, ny, nk = CONFIGS[name][0], CONFIGS[name][1], CONFIGS[name][2]
    shape = (nk, ny, nx)
    tmp = tempfile.mkdtemp(prefix="tbl_refloop_")
    try:
        fin = os.path.join(tmp, "in.bin")
        fout = os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            for k in LOOP_FIELDS:
                a = fields.get(k)
                a = np.zeros(shape, np.complex128) if a is None else a
                f.write(np.ascontiguousarray(a, dtype="<c16").tobytes())
            f.write(np.ascontiguousarray(d2uydx2, dtype="<c16").tobytes())
            f.write(np.ascontiguousarray(bdfc, dtype="<f8").tobytes())
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run(["bash", "-c", 'ulimit -s unlimited && exec "$@"', "ref",
                            loop_exe_path(name), coefs_path(name), fin, fout, repr(float(erro)),
                            what], capture_output=True, text=True, timeout=timeout, env=env,
                           cwd=tmp)
        if r.returncode != 0:
            raise RuntimeError(f"ref_loop failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        b = np.fromfile(fout, dtype="<c16")
        n3 = nk * ny * nx
        out = {k: b[q * n3:(q + 1) * n3].reshape(shape).copy()
               for q, k in enumerate(LOOP_FIELDS)}
        rest = b[len(LOOP_FIELDS) * n3:]
        if what in ("poi", "outuy"):
            out["duydy" if what == "poi" else "dwzdx"] = rest.reshape(shape).copy()
        elif what == "wall":
            out["lapv"] = rest.reshape(nk, nx).copy()
        cycles = np.zeros(nk, np.int32)
        res = np.zeros(nk, np.float64)
        for line in r.stdout.splitlines():
            t = line.split()
            if len(t) == 3 and t[0].isdigit():
                cycles[int(t[0]) - 1] = int(t[1])
                res[int(t[0]) - 1] = float(t[2].replace("D", "E"))
        out["cycles"] = cycles
        out["residual"] = res
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


BASE_FIELDS = ("uxb", "uyb", "wzb", "varg", "dvargdx")


def run_rhs(name: str, fields: dict, base: dict, d2uydx2, bdfc, fcx2, fcllx2, dt: float,
            what: str, t: int = 1, timeout: float = 600.0):
    """drv_gv ("drv": dvx, dvy, dvz), nlterms_gv ("nlterms": a, b, c, d),
    drv_fi ("drvfi"), nlterms_fi ("nltfi": a, b, c), lterms_fi ("ltfi") or
    one RK4 time step at step t ("rk4": ux .. wz and d2uydx2) by the
    reference's code (ref_rhs.f).  fields: complex (nk, ny, nx) ux .. wz;
    base: real (ny, nx) uxb, uyb, wzb, varg, dvargdx; fcx2, fcllx2: the
    disturbance shapes of gv_pert (length i2).  Also returns "params" (Re,
    i1, i2, stpp of the configured par.for) and poi_uy's cycle lines."""
    if not available(name):
        raise RuntimeError(f"reference driver {name} is not built")
    nx, ny, nk = CONFIGS[name][0], CONFIGS[name][1], CONFIGS[name][2]
    shape = (nk, ny, nx)
    tmp = tempfile.mkdtemp(prefix="tbl_refrhs_")
    try:
        fin = os.path.join(tmp, "in.bin")
        fout = os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            for k in ("ux", "uy", "uz", "wx", "wy", "wz"):
                f.write(np.ascontiguousarray(fields[k], dtype="<c16").tobytes())
            for k in BASE_FIELDS:
                f.write(np.ascontiguousarray(base[k], dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(d2uydx2, dtype="<c16").tobytes())
            f.write(np.ascontiguousarray(bdfc, dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(fcx2, dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(fcllx2, dtype="<f8").tobytes())
            f.write(np.array([dt], dtype="<f8").tobytes())
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run(["bash", "-c", 'ulimit -s unlimited && exec "$@"', "ref",
                            rhs_exe_path(name), coefs_path(name), fin, fout, what, str(t)],
                           capture_output=True, text=True, timeout=timeout, env=env, cwd=tmp)
        if r.returncode != 0:
            raise RuntimeError(f"ref_rhs failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        b = np.fromfile(fout, dtype="<c16")
        n3 = nk * ny * nx
        names = {"drv": ("dvx", "dvy", "dvz"), "nlterms": ("a", "b", "c", "d"),
                 "drvfi": ("dvx", "dvy", "dvz"), "nltfi": ("a", "b", "c"),
                 "ltfi": ("a", "b", "c"),
                 "rk4": ("ux", "uy", "uz", "wx", "wy", "wz")}[what]
        out = {k: b[q * n3:(q + 1) * n3].reshape(shape).copy() for q, k in enumerate(names)}
        if what == "rk4":
            out["d2uydx2"] = b[len(names) * n3:].reshape(nk, nx).copy()
        cycles = []
        for line in r.stdout.splitlines():
            tk = line.split()
            if tk and tk[0] == "params":
                out["params"] = {"Re": float(tk[1].replace("D", "E")), "i1": int(tk[2]),
                                 "i2": int(tk[3]), "stpp": int(tk[4])}
            elif len(tk) == 3 and tk[0].isdigit():
                cycles.append(int(tk[1]))
        out["cycles"] = np.array(cycles, np.int32)
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
