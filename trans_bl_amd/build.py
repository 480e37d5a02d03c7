"""Build the trans_bl_amd HIP library in-tree for gfx950.

    python -m trans_bl_amd.build        (or trans_bl_amd.build.build())

Compiles every source under csrc/ with hipcc into csrc/_obj/ and links
trans_bl_amd/libtrans_bl_amd.so.  Always recompiles (it is the build check).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(CSRC, "_obj")
LIB = os.path.join(HERE, "libtrans_bl_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TBL_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no fused multiply-add: FP64 results follow the reference's rounding
    "-ffp-contract=off",
    # packed FP32 instructions stay disabled
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

SOURCES = [
    "tbl_host.cpp",
    "tbl_solver.cpp",
    "tbl_mg_kernels.hip",
    "tbl_smooth_phase.hip",
    "tbl_smooth_strip.hip",
    "tbl_smooth_coarse.hip",
    "tbl_smooth_sweep.hip",
    "tbl_deriv_kernels.hip",
    "tbl_unet.hip",
    "tbl_net.cpp",
    "tbl_guess.hip",
    "tbl_batch.cpp",
    "tbl_noalloc.hip",
    "tbl_files.cpp",
    "tbl_xline.hip",
    "tbl_loop.cpp",
    "tbl_rhs.hip",
    "tbl_rk4.cpp",
]


HOST_ONLY = {"tbl_host.cpp", "tbl_files.cpp"}


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src in HOST_ONLY:
        cmd = [HIPCC, "-x", "c++", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
               "-c", path, "-o", out]
    else:
        cmd = [HIPCC, "-x", "hip", *COMMON, "-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return out


FORT = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FORT, "_build")
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")


def build_fortran(verbose: bool = True) -> str:
    """ISO_C_BINDING module, the drop-in subroutines and the Fortran
    test drivers rk_stage, rk_loop and rk_drv (include path: the test's own
    par.for/comm.*)."""
    os.makedirs(FBUILD, exist_ok=True)
    inc = os.path.join(os.path.dirname(HERE), "tests", "fortran")
    steps = [
        [FC, "-O2", "-c", os.path.join(FORT, "tbl_iface.f90"), "-J", FBUILD,
         "-o", os.path.join(FBUILD, "tbl_iface.o")],
        [FC, "-O2", "-ffixed-form", "-I", inc, "-J", FBUILD, "-c",
         os.path.join(FORT, "tbl_dropin.f"), "-o", os.path.join(FBUILD, "tbl_dropin.o")],
        [FC, "-O2", "-ffixed-form", "-I", inc, "-J", FBUILD, "-c",
         os.path.join(inc, "rk_stage.f"), "-o", os.path.join(FBUILD, "rk_stage.o")],
        [FC, os.path.join(FBUILD, "rk_stage.o"), os.path.join(FBUILD, "tbl_dropin.o"),
         os.path.join(FBUILD, "tbl_iface.o"), "-L", HERE, "-ltrans_bl_amd",
         "-Wl,-rpath,$ORIGIN/../..", "-o", os.path.join(FBUILD, "rk_stage")],
        [FC, "-O2", "-ffixed-form", "-I", inc, "-J", FBUILD, "-c",
         os.path.join(inc, "rk_loop.f"), "-o", os.path.join(FBUILD, "rk_loop.o")],
        [FC, os.path.join(FBUILD, "rk_loop.o"), os.path.join(FBUILD, "tbl_dropin.o"),
         os.path.join(FBUILD, "tbl_iface.o"), "-L", HERE, "-ltrans_bl_amd",
         "-Wl,-rpath,$ORIGIN/../..", "-o", os.path.join(FBUILD, "rk_loop")],
        [FC, "-O2", "-ffixed-form", "-I", inc, "-J", FBUILD, "-c",
         os.path.join(inc, "rk_drv.f"), "-o", os.path.join(FBUILD, "rk_drv.o")],
        [FC, os.path.join(FBUILD, "rk_drv.o"), os.path.join(FBUILD, "tbl_dropin.o"),
         os.path.join(FBUILD, "tbl_iface.o"), "-L", HERE, "-ltrans_bl_amd",
         "-Wl,-rpath,$ORIGIN/../..", "-o", os.path.join(FBUILD, "rk_drv")],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed: {' '.join(cmd)}\n{r.stderr[-3000:]}")
    exe = os.path.join(FBUILD, "rk_stage")
    if verbose:
        print(f"built {exe}")
    return exe


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    if os.path.exists(FC):
        build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
