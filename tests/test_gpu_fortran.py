"""The Fortran drop-in layer (trans_bl_amd/fortran) driven by a Fortran
RK-stage program: derparx, derparxx, dery, deryy of uy, then poi_uy, with
the reference's subroutine names, COMMON blocks and PARAMETERs (the test's
own par.for, 129 x 33 x 4 modes, 3 levels), against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from trans_bl_amd import workload as wl
from conftest import rel

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(__file__)), "trans_bl_amd", "fortran",
                   "_build", "rk_stage")
G = wl.Grid(129, 33, 4, 3, 0.03, 1e-3, 1.03, 0.2, 0.6)


def test_fortran_stage_matches_oracle(tmp_path, gpu_ready):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    coefs = oracle.coefgen.generate(G.stf, G.levels)
    f = wl.wavepacket_rhs(G, tag="fortran")
    wx = 0.05 * wl.wavepacket_rhs(G, tag="fortran-wx")
    rng = np.random.default_rng(3)
    uy = np.zeros_like(f)
    uy[:, 1:, 1:-1] = 1e-3 * (rng.normal(size=uy[:, 1:, 1:-1].shape))
    erromax = 1e-9
    with open(tmp_path / "stage_in.bin", "wb") as fh:
        for name in ("fp_fd_e", "sp_fd", "cp_fd", "pp_fd", "lp_fd", "fp_sd", "sp_sd", "cp_sd",
                     "pp_sd", "lp_sd", "sp_poi", "cp_poi", "pp_poi", "lp_poi"):
            np.ascontiguousarray(coefs[name], dtype="<f8").tofile(fh)
        np.array([erromax], dtype="<f8").tofile(fh)
        for a in (wx, -f, uy):
            np.ascontiguousarray(a, dtype="<c16").tofile(fh)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(tmp_path / "stage_out.bin", dtype="<c16").reshape((5,) + f.shape)
    p = oracle.make_params(G.nx, G.ny, G.nmodes, G.levels, G.dx, G.dy0, G.stf, G.alpha, G.beta,
                           coefs=coefs)
    for i, kind in enumerate(("x", "xx", "y", "yy")):
        ref = oracle.deriv(p, kind, uy)
        assert rel(out[i], ref) < 1e-12, (kind, rel(out[i], ref))
    ref, cyc, res = oracle.poi_uy(p, wx, -f, uy, erromax, 11, order=1)
    assert rel(out[4], ref) < 1e-12, rel(out[4], ref)
    # poi_uy prints (k, V-cycles, residual) per mode like the reference
    lines = [l.split() for l in r.stdout.strip().splitlines() if l.strip()]
    assert [int(l[1]) for l in lines[-G.nmodes:]] == cyc.tolist()
