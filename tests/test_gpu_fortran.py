"""The Fortran drop-in layer (trans_bl_amd/fortran) driven by a Fortran
RK-stage program: derparx, derparxx, dery, deryy of uy, then poi_uy, with
the reference's subroutine names, COMMON blocks and PARAMETERs (the test's
own par.for, 129 x 33 x 4 modes, 3 levels), against the CPU oracle; a
Fortran program calling the drop-in `loop(erro)`, the whole RK-stage update;
and one calling the drop-in `drv_gv(dvx, dvy, dvz)` or `drv_fi(dvx, dvy,
dvz)`, the RK right-hand side of the Goertler or the flat-plate form."""
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


LOOP_EXE = os.path.join(os.path.dirname(EXE), "rk_loop")


@pytest.mark.parametrize("erro", [1e-5, 1e-6])
def test_fortran_loop_matches_oracle(tmp_path, gpu_ready, erro):
    """The drop-in `loop(erro)` (main_par.f:1762-1785) from a Fortran
    program with the reference's COMMON blocks /var/, /coefs/, /bd/, /derw/:
    every field against the oracle's loop (colour-ordered poi_uy, as on the
    GPU), and poi_uy's printed cycle counts."""
    import trans_bl_amd as tb
    if not os.path.exists(LOOP_EXE):
        pytest.fail(f"{LOOP_EXE} not built (run __graft_entry__.build())")
    coefs = oracle.coefgen.generate(G.stf, G.levels)
    F = {k: wl.wavepacket_rhs(G, tag="floop-" + k) * (0.1 if k[0] == "w" else 1.0)
         for k in ("ux", "uy", "uz", "wx", "wy", "wz")}
    rng = np.random.default_rng(11)
    d2 = rng.standard_normal((G.nmodes, G.nx)) + 1j * rng.standard_normal((G.nmodes, G.nx))
    b = tb.bdfc_generate(G.nx, 10, 101, 113)
    with open(tmp_path / "loop_in.bin", "wb") as fh:
        for name in ("fp_fd_e", "sp_fd", "cp_fd", "pp_fd", "lp_fd", "fp_sd", "sp_sd", "cp_sd",
                     "pp_sd", "lp_sd", "sp_poi", "cp_poi", "pp_poi", "lp_poi", "w_at_w",
                     "dwydy"):
            np.ascontiguousarray(coefs[name], dtype="<f8").tofile(fh)
        np.array([erro], dtype="<f8").tofile(fh)
        for k in ("ux", "uy", "uz", "wx", "wy", "wz"):
            np.ascontiguousarray(F[k], dtype="<c16").tofile(fh)
        np.ascontiguousarray(d2, dtype="<c16").tofile(fh)
        np.ascontiguousarray(b, dtype="<f8").tofile(fh)
    r = subprocess.run([LOOP_EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shp = (G.nmodes, G.ny, G.nx)
    out = np.fromfile(tmp_path / "loop_out.bin", dtype="<c16").reshape((6,) + shp)
    p = oracle.make_params(G.nx, G.ny, G.nmodes, G.levels, G.dx, G.dy0, G.stf, G.alpha, G.beta,
                           coefs=coefs)
    ref, cyc, _ = oracle.loop(p, erro, b, d2, F, order=1)
    for i, k in enumerate(("ux", "uy", "uz", "wx", "wy", "wz")):
        assert rel(out[i], ref[k]) < 1e-12, (k, rel(out[i], ref[k]))
    lines = [l.split() for l in r.stdout.strip().splitlines() if l.strip()]
    assert [int(l[1]) for l in lines[-G.nmodes:]] == cyc.tolist()


DRV_EXE = os.path.join(os.path.dirname(EXE), "rk_drv")
RE_PARFOR = 2.5e3           # Re of tests/fortran/par.for
KPHYS_PARFOR = 8            # its kphys


@pytest.mark.parametrize("form", ["gv", "fi"])
def test_fortran_drv_matches_oracle(tmp_path, gpu_ready, form):
    """The drop-in `drv_gv(dvx, dvy, dvz)` (main_par.f:1339-1403, with
    lterms_gv) or `drv_fi(dvx, dvy, dvz)` (1264-1336, with nlterms_fi at
    par.for's kphys) from a Fortran program with the reference's COMMON
    blocks /var/, /coefs/, /blas/, /vc/ and par.for's Re: the three results
    on rows 2..jmax against the oracle at 1e-12 of each field's largest
    value, as the device entry points are checked; row 1 stays the caller's
    zero."""
    if not os.path.exists(DRV_EXE):
        pytest.fail(f"{DRV_EXE} not built (run __graft_entry__.build())")
    coefs = oracle.coefgen.generate(G.stf, G.levels)
    F = {k: wl.wavepacket_rhs(G, tag="fdrv-" + k) * (0.1 if k[0] == "w" else 1.0)
         for k in ("ux", "uy", "uz", "wx", "wy", "wz")}
    X, Y = np.meshgrid(G.x(), G.y())
    d = 0.002 * np.sqrt(X + 1)
    base = {"uxb": np.tanh(Y / d), "uyb": 0.01 * (1 - np.exp(-Y / d)),
            "wzb": -(1 - np.tanh(Y / d) ** 2) / d, "varg": 0.05 + 0.01 * X,
            "dvargdx": 0.01 + 0 * X}
    with open(tmp_path / "drv_in.bin", "wb") as fh:
        for name in ("fp_fd_e", "sp_fd", "cp_fd", "pp_fd", "lp_fd", "fp_sd", "sp_sd", "cp_sd",
                     "pp_sd", "lp_sd", "sp_poi", "cp_poi", "pp_poi", "lp_poi", "w_at_w",
                     "dwydy"):
            np.ascontiguousarray(coefs[name], dtype="<f8").tofile(fh)
        for k in ("ux", "uy", "uz", "wx", "wy", "wz"):
            np.ascontiguousarray(F[k], dtype="<c16").tofile(fh)
        for k in ("uxb", "uyb", "wzb", "varg", "dvargdx"):
            np.ascontiguousarray(base[k], dtype="<f8").tofile(fh)
    r = subprocess.run([DRV_EXE, form], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shp = (G.nmodes, G.ny, G.nx)
    out = np.fromfile(tmp_path / "drv_out.bin", dtype="<c16").reshape((3,) + shp)
    p = oracle.make_params(G.nx, G.ny, G.nmodes, G.levels, G.dx, G.dy0, G.stf, G.alpha, G.beta,
                           coefs=coefs)
    ref = (oracle.drv_gv(p, RE_PARFOR, 0, base, F) if form == "gv"
           else oracle.drv_fi(p, RE_PARFOR, KPHYS_PARFOR, None, F))
    for q, name in enumerate(("dvx", "dvy", "dvz")):
        e = rel(out[q][:, 1:], ref[q][:, 1:])
        assert e < 1e-12, (name, e)
        assert not np.any(out[q][:, 0]), name

