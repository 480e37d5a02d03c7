"""ctypes front end of the CPU oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the reference that results are
compared against (never as the thing measured or shipped).  It wraps
``libtbl_oracle.so`` (built from ``tbl_oracle.c`` by ``oracle/Makefile``) and
the numpy coefficient generator in ``coefgen.py``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import coefgen

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libtbl_oracle.so")
MAXLVL = 8


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nk", C.c_int), ("msh", C.c_int),
        ("dx", C.c_double), ("dy0", C.c_double), ("stf", C.c_double),
        ("alpha", C.c_double), ("beta", C.c_double), ("fac_y", C.c_double),
        ("fp_fd_e", C.c_double * 6), ("sp_fd", C.c_double * 9),
        ("cp_fd", C.c_double * 8), ("pp_fd", C.c_double * 9), ("lp_fd", C.c_double * 7),
        ("fp_sd", C.c_double * 8), ("sp_sd", C.c_double * 10),
        ("cp_sd", C.c_double * 8), ("pp_sd", C.c_double * 10), ("lp_sd", C.c_double * 8),
        ("sp_poi", (C.c_double * 7) * MAXLVL), ("cp_poi", (C.c_double * 8) * MAXLVL),
        ("pp_poi", (C.c_double * 4) * MAXLVL), ("lp_poi", (C.c_double * 5) * MAXLVL),
        ("w_at_w", C.c_double * 9), ("dwydy", C.c_double * 8),
    ]


class Base(C.Structure):
    """or_base: COMMON /blas/ uxb, uyb, wzb and /vc/ varg, dvargdx."""
    _fields_ = [("uxb", C.c_void_p), ("uyb", C.c_void_p), ("wzb", C.c_void_p),
                ("varg", C.c_void_p), ("dvargdx", C.c_void_p)]


class Metrics(C.Structure):
    """or_metrics: create_ctes (main_par.f:1828-1892) per level."""
    _fields_ = [
        ("ptsx", C.c_int * MAXLVL), ("ptsy", C.c_int * MAXLVL),
        ("dy0", C.c_double * MAXLVL), ("stf", C.c_double * MAXLVL),
        ("dx2", C.c_double * MAXLVL),
        ("qdy", C.c_double * 1024),
        ("qdy2", (C.c_double * 1024) * MAXLVL),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.POINTER(Params)
        vp = C.c_void_p
        for name in ("or_derparx", "or_derparxx", "or_dery", "or_deryy", "or_deryfv"):
            getattr(L, name).argtypes = [P, vp, vp]
            getattr(L, name).restype = C.c_int
        L.or_poi_uy.argtypes = [P, vp, vp, vp, C.c_double, C.c_int, C.c_int, vp, vp]
        L.or_poi_uy.restype = C.c_int
        L.or_poi_uy_modes.argtypes = [P, vp, vp, vp, C.c_double, C.c_int, C.c_int, vp, vp, vp,
                                      C.c_int]
        L.or_poi_uy_modes.restype = C.c_int
        L.or_apply_L_level.argtypes = [P, C.c_int, C.c_int, vp, vp]
        L.or_apply_L_level.restype = C.c_int
        L.or_band5_factor.argtypes = [vp, C.c_int, vp]
        L.or_band5_solve.argtypes = [vp, C.c_int, vp, vp]
        M = C.POINTER(Metrics)
        L.or_create_ctes.argtypes = [P, M]
        L.or_create_ctes.restype = C.c_int
        L.or_restrict.argtypes = [P, M, C.c_int, vp, vp, vp, vp, vp]
        L.or_restrict.restype = C.c_int
        L.or_correct_prolong.argtypes = [P, M, C.c_int, vp, vp, vp]
        L.or_correct_prolong.restype = C.c_int
        L.or_source.argtypes = [P, C.c_int, vp, vp, vp]
        L.or_source.restype = C.c_int
        L.or_sor.argtypes = [P, M, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int]
        L.or_sor.restype = C.c_int
        L.or_lhs5.argtypes = [C.c_int, C.c_double, vp]
        L.or_ludecomp.argtypes = [C.c_int, vp, vp]
        L.or_lusolve.argtypes = [C.c_int, vp, vp, C.c_size_t, C.c_int]
        for name, n in (("or_poi_ux", 4), ("or_poi_uz", 3), ("or_wx_wall", 5),
                        ("or_wz_wall", 3), ("or_filter_trid", 3), ("or_outuy", 3)):
            getattr(L, name).argtypes = [P] + [vp] * n
            getattr(L, name).restype = C.c_int
        L.or_der2wydxdy.argtypes = [P, vp, C.c_int, vp]
        L.or_der2wydxdy.restype = C.c_int
        L.or_bzone.argtypes = [P, vp, C.c_int, vp, vp, vp, vp]
        L.or_bzone.restype = C.c_int
        L.or_bdfc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp]
        L.or_loop.argtypes = [P, C.c_double, vp, vp] + [vp] * 7 + [C.c_int, vp, vp]
        L.or_loop.restype = C.c_int
        B = C.POINTER(Base)
        L.or_f_to_p.argtypes = [P, C.c_int, vp, vp]
        L.or_f_to_p.restype = C.c_int
        L.or_p_to_f.argtypes = [P, C.c_int, vp, vp]
        L.or_p_to_f.restype = C.c_int
        L.or_terms.argtypes = [P, C.c_int, B] + [vp] * 10
        L.or_terms.restype = C.c_int
        L.or_drv_gv.argtypes = [P, C.c_double, C.c_int, B] + [vp] * 9
        L.or_drv_gv.restype = C.c_int
        L.or_terms_fi.argtypes = [P, C.c_int, B] + [vp] * 9
        L.or_terms_fi.restype = C.c_int
        L.or_drv_fi.argtypes = [P, C.c_double, C.c_int, B] + [vp] * 9
        L.or_drv_fi.restype = C.c_int
        L.or_rk4_step.argtypes = [P, C.c_double, C.c_double, C.c_int, B, vp, vp, vp, vp,
                                  C.c_int, C.c_int, vp, C.c_int, vp]
        L.or_rk4_step.restype = C.c_int
        _lib = L
    return _lib


def make_params(nx, ny, nk, msh, dx, dy0, stf, alpha=0.0, beta=0.2, fac_y=1.0,
                coefs=None):
    """Fill the oracle's parameter block; coefs defaults to the oracle's own
    generator at (stf, msh)."""
    if coefs is None:
        coefs = coefgen.generate(stf, msh)
    p = Params()
    p.nx, p.ny, p.nk, p.msh = nx, ny, nk, msh
    p.dx, p.dy0, p.stf = dx, dy0, stf
    p.alpha, p.beta, p.fac_y = alpha, beta, fac_y
    for name in ("fp_fd_e", "sp_fd", "cp_fd", "pp_fd", "lp_fd",
                 "fp_sd", "sp_sd", "cp_sd", "pp_sd", "lp_sd"):
        arr = getattr(p, name)
        for i, v in enumerate(np.asarray(coefs[name], dtype=np.float64)):
            arr[i] = v
    for name in ("sp_poi", "cp_poi", "pp_poi", "lp_poi"):
        tab = np.asarray(coefs[name], dtype=np.float64)
        arr = getattr(p, name)
        for l in range(msh):
            for i, v in enumerate(tab[l]):
                arr[l][i] = v
    for name in ("w_at_w", "dwydy"):
        if name in coefs:
            arr = getattr(p, name)
            for i, v in enumerate(np.asarray(coefs[name], dtype=np.float64)):
                arr[i] = v
    return p


def _c(a):
    a = np.ascontiguousarray(a)
    return a, a.ctypes.data_as(C.c_void_p)


def _field(p):
    return (p.nk, p.ny, p.nx)


def deriv(p, kind, f):
    """kind in {'x','xx','y','yy','yfv'}; f complex128 of shape (nk, ny, nx)
    (C view of the Fortran (nx, ny, nk) array)."""
    f = np.ascontiguousarray(f, dtype=np.complex128)
    assert f.shape == _field(p)
    out = np.zeros_like(f)
    fn = {"x": "or_derparx", "xx": "or_derparxx", "y": "or_dery", "yy": "or_deryy",
          "yfv": "or_deryfv"}[kind]
    rc = getattr(lib(), fn)(C.byref(p), f.ctypes.data_as(C.c_void_p),
                            out.ctypes.data_as(C.c_void_p))
    if rc:
        raise ValueError(f"{fn} rejected the grid (rc={rc})")
    return out


def poi_uy(p, wx, dwzdx, uy, erromax, max_cycles=11, order=0):
    """Run the oracle's poi_uy.  order 0 = reference lexicographic smoother,
    1 = the GPU's colour order.  Returns (uy_out, cycles, residuals)."""
    wx = np.ascontiguousarray(wx, dtype=np.complex128)
    dz = np.ascontiguousarray(dwzdx, dtype=np.complex128)
    out = np.array(uy, dtype=np.complex128, order="C", copy=True)
    cyc = np.zeros(p.nk, dtype=np.int32)
    res = np.zeros(p.nk, dtype=np.float64)
    vp = C.c_void_p
    rc = lib().or_poi_uy(C.byref(p), wx.ctypes.data_as(vp), dz.ctypes.data_as(vp),
                         out.ctypes.data_as(vp), erromax, max_cycles, order,
                         cyc.ctypes.data_as(vp), res.ctypes.data_as(vp))
    if rc:
        raise ValueError(f"or_poi_uy rejected the grid (rc={rc})")
    return out, cyc, res


def poi_uy_modes(p, wx, dwzdx, uy, erromax, modes, max_cycles=11, order=0):
    """or_poi_uy on the listed modes only (0-based mode indices, as the
    fields' first axis).  Returns (uy_out, cycles, residuals) for all modes;
    entries of unlisted modes are the inputs' (cycles 0)."""
    wx = np.ascontiguousarray(wx, dtype=np.complex128)
    dz = np.ascontiguousarray(dwzdx, dtype=np.complex128)
    out = np.array(uy, dtype=np.complex128, order="C", copy=True)
    cyc = np.zeros(p.nk, dtype=np.int32)
    res = np.zeros(p.nk, dtype=np.float64)
    sel = np.ascontiguousarray(np.asarray(modes, dtype=np.int32) + 1)
    vp = C.c_void_p
    rc = lib().or_poi_uy_modes(C.byref(p), wx.ctypes.data_as(vp), dz.ctypes.data_as(vp),
                               out.ctypes.data_as(vp), erromax, max_cycles, order,
                               cyc.ctypes.data_as(vp), res.ctypes.data_as(vp),
                               sel.ctypes.data_as(vp), len(sel))
    if rc:
        raise ValueError(f"or_poi_uy_modes rejected the call (rc={rc})")
    return out, cyc, res


def apply_L(p, k, v, lvl=1):
    """Operator L of mode k (1-based) on level lvl; v complex (ny_l, nx_l)."""
    v = np.ascontiguousarray(v, dtype=np.complex128)
    out = np.zeros_like(v)
    lib().or_apply_L_level(C.byref(p), lvl, k, v.ctypes.data_as(C.c_void_p),
                           out.ctypes.data_as(C.c_void_p))
    return out


def band5_factor(a):
    a = np.array(a, dtype=np.float64, order="C", copy=True)
    n = a.shape[0]
    al = np.zeros((n, 2), dtype=np.float64)
    lib().or_band5_factor(a.ctypes.data_as(C.c_void_p), n, al.ctypes.data_as(C.c_void_p))
    return a, al


def band5_solve(a, al, rhs):
    rhs = np.array(rhs, dtype=np.complex128, copy=True)
    a = np.ascontiguousarray(a)
    al = np.ascontiguousarray(al)
    lib().or_band5_solve(a.ctypes.data_as(C.c_void_p), a.shape[0],
                         al.ctypes.data_as(C.c_void_p), rhs.ctypes.data_as(C.c_void_p))
    return rhs


def metrics(p):
    """create_ctes of the oracle: level sizes, spacings and metrics."""
    m = Metrics()
    if lib().or_create_ctes(C.byref(p), C.byref(m)):
        raise ValueError("or_create_ctes rejected the grid")
    return m


def level_shape(p, lvl):
    """(ny_l, nx_l) of multigrid level lvl (1-based)."""
    m = metrics(p)
    return m.ptsy[lvl - 1], m.ptsx[lvl - 1]


def restrict(p, lvl, vf, rf, vc=None, wc=None, Sc=None):
    """restrict (poisson_par.f:796-874) from level lvl to lvl+1 on one mode:
    vf, rf complex (ny_l, nx_l); returns (vc, wc, Sc) of the coarse level
    (entries the routine does not write keep the given arrays' values, zero
    by default)."""
    m = metrics(p)
    shp = (m.ptsy[lvl], m.ptsx[lvl])
    out = []
    for a in (vc, wc, Sc):
        out.append(np.zeros(shp, dtype=np.complex128) if a is None
                   else np.array(a, dtype=np.complex128, order="C", copy=True))
    vf, pvf = _c(np.asarray(vf, dtype=np.complex128))
    rf, prf = _c(np.asarray(rf, dtype=np.complex128))
    vp = C.c_void_p
    lib().or_restrict(C.byref(p), C.byref(m), lvl, pvf, prf,
                      *[a.ctypes.data_as(vp) for a in out])
    return tuple(out)


def correct_prolong(p, lvl, vc, wc, vf):
    """correction + intpol (poisson_par.f:877-942): lvl is the coarse level;
    returns (vf + P(vc - wc), vc after the correction step)."""
    m = metrics(p)
    vc = np.array(vc, dtype=np.complex128, order="C", copy=True)
    vf = np.array(vf, dtype=np.complex128, order="C", copy=True)
    wc, pwc = _c(np.asarray(wc, dtype=np.complex128))
    vp = C.c_void_p
    lib().or_correct_prolong(C.byref(p), C.byref(m), lvl, vc.ctypes.data_as(vp), pwc,
                             vf.ctypes.data_as(vp))
    return vf, vc


def source(p, k, wx, dwzdx):
    """function_source (poisson_par.f:111-164) of mode k (1-based): wx, dwzdx
    complex (nk, ny, nx); returns S (ny, nx)."""
    wx, pw = _c(np.asarray(wx, dtype=np.complex128))
    dz, pd = _c(np.asarray(dwzdx, dtype=np.complex128))
    S = np.zeros((p.ny, p.nx), dtype=np.complex128)
    lib().or_source(C.byref(p), k, pw, pd, S.ctypes.data_as(C.c_void_p))
    return S


def sor(p, lvl, k, v, S, itera, order=0):
    """itera line-relaxation sweeps (sor, poisson_par.f:167-226) of mode k on
    level lvl; returns the relaxed v."""
    m = metrics(p)
    v = np.array(v, dtype=np.complex128, order="C", copy=True)
    S, pS = _c(np.asarray(S, dtype=np.complex128))
    lib().or_sor(C.byref(p), C.byref(m), lvl, k, v.ctypes.data_as(C.c_void_p), pS, itera, order)
    return v


# ---- the rest of loop (tbl_oracle_loop.c) -----------------------------------
def _cf(a, shape):
    a = np.array(a, dtype=np.complex128, order="C", copy=True)
    assert a.shape == shape, (a.shape, shape)
    return a


def _vp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _rc(rc, name):
    if rc:
        raise ValueError(f"{name} rejected the grid (rc={rc})")


def lhs5(n, k2b2):
    a = np.zeros((n, 5))
    lib().or_lhs5(n, k2b2, _vp(a))
    return a


def ludecomp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    lu = np.zeros_like(a)
    lib().or_ludecomp(a.shape[0], _vp(a), _vp(lu))
    return lu


def lusolve(lu, rhs, variant=0):
    lu = np.ascontiguousarray(lu, dtype=np.float64)
    r = np.array(rhs, dtype=np.complex128, copy=True)
    lib().or_lusolve(lu.shape[0], _vp(lu), _vp(r), 1, variant)
    return r


def poi_ux(p, uy, wy, ux):
    """poi_ux (loop_par.f:428-556): returns (ux, duydy)."""
    shp = _field(p)
    uy, wy, ux = _cf(uy, shp), _cf(wy, shp), _cf(ux, shp)
    duydy = np.zeros(shp, np.complex128)
    _rc(lib().or_poi_ux(C.byref(p), _vp(uy), _vp(wy), _vp(ux), _vp(duydy)), "or_poi_ux")
    return ux, duydy


def poi_uz(p, wy, duydy, uz):
    shp = _field(p)
    wy, du, uz = _cf(wy, shp), _cf(duydy, shp), _cf(uz, shp)
    _rc(lib().or_poi_uz(C.byref(p), _vp(wy), _vp(du), _vp(uz)), "or_poi_uz")
    return uz


def der2wydxdy(p, wy, k):
    wy = _cf(wy, _field(p))
    d = np.zeros(p.nx, np.complex128)
    _rc(lib().or_der2wydxdy(C.byref(p), _vp(wy), k, _vp(d)), "or_der2wydxdy")
    return d


def wx_wall(p, uy, wy, d2uydx2, wx):
    """wx_wall (loop_par.f:174-255): returns (wx, lapv (nk, nx))."""
    shp = _field(p)
    uy, wy, wx = _cf(uy, shp), _cf(wy, shp), _cf(wx, shp)
    d2 = _cf(d2uydx2, (p.nk, p.nx))
    lapv = np.zeros((p.nk, p.nx), np.complex128)
    _rc(lib().or_wx_wall(C.byref(p), _vp(uy), _vp(wy), _vp(d2), _vp(wx), _vp(lapv)),
        "or_wx_wall")
    return wx, lapv


def wz_wall(p, wx, lapv, wz):
    shp = _field(p)
    wx, wz = _cf(wx, shp), _cf(wz, shp)
    lapv = _cf(lapv, (p.nk, p.nx))
    _rc(lib().or_wz_wall(C.byref(p), _vp(wx), _vp(lapv), _vp(wz)), "or_wz_wall")
    return wz


def bzone(p, bdfc, wx, wy, wz, th=None, wall_row=False):
    shp = _field(p)
    f = [_cf(a, shp) for a in (wx, wy, wz)]
    th = None if th is None else _cf(th, shp)
    b = np.ascontiguousarray(bdfc, dtype=np.float64)
    assert b.shape == (p.nx, 2)
    lib().or_bzone(C.byref(p), _vp(b), int(wall_row), *[_vp(a) for a in f], _vp(th))
    return (*f, th)


def filter_trid(p, wx, wy, wz):
    shp = _field(p)
    f = [_cf(a, shp) for a in (wx, wy, wz)]
    _rc(lib().or_filter_trid(C.byref(p), *[_vp(a) for a in f]), "or_filter_trid")
    return tuple(f)


def outuy(p, wz, uy):
    """outuy (loop_par.f:48-119): returns (uy, dwzdx)."""
    shp = _field(p)
    wz, uy = _cf(wz, shp), _cf(uy, shp)
    d = np.zeros(shp, np.complex128)
    _rc(lib().or_outuy(C.byref(p), _vp(wz), _vp(uy), _vp(d)), "or_outuy")
    return uy, d


def bdfc(nx, i0, i3, i4):
    """Buffer-zone table (main_par.f:738-755) as (nx, 2): C view of the
    Fortran bdfc(2, imax)."""
    b = np.zeros((nx, 2))
    lib().or_bdfc(nx, i0, i3, i4, _vp(b))
    return b


FIELDS = ("ux", "uy", "uz", "wx", "wy", "wz", "th")


def loop(p, erro, bdfc_, d2uydx2, fields, order=0):
    """loop(erro) (main_par.f:1762-1785): fields is a dict of the complex
    (nk, ny, nx) arrays ux, uy, uz, wx, wy, wz (th optional: my_form 2).
    Returns (new fields dict, cycles, residuals) of the poi_uy call."""
    shp = _field(p)
    f = {k: _cf(fields[k], shp) for k in FIELDS if fields.get(k) is not None}
    b = np.ascontiguousarray(bdfc_, dtype=np.float64)
    d2 = _cf(d2uydx2, (p.nk, p.nx))
    cyc = np.zeros(p.nk, np.int32)
    res = np.zeros(p.nk)
    _rc(lib().or_loop(C.byref(p), erro, _vp(b), _vp(d2), *[_vp(f.get(k)) for k in FIELDS],
                      order, _vp(cyc), _vp(res)), "or_loop")
    return f, cyc, res



# ---- right-hand side and RK4 step (tbl_oracle_rhs.c) -----------------------
BASE_FIELDS = ("uxb", "uyb", "wzb", "varg", "dvargdx")


def _base(p, base):
    arrs = [np.ascontiguousarray(base[k], dtype=np.float64) for k in BASE_FIELDS]
    for a in arrs:
        assert a.shape == (p.ny, p.nx)
    return Base(*[a.ctypes.data_as(C.c_void_p) for a in arrs]), arrs


def f_to_p(p, kphys, f):
    f = _cf(f, _field(p))
    out = np.zeros((kphys, p.ny, p.nx))
    _rc(lib().or_f_to_p(C.byref(p), kphys, _vp(f), _vp(out)), "or_f_to_p")
    return out


def p_to_f(p, kphys, phys):
    phys = np.ascontiguousarray(phys, dtype=np.float64)
    out = np.zeros(_field(p), np.complex128)
    _rc(lib().or_p_to_f(C.byref(p), kphys, _vp(phys), _vp(out)), "or_p_to_f")
    return out


def terms(p, kphys, base, F):
    """lterms_gv (kphys 0) or nlterms_gv: returns (a, b, c, d)."""
    shp = _field(p)
    b, keep = _base(p, base)
    f = [_cf(F[k], shp) for k in ("ux", "uy", "uz", "wx", "wy", "wz")]
    out = [np.zeros(shp, np.complex128) for _ in range(4)]
    _rc(lib().or_terms(C.byref(p), kphys, C.byref(b), *[_vp(a) for a in f + out]), "or_terms")
    return tuple(out)


def drv_gv(p, Re, kphys, base, F):
    """drv_gv (main_par.f:1339-1403): returns (dvx, dvy, dvz); row 1 zero."""
    shp = _field(p)
    b, keep = _base(p, base)
    f = [_cf(F[k], shp) for k in ("ux", "uy", "uz", "wx", "wy", "wz")]
    out = [np.zeros(shp, np.complex128) for _ in range(3)]
    _rc(lib().or_drv_gv(C.byref(p), Re, kphys, C.byref(b), *[_vp(a) for a in f + out]),
        "or_drv_gv")
    return tuple(out)


def terms_fi(p, kphys, base, F):
    """lterms_fi (kphys 0, base needed) or nlterms_fi (base may be None):
    returns (a, b, c)."""
    shp = _field(p)
    b, keep = _base(p, base) if base is not None else (None, None)
    f = [_cf(F[k], shp) for k in ("ux", "uy", "uz", "wx", "wy", "wz")]
    out = [np.zeros(shp, np.complex128) for _ in range(3)]
    _rc(lib().or_terms_fi(C.byref(p), kphys, C.byref(b) if b is not None else None,
                          *[_vp(a) for a in f + out]), "or_terms_fi")
    return tuple(out)


def drv_fi(p, Re, kphys, base, F):
    """drv_fi (main_par.f:1264-1336): returns (dvx, dvy, dvz); row 1 zero."""
    shp = _field(p)
    b, keep = _base(p, base) if base is not None else (None, None)
    f = [_cf(F[k], shp) for k in ("ux", "uy", "uz", "wx", "wy", "wz")]
    out = [np.zeros(shp, np.complex128) for _ in range(3)]
    _rc(lib().or_drv_fi(C.byref(p), Re, kphys, C.byref(b) if b is not None else None,
                        *[_vp(a) for a in f + out]), "or_drv_fi")
    return tuple(out)


def rk4_step(p, Re, dt, kphys, base, bdfc_, d2uydx2, F, pert=None, order=0):
    """One RK4 time step (main_par.f:266-349).  pert = (pert_uy, pert_d2,
    i0, i1) replaces uy(i,1,k) and d2uydx2(i,k), i0 <= i <= i1 (1-based),
    k >= 2, before the loops of stages 1 and 3.  Returns (fields dict,
    d2uydx2, cycles (4, nk))."""
    shp = _field(p)
    b, keep = _base(p, base)
    f = {k: _cf(F[k], shp) for k in FIELDS if F.get(k) is not None}
    ptrs = (C.c_void_p * 7)(*[_vp(f.get(k)) for k in FIELDS])
    d2 = _cf(d2uydx2, (p.nk, p.nx))
    bd = np.ascontiguousarray(bdfc_, dtype=np.float64)
    cyc = np.zeros((4, p.nk), np.int32)
    if pert is None:
        pu = pd = None
        i0 = i1 = 0
    else:
        pu, pd = _cf(pert[0], (p.nk, p.nx)), _cf(pert[1], (p.nk, p.nx))
        i0, i1 = pert[2], pert[3]
    _rc(lib().or_rk4_step(C.byref(p), Re, dt, kphys, C.byref(b), _vp(bd), _vp(d2), _vp(pu),
                          _vp(pd), i0, i1, ptrs, order, _vp(cyc)), "or_rk4_step")
    return f, d2, cyc
