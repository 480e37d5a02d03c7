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
        L.or_apply_L_level.argtypes = [P, C.c_int, C.c_int, vp, vp]
        L.or_apply_L_level.restype = C.c_int
        L.or_band5_factor.argtypes = [vp, C.c_int, vp]
        L.or_band5_solve.argtypes = [vp, C.c_int, vp, vp]
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
