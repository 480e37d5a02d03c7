"""trans_bl_amd: MI355X-native v-Poisson multigrid and compact derivatives.

Python mirror of the reference's hot-path interface (poi_uy, derparx,
derparxx, dery, deryy, deryfv; poisson_par.f:6, derivs_par.f:80-177) over
the C ABI of ``libtrans_bl_amd.so`` (include/trans_bl_amd.h).  Host entry
points take numpy complex128 arrays of shape (nmodes, ny, nx) -- the C view
of the reference's Fortran (ptsx, jmax, kfour) arrays; ``*_dev`` entry points
take device pointers (for instance ``torch.Tensor.data_ptr()`` of a complex128
CUDA tensor of that shape).

The HIP library is required: there is no CPU fallback.  Importing works on a
machine without a GPU (for the host-only routines); device calls fail loudly
if the library is missing or no GPU is present.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .workload import Grid  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtrans_bl_amd.so")
MAX_LEVELS = 8

STATUS = {0: "OK", 1: "bad argument", 2: "grid not usable by the multigrid", 3: "HIP error",
          4: "allocation failed", 5: "invalid state", 6: "not converged"}


class TblError(RuntimeError):
    pass


class CBase(C.Structure):
    """tbl_base: device pointers of uxb, uyb, wzb, varg, dvargdx."""
    _fields_ = [("uxb", C.c_void_p), ("uyb", C.c_void_p), ("wzb", C.c_void_p),
                ("varg", C.c_void_p), ("dvargdx", C.c_void_p)]


class CGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nmodes", C.c_int), ("levels", C.c_int),
                ("dx", C.c_double), ("dy0", C.c_double), ("stf", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("fac_y", C.c_double)]


class CCoefs(C.Structure):
    _fields_ = [
        ("fp_fd_e", C.c_double * 6), ("sp_fd", C.c_double * 9), ("cp_fd", C.c_double * 8),
        ("pp_fd", C.c_double * 9), ("lp_fd", C.c_double * 7),
        ("fp_sd", C.c_double * 8), ("sp_sd", C.c_double * 10), ("cp_sd", C.c_double * 8),
        ("pp_sd", C.c_double * 10), ("lp_sd", C.c_double * 8),
        ("sp_poi", (C.c_double * 7) * MAX_LEVELS), ("cp_poi", (C.c_double * 8) * MAX_LEVELS),
        ("pp_poi", (C.c_double * 4) * MAX_LEVELS), ("lp_poi", (C.c_double * 5) * MAX_LEVELS),
        ("w_at_w", C.c_double * 9), ("dwydy", C.c_double * 8),
    ]


class SolveOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("max_cycles", C.c_int), ("nu1", C.c_int),
                ("nu2", C.c_int), ("ncoarse", C.c_int), ("use_graph", C.c_int),
                ("guess", C.c_int), ("fmg", C.c_int), ("reserved", C.c_int * 4)]


class SolveStats(C.Structure):
    _fields_ = [("cycles_max", C.c_int), ("cycles_total", C.c_int), ("nconverged", C.c_int),
                ("fallback", C.c_int), ("residual_max", C.c_double),
                ("guess_residual_max", C.c_double), ("device_ms", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# every extern "C" symbol include/trans_bl_amd.h declares
EXPORTS = [
    "tbl_version", "tbl_last_error", "tbl_coefs_generate", "tbl_coefs_read", "tbl_create", "tbl_destroy",
    "tbl_set_stream", "tbl_synchronize", "tbl_poi_uy", "tbl_poi_uy_dev", "tbl_solve_rhs_dev", "tbl_derparx",
    "tbl_derparxx", "tbl_dery", "tbl_deryy", "tbl_deryfv", "tbl_deriv_dev", "tbl_derivs_dev", "tbl_residual_dev",
    "tbl_residual_dev_noalloc", "tbl_set_line_shape", "tbl_get_line_shape", "tbl_time_smoother",
    "tbl_net_create", "tbl_net_destroy", "tbl_net_num_params", "tbl_net_get_params",
    "tbl_net_set_params", "tbl_net_forward", "tbl_net_gradients", "tbl_net_train_step",
    "tbl_net_synchronize", "tbl_mfma_gemm_bf16", "tbl_guess_attach", "tbl_guess_shape",
    "tbl_guess_pairs", "tbl_batch_create", "tbl_batch_destroy", "tbl_batch_context",
    "tbl_batch_stage", "tbl_batch_stage_converge", "tbl_batch_results", "tbl_checkpoint_write", "tbl_checkpoint_read",
    "tbl_baseflow_write", "tbl_baseflow_read", "tbl_poi_ux_dev", "tbl_poi_uz_dev",
    "tbl_wx_wall_dev", "tbl_wz_wall_dev", "tbl_bzone_dev", "tbl_filter_trid_dev",
    "tbl_outuy_dev", "tbl_loop_dev", "tbl_loop", "tbl_bdfc_generate", "tbl_terms_dev",
    "tbl_drv_gv_dev", "tbl_drv_gv", "tbl_terms_fi_dev", "tbl_drv_fi_dev", "tbl_drv_fi",
    "tbl_rk4_step_dev",
]

_lib = None


def lib():
    """Load libtrans_bl_amd.so (fails loudly when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TblError(f"{LIB_PATH} is missing: run `python -m trans_bl_amd.build` first")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.c_int, C.c_double
    sig = [vp, vp, vp, vp, dp, C.POINTER(SolveOpts), vp, vp, C.POINTER(SolveStats)]
    types = {
        "tbl_version": ([], C.c_char_p),
        "tbl_last_error": ([], C.c_char_p),
        "tbl_coefs_generate": ([dp, ip, C.POINTER(CCoefs)], C.c_int),
        "tbl_coefs_read": ([C.c_char_p, ip, C.POINTER(CCoefs)], C.c_int),
        "tbl_create": ([C.POINTER(CGrid), C.POINTER(CCoefs), C.POINTER(vp)], C.c_int),
        "tbl_destroy": ([vp], C.c_int),
        "tbl_set_stream": ([vp, vp], C.c_int),
        "tbl_synchronize": ([vp], C.c_int),
        "tbl_poi_uy": (sig, C.c_int),
        "tbl_poi_uy_dev": (sig, C.c_int),
        "tbl_solve_rhs_dev": ([vp, vp, vp, dp, C.POINTER(SolveOpts), vp, vp,
                               C.POINTER(SolveStats)], C.c_int),
        "tbl_deriv_dev": ([vp, ip, vp, vp], C.c_int),
        "tbl_derivs_dev": ([vp, vp, vp, vp, vp, vp], C.c_int),
        "tbl_residual_dev": ([vp, vp, vp, vp, vp, vp], C.c_int),
        "tbl_residual_dev_noalloc": ([vp, vp, vp, vp, vp, vp], C.c_int),
        "tbl_set_line_shape": ([vp, ip, ip, ip], C.c_int),
        "tbl_get_line_shape": ([vp, ip, ip], C.c_int),
        "tbl_time_smoother": ([vp, ip, ip, ip, C.POINTER(dp), C.POINTER(ip)], C.c_int),
        "tbl_poi_ux_dev": ([vp, vp, vp, vp, vp], C.c_int),
        "tbl_poi_uz_dev": ([vp, vp, vp, vp], C.c_int),
        "tbl_wx_wall_dev": ([vp, vp, vp, vp, vp, vp], C.c_int),
        "tbl_wz_wall_dev": ([vp, vp, vp, vp], C.c_int),
        "tbl_bzone_dev": ([vp, vp, ip, vp, vp, vp, vp], C.c_int),
        "tbl_filter_trid_dev": ([vp, vp, vp, vp], C.c_int),
        "tbl_outuy_dev": ([vp, vp, vp, vp], C.c_int),
        "tbl_loop_dev": ([vp, dp, vp, vp] + [vp] * 7 + [C.POINTER(SolveOpts), vp, vp,
                                                     C.POINTER(SolveStats)], C.c_int),
        "tbl_loop": ([vp, dp, vp, vp] + [vp] * 7 + [C.POINTER(SolveOpts), vp, vp,
                                                 C.POINTER(SolveStats)], C.c_int),
        "tbl_bdfc_generate": ([ip, ip, ip, ip, vp], C.c_int),
        "tbl_terms_dev": ([vp, ip, C.POINTER(CBase)] + [vp] * 10, C.c_int),
        "tbl_drv_gv_dev": ([vp, dp, ip, C.POINTER(CBase)] + [vp] * 9, C.c_int),
        "tbl_drv_gv": ([vp, dp, ip] + [vp] * 14, C.c_int),
        "tbl_terms_fi_dev": ([vp, ip, C.POINTER(CBase)] + [vp] * 9, C.c_int),
        "tbl_drv_fi_dev": ([vp, dp, ip, C.POINTER(CBase)] + [vp] * 9, C.c_int),
        "tbl_drv_fi": ([vp, dp, ip] + [vp] * 12, C.c_int),
        "tbl_rk4_step_dev": ([vp, dp, dp, ip, C.POINTER(CBase), vp, vp, vp, vp, ip, ip]
                             + [vp] * 7 + [C.POINTER(SolveOpts), vp, vp], C.c_int),
    }
    for n in ("tbl_derparx", "tbl_derparxx", "tbl_dery", "tbl_deryy", "tbl_deryfv"):
        types[n] = ([vp, vp, vp], C.c_int)
    # a symbol missing from an older build fails when it is called, not here
    for name, (args, res) in types.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        msg = lib().tbl_last_error().decode(errors="replace")
        raise TblError(f"{what} failed: {STATUS.get(rc, rc)} {msg}".strip())


def coefs_generate(stf: float, levels: int) -> dict:
    """Coefficient tables (the reference's coefs.bin content) from the
    library's host generator."""
    c = CCoefs()
    _check(lib().tbl_coefs_generate(stf, levels, C.byref(c)), "tbl_coefs_generate")
    return coefs_to_dict(c, levels)


def coefs_read(path: str, levels: int) -> dict:
    """Tables of a coefs.bin file (the reference's coefficient file)."""
    c = CCoefs()
    _check(lib().tbl_coefs_read(path.encode(), levels, C.byref(c)), "tbl_coefs_read")
    return coefs_to_dict(c, levels)


def checkpoint_write(path: str, t: int, fields) -> None:
    """Write the reference's checkpoint file data_RR.bin (escreve,
    escreve_par.f:6-33): t and 6 (or 7) complex fields of shape (nmodes, ny, nx)."""
    arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in fields]
    nk, ny, nx = arrs[0].shape
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    L = lib()
    L.tbl_checkpoint_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_void_p]
    _check(L.tbl_checkpoint_write(path.encode(), nx, ny, nk, t, len(arrs), ptrs),
           "tbl_checkpoint_write")


def checkpoint_read(path: str, nx: int, ny: int, nmodes: int, nfields: int = 6):
    """Read data_RR.bin: (t, list of nfields complex (nmodes, ny, nx) arrays)."""
    arrs = [np.zeros((nmodes, ny, nx), np.complex128) for _ in range(nfields)]
    ptrs = (C.c_void_p * nfields)(*[a.ctypes.data for a in arrs])
    t = C.c_int()
    L = lib()
    L.tbl_checkpoint_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(C.c_int), C.c_int, C.c_void_p]
    _check(L.tbl_checkpoint_read(path.encode(), nx, ny, nmodes, C.byref(t), nfields, ptrs),
           "tbl_checkpoint_read")
    return t.value, arrs


def bdfc_generate(nx: int, i0: int, i3: int, i4: int):
    """Buffer-zone table (main_par.f:738-755) as (nx, 2): the C view of the
    Fortran bdfc(2, imax)."""
    b = np.zeros((nx, 2), np.float64)
    _check(lib().tbl_bdfc_generate(nx, i0, i3, i4, _ptr(b)), "tbl_bdfc_generate")
    return b


def baseflow_write(path: str, ux, uy, wz) -> None:
    """Write based_RR.bin (ux, uy, wz real (ny, nx))."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (ux, uy, wz)]
    ny, nx = arrs[0].shape
    L = lib()
    L.tbl_baseflow_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
    _check(L.tbl_baseflow_write(path.encode(), nx, ny, *[a.ctypes.data for a in arrs]),
           "tbl_baseflow_write")


def baseflow_read(path: str, nx: int, ny: int):
    """Read based_RR.bin: (ux, uy, wz) real (ny, nx)."""
    arrs = [np.zeros((ny, nx), np.float64) for _ in range(3)]
    L = lib()
    L.tbl_baseflow_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p]
    _check(L.tbl_baseflow_read(path.encode(), nx, ny, *[a.ctypes.data for a in arrs]),
           "tbl_baseflow_read")
    return tuple(arrs)


def coefs_to_dict(c: CCoefs, levels: int) -> dict:
    out = {}
    for name, _ in CCoefs._fields_:
        arr = np.ctypeslib.as_array(getattr(c, name)).copy()
        if name.endswith("_poi"):
            arr = arr[:levels]
        out[name] = arr
    return out


def coefs_from_dict(d: dict) -> CCoefs:
    c = CCoefs()
    for name, _ in CCoefs._fields_:
        if name not in d:
            continue
        src = np.asarray(d[name], dtype=np.float64)
        dst = np.ctypeslib.as_array(getattr(c, name))
        if name.endswith("_poi"):
            dst[:src.shape[0], :src.shape[1]] = src
        else:
            dst[:src.shape[0]] = src
    return c


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _cbase(base):
    """tbl_base of device pointers (missing keys: NULL), or None."""
    if base is None:
        return None
    return CBase(*[C.c_void_p(base[k]) if base.get(k) else None
                   for k in ("uxb", "uyb", "wzb", "varg", "dvargdx")])


class Context:
    """One grid's solver state on the current GPU (runtime par.for + the
    COMMON blocks the hot path reads)."""

    def __init__(self, grid: Grid, coefs: dict | None = None):
        self.grid = grid
        g = CGrid(grid.nx, grid.ny, grid.nmodes, grid.levels, grid.dx, grid.dy0, grid.stf,
                  grid.alpha, grid.beta, grid.fac_y)
        self._cc = coefs_from_dict(coefs) if coefs is not None else None
        h = C.c_void_p()
        _check(lib().tbl_create(C.byref(g), C.byref(self._cc) if self._cc else None,
                                C.byref(h)), "tbl_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().tbl_destroy(self.h)
            self.h = None
        self._guess_net = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def shape(self):
        return (self.grid.nmodes, self.grid.ny, self.grid.nx)

    def _arr(self, a, name):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != self.shape:
            raise ValueError(f"{name} must have shape {self.shape}, got {a.shape}")
        return a

    @staticmethod
    def opts(precision=0, max_cycles=11, nu1=0, nu2=0, ncoarse=0, use_graph=0, guess=0, fmg=0):
        return SolveOpts(precision, max_cycles, nu1, nu2, ncoarse, use_graph, guess, fmg)

    # ---- the reference's subroutines on host arrays ----------------------
    def poi_uy(self, wx, dwzdx, uy, erromax, **kw):
        """poi_uy(dwzdx, erromax) with the COMMON inputs wx, uy explicit.
        Returns (uy_new, cycles, residuals, stats)."""
        wx = None if wx is None else self._arr(wx, "wx")
        dz = self._arr(dwzdx, "dwzdx")
        out = np.array(self._arr(uy, "uy"), copy=True)
        cyc = np.zeros(self.grid.nmodes, np.int32)
        res = np.zeros(self.grid.nmodes, np.float64)
        st = SolveStats()
        o = self.opts(**kw)
        _check(lib().tbl_poi_uy(self.h, None if wx is None else _ptr(wx), _ptr(dz), _ptr(out),
                                erromax, C.byref(o), _ptr(cyc), _ptr(res), C.byref(st)),
               "tbl_poi_uy")
        return out, cyc, res, st.as_dict()

    def _deriv_host(self, fn, f):
        f = self._arr(f, "fc")
        out = np.zeros_like(f)
        _check(getattr(lib(), fn)(self.h, _ptr(f), _ptr(out)), fn)
        return out

    def derparx(self, f):
        return self._deriv_host("tbl_derparx", f)

    def derparxx(self, f):
        return self._deriv_host("tbl_derparxx", f)

    def dery(self, f):
        return self._deriv_host("tbl_dery", f)

    def deryy(self, f):
        return self._deriv_host("tbl_deryy", f)

    def deryfv(self, f):
        return self._deriv_host("tbl_deryfv", f)

    # ---- device-pointer entry points --------------------------------------
    def set_stream(self, handle):
        _check(lib().tbl_set_stream(self.h, C.c_void_p(handle) if handle else None),
               "tbl_set_stream")

    def synchronize(self):
        _check(lib().tbl_synchronize(self.h), "tbl_synchronize")

    def solve_rhs_dev(self, f_ptr, uy_ptr, erromax, **kw):
        cyc = np.zeros(self.grid.nmodes, np.int32)
        res = np.zeros(self.grid.nmodes, np.float64)
        st = SolveStats()
        o = self.opts(**kw)
        _check(lib().tbl_solve_rhs_dev(self.h, C.c_void_p(f_ptr), C.c_void_p(uy_ptr), erromax,
                                       C.byref(o), _ptr(cyc), _ptr(res), C.byref(st)),
               "tbl_solve_rhs_dev")
        return cyc, res, st.as_dict()

    def poi_uy_dev(self, wx_ptr, dwzdx_ptr, uy_ptr, erromax, **kw):
        cyc = np.zeros(self.grid.nmodes, np.int32)
        res = np.zeros(self.grid.nmodes, np.float64)
        st = SolveStats()
        o = self.opts(**kw)
        _check(lib().tbl_poi_uy_dev(self.h, C.c_void_p(wx_ptr) if wx_ptr else None,
                                    C.c_void_p(dwzdx_ptr), C.c_void_p(uy_ptr), erromax,
                                    C.byref(o), _ptr(cyc), _ptr(res), C.byref(st)),
               "tbl_poi_uy_dev")
        return cyc, res, st.as_dict()

    def deriv_dev(self, kind, in_ptr, out_ptr):
        _check(lib().tbl_deriv_dev(self.h, kind, C.c_void_p(in_ptr), C.c_void_p(out_ptr)),
               "tbl_deriv_dev")

    def derivs_dev(self, in_ptr, dx_ptr=0, dxx_ptr=0, dy_ptr=0, dyy_ptr=0):
        """d/dx, d2/dx2, d/dy, d2/dy2 of one field (device pointers; 0 skips
        an output)."""
        p = [C.c_void_p(x) if x else None for x in (dx_ptr, dxx_ptr, dy_ptr, dyy_ptr)]
        _check(lib().tbl_derivs_dev(self.h, C.c_void_p(in_ptr), *p), "tbl_derivs_dev")

    def residual_dev(self, wx_ptr, dwzdx_ptr, uy_ptr, out_ptr):
        res = np.zeros(self.grid.nmodes, np.float64)
        _check(lib().tbl_residual_dev(self.h, C.c_void_p(wx_ptr) if wx_ptr else None,
                                      C.c_void_p(dwzdx_ptr), C.c_void_p(uy_ptr),
                                      C.c_void_p(out_ptr), _ptr(res)), "tbl_residual_dev")
        return res

    def residual_dev_noalloc(self, wx_ptr, dwzdx_ptr, uy_ptr, out_ptr):
        """residual_dev without device allocation (the library's fixed work
        array)."""
        res = np.zeros(self.grid.nmodes, np.float64)
        _check(lib().tbl_residual_dev_noalloc(self.h, C.c_void_p(wx_ptr) if wx_ptr else None,
                                              C.c_void_p(dwzdx_ptr), C.c_void_p(uy_ptr),
                                              C.c_void_p(out_ptr), _ptr(res)),
               "tbl_residual_dev_noalloc")
        return res

    # ---- the rest of the RK-stage loop (device pointers, user layout) ----
    def poi_ux_dev(self, uy, wy, ux, duydy):
        _check(lib().tbl_poi_ux_dev(self.h, *[C.c_void_p(x) for x in (uy, wy, ux, duydy)]),
               "tbl_poi_ux_dev")

    def poi_uz_dev(self, wy, duydy, uz):
        _check(lib().tbl_poi_uz_dev(self.h, *[C.c_void_p(x) for x in (wy, duydy, uz)]),
               "tbl_poi_uz_dev")

    def wx_wall_dev(self, uy, wy, d2uydx2, wx, lapv):
        _check(lib().tbl_wx_wall_dev(self.h, *[C.c_void_p(x)
                                               for x in (uy, wy, d2uydx2, wx, lapv)]),
               "tbl_wx_wall_dev")

    def wz_wall_dev(self, wx, lapv, wz):
        _check(lib().tbl_wz_wall_dev(self.h, *[C.c_void_p(x) for x in (wx, lapv, wz)]),
               "tbl_wz_wall_dev")

    def bzone_dev(self, bdfc, wx, wy, wz, th=0, wall_row=False):
        _check(lib().tbl_bzone_dev(self.h, C.c_void_p(bdfc), int(wall_row),
                                   *[C.c_void_p(x) if x else None for x in (wx, wy, wz, th)]),
               "tbl_bzone_dev")

    def filter_trid_dev(self, wx, wy, wz):
        _check(lib().tbl_filter_trid_dev(self.h, *[C.c_void_p(x) for x in (wx, wy, wz)]),
               "tbl_filter_trid_dev")

    def outuy_dev(self, wz, uy, dwzdx):
        _check(lib().tbl_outuy_dev(self.h, *[C.c_void_p(x) for x in (wz, uy, dwzdx)]),
               "tbl_outuy_dev")

    def loop_dev(self, erro, bdfc, d2uydx2, ux, uy, uz, wx, wy, wz, th=0, **kw):
        """loop(erro) on device fields; returns (cycles, residuals, stats)
        of its poi_uy call."""
        cyc = np.zeros(self.grid.nmodes, np.int32)
        res = np.zeros(self.grid.nmodes, np.float64)
        st = SolveStats()
        o = self.opts(**kw)
        ptrs = [C.c_void_p(x) if x else None for x in (ux, uy, uz, wx, wy, wz, th)]
        _check(lib().tbl_loop_dev(self.h, erro, C.c_void_p(bdfc), C.c_void_p(d2uydx2), *ptrs,
                                  C.byref(o), _ptr(cyc), _ptr(res), C.byref(st)),
               "tbl_loop_dev")
        return cyc, res, st.as_dict()

    def loop(self, erro, bdfc, d2uydx2, fields: dict, **kw):
        """loop(erro) on host arrays: fields is a dict of complex (nk, ny, nx)
        arrays ux, uy, uz, wx, wy, wz (th optional), updated in place."""
        arrs = {}
        for k in ("ux", "uy", "uz", "wx", "wy", "wz", "th"):
            a = fields.get(k)
            if a is None:
                continue
            if not (a.dtype == np.complex128 and a.flags.c_contiguous):
                raise TblError(f"field {k} must be a C-contiguous complex128 array")
            arrs[k] = a
        b = np.ascontiguousarray(bdfc, dtype=np.float64)
        d2 = np.ascontiguousarray(d2uydx2, dtype=np.complex128)
        cyc = np.zeros(self.grid.nmodes, np.int32)
        res = np.zeros(self.grid.nmodes, np.float64)
        st = SolveStats()
        o = self.opts(**kw)
        ptrs = [_ptr(arrs[k]) if k in arrs else None
                for k in ("ux", "uy", "uz", "wx", "wy", "wz", "th")]
        _check(lib().tbl_loop(self.h, erro, _ptr(b), _ptr(d2), *ptrs, C.byref(o), _ptr(cyc),
                              _ptr(res), C.byref(st)), "tbl_loop")
        return cyc, res, st.as_dict()

    # ---- right-hand side and RK4 step (device pointers) ----
    def terms_dev(self, kphys, base, ux, uy, uz, wx, wy, wz, a, b, c, d):
        """base: dict of device pointers uxb, uyb, wzb, varg, dvargdx."""
        cb = CBase(*[C.c_void_p(base[k]) for k in ("uxb", "uyb", "wzb", "varg", "dvargdx")])
        _check(lib().tbl_terms_dev(self.h, kphys, C.byref(cb),
                                   *[C.c_void_p(x) for x in (ux, uy, uz, wx, wy, wz, a, b, c, d)]),
               "tbl_terms_dev")

    def drv_gv_dev(self, Re, kphys, base, ux, uy, uz, wx, wy, wz, dvx, dvy, dvz):
        cb = CBase(*[C.c_void_p(base[k]) for k in ("uxb", "uyb", "wzb", "varg", "dvargdx")])
        _check(lib().tbl_drv_gv_dev(self.h, Re, kphys, C.byref(cb),
                                    *[C.c_void_p(x) for x in (ux, uy, uz, wx, wy, wz, dvx, dvy,
                                                              dvz)]), "tbl_drv_gv_dev")

    def drv_gv(self, Re, kphys, base: dict, fields: dict):
        """drv_gv on host arrays: base is a dict of real (ny, nx) arrays uxb,
        uyb, wzb, varg, dvargdx; fields of complex (nk, ny, nx) arrays ux ..
        wz.  Returns dvx, dvy, dvz (row 1 zero: drv_gv does not write it)."""
        g = self.grid
        shp = (g.nmodes, g.ny, g.nx)
        bs = [np.ascontiguousarray(base[k], dtype=np.float64)
              for k in ("uxb", "uyb", "wzb", "varg", "dvargdx")]
        fs = [np.ascontiguousarray(fields[k], dtype=np.complex128)
              for k in ("ux", "uy", "uz", "wx", "wy", "wz")]
        if any(b.shape != shp[1:] for b in bs) or any(f.shape != shp for f in fs):
            raise TblError(f"drv_gv: base (ny, nx) and fields {shp} expected")
        out = [np.zeros(shp, np.complex128) for _ in range(3)]
        _check(lib().tbl_drv_gv(self.h, Re, kphys, *[_ptr(a) for a in bs + fs + out]),
               "tbl_drv_gv")
        return tuple(out)

    def terms_fi_dev(self, kphys, base, ux, uy, uz, wx, wy, wz, a, b, c):
        """Flat-plate terms (lterms_fi for kphys 0, else nlterms_fi) on device
        pointers; base: dict of device pointers uxb, uyb, wzb, or None."""
        cb = _cbase(base)
        _check(lib().tbl_terms_fi_dev(self.h, kphys, C.byref(cb) if cb else None,
                                      *[C.c_void_p(x) for x in (ux, uy, uz, wx, wy, wz, a, b,
                                                                c)]), "tbl_terms_fi_dev")

    def drv_fi_dev(self, Re, kphys, base, ux, uy, uz, wx, wy, wz, dvx, dvy, dvz):
        cb = _cbase(base)
        _check(lib().tbl_drv_fi_dev(self.h, Re, kphys, C.byref(cb) if cb else None,
                                    *[C.c_void_p(x) for x in (ux, uy, uz, wx, wy, wz, dvx, dvy,
                                                              dvz)]), "tbl_drv_fi_dev")

    def drv_fi(self, Re, kphys, base, fields: dict):
        """drv_fi on host arrays: base (for kphys 0) a dict of real (ny, nx)
        arrays uxb, uyb, wzb, or None; returns dvx, dvy, dvz (row 1 zero)."""
        g = self.grid
        shp = (g.nmodes, g.ny, g.nx)
        bs = [np.ascontiguousarray(base[k], dtype=np.float64) for k in ("uxb", "uyb", "wzb")] \
            if base is not None else []
        fs = [np.ascontiguousarray(fields[k], dtype=np.complex128)
              for k in ("ux", "uy", "uz", "wx", "wy", "wz")]
        if any(b.shape != shp[1:] for b in bs) or any(f.shape != shp for f in fs):
            raise TblError(f"drv_fi: base (ny, nx) and fields {shp} expected")
        out = [np.zeros(shp, np.complex128) for _ in range(3)]
        bp = [_ptr(a) for a in bs] if bs else [None] * 3
        _check(lib().tbl_drv_fi(self.h, Re, kphys, *bp, *[_ptr(a) for a in fs + out]),
               "tbl_drv_fi")
        return tuple(out)

    def rk4_step_dev(self, Re, dt, kphys, base, bdfc, d2uydx2, fields, pert=None, **kw):
        """One RK4 time step on device fields (dict of pointers ux .. wz, th
        optional); pert = (pert_uy_ptr, pert_d2_ptr, i0, i1) or None.
        Returns cycles (4, nmodes) and residuals (4, nmodes)."""
        cb = CBase(*[C.c_void_p(base[k]) for k in ("uxb", "uyb", "wzb", "varg", "dvargdx")])
        cyc = np.zeros((4, self.grid.nmodes), np.int32)
        res = np.zeros((4, self.grid.nmodes), np.float64)
        pu, pd, i0, i1 = pert if pert is not None else (0, 0, 0, 0)
        o = self.opts(**kw)
        ptrs = [C.c_void_p(fields[k]) if fields.get(k) else None
                for k in ("ux", "uy", "uz", "wx", "wy", "wz", "th")]
        _check(lib().tbl_rk4_step_dev(self.h, Re, dt, kphys, C.byref(cb), C.c_void_p(bdfc),
                                      C.c_void_p(d2uydx2), C.c_void_p(pu) if pu else None,
                                      C.c_void_p(pd) if pd else None, i0, i1, *ptrs,
                                      C.byref(o), _ptr(cyc), _ptr(res)), "tbl_rk4_step_dev")
        return cyc, res

    def set_line_shape(self, prec, level, lanes):
        _check(lib().tbl_set_line_shape(self.h, prec, level, lanes), "tbl_set_line_shape")

    def time_smoother(self, prec, level, sweeps):
        """(elapsed ms, kernel launches) of `sweeps` sweeps on one level."""
        ms = C.c_double()
        n = C.c_int()
        _check(lib().tbl_time_smoother(self.h, prec, level, sweeps, C.byref(ms), C.byref(n)),
               "tbl_time_smoother")
        return ms.value, n.value

    def line_shape(self, prec, level):
        return lib().tbl_get_line_shape(self.h, prec, level)


DERIV_KINDS = {"x": 0, "xx": 1, "y": 2, "yy": 3, "yfv": 4}
