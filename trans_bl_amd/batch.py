"""Uneven batches of independent problems and the captured RK4 stage
(tbl_batch_* of the C ABI)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import CGrid, Context, lib, _check


def _setup():
    L = lib()
    vp, ip, dp = C.c_void_p, C.c_int, C.c_double
    types = {
        "tbl_batch_create": ([ip, C.POINTER(CGrid), C.POINTER(vp)], C.c_int),
        "tbl_batch_destroy": ([vp], C.c_int),
        "tbl_batch_context": ([vp, ip], vp),
        "tbl_batch_stage": ([vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), dp, ip, ip, ip,
                             C.POINTER(dp)], C.c_int),
        "tbl_batch_stage_converge": ([vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), dp, vp,
                                      C.POINTER(dp), C.POINTER(ip)], C.c_int),
        "tbl_batch_results": ([vp, ip, vp, vp], C.c_int),
    }
    for name, (args, res) in types.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
    return L


class _Borrowed(Context):
    """A batch-owned context (not destroyed by Python)."""

    def __init__(self, grid, handle):  # noqa: D401 - no tbl_create here
        self.grid = grid
        self.h = handle

    def close(self):
        self.h = None


class Batch:
    def __init__(self, grids):
        self.L = _setup()
        self.grids = list(grids)
        arr = (CGrid * len(self.grids))(*[
            CGrid(g.nx, g.ny, g.nmodes, g.levels, g.dx, g.dy0, g.stf, g.alpha, g.beta, g.fac_y)
            for g in self.grids])
        h = C.c_void_p()
        _check(self.L.tbl_batch_create(len(self.grids), arr, C.byref(h)), "tbl_batch_create")
        self.h = h
        self.contexts = [_Borrowed(g, C.c_void_p(self.L.tbl_batch_context(h, p)))
                         for p, g in enumerate(self.grids)]

    def close(self):
        if getattr(self, "h", None):
            self.L.tbl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stage(self, f_ptrs, v_ptrs, d_ptrs, tol, cycles, precision=0, use_graph=1):
        n = len(self.grids)
        fa = (C.c_void_p * n)(*f_ptrs)
        va = (C.c_void_p * n)(*v_ptrs)
        da = (C.c_void_p * (4 * n))(*d_ptrs)
        ms = C.c_double()
        _check(self.L.tbl_batch_stage(self.h, fa, va, da, tol, cycles, precision, use_graph,
                                      C.byref(ms)), "tbl_batch_stage")
        return ms.value

    def stage_converge(self, f_ptrs, v_ptrs, d_ptrs, tol, max_cycles=12, **form):
        """The stage solved to tol (tbl_batch_stage_converge); form: precision,
        nu1, nu2, ncoarse, fmg.  Returns (ms, cycles of the slowest problem)."""
        n = len(self.grids)
        fa = (C.c_void_p * n)(*f_ptrs)
        va = (C.c_void_p * n)(*v_ptrs)
        da = (C.c_void_p * (4 * n))(*d_ptrs)
        o = Context.opts(max_cycles=max_cycles, **form)
        ms, run = C.c_double(), C.c_int()
        _check(self.L.tbl_batch_stage_converge(self.h, fa, va, da, tol, C.byref(o), C.byref(ms),
                                               C.byref(run)), "tbl_batch_stage_converge")
        return ms.value, run.value

    def results(self, p):
        K = self.grids[p].nmodes
        cyc = np.zeros(K, np.int32)
        res = np.zeros(K, np.float64)
        _check(self.L.tbl_batch_results(self.h, p, cyc.ctypes.data_as(C.c_void_p),
                                        res.ctypes.data_as(C.c_void_p)), "tbl_batch_results")
        return cyc, res
