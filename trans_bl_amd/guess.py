"""Learned initial guess: the BF16 U-Net (tbl_net_* of the C ABI), its
training on (RHS, v) pairs that the solver itself generates, and the solve
that starts from the network's prediction.

The network runs on one coarse multigrid level (default: the coarsest); its
prediction is prolongated to the fine grid and multigrid runs to the same
tolerance as without it (guess=3), or, with guess=2, only modes whose
prediction passes a residual test on that level start from it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import lib, _check, Context

HERE = os.path.dirname(os.path.abspath(__file__))


def _setup():
    L = lib()
    vp, ip, fp = C.c_void_p, C.c_int, C.c_float
    types = {
        "tbl_net_create": ([ip, ip, ip, C.c_uint, C.POINTER(vp)], C.c_int),
        "tbl_net_destroy": ([vp], C.c_int),
        "tbl_net_num_params": ([vp], C.c_long),
        "tbl_net_get_params": ([vp, vp], C.c_int),
        "tbl_net_set_params": ([vp, vp], C.c_int),
        "tbl_net_forward": ([vp, vp, vp, ip], C.c_int),
        "tbl_net_gradients": ([vp, vp, vp, ip, C.POINTER(fp), vp, vp], C.c_int),
        "tbl_net_train_step": ([vp, vp, vp, ip, fp, C.POINTER(fp)], C.c_int),
        "tbl_net_synchronize": ([vp], C.c_int),
        "tbl_mfma_gemm_bf16": ([ip, ip, ip, vp, vp, vp, ip, ip, ip], C.c_int),
        "tbl_guess_attach": ([vp, vp, ip], C.c_int),
        "tbl_guess_shape": ([vp, C.POINTER(ip), C.POINTER(ip)], C.c_int),
        "tbl_guess_pairs": ([vp, vp, vp, vp, vp], C.c_int),
    }
    for name, (args, res) in types.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
    return L


class Net:
    """BF16 U-Net on H x W images (multiples of 8), batches up to max_batch."""

    def __init__(self, H: int, W: int, max_batch: int, seed: int = 0):
        self.L = _setup()
        self.H, self.W, self.max_batch = H, W, max_batch
        h = C.c_void_p()
        _check(self.L.tbl_net_create(H, W, max_batch, seed, C.byref(h)), "tbl_net_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.tbl_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        return int(self.L.tbl_net_num_params(self.h))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        _check(self.L.tbl_net_get_params(self.h, out.ctypes.data_as(C.c_void_p)), "get_params")
        return out

    def set_params(self, theta) -> None:
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        assert theta.size == self.num_params
        _check(self.L.tbl_net_set_params(self.h, theta.ctypes.data_as(C.c_void_p)), "set_params")

    def forward(self, x_ptr: int, y_ptr: int, n: int) -> None:
        _check(self.L.tbl_net_forward(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), n),
               "tbl_net_forward")

    def gradients(self, x_ptr: int, t_ptr: int, n: int, xgrad_ptr: int = 0):
        loss = C.c_float()
        g = np.zeros(self.num_params, np.float32)
        _check(self.L.tbl_net_gradients(self.h, C.c_void_p(x_ptr), C.c_void_p(t_ptr), n,
                                        C.byref(loss), g.ctypes.data_as(C.c_void_p),
                                        C.c_void_p(xgrad_ptr) if xgrad_ptr else None),
               "tbl_net_gradients")
        return loss.value, g

    def train_step(self, x_ptr: int, t_ptr: int, n: int, lr: float) -> float:
        loss = C.c_float()
        _check(self.L.tbl_net_train_step(self.h, C.c_void_p(x_ptr), C.c_void_p(t_ptr), n, lr,
                                         C.byref(loss)), "tbl_net_train_step")
        return loss.value

    def save(self, path: str) -> None:
        hdr = np.array([0x54424C4E, self.H, self.W, self.num_params], dtype=np.int64)
        with open(path, "wb") as f:
            hdr.tofile(f)
            self.get_params().tofile(f)

    def load(self, path: str) -> None:
        with open(path, "rb") as f:
            hdr = np.fromfile(f, dtype=np.int64, count=4)
            if hdr[0] != 0x54424C4E or hdr[3] != self.num_params:
                raise ValueError(f"{path}: not a matching network file")
            self.set_params(np.fromfile(f, dtype=np.float32, count=int(hdr[3])))


def mfma_gemm(a_ptr, b_ptr, c_ptr, M, N, K, transpose_a=0, transpose_b=0, splitk=1):
    L = _setup()
    _check(L.tbl_mfma_gemm_bf16(M, N, K, C.c_void_p(a_ptr), C.c_void_p(b_ptr),
                                C.c_void_p(c_ptr), transpose_a, transpose_b, splitk),
           "tbl_mfma_gemm_bf16")


def default_level(ctx: Context) -> int:
    """The coarsest multigrid level: the smooth part of v is what the
    network must supply (the V-cycles remove the rest), and the image there
    is 4**(levels-1) times smaller than the fine grid."""
    return ctx.grid.levels


def attach(ctx: Context, net: Net, level: int) -> None:
    """Attach `net` to `ctx` on multigrid `level`.  The context keeps a
    reference: the library uses the network's buffers in every guessed solve,
    so the network must outlive the attachment."""
    L = _setup()
    _check(L.tbl_guess_attach(ctx.h, net.h, level), "tbl_guess_attach")
    ctx._guess_net = net


def image_shape(ctx: Context, level: int):
    nx, ny = ctx.grid.level_shape(level)
    return (ny + 7) // 8 * 8, (nx + 7) // 8 * 8


def pairs(ctx: Context, f_ptr: int, v_ptr: int, x_ptr: int, t_ptr: int) -> None:
    L = _setup()
    _check(L.tbl_guess_pairs(ctx.h, C.c_void_p(f_ptr), C.c_void_p(v_ptr), C.c_void_p(x_ptr),
                             C.c_void_p(t_ptr)), "tbl_guess_pairs")


def weights_path(grid) -> str:
    return os.path.join(HERE, "weights", f"guess_{grid.nx}x{grid.ny}x{grid.nmodes}.bin")


class Guess:
    """A trained network attached to a context."""

    def __init__(self, ctx: Context, net: Net, level: int):
        self.net, self.level = net, level
        attach(ctx, net, level)

    def solve(self, ctx: Context, f_dev, u_dev, tol, test=False, **kw):
        """Solve L v = B f starting from the network's prediction (every
        mode; with test=True only modes whose prediction passes the
        residual test, guess=2)."""
        kw = dict(kw)
        kw["guess"] = 2 if test else 3
        return ctx.solve_rhs_dev(f_dev.data_ptr(), u_dev.data_ptr(), tol, **kw)


def load_default(ctx: Context):
    """The committed network for this grid, or None."""
    path = weights_path(ctx.grid)
    if not os.path.exists(path):
        return None
    lvl = default_level(ctx)
    H, W = image_shape(ctx, lvl)
    net = Net(H, W, ctx.grid.nmodes)
    net.load(path)
    return Guess(ctx, net, lvl)
