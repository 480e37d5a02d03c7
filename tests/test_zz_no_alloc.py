"""tbl_residual_dev_noalloc (the residual entry point that allocates
nothing: its work fields live in the library's fixed device array) against
tbl_residual_dev on the same inputs: identical bits, on two calls in a row,
the second on a smaller grid (the array keeps the first call's contents),
and grids larger than the array are rejected."""
import numpy as np
import pytest

import trans_bl_amd as tb
from trans_bl_amd import workload as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod(gpu_ready):
    import torch
    return torch


def _inputs(g, tag):
    rng = np.random.default_rng(abs(hash(tag)) % 2**32)
    f = wl.wavepacket_rhs(g, tag=tag)
    wx = 0.1 * wl.wavepacket_rhs(g, tag=tag + "-wx")
    uy = rng.normal(size=f.shape) + 1j * rng.normal(size=f.shape)
    return wx, -f, uy


def _both(torch, ctx, g, tag):
    wx, dz, uy = (torch.from_numpy(np.ascontiguousarray(a)).to("cuda") for a in _inputs(g, tag))
    out_a = torch.full_like(dz, 7.0)
    out_b = torch.full_like(dz, -3.0)
    torch.cuda.synchronize()
    ra = ctx.residual_dev(wx.data_ptr(), dz.data_ptr(), uy.data_ptr(), out_a.data_ptr())
    rb = ctx.residual_dev_noalloc(wx.data_ptr(), dz.data_ptr(), uy.data_ptr(), out_b.data_ptr())
    torch.cuda.synchronize()
    return out_a.cpu().numpy(), ra, out_b.cpu().numpy(), rb


def test_noalloc_matches_residual_entry_on_two_calls(torch_mod):
    big = wl.SMALL_385x97x8
    small = wl.Grid(129, 33, 4, 3, 0.03, 1e-3, 1.03, 0.2, 0.6)
    for g, tag in ((big, "noalloc-1"), (small, "noalloc-2")):
        with tb.Context(g) as ctx:
            a, ra, b, rb = _both(torch_mod, ctx, g, tag)
        assert np.array_equal(a, b), g
        assert np.array_equal(ra, rb), g
        assert np.all(ra > 0)


def test_noalloc_rejects_grids_beyond_the_array(torch_mod):
    g = wl.Grid(1025, 193, 16, 4, 0.02, 2e-4, 1.015, 0.0, 0.1)   # 3 x 52 MB of work fields
    with tb.Context(g) as ctx:
        z = torch_mod.zeros((g.nmodes, g.ny, g.nx), dtype=torch_mod.complex128, device="cuda")
        o = torch_mod.zeros_like(z)
        with pytest.raises(tb.TblError, match="scratch array"):
            ctx.residual_dev_noalloc(0, z.data_ptr(), z.data_ptr(), o.data_ptr())
