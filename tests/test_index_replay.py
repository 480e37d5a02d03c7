"""Host-side replay of the index and bounds math of the network's direct 3x3
convolution kernels (conv3x3_kernel and conv3x3s_kernel, tbl_unet.hip).

Both tile the flattened (image, row, column) pixel index, so the last tile
of a layer whose pixel count is not a multiple of the tile reaches past the
image batch.  The replay enumerates every lane's pixel and tap for every
layer shape of the U-Net at the image sizes the solver and the tests use,
and checks that every global read that is not masked off and every store
lies inside its tensor -- and that the pixel guard (pa < NP) is what keeps
the tail tile in bounds (without it the replay finds reads past the end).
"""
import numpy as np
import pytest

from oracle import unet_ref as U


def _layer_shapes(H, W):
    """(h, w, cin, cout) of every 3x3 convolution of the U-Net on H x W."""
    out = []
    for name, cin, cout, ks, lvl, relu in U.layers():
        if ks == 3:
            out.append((H >> lvl, W >> lvl, cin, cout))
    return out


def _fallback_reads(NP, H, W, cin, CT, guard=True):
    """Largest element offset read by conv3x3_kernel<CT> and the largest
    stored pixel, over all workgroups (lane pixel pa = p0 + wm + lane%32)."""
    PX = 64 if CT == 64 else 128
    nblk = (NP + PX - 1) // PX
    wm = np.array([0, 32] if CT == 64 else [0, 32, 64, 96])
    p0 = np.arange(nblk)[:, None, None] * PX
    pa = (p0 + wm[None, :, None] + np.arange(32)[None, None, :]).ravel()
    wa, ha, na = pa % W, (pa // W) % H, pa // (W * H)
    hi = -1
    for t in range(9):
        hh, ww = ha + t // 3 - 1, wa + t % 3 - 1
        ok = (hh >= 0) & (hh < H) & (ww >= 0) & (ww < W)
        if guard:
            ok &= pa < NP
        if ok.any():
            off = ((na * H + hh) * W + ww)[ok] * cin + (cin - 4)   # last 4-channel piece
            hi = max(hi, int(off.max()) + 3)
    # stores: p = p0 + wm + 8*(i/4) + 4*(lane/32) + i%4 < NP
    i = np.arange(16)
    ps = (p0 + wm[None, :, None] + (8 * (i // 4) + (i % 4))[None, None, :]).ravel()
    ps = np.concatenate([ps, ps + 4])
    stored = ps[ps < NP]
    return hi, int(stored.max())


@pytest.mark.parametrize("H, W, N", [(16, 24, 2), (32, 136, 1), (32, 136, 64), (16, 24, 4)])
def test_fallback_conv_stays_in_bounds(H, W, N):
    for h, w, cin, cout in _layer_shapes(H, W):
        NP = N * h * w
        for CT in (64, 32):
            if cout % CT:
                continue
            hi, smax = _fallback_reads(NP, h, w, cin, CT)
            assert hi < NP * cin, (h, w, cin, cout, CT)
            assert smax < NP


def test_pixel_guard_is_what_keeps_the_tail_tile_in_bounds():
    # a layer whose pixel count leaves a partial last tile (4 x 6 x 2 = 48 < 64)
    NP, h, w, cin = 48, 4, 6, 128
    hi, _ = _fallback_reads(NP, h, w, cin, 64, guard=False)
    assert hi >= NP * cin          # reads past the activation tensor
    hi, _ = _fallback_reads(NP, h, w, cin, 64, guard=True)
    assert hi < NP * cin


def _staged_reach(NP, W, MT):
    """conv3x3s_kernel<MT,...>: the LDS stage holds pixels q0 .. q0+NQ-1 with
    q0 = p0 - W - 1 and NQ = PX + 2W + 2; every tap of every lane pixel must
    land inside it (lq + (t/3-1)*W + (t%3-1) in [0, NQ))."""
    PX = 4 * 32 * MT
    NQ = PX + 2 * W + 2
    lp = np.arange(PX)
    lq = lp + W + 1
    lo, hi = np.inf, -np.inf
    for t in range(9):
        q = lq + (t // 3 - 1) * W + (t % 3 - 1)
        lo, hi = min(lo, q.min()), max(hi, q.max())
    return lo, hi, NQ


@pytest.mark.parametrize("W", [3, 6, 12, 17, 24, 34, 68, 136, 143])
@pytest.mark.parametrize("MT", [1, 2])
def test_staged_conv_taps_stay_in_the_lds_stage(W, MT):
    lo, hi, NQ = _staged_reach(1 << 20, W, MT)
    assert lo >= 0 and hi < NQ
