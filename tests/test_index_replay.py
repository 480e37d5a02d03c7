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


# ---- derivative kernels (tbl_deriv_kernels.hip) ------------------------------
def _x_chunk(nx):
    """deriv_x_chunk: multiple of 64, workgroup LDS (4 lines + 4 chunks of
    16-byte points) within 80 KB, at most the line rounded up to 64."""
    room = 80 * 1024 // 16 // 4 - (nx | 1)
    return max(64, min((nx + 63) // 64 * 64, room // 64 * 64))


def _x_kernel_reach(nx, ny, K):
    """deriv_x_kernel<R>: every wavefront (one x-line) stages the line in its
    LDS slice of ldx = nx|1 points, reads clamped windows, writes the first
    derivative's solved points chunk by chunk into its CH-point output chunk
    (after the 4 input slices) and the second's into its own input slice,
    and stores from there; global loads and stores stay on its line.
    Returns (R, max LDS index, LDS elements, max global index, elements)."""
    R = next(r for r in (1, 2, 3, 5, 7, 9, 17) if 64 * r >= nx)
    ldx = nx | 1
    lines = ny * K
    blocks = (lines + 3) // 4
    w = np.arange(4)
    line = np.arange(blocks)[:, None] * 4 + w[None, :]
    live = line < lines
    lane = np.arange(64)
    gi = (lane[:, None] + 64 * np.arange(R)[None, :])            # loads, i < nx
    gi = gi[gi < nx]
    i0 = lane * R
    q = np.arange(R + 4)
    win = np.clip(i0[:, None] + q[None, :] - 2, 0, nx - 1)        # LDS window
    seg = (lane[:, None] * R + np.arange(R)[None, :])             # lane-segment points
    seg = seg[seg < nx]
    CH = _x_chunk(nx)
    assert CH % 64 == 0 and CH >= 64
    e_max = -1
    for c0 in range(0, nx, CH):
        e = seg - c0                                              # chunk writes
        e = e[(e >= 0) & (e < CH)]
        st = c0 + np.arange(CH)                                   # chunk reads, all lanes
        assert np.all(np.isin(st[st < nx], seg))                  # every stored point written
        e_max = max(e_max, e.max() if e.size else -1, (st[st < nx] - c0).max())
    lds_in = 3 * ldx + max(gi.max(), win.max(), seg.max())      # incl. last-derivative line
    lds_out = 4 * ldx + 3 * CH + e_max
    gmax = (line[live].max() * nx + nx - 1)
    return R, max(lds_in, lds_out), 4 * (ldx + CH), gmax, lines * nx


def _y_kernel_reach(nx, ny, K):
    G = 32 if (ny + 31) // 32 <= 9 else 64
    R = next(r for r in (1, 2, 3, 4, 5, 7, 9) if G * r >= ny)
    TX = 256 // G
    ldy = ny | 1
    e = np.arange(TX * ny)
    x, j = e % TX, e // TX
    tile_max = (x * ldy + j).max()
    lane = np.arange(G)
    win = np.clip(lane[:, None] * R + np.arange(R + 4)[None, :] - 2, 0, ny - 1)
    col_max = (TX - 1) * ldy + win.max()
    # global: x0 + x < nx masked, rows j < ny, modes k < K
    gmax = (K - 1) * nx * ny + (ny - 1) * nx + (nx - 1)
    return G, R, max(tile_max, col_max), TX * ldy, gmax, K * nx * ny


@pytest.mark.parametrize("nx, ny, K", [(129, 33, 4), (385, 97, 8), (1025, 193, 64),
                                       (1088, 33, 2), (65, 545, 2), (257, 576, 2)])
def test_derivative_kernels_stay_in_bounds(nx, ny, K):
    if (nx + 63) // 64 <= 17:
        R, lds_max, lds_n, gmax, n = _x_kernel_reach(nx, ny, K)
        assert lds_max < lds_n and gmax < n and 64 * R >= nx
        # two workgroups per CU up to the 1088-point lines of R = 17
        assert 2 * lds_n * 16 <= 160 * 1024 and _x_chunk(nx) >= 128
    if ny <= 64 * 9:
        G, R, lds_max, lds_n, gmax, n = _y_kernel_reach(nx, ny, K)
        assert lds_max < lds_n and gmax < n and G * R >= ny
        assert lds_n * 16 <= 160 * 1024


# ---- line smoother rows (tbl_relax.h, tbl_smooth_phase.hip, tbl_mg_kernels.hip)
K_RS = (1, 2, 3, 4, 5, 6, 7, 9, 13, 17)    # kRs


def _line_rows_for(G, ny):
    """line_rows_for: rows per lane for rows 1..ny-1 (row 0 is not relaxed)."""
    need = (ny - 1 + G - 1) // G
    for r in K_RS:
        if r < need:
            continue
        return -1 if G == 32 and r > 9 else r
    return -1


def _lu_stride(ny):
    s = ny
    for G in (16, 32, 64):
        R = _line_rows_for(G, ny)
        if R > 0 and 1 + G * R > s:
            s = 1 + G * R
    return (s + 7) // 8 * 8


@pytest.mark.parametrize("ny", [9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 577])
@pytest.mark.parametrize("G", [16, 32, 64])
def test_line_rows_cover_rows_1_to_ny_once(G, ny):
    """relax_core's lanes own rows line_j0(lane, R) = 1 + lane R .. + R - 1:
    together rows 1..ny-1, each once; the staged kernels' lane-strided
    loads (rows 1 + lane + s G, s < R) reach the same rows; the LU-table
    reads stay inside the padded table, the x-term reads inside the line."""
    R = _line_rows_for(G, ny)
    if R < 0:
        return
    lane = np.arange(G)
    rows = (1 + lane[:, None] * R + np.arange(R)[None, :]).ravel()
    rel_rows = rows[rows < ny]
    assert np.array_equal(np.sort(rel_rows), np.arange(1, ny))
    staged = (1 + lane[:, None] + G * np.arange(R)[None, :]).ravel()
    assert np.array_equal(np.sort(staged[staged < ny]), np.arange(1, ny))
    assert rows.max() < _lu_stride(ny)              # identity rows past ny
    xr = (rows[:, None] + np.array([-1, 0, 1])[None, :]).ravel()
    xr = xr[(xr >= 0) & (xr < ny)]
    assert xr.min() >= 0 and xr.max() <= ny - 1
    # 193 = 1 + 64 * 3: no idle lane slot at the benchmark's finest level
    if ny == 193 and G == 64:
        assert R == 3


# ---- restricting defect (defect_march_kernel<float, float, true>, tbl_mg_kernels.hip)
K_DEF_ROWS_RS = 58    # kDefRowsRs: rows per wavefront, three halo rows each side


def _defect_strip(nx):
    """defect_strip<float, float>: columns per strip (even)."""
    return 32 if nx >= 512 else 16


@pytest.mark.parametrize("nx,ny", [(1025, 193), (513, 97), (257, 49), (129, 25), (385, 97),
                                   (193, 49), (129, 33), (2049, 257), (257, 65), (65, 17)])
def test_restricting_defect_writes_every_coarse_source_point_once(nx, ny):
    """The restricting defect's wavefronts own rows 1 + 58 c + 0..57 (lanes
    3..60) and form the defect on lanes 2..61; each strip of columns starts
    on an odd column and walks one column past its end.  On odd columns i
    (from the strip's third) a lane of an owned even row writes coarse point
    ((i-1)/2, j/2) from the defects of rows j-1..j+1 (the top row: j alone)
    at columns i-2..i.  Replayed: every interior coarse point (columns
    1..cnx-2, rows 1..cny-1, the points restrict_kernel writes) is written
    exactly once, from lanes that formed their defect and columns the strip
    walked, and no load leaves the field."""
    kH = 3
    strip = _defect_strip(nx)
    cnx, cny = (nx - 1) // 2 + 1, (ny - 1) // 2 + 1
    nstrips = (nx - 2 + strip - 1) // strip
    jchunks = (ny - 1 + K_DEF_ROWS_RS - 1) // K_DEF_ROWS_RS
    count = np.zeros((cnx, cny), int)
    lane = np.arange(64)
    for sc in range(nstrips):
        ia = 1 + sc * strip
        ib0 = min(ia + strip, nx - 1)
        ib = ib0 + 1 if ib0 < nx - 1 else ib0
        assert ib <= nx - 1                       # S loads: columns < nx
        for jc in range(jchunks):
            j = 1 + jc * K_DEF_ROWS_RS - kH + lane
            own = (lane >= kH) & (lane < kH + K_DEF_ROWS_RS) & (j >= 1) & (j < ny)
            calc = (lane >= kH - 1) & (lane <= kH + K_DEF_ROWS_RS) & (j >= 1) & (j < ny)
            for i in range(ia + 2, ib, 2):
                assert i & 1
                ic = (i - 1) >> 1
                if not 1 <= ic <= cnx - 2:
                    continue
                for L in np.nonzero(own & (j % 2 == 0))[0]:
                    jcr = j[L] >> 1
                    if not 1 <= jcr <= cny - 1:
                        continue
                    used = [L] if j[L] == ny - 1 else [L - 1, L, L + 1]
                    assert all(0 <= u < 64 and calc[u] for u in used)
                    count[ic, jcr] += 1
    expect = np.zeros_like(count)
    expect[1:cnx - 1, 1:cny] = 1
    assert np.array_equal(count, expect)
