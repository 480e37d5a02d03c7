"""GPU parity: the HIP path against the CPU oracle on the same inputs.

* derivatives (derivs_par.f:80-177) to 1e-12 relative;
* the FP64 multigrid iterate by iterate against the oracle run with the same
  (colour) line ordering, to 1e-12 relative;
* the converged FP64 solution against the oracle's reference-ordered
  (lexicographic) fixed point, 1e-12 relative;
* mixed precision (FP32 smoothing, FP64 defect correction) against the FP64
  oracle at the 1e-9 residual, 1e-8 relative (the stated tolerance).
"""
import numpy as np
import pytest

import oracle
import trans_bl_amd as tb
from trans_bl_amd import workload as wl
from conftest import rel, rel_l2

pytestmark = pytest.mark.gpu

SMALL = wl.Grid(129, 33, 4, 3, 0.03, 1e-3, 1.03, 0.2, 0.6)
CFG1 = wl.SMALL_385x97x8


def oparams(g):
    return oracle.make_params(g.nx, g.ny, g.nmodes, g.levels, g.dx, g.dy0, g.stf, g.alpha,
                              g.beta, g.fac_y, coefs=coefs_of(g))


def coefs_of(g):
    """The coefficient tables both sides use (same inputs for parity)."""
    return oracle.coefgen.generate(g.stf, g.levels)


def context(g):
    return tb.Context(g, coefs=coefs_of(g))


@pytest.fixture(scope="module")
def torch_mod(gpu_ready):
    import torch
    return torch


def _dev(torch, a):
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    torch.cuda.synchronize()
    return t


@pytest.mark.parametrize("grid", [SMALL, CFG1], ids=["small", "385x97x8"])
@pytest.mark.parametrize("kind", ["x", "xx", "y", "yy", "yfv"])
def test_derivatives_match_oracle(torch_mod, grid, kind):
    f = wl.wavepacket_rhs(grid, tag="deriv")
    p = oparams(grid)
    ref = oracle.deriv(p, kind, f)
    with context(grid) as ctx:
        got = getattr(ctx, {"x": "derparx", "xx": "derparxx", "y": "dery", "yy": "deryy",
                            "yfv": "deryfv"}[kind])(f)
        # device-pointer form on torch memory gives the same bits
        fi = _dev(torch_mod, f)
        fo = torch_mod.zeros_like(fi)
        ctx.deriv_dev(tb.DERIV_KINDS[kind], fi.data_ptr(), fo.data_ptr())
        ctx.synchronize()
        dev = fo.cpu().numpy()
    assert rel(got, ref) < 1e-12, rel(got, ref)
    assert np.array_equal(dev, got)


@pytest.mark.parametrize("grid", [
    SMALL,
    pytest.param(wl.Grid(2049, 33, 2, 3, 0.01, 1e-3, 1.03, 0.0, 0.3), marks=pytest.mark.skip(
        reason="an LDS-staged x kernel for lines above 1088 points faulted the GPU once and gave "
               "run-to-run different bits once; that form is removed (such lines take the "
               "long-line kernel) and this case stays off the GPU until the cause is shown")),
    wl.Grid(65, 545, 2, 3, 0.05, 2e-4, 1.005, 0.0, 0.3)],
    ids=["small", "2049x33", "65x545"])
def test_fused_derivatives_match_single_and_oracle(torch_mod, grid):
    """tbl_derivs_dev (the four derivatives of one field, two passes over it)
    gives the single-derivative entry points' bits, and the oracle's values,
    also on the widest x-lines (2049 points) and long y-lines (545 rows)."""
    torch = torch_mod
    g = grid
    f = wl.wavepacket_rhs(g, tag="fused")
    p = oparams(g)
    with context(g) as ctx:
        fi = _dev(torch, f)
        outs = [torch.zeros_like(fi) for _ in range(4)]
        ctx.derivs_dev(fi.data_ptr(), *[o.data_ptr() for o in outs])
        ctx.synchronize()
        fused = [o.cpu().numpy() for o in outs]
        single = []
        for kind in range(4):
            o = torch.zeros_like(fi)
            ctx.deriv_dev(kind, fi.data_ptr(), o.data_ptr())
            ctx.synchronize()
            single.append(o.cpu().numpy())
    for kind, name in enumerate(("x", "xx", "y", "yy")):
        assert np.array_equal(fused[kind], single[kind]), name
        ref = oracle.deriv(p, name, f)
        assert rel(fused[kind], ref) < 1e-12, (name, rel(fused[kind], ref))


@pytest.mark.parametrize("cycles", [1, 3])
def test_fp64_iterates_match_colour_oracle(torch_mod, cycles):
    g = SMALL
    f = wl.wavepacket_rhs(g, tag="iter")
    z = np.zeros_like(f)
    ref, rc, _ = oracle.poi_uy(oparams(g), z, -f, z, 0.0, max_cycles=cycles, order=1)
    with context(g) as ctx:
        got, cyc, res, st = ctx.poi_uy(None, -f, z, 0.0, max_cycles=cycles)
    assert np.all(cyc == cycles) and np.all(rc == cycles)
    assert rel(got, ref) < 1e-12, rel(got, ref)


def test_fp64_iterates_with_boundary_data_and_wx(torch_mod):
    """Nonzero wx (source g = kb*wx - dwzdx) and nonzero Dirichlet data on the
    boundary columns, which the prolongation carries into columns 2 and nx-1
    exactly as intpol does (poisson_par.f:877-942)."""
    g = SMALL
    rng = np.random.default_rng(7)
    f = wl.wavepacket_rhs(g, tag="bd")
    wx = 0.1 * wl.wavepacket_rhs(g, tag="wx")
    uy = 0.01 * (rng.normal(size=f.shape) + 1j * rng.normal(size=f.shape))
    ref, _, _ = oracle.poi_uy(oparams(g), wx, -f, uy, 0.0, max_cycles=2, order=1)
    with context(g) as ctx:
        got, _, _, _ = ctx.poi_uy(wx, -f, uy, 0.0, max_cycles=2)
    assert rel(got, ref) < 1e-12, rel(got, ref)
    # untouched entries: wall row and the two boundary columns
    assert np.array_equal(got[:, 0, :], uy[:, 0, :])
    assert np.array_equal(got[:, :, 0], uy[:, :, 0])
    assert np.array_equal(got[:, :, -1], uy[:, :, -1])


@pytest.mark.parametrize("grid", [SMALL, CFG1], ids=["small", "385x97x8"])
def test_fp64_fixed_point_matches_reference_ordering(torch_mod, grid):
    f = wl.wavepacket_rhs(grid, tag="fixed")
    z = np.zeros_like(f)
    ref, rc, rres = oracle.poi_uy(oparams(grid), z, -f, z, 0.0, max_cycles=25, order=0)
    with context(grid) as ctx:
        got, cyc, res, st = ctx.poi_uy(None, -f, z, 0.0, max_cycles=25)
    assert rel_l2(got, ref) < 1e-12, rel_l2(got, ref)


def test_fp64_stop_rule(torch_mod):
    g = CFG1
    f = wl.wavepacket_rhs(g, tag="stop")
    z = np.zeros_like(f)
    ref, rc, rres = oracle.poi_uy(oparams(g), z, -f, z, 1e-9, max_cycles=11, order=1)
    with context(g) as ctx:
        got, cyc, res, st = ctx.poi_uy(None, -f, z, 1e-9)
        # a converged start stops at the first check with zero cycles (after
        # the level-1 pre-smoothing that precedes the check, poisson_par.f:35-40)
        again, cyc2, res2, _ = ctx.poi_uy(None, -f, got, 1e-9)
        # an unreachable tolerance runs exactly to the cap
        _, cyc3, _, st3 = ctx.poi_uy(None, -f, z, 1e-30, max_cycles=4)
    assert np.array_equal(cyc, rc), (cyc, rc)
    assert np.all(res < 1e-9) and st["nconverged"] == g.nmodes
    assert rel(got, ref) < 1e-12
    assert np.all(cyc2 == 0) and np.all(res2 < 1e-9) and rel(again, got) < 1e-9
    assert np.all(cyc3 == 4) and st3["nconverged"] == 0


def test_mixed_precision_matches_fp64_oracle(torch_mod):
    g = CFG1
    f = wl.wavepacket_rhs(g, tag="mixed")
    z = np.zeros_like(f)
    ref, _, _ = oracle.poi_uy(oparams(g), z, -f, z, 0.0, max_cycles=25, order=0)
    with context(g) as ctx:
        got, cyc, res, st = ctx.poi_uy(None, -f, z, 1e-9, precision=1, max_cycles=30)
    assert st["nconverged"] == g.nmodes and np.all(res < 1e-9)
    # stated tolerance of the mixed-precision path: 1e-8 relative (L2) in v
    assert rel_l2(got, ref) < 1e-8, rel_l2(got, ref)


@pytest.mark.parametrize("nu1", [0, 3])
@pytest.mark.parametrize("start", ["zero", "warm"])
def test_fmg_cycles_match_fp64_oracle(torch_mod, nu1, start):
    # full-multigrid corrections (first cycle, every cycle) from zero and
    # from a warm start: same stopping rule and tolerance as the V-cycle
    # path, fewer cycles, graph replay bit for bit
    g = CFG1
    f = wl.wavepacket_rhs(g, tag="mixed")
    z = np.zeros_like(f)
    ref, _, _ = oracle.poi_uy(oparams(g), z, -f, z, 0.0, max_cycles=25, order=0)
    with context(g) as ctx:
        u0 = z
        if start == "warm":
            # the solution for the packets 0.4 upstream
            fp = wl.wavepacket_rhs(g, tag="mixed", shift=-0.4)
            u0, _, _, _ = ctx.poi_uy(None, -fp, z, 1e-9, precision=1, max_cycles=30)
        kw = dict(precision=1, max_cycles=30, guess=1 if start == "zero" else 0, nu1=nu1)
        vc, cv, _, _ = ctx.poi_uy(None, -f, u0, 1e-9, **kw)
        got, cyc, res, st = ctx.poi_uy(None, -f, u0, 1e-9, fmg=1, **kw)
        graph, cg, _, _ = ctx.poi_uy(None, -f, u0, 1e-9, fmg=1, use_graph=1, **kw)
        again, ca, _, _ = ctx.poi_uy(None, -f, u0, 1e-9, fmg=1, use_graph=1, **kw)
        every, ce, re, se = ctx.poi_uy(None, -f, u0, 1e-9, fmg=2, use_graph=1, **kw)
        # ignored by the FP64 cycle
        d0, cd0, _, _ = ctx.poi_uy(None, -f, u0, 1e-9, max_cycles=30)
        d1, cd1, _, _ = ctx.poi_uy(None, -f, u0, 1e-9, max_cycles=30, fmg=2)
    for out, c, r, s in ((got, cyc, res, st), (every, ce, re, se)):
        assert s["nconverged"] == g.nmodes and np.all(r < 1e-9)
        assert rel_l2(out, ref) < 1e-8, rel_l2(out, ref)
        assert c.sum() < cv.sum(), (c, cv)
    assert ce.sum() <= cyc.sum(), (ce, cyc)
    assert np.array_equal(got, graph) and np.array_equal(got, again)
    assert np.array_equal(cyc, cg) and np.array_equal(cyc, ca)
    assert np.array_equal(d0, d1) and np.array_equal(cd0, cd1)


@pytest.mark.parametrize("g", [
    wl.Grid(129, 33, 4, 2, 0.03, 1e-3, 1.03, 0.2, 0.6),      # two levels
    SMALL,                                                    # three
    wl.Grid(257, 65, 6, 4, 0.02, 5e-4, 1.02, 0.0, 0.3),      # four
    wl.Grid(193, 49, 3, 3, 0.025, 8e-4, 1.025, 0.0, 0.4)],   # odd mode count
    ids=["129x33x4_L2", "129x33x4_L3", "257x65x6_L4", "193x49x3_L3"])
def test_fmg_on_other_level_counts(torch_mod, g):
    # full-multigrid corrections (first cycle, every cycle) on level counts
    # 2-4: the FP64 stopping rule met, the answer within the mixed path's
    # 1e-8 of the FP64 cycle's, at most the V-cycle form's largest cycle count
    f = wl.wavepacket_rhs(g, tag="fmg")
    z = np.zeros_like(f)
    with context(g) as ctx:
        ref, _, _, st64 = ctx.poi_uy(None, -f, z, 1e-10, max_cycles=30)
        vc, cv, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=1, max_cycles=30, guess=1)
        for fmg in (1, 2):
            got, cyc, res, st = ctx.poi_uy(None, -f, z, 1e-9, precision=1, max_cycles=30,
                                           guess=1, fmg=fmg, use_graph=1)
            assert st["nconverged"] == g.nmodes and np.all(res < 1e-9), (fmg, res)
            assert rel_l2(got, ref) < 1e-8, (fmg, rel_l2(got, ref))
            assert cyc.max() <= cv.max(), (fmg, cyc, cv)
    assert st64["nconverged"] == g.nmodes


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_equals_stream_launches(torch_mod, precision):
    g = SMALL
    f = wl.wavepacket_rhs(g, tag="graph")
    z = np.zeros_like(f)
    with context(g) as ctx:
        a, ca, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=precision, max_cycles=20)
        b, cb, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=precision, max_cycles=20,
                                 use_graph=1)
        # replay the captured graph again (eager work in between)
        c, cc, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=precision, max_cycles=20,
                                 use_graph=1)
    assert np.array_equal(a, b) and np.array_equal(a, c)
    assert np.array_equal(ca, cb) and np.array_equal(ca, cc)


def test_device_entry_matches_host_entry(torch_mod):
    torch = torch_mod
    g = SMALL
    f = wl.wavepacket_rhs(g, tag="dev")
    z = np.zeros_like(f)
    with context(g) as ctx:
        host, hc, _, _ = ctx.poi_uy(None, -f, z, 1e-9)
        fd = _dev(torch, f)
        ud = _dev(torch, z)
        dc, dres, st = ctx.solve_rhs_dev(fd.data_ptr(), ud.data_ptr(), 1e-9)
        dev = ud.cpu().numpy()
        # residual entry point agrees with the solver's own check
        rd = torch.zeros_like(fd)
        nd = _dev(torch, -f)
        rmax = ctx.residual_dev(0, nd.data_ptr(), ud.data_ptr(), rd.data_ptr())
    assert np.array_equal(host, dev) and np.array_equal(hc, dc)
    assert np.all(rmax < 1e-9)


def test_grid_rejection(gpu_ready):
    with pytest.raises(tb.TblError):
        tb.Context(wl.Grid(130, 33, 2, 3))      # (nx-1) not divisible by 4
    with pytest.raises(tb.TblError):
        tb.Context(wl.Grid(129, 33, 2, 6))      # coarsest level too small


def _lanes(code):
    if code > 3000:
        return (code - 3000) % 100
    return code - 2000 if code > 2000 else code - 1000 if code > 1000 else abs(code)


@pytest.mark.parametrize("grid", [SMALL, CFG1], ids=["small", "385x97x8"])
@pytest.mark.parametrize("wsel", [0, 1, 5])
def test_one_launch_sweep_matches_direct(torch_mod, wsel, grid):
    """The one-launch out-of-place sweep (all four colour phases of a strip
    in one workgroup, halo lines recomputed) relaxes the same lines in the
    same order with the same arithmetic as the phase-per-launch forms: the
    correction cycle's solve is bit-identical, the zero-start first sweep and
    the fused FP64 update included."""
    g = grid
    f = wl.wavepacket_rhs(g, tag="sweep")
    z = np.zeros_like(f)
    with context(g) as ctx:
        lanes = [_lanes(ctx.line_shape(1, lvl)) for lvl in range(1, g.levels)]
        for lvl in range(1, g.levels):
            ctx.set_line_shape(1, lvl, lanes[lvl - 1])
        a, ca, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=1, max_cycles=20)
        a3, _, _, _ = ctx.poi_uy(None, -f, z, 0.0, precision=1, max_cycles=3, use_graph=1)
        for lvl in range(1, g.levels):
            ctx.set_line_shape(1, lvl, 3000 + 100 * wsel + lanes[lvl - 1])
        b, cb, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=1, max_cycles=20)
        b3, _, _, _ = ctx.poi_uy(None, -f, z, 0.0, precision=1, max_cycles=3, use_graph=1)
        # FP64 cycles do not take the one-launch form
        with pytest.raises(tb.TblError):
            ctx.set_line_shape(0, 1, 3000 + lanes[0])
    assert np.array_equal(a, b) and np.array_equal(ca, cb)
    assert np.array_equal(a3, b3)


@pytest.mark.parametrize("pattern", ["sweep-direct", "direct-sweep", "sweep-staged"])
def test_one_launch_sweep_interleaved_with_other_forms(torch_mod, pattern):
    """Levels on the out-of-place sweep (e moving between two buffers) next to
    levels smoothed in place: same bits as all-direct."""
    g = CFG1
    f = wl.wavepacket_rhs(g, tag="sweep-mix")
    z = np.zeros_like(f)
    with context(g) as ctx:
        lanes = [_lanes(ctx.line_shape(1, lvl)) for lvl in range(1, g.levels)]
        for lvl in range(1, g.levels):
            ctx.set_line_shape(1, lvl, lanes[lvl - 1])
        a, ca, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=1, max_cycles=20)
        first, other = pattern.split("-")
        for lvl in range(1, g.levels):
            form = first if (lvl - 1) % 2 == 0 else other
            code = {"sweep": 3000 + lanes[lvl - 1], "direct": lanes[lvl - 1],
                    "staged": -lanes[lvl - 1]}[form]
            ctx.set_line_shape(1, lvl, code)
        b, cb, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=1, max_cycles=20)
    assert np.array_equal(a, b) and np.array_equal(ca, cb)


@pytest.mark.parametrize("grid", [SMALL, CFG1], ids=["small", "385x97x8"])
@pytest.mark.parametrize("form", ["staged", "xstaged"])
@pytest.mark.parametrize("precision", [0, 1])
def test_staged_smoother_matches_direct(torch_mod, precision, form, grid):
    """The LDS-staged (coalesced) smoothers -- x terms, v and S staged, or
    only the x terms -- do the same arithmetic as the direct one: identical
    bits."""
    g = grid
    f = wl.wavepacket_rhs(g, tag="staged")
    z = np.zeros_like(f)
    with context(g) as ctx:
        # (the coarsest level keeps its own kernel)
        for lvl in range(1, g.levels):
            for prec in (0, 1):
                ctx.set_line_shape(prec, lvl, _lanes(ctx.line_shape(prec, lvl)))
        a, ca, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=precision, max_cycles=20)
        for lvl in range(1, g.levels):
            for prec in (0, 1):
                lanes = _lanes(ctx.line_shape(prec, lvl))
                ctx.set_line_shape(prec, lvl, -lanes if form == "staged" else 2000 + lanes)
        b, cb, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=precision, max_cycles=20)
    assert np.array_equal(a, b) and np.array_equal(ca, cb)


@pytest.mark.parametrize("pattern", ["staged-direct", "direct-staged"])
def test_mixed_cycle_level_forms_interleaved(torch_mod, pattern):
    """The correction cycle starts e from zero without clearing it on the
    staged levels and clears it (fill or restriction) on the others, and
    folds the FP64 update into the finest staged sweep: with staged and
    direct forms alternating over the levels, the solve is bit-identical to
    the all-direct one."""
    g = CFG1
    f = wl.wavepacket_rhs(g, tag="interleaved")
    z = np.zeros_like(f)
    with context(g) as ctx:
        lanes = [_lanes(ctx.line_shape(1, lvl)) for lvl in range(1, g.levels)]
        for lvl in range(1, g.levels):
            ctx.set_line_shape(1, lvl, lanes[lvl - 1])
        a, ca, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=1, max_cycles=20)
        first = 0 if pattern == "staged-direct" else 1
        for lvl in range(1, g.levels):
            staged = (lvl - 1) % 2 == first
            ctx.set_line_shape(1, lvl, -lanes[lvl - 1] if staged else lanes[lvl - 1])
        b, cb, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=1, max_cycles=20)
    assert np.array_equal(a, b) and np.array_equal(ca, cb)


@pytest.mark.parametrize("grid", [SMALL, CFG1], ids=["small", "385x97x8"])
@pytest.mark.parametrize("lanes", [16, 32])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_coarsest_scan_matches_sequential(torch_mod, grid, lanes, precision):
    """The coarsest level's lane-scan kernel (shape G) and its sequential
    kernels (shape 1) solve the same lines: same converged solution, iterates
    within rounding -- in the FP64 cycle and in the FP32 correction cycle
    (whose coarsest sweeps run in FP32: iterates within FP32 rounding of the
    corrections, solutions within the mixed path's 1e-8)."""
    f = wl.wavepacket_rhs(grid, tag="coarse")
    z = np.zeros_like(f)
    kw = dict(precision=precision)
    with context(grid) as ctx:
        assert ctx.line_shape(precision, grid.levels) == 1
        a, ca, _, _ = ctx.poi_uy(None, -f, z, 1e-9, max_cycles=20, **kw)
        a2, _, _, _ = ctx.poi_uy(None, -f, z, 0.0, max_cycles=2, **kw)
        ctx.set_line_shape(precision, grid.levels, lanes)
        b, cb, _, _ = ctx.poi_uy(None, -f, z, 1e-9, max_cycles=20, **kw)
        b2, _, _, _ = ctx.poi_uy(None, -f, z, 0.0, max_cycles=2, **kw)
    if precision == 0:
        assert rel(b2, a2) < 1e-12, rel(b2, a2)
        assert np.array_equal(ca, cb)
        assert rel(b, a) < 1e-12, rel(b, a)
    else:
        assert rel(b2, a2) < 1e-5, rel(b2, a2)
        assert abs(int(cb.max()) - int(ca.max())) <= 1
        assert rel(b, a) < 1e-8, rel(b, a)


def _strip_shapes(ctx, g):
    """Strip form on every level and precision that accepts one (not the
    coarsest level; rows per lane bounded)."""
    used = 0
    for lvl in range(1, g.levels):
        for prec in (0, 1):
            for lanes in (1064, 1032, 1016):
                try:
                    ctx.set_line_shape(prec, lvl, lanes)
                    used += 1
                    break
                except tb.TblError:
                    continue
    assert used >= 2


@pytest.mark.parametrize("grid", [SMALL, CFG1], ids=["small", "385x97x8"])
@pytest.mark.parametrize("precision", [0, 1])
def test_strip_smoother_matches_direct(torch_mod, precision, grid):
    """The fused-sweep strip smoother (a whole sweep per workgroup in LDS, the
    halo lines recomputed from a snapshot) gives the phase-by-phase sweep."""
    f = wl.wavepacket_rhs(grid, tag="strip")
    z = np.zeros_like(f)
    with context(grid) as ctx:
        a, ca, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=precision, max_cycles=20)
        _strip_shapes(ctx, grid)
        b, cb, _, _ = ctx.poi_uy(None, -f, z, 1e-9, precision=precision, max_cycles=20)
    assert np.array_equal(ca, cb)
    assert rel(b, a) < (1e-12 if precision == 0 else 1e-8), rel(b, a)


def test_strip_smoother_iterates_match_colour_oracle(torch_mod):
    """Strip smoother with nonzero boundary data and wx, iterate by iterate
    against the oracle (the halo snapshot must carry the boundary columns)."""
    g = SMALL
    rng = np.random.default_rng(11)
    f = wl.wavepacket_rhs(g, tag="stripbd")
    wx = 0.1 * wl.wavepacket_rhs(g, tag="stripwx")
    uy = 0.01 * (rng.normal(size=f.shape) + 1j * rng.normal(size=f.shape))
    ref, _, _ = oracle.poi_uy(oparams(g), wx, -f, uy, 0.0, max_cycles=2, order=1)
    with context(g) as ctx:
        _strip_shapes(ctx, g)
        got, _, _, _ = ctx.poi_uy(wx, -f, uy, 0.0, max_cycles=2)
    assert rel(got, ref) < 1e-12, rel(got, ref)


def test_autotuned_shapes_keep_parity(torch_mod):
    from trans_bl_amd import autotune
    g = SMALL
    f = wl.wavepacket_rhs(g, tag="tuned")
    z = np.zeros_like(f)
    ref, rc, _ = oracle.poi_uy(oparams(g), z, -f, z, 0.0, max_cycles=3, order=1)
    with context(g) as ctx:
        table = autotune.tune(ctx, sweeps=2)
        assert len(table) == g.levels
        got, cyc, _, _ = ctx.poi_uy(None, -f, z, 0.0, max_cycles=3)
    assert rel(got, ref) < 1e-12
