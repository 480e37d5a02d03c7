"""Learned initial guess inside the solve.

* a far-off prediction is rejected for every mode and the solve is then
  bit-identical to the plain zero-start solve (the fallback path cannot
  change the answer);
* guess=3 uses the prediction everywhere and keeps the warm start's
  boundary data;
* the committed trained network (trans_bl_amd/weights/) for the 1025x193x64
  grid, on right-hand sides it was not trained on, needs fewer V-cycles than
  the zero start and reaches the same 1e-9 residual with v within 1e-8 of
  the plain solve.
"""
import os

import numpy as np
import pytest

import trans_bl_amd as tb
from trans_bl_amd import guess as G
from trans_bl_amd import workload as wl
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod(gpu_ready):
    import torch
    return torch


def _solve(ctx, torch, f, **kw):
    fd = torch.from_numpy(f).cuda()
    ud = torch.zeros_like(fd)
    torch.cuda.synchronize()
    cyc, res, st = ctx.solve_rhs_dev(fd.data_ptr(), ud.data_ptr(), 1e-9, max_cycles=40, **kw)
    return ud.cpu().numpy(), cyc, res, st


@pytest.mark.parametrize("precision", [0, 1])
def test_wild_guess_falls_back_to_warm_start(torch_mod, precision):
    """A network whose predictions are far off (untrained, weights scaled up)
    is rejected for every mode, and the solve is then bit-identical to the
    plain solve from the same (zero) warm start."""
    torch = torch_mod
    g = wl.Grid(129, 33, 4, 3, 0.03, 1e-3, 1.03, 0.0, 0.6)
    f = wl.wavepacket_rhs(g, tag="fallback")
    with tb.Context(g) as ctx:
        ref, rc, rr, _ = _solve(ctx, torch, f, precision=precision, guess=1)
        lvl = G.default_level(ctx)
        H, W = G.image_shape(ctx, lvl)
        net = G.Net(H, W, g.nmodes, seed=1)
        net.set_params(net.get_params() * 30.0)
        G.attach(ctx, net, lvl)
        got, gc, gr, st = _solve(ctx, torch, f, precision=precision, guess=2)
    assert st["fallback"] == g.nmodes
    assert st["guess_residual_max"] > 0
    assert np.array_equal(got, ref) and np.array_equal(gc, rc)


@pytest.mark.parametrize("precision", [1])
def test_forced_guess_keeps_boundary_data(torch_mod, precision):
    """guess=3 uses the prediction for every mode; the warm start's boundary
    columns and wall row are kept and the solve reaches the tolerance.
    (Mixed precision only: the reference's FAS cycle adds the coarse grid's
    boundary-column values themselves as a correction (intpol,
    poisson_par.f:877-942), so with nonzero data on the x boundaries it
    stalls -- the oracle does the same.)"""
    torch = torch_mod
    g = wl.Grid(129, 33, 4, 3, 0.03, 1e-3, 1.03, 0.0, 0.6)
    rng = np.random.default_rng(3)
    f = wl.wavepacket_rhs(g, tag="forced")
    uy = np.zeros_like(f)
    uy[:, :, 0] = 0.01 * rng.normal(size=uy[:, :, 0].shape)
    uy[:, :, -1] = 0.01 * rng.normal(size=uy[:, :, -1].shape)
    with tb.Context(g) as ctx:
        fd = torch.from_numpy(f).cuda()
        out = {}
        for guess in (0, 3):
            if guess == 3:
                lvl = G.default_level(ctx)
                H, W = G.image_shape(ctx, lvl)
                G.attach(ctx, G.Net(H, W, g.nmodes, seed=2), lvl)
            ud = torch.from_numpy(uy).cuda()
            torch.cuda.synchronize()
            cyc, res, st = ctx.solve_rhs_dev(fd.data_ptr(), ud.data_ptr(), 1e-9,
                                             precision=precision, max_cycles=40, guess=guess)
            out[guess] = (ud.cpu().numpy(), res, st)
    got, res, st = out[3]
    assert st["fallback"] == 0
    assert np.all(res < 1e-9) and np.all(out[0][1] < 1e-9)
    # the boundary data enters the solution: same answer as the warm start
    assert rel_l2(got, out[0][0]) < 1e-8


def test_trained_guess_saves_cycles(torch_mod):
    torch = torch_mod
    g = wl.BATCH_1025x193x64
    path = G.weights_path(g)
    if not os.path.exists(path):
        pytest.skip(f"no trained network committed at {path}")
    cyc_plain, cyc_learned = [], []
    with tb.Context(g) as ctx:
        guess = G.load_default(ctx)
        for seed in range(3):
            f = wl.wavepacket_rhs(g, tag=f"unseen-{seed}")
            ref, rc, rr, _ = _solve(ctx, torch, f, precision=1, guess=1)
            got, gc, gr, st = _solve(ctx, torch, f, precision=1, guess=2)
            assert np.all(gr < 1e-9) and np.all(rr < 1e-9)
            assert rel_l2(got, ref) < 1e-8
            cyc_plain.append(rc.mean())
            cyc_learned.append(gc.mean())
    assert np.mean(cyc_learned) < np.mean(cyc_plain), (cyc_learned, cyc_plain)
