"""Learned initial guess inside the solve.

* an untrained network's prediction has a larger residual than the zero
  start: every mode falls back, and the solve is bit-identical to the plain
  zero-start solve (the fallback path cannot change the answer);
* the committed trained network (trans_bl_amd/weights/) for the 385x97x8
  grid, on right-hand sides it was not trained on, starts from a smaller
  residual and needs fewer V-cycles than the zero start, and reaches the same
  1e-9 residual with v within 1e-8 of the plain solve.
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
def test_untrained_guess_falls_back_to_zero_start(torch_mod, precision):
    torch = torch_mod
    g = wl.Grid(129, 33, 4, 3, 0.03, 1e-3, 1.03, 0.0, 0.6)
    f = wl.wavepacket_rhs(g, tag="fallback")
    with tb.Context(g) as ctx:
        ref, rc, rr, _ = _solve(ctx, torch, f, precision=precision, guess=1)
        lvl = G.default_level(ctx)
        H, W = G.image_shape(ctx, lvl)
        net = G.Net(H, W, g.nmodes, seed=1)
        G.attach(ctx, net, lvl)
        got, gc, gr, st = _solve(ctx, torch, f, precision=precision, guess=2)
    assert st["fallback"] == g.nmodes
    assert st["guess_residual_max"] > 0
    assert np.array_equal(got, ref) and np.array_equal(gc, rc)


def test_trained_guess_saves_cycles(torch_mod):
    torch = torch_mod
    g = wl.SMALL_385x97x8
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
