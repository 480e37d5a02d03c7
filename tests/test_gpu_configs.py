"""The larger workloads as parity tests (one GPU).

* uneven batch: the 32 problems of bench_stage.uneven_grids() (nx 257-2049,
  ny 65-257, 16-128 modes; 1632 modes, 241 M points) in one captured RK4
  stage (four derivatives and the solve of every problem) with the FP32
  correction cycle: every mode below the 1e-9 residual, and the smallest
  problem's derivatives (1e-12) and solution (1e-8 of the FP64 oracle's
  converged solution) against the oracle.
* large: one 2049x257x128 stage (dy0 2e-4, stretch 1.012, kz = 0.05 k):
  the FP32 correction cycle and the FP64 cycle both to 1e-9, within 1e-8 of
  each other; d/dx (the long-line kernel) and d/dy of the solution against
  the oracle; the next stage warm-started from it.
* learned guess on 385x97x8 with the committed network for that grid: every
  mode converges to 1e-9 from the network's prediction (accepted or not by
  the residual test), within 1e-8 of the FP64 oracle.
"""
import numpy as np
import pytest

import oracle
import trans_bl_amd as tb
from trans_bl_amd import workload as wl
from conftest import rel, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod(gpu_ready):
    import torch
    return torch


def _oparams(g):
    return oracle.make_params(g.nx, g.ny, g.nmodes, g.levels, g.dx, g.dy0, g.stf, g.alpha,
                              g.beta, g.fac_y)


def test_uneven_batch_full_size(torch_mod):
    torch = torch_mod
    from trans_bl_amd import batch as B
    from trans_bl_amd.bench_stage import uneven_grids
    from trans_bl_amd.train_guess import packet_params, rhs_torch
    grids = uneven_grids()
    sizes = [g.nx * g.ny * g.nmodes for g in grids]
    ps = int(np.argmin(sizes))
    gs = grids[ps]
    print(f"uneven batch: {len(grids)} problems, {sum(g.nmodes for g in grids)} modes; "
          f"sample problem {ps}: {gs.nx}x{gs.ny}x{gs.nmodes}", flush=True)
    bt = B.Batch(grids)
    fs, vs, ds = [], [], []
    for p, g in enumerate(grids):
        fs.append(rhs_torch(torch, g, packet_params(g, 7000 + p), 0.0))
        vs.append(torch.zeros_like(fs[-1]))
        ds.extend(torch.zeros_like(fs[-1]) for _ in range(4))
    v0 = 0.01 * rhs_torch(torch, gs, packet_params(gs, 123), 0.2)
    v0[:, 0, :] = 0                          # the wall row is not solved for
    vs[ps].copy_(v0)
    torch.cuda.synchronize()
    ms = bt.stage([t.data_ptr() for t in fs], [t.data_ptr() for t in vs],
                  [t.data_ptr() for t in ds], 1e-9, 12, 1, 1)
    torch.cuda.synchronize()
    worst = 0.0
    for p in range(len(grids)):
        cyc, res = bt.results(p)
        worst = max(worst, float(res.max()))
        assert np.all(res < 1e-9), (p, res.max())
    print(f"  stage {ms:.1f} ms, max residual {worst:.2e}")
    f_h = fs[ps].cpu().numpy()
    v0_h = v0.cpu().numpy()
    v_h = vs[ps].cpu().numpy()
    d_h = [ds[4 * ps + q].cpu().numpy() for q in range(4)]
    bt.close()
    p = _oparams(gs)
    for q, kind in enumerate(("x", "xx", "y", "yy")):
        e = rel(d_h[q], oracle.deriv(p, kind, v0_h))
        print(f"  d/{kind}: rel {e:.2e}")
        assert e < 1e-12, kind
    ref, _, rres = oracle.poi_uy(p, np.zeros_like(f_h), -f_h, v0_h, 1e-12, 30, order=1)
    e = rel_l2(v_h, ref)
    print(f"  v vs FP64 oracle (residual {rres.max():.1e}): rel l2 {e:.2e}")
    assert e < 1e-8


def test_large_stage_mixed_against_fp64(torch_mod):
    torch = torch_mod
    from trans_bl_amd.train_guess import packet_params, rhs_torch
    g = wl.LARGE_2049x257x128
    print("large stage 2049x257x128", flush=True)
    params = packet_params(g, 9000)
    f = rhs_torch(torch, g, params, 0.0)
    with tb.Context(g) as ctx:
        v32 = torch.zeros_like(f)
        v64 = torch.zeros_like(f)
        torch.cuda.synchronize()
        cyc32, res32, _ = ctx.solve_rhs_dev(f.data_ptr(), v32.data_ptr(), 1e-9, precision=1,
                                            max_cycles=40, use_graph=1)
        cyc64, res64, _ = ctx.solve_rhs_dev(f.data_ptr(), v64.data_ptr(), 1e-9, precision=0,
                                            max_cycles=40, use_graph=1)
        torch.cuda.synchronize()
        e = rel_l2(v32.cpu().numpy(), v64.cpu().numpy())
        print(f"  mixed: cycles {cyc32.mean():.2f} (max {cyc32.max()}), residual "
              f"{res32.max():.2e}; FP64: cycles {cyc64.mean():.2f}, residual {res64.max():.2e}; "
              f"rel l2 {e:.2e}")
        assert res32.max() < 1e-9 and res64.max() < 1e-9
        assert e < 1e-8
        # the full-multigrid first correction with three pre-sweeps (the
        # coarsest sweeps by the default rule: coarsest ny 33 here)
        vf = torch.zeros_like(f)
        cycf, resf, _ = ctx.solve_rhs_dev(f.data_ptr(), vf.data_ptr(), 1e-9, precision=1,
                                          max_cycles=40, use_graph=1, guess=1, fmg=1, nu1=3)
        torch.cuda.synchronize()
        ef = rel_l2(vf.cpu().numpy(), v64.cpu().numpy())
        print(f"  mixed, fmg, nu1 3: cycles {cycf.mean():.2f} (max {cycf.max()}), residual "
              f"{resf.max():.2e}; rel l2 {ef:.2e}")
        assert resf.max() < 1e-9 and ef < 1e-8
        assert cycf.max() <= cyc32.max()
        del vf
        # derivatives of the solution: the long x-lines (2049 points) and y
        dx = torch.zeros_like(f)
        dy = torch.zeros_like(f)
        ctx.derivs_dev(v64.data_ptr(), dx.data_ptr(), 0, dy.data_ptr(), 0)
        torch.cuda.synchronize()
        vh = v64.cpu().numpy()
        p = _oparams(g)
        for kind, t in (("x", dx), ("y", dy)):
            er = rel(t.cpu().numpy(), oracle.deriv(p, kind, vh))
            print(f"  d/{kind}: rel {er:.2e}")
            assert er < 1e-12, kind
        del dx, dy
        # the next stage (packets 0.4 downstream), warm-started from v
        f1 = rhs_torch(torch, g, params, 0.4)
        cyc, res, _ = ctx.solve_rhs_dev(f1.data_ptr(), v32.data_ptr(), 1e-9, precision=1,
                                        max_cycles=40, use_graph=1)
        print(f"  next stage from the previous solution: cycles {cyc.mean():.2f}, "
              f"residual {res.max():.2e}")
        assert res.max() < 1e-9


def test_learned_guess_on_385x97x8(torch_mod):
    torch = torch_mod
    from trans_bl_amd import guess as G
    from trans_bl_amd.train_guess import packet_params, rhs_torch
    g = wl.SMALL_385x97x8
    with tb.Context(g) as ctx:
        gs = G.load_default(ctx)
        if gs is None:
            pytest.fail("no committed network for 385x97x8")
        f = rhs_torch(torch, g, packet_params(g, 4242), 0.0)
        out = {}
        for name, guess in (("zero", 0), ("learned", 2)):
            v = torch.zeros_like(f)
            torch.cuda.synchronize()
            cyc, res, st = ctx.solve_rhs_dev(f.data_ptr(), v.data_ptr(), 1e-9, precision=1,
                                             max_cycles=20, use_graph=1, guess=guess)
            out[name] = (cyc, res, st, v.cpu().numpy())
            print(f"  {name}: cycles mean {cyc.mean():.2f} max {cyc.max()}, residual "
                  f"{res.max():.2e}, fallback {st['fallback']}")
            assert res.max() < 1e-9
    fh = f.cpu().numpy()
    ref, _, _ = oracle.poi_uy(_oparams(g), np.zeros_like(fh), -fh, np.zeros_like(fh), 1e-12,
                              30, order=1)
    for name in out:
        e = rel_l2(out[name][3], ref)
        print(f"  {name} vs FP64 oracle: rel l2 {e:.2e}")
        assert e < 1e-8
