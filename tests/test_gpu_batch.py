"""Uneven batch, one RK4 stage per problem (4 derivatives + the v-Poisson
solve of every mode), issued concurrently and captured as one hipGraph,
against the single-problem entry points on the same inputs."""
import numpy as np
import pytest

import trans_bl_amd as tb
from trans_bl_amd import batch as B
from trans_bl_amd import workload as wl

pytestmark = pytest.mark.gpu

GRIDS = [wl.Grid(129, 33, 4, 3, 0.03, 1e-3, 1.03, 0.0, 0.6),
         wl.Grid(257, 65, 2, 4, 0.02, 5e-4, 1.02, 0.0, 0.3),
         wl.Grid(65, 33, 3, 3, 0.05, 2e-3, 1.03, 0.1, 0.9)]


@pytest.mark.parametrize("use_graph", [0, 1])
def test_batch_stage_matches_single_problem_calls(gpu_ready, use_graph):
    import torch
    rng = np.random.default_rng(2)
    fs, v0s = [], []
    for p, g in enumerate(GRIDS):
        fs.append(wl.wavepacket_rhs(g, tag=f"batch{p}"))
        v = np.zeros_like(fs[-1])
        v[:, 1:, 1:-1] = 1e-3 * (rng.normal(size=v[:, 1:, 1:-1].shape)
                                 + 1j * rng.normal(size=v[:, 1:, 1:-1].shape))
        v0s.append(v)
    bt = B.Batch(GRIDS)
    fd = [torch.from_numpy(f).cuda() for f in fs]
    vd = [torch.from_numpy(v.copy()).cuda() for v in v0s]
    dd = [torch.zeros_like(vd[p]) for p in range(len(GRIDS)) for _ in range(4)]
    torch.cuda.synchronize()
    for rep in range(2):                 # second call replays the graph
        for p in range(len(GRIDS)):
            vd[p].copy_(torch.from_numpy(v0s[p]).cuda())
        torch.cuda.synchronize()
        ms = bt.stage([t.data_ptr() for t in fd], [t.data_ptr() for t in vd],
                      [t.data_ptr() for t in dd], 1e-9, 12, 0, use_graph)
        assert ms > 0
        for p, g in enumerate(GRIDS):
            with tb.Context(g) as ctx:
                refd = [ctx.derparx(v0s[p]), ctx.derparxx(v0s[p]), ctx.dery(v0s[p]),
                        ctx.deryy(v0s[p])]
                refv, rc, rr, _ = ctx.poi_uy(None, -fs[p], v0s[p], 1e-9, max_cycles=12)
            for q in range(4):
                assert np.array_equal(dd[4 * p + q].cpu().numpy(), refd[q]), (p, q)
            assert np.array_equal(vd[p].cpu().numpy(), refv), p
            cyc, res = bt.results(p)
            assert np.array_equal(cyc, rc) and np.all(res < 1e-9)
    bt.close()


def _inputs(torch, v_scale=1e-3, seed=2):
    rng = np.random.default_rng(seed)
    fs, v0s = [], []
    for p, g in enumerate(GRIDS):
        fs.append(wl.wavepacket_rhs(g, tag=f"batch{p}"))
        v = np.zeros_like(fs[-1])
        v[:, 1:, 1:-1] = v_scale * (rng.normal(size=v[:, 1:, 1:-1].shape)
                                    + 1j * rng.normal(size=v[:, 1:, 1:-1].shape))
        v0s.append(v)
    return fs, v0s


@pytest.mark.parametrize("precision", [0, 1])
def test_converging_stage_matches_fixed_stage(gpu_ready, precision):
    """tbl_batch_stage_converge against tbl_batch_stage with max_cycles fixed
    cycles: the same derivatives, solutions and per-mode cycle counts, bit
    for bit -- on its first call (the graph holds max_cycles), on later calls
    (the graph holds what the previous call needed, then replays), and when
    a call needs more cycles than the previous one (the host-checked extra
    cycles: a converged start first, then the random one)."""
    import torch
    fs, v0s = _inputs(torch)
    n = len(GRIDS)
    bt = B.Batch(GRIDS)
    fd = [torch.from_numpy(f).cuda() for f in fs]
    vd = [torch.from_numpy(v.copy()).cuda() for v in v0s]
    dd = [torch.zeros_like(vd[p]) for p in range(n) for _ in range(4)]
    ptrs = lambda: ([t.data_ptr() for t in fd], [t.data_ptr() for t in vd],  # noqa: E731
                    [t.data_ptr() for t in dd])

    def load(vs):
        for p in range(n):
            vd[p].copy_(torch.from_numpy(vs[p]).cuda())
        torch.cuda.synchronize()

    def state():
        return ([t.cpu().numpy() for t in vd], [t.cpu().numpy() for t in dd],
                [bt.results(p)[0] for p in range(n)], [bt.results(p)[1] for p in range(n)])

    load(v0s)
    bt.stage(*ptrs(), 1e-9, 12, precision, 1)
    ref_v, ref_d, ref_c, ref_r = state()
    solved = [v.copy() for v in ref_v]
    load(solved)
    bt.stage(*ptrs(), 1e-9, 12, precision, 1)
    sol_v, sol_d, sol_c, _ = state()
    runs = []
    for start in ("random", "random", "random", "solved", "random"):
        load(v0s if start == "random" else solved)
        ms, run = bt.stage_converge(*ptrs(), 1e-9, max_cycles=12, precision=precision)
        assert ms > 0
        runs.append(run)
        got_v, got_d, got_c, got_r = state()
        want_v, want_d, want_c = (ref_v, ref_d, ref_c) if start == "random" else \
            (sol_v, sol_d, sol_c)
        for p in range(n):
            assert np.array_equal(got_v[p], want_v[p]), (start, p)
            assert np.array_equal(got_c[p], want_c[p]), (start, p, got_c[p], want_c[p])
            for q in range(4):
                assert np.array_equal(got_d[4 * p + q], want_d[4 * p + q]), (start, p, q)
        if start == "random":
            assert all(np.all(r < 1e-9) for r in got_r)
    # the first call runs the cap in its graph; later ones the slowest
    # problem's cycles: its modes' count plus the check that found them
    # converged (the last call reaching it through host-checked cycles)
    need = max(int(c.max()) for c in ref_c)
    assert runs[0] == 12 and runs[1] == runs[2] == runs[4] == min(need + 1, 12), runs
    bt.close()


def test_converging_stage_full_multigrid_form(gpu_ready):
    """The converging stage with the full-multigrid first correction and three
    pre-sweeps: converged, within the mixed path's 1e-8 of the FP64 stage,
    in no more cycles per mode, and the same on a replay."""
    import torch
    fs, v0s = _inputs(torch)
    n = len(GRIDS)
    bt = B.Batch(GRIDS)
    fd = [torch.from_numpy(f).cuda() for f in fs]
    vd = [torch.from_numpy(v.copy()).cuda() for v in v0s]
    dd = [torch.zeros_like(vd[p]) for p in range(n) for _ in range(4)]
    ptrs = ([t.data_ptr() for t in fd], [t.data_ptr() for t in vd], [t.data_ptr() for t in dd])

    def load():
        for p in range(n):
            vd[p].copy_(torch.from_numpy(v0s[p]).cuda())
        torch.cuda.synchronize()

    load()
    bt.stage(*ptrs, 1e-11, 20, 0, 1)
    ref = [t.cpu().numpy() for t in vd]
    load()
    bt.stage_converge(*ptrs, 1e-9, max_cycles=20, precision=1)
    plain = [bt.results(p)[0] for p in range(n)]
    out = []
    for rep in range(2):
        load()
        _, run = bt.stage_converge(*ptrs, 1e-9, max_cycles=20, precision=1, fmg=1, nu1=3)
        out.append([t.cpu().numpy() for t in vd])
        for p in range(n):
            cyc, res = bt.results(p)
            assert np.all(res < 1e-9), (rep, p, res)
            assert cyc.max() <= plain[p].max(), (rep, p, cyc, plain[p])
            e = np.linalg.norm(out[-1][p] - ref[p]) / np.linalg.norm(ref[p])
            assert e < 1e-8, (rep, p, e)
        assert run <= 20
    for p in range(n):
        assert np.array_equal(out[0][p], out[1][p])
    bt.close()
