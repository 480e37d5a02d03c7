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
