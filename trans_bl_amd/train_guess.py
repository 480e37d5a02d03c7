"""Train the learned initial guess on pairs the solver itself generates.

    python -m trans_bl_amd.train_guess --grid 1025x193x64 --pairs 20000 --out PATH

Right-hand sides are seeded wavepacket sequences drifting downstream by 0.4
per step (the packets' centres move +0.4 in x between consecutive batches);
each batch of modes is solved on the GPU (FP32 smoothing, FP64 correction)
to the training tolerance, the (features, target) pairs of the network's
multigrid level are extracted on the GPU, and the BF16 U-Net is trained with
its own forward/backward kernels and Adam.  Prints one JSON line per epoch.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import trans_bl_amd as tb  # noqa: E402
from trans_bl_amd import guess as G  # noqa: E402
from trans_bl_amd import workload as wl  # noqa: E402

GRIDS = {
    "1025x193x64": wl.BATCH_1025x193x64,
    "385x97x8": wl.SMALL_385x97x8,
    "2049x257x128": wl.LARGE_2049x257x128,
}


def packet_params(grid, seed, npk=6):
    """Per (mode, part, packet): amplitude, width, wavenumber, xc, yc."""
    rng = np.random.default_rng(seed)
    lx, ytop = grid.x()[-1], grid.y()[-1]
    shape = (grid.nmodes, 2, npk)
    return (rng.normal(size=shape), rng.uniform(0.05, 0.3, shape), rng.uniform(10, 30, shape),
            rng.uniform(0.1 * lx, 0.9 * lx, shape), rng.uniform(0, 0.5 * ytop, shape))


def rhs_torch(torch, grid, params, shift):
    """Complex (nmodes, ny, nx) wavepacket RHS on the GPU."""
    amp, sig, alp, xc, yc = (torch.as_tensor(p, device="cuda") for p in params)
    x = torch.as_tensor(grid.x(), device="cuda")
    y = torch.as_tensor(grid.y(), device="cuda")
    lx = float(grid.x()[-1])
    xc = 0.1 * lx + torch.remainder(xc + shift - 0.1 * lx, 0.8 * lx)
    out = torch.zeros((grid.nmodes, 2, grid.ny, grid.nx), dtype=torch.float64, device="cuda")
    for p in range(amp.shape[2]):
        a = amp[:, :, p, None, None]
        s = sig[:, :, p, None, None]
        gx = torch.exp(-(x[None, None, None, :] - xc[:, :, p, None, None]) ** 2 / (2 * s * s)) \
            * torch.cos(alp[:, :, p, None, None] * x[None, None, None, :])
        gy = torch.exp(-(y[None, None, :, None] - yc[:, :, p, None, None]) ** 2 / (2 * s * s))
        out += a * gx * gy
    return torch.complex(out[:, 0], out[:, 1]).contiguous()


def generate(torch, ctx, grid, npairs, H, W, tol, seed0=1000, steps_per_seq=8):
    K = grid.nmodes
    nb = math.ceil(npairs / K)
    X = torch.zeros((nb * K, H, W, 4), dtype=torch.float32, device="cuda")
    T = torch.zeros((nb * K, H, W, 2), dtype=torch.float32, device="cuda")
    u = torch.zeros((K, grid.ny, grid.nx), dtype=torch.complex128, device="cuda")
    cyc_total = 0
    for b in range(nb):
        seq, step = divmod(b, steps_per_seq)
        f = rhs_torch(torch, grid, packet_params(grid, seed0 + seq), 0.4 * step)
        u.zero_()
        torch.cuda.synchronize()
        cyc, res, st = ctx.solve_rhs_dev(f.data_ptr(), u.data_ptr(), tol, precision=1,
                                         max_cycles=40, guess=1)
        cyc_total += int(cyc.sum())
        if b % 25 == 0:
            print(json.dumps({"phase": "data", "batch": b, "of": nb, "cycles_mean": float(cyc.mean()),
                              "residual_max": float(res.max())}), flush=True)
        G.pairs(ctx, f.data_ptr(), u.data_ptr(), X[b * K:].data_ptr(), T[b * K:].data_ptr())
    return X[:npairs], T[:npairs], cyc_total


def main():
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", default="1025x193x64", choices=sorted(GRIDS))
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--minutes", type=float, default=0.0, help="stop training after this long")
    a = ap.parse_args()
    grid = GRIDS[a.grid]
    ctx = tb.Context(grid)
    lvl = G.default_level(ctx)
    H, W = G.image_shape(ctx, lvl)
    net = G.Net(H, W, max(a.batch, grid.nmodes), seed=a.seed)
    G.attach(ctx, net, lvl)
    t0 = time.time()
    X, T, cyc = generate(torch, ctx, grid, a.pairs, H, W, a.tol)
    torch.cuda.synchronize()
    tgen = time.time() - t0
    print(json.dumps({"phase": "data", "pairs": int(X.shape[0]), "level": lvl, "H": H, "W": W,
                      "solver_cycles": cyc, "seconds": round(tgen, 1)}), flush=True)
    n = X.shape[0]
    nval = max(a.batch, n // 20)
    perm = torch.randperm(n, device="cuda", generator=torch.Generator(device="cuda").manual_seed(0))
    Xv, Tv = X[perm[:nval]].contiguous(), T[perm[:nval]].contiguous()
    Xt, Tt = X[perm[nval:]].contiguous(), T[perm[nval:]].contiguous()
    ntr = Xt.shape[0]
    steps = a.epochs * (ntr // a.batch)
    step = 0
    for ep in range(a.epochs):
        t1 = time.time()
        order = torch.randperm(ntr, device="cuda")
        losses = []
        for i in range(ntr // a.batch):
            idx = order[i * a.batch:(i + 1) * a.batch]
            xb, tb_ = Xt[idx].contiguous(), Tt[idx].contiguous()
            torch.cuda.synchronize()
            lr = a.lr * 0.5 * (1 + math.cos(math.pi * step / steps))
            losses.append(net.train_step(xb.data_ptr(), tb_.data_ptr(), a.batch, lr))
            step += 1
            if step % 200 == 0:
                print(json.dumps({"phase": "train", "step": step, "of": steps,
                                  "loss": float(np.mean(losses[-50:]))}), flush=True)
            if a.minutes and time.time() - t0 > 60 * a.minutes:
                break
        vl = []
        for i in range(0, nval - a.batch + 1, a.batch):
            xb, tb_ = Xv[i:i + a.batch].contiguous(), Tv[i:i + a.batch].contiguous()
            torch.cuda.synchronize()
            vl.append(net.gradients(xb.data_ptr(), tb_.data_ptr(), a.batch)[0])
        stop = a.minutes and time.time() - t0 > 60 * a.minutes
        print(json.dumps({"phase": "train", "epoch": ep, "loss": float(np.mean(losses[-50:])),
                          "val_loss": float(np.mean(vl)), "target_var": float(Tv.var()),
                          "seconds": round(time.time() - t1, 1)}), flush=True)
        if stop:
            break
    out = a.out or G.weights_path(grid)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    net.save(out)
    print(json.dumps({"phase": "saved", "path": out}), flush=True)


if __name__ == "__main__":
    main()
