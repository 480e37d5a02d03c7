"""CPU oracle of the learned-guess U-Net (test infrastructure only).

A float64 numpy restatement of the network in trans_bl_amd/csrc/tbl_net.cpp
(same topology, parameter layout and im2col ordering) with a hand-written
backward pass.  The backward pass itself is pinned by central finite
differences in tests/test_oracle_unet.py.  Used only as the reference the
GPU forward/backward kernels are compared against.

Parameter layout: for each convolution in order E0A..OUT, the weights
W[ks*ks*cin, cout] (row = tap*cin + ci, tap = (dh+1)*3 + (dw+1)), then the
bias[cout].
"""
from __future__ import annotations

import numpy as np

B = 32


def layers(base=B):
    # (name, cin, cout, ks, level, relu)
    return [
        ("E0A", 4, base, 3, 0, 1), ("E0B", base, base, 3, 0, 1),
        ("E1A", base, 2 * base, 3, 1, 1), ("E1B", 2 * base, 2 * base, 3, 1, 1),
        ("E2A", 2 * base, 4 * base, 3, 2, 1), ("E2B", 4 * base, 4 * base, 3, 2, 1),
        ("E3A", 4 * base, 8 * base, 3, 3, 1), ("E3B", 8 * base, 8 * base, 3, 3, 1),
        ("D2A", 12 * base, 4 * base, 3, 2, 1), ("D2B", 4 * base, 4 * base, 3, 2, 1),
        ("D1A", 6 * base, 2 * base, 3, 1, 1), ("D1B", 2 * base, 2 * base, 3, 1, 1),
        ("D0A", 3 * base, base, 3, 0, 1), ("D0B", base, base, 3, 0, 1),
        ("OUT", base, 2, 1, 0, 0),
    ]


def unpack(theta, base=B):
    out, off = {}, 0
    for name, cin, cout, ks, lvl, relu in layers(base):
        nw = ks * ks * cin * cout
        out[name] = (theta[off:off + nw].reshape(ks * ks * cin, cout),
                     theta[off + nw:off + nw + cout], ks, relu)
        off += nw + cout
    return out


def num_params(base=B):
    return sum(ks * ks * cin * cout + cout for _, cin, cout, ks, _, _ in layers(base))


def im2col(x):
    n, h, w, c = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    cols = []
    for t in range(9):
        dh, dw = t // 3 - 1, t % 3 - 1
        cols.append(xp[:, 1 + dh:1 + dh + h, 1 + dw:1 + dw + w, :])
    return np.stack(cols, axis=3).reshape(n * h * w, 9 * c)


def col2im(dcol, shape):
    n, h, w, c = shape
    d = dcol.reshape(n, h, w, 9, c)
    dxp = np.zeros((n, h + 2, w + 2, c))
    for t in range(9):
        dh, dw = t // 3 - 1, t % 3 - 1
        dxp[:, 1 + dh:1 + dh + h, 1 + dw:1 + dw + w, :] += d[:, :, :, t, :]
    return dxp[:, 1:-1, 1:-1, :]


def pool(x):
    n, h, w, c = x.shape
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])


def pool_back(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_back(dy):
    return dy[:, 0::2, 0::2] + dy[:, 0::2, 1::2] + dy[:, 1::2, 0::2] + dy[:, 1::2, 1::2]


class UNet:
    """bf16=True rounds where the GPU stores bfloat16: the input, every
    post-activation and pooled activation, and each convolution's output
    gradient before its GEMMs; accumulation stays float64."""

    def __init__(self, theta, base=B, bf16=False):
        self.p = unpack(np.asarray(theta, dtype=np.float64), base)
        self.base = base
        self.r = to_bf16 if bf16 else (lambda a: a)

    def conv(self, name, x):
        W, b, ks, relu = self.p[name]
        n, h, w, _ = x.shape
        col = im2col(x) if ks == 3 else x.reshape(n * h * w, -1)
        z = col @ W + b
        y = np.maximum(z, 0.0) if relu else z
        if relu:
            y = self.r(y)
        return y.reshape(n, h, w, -1), (x, col, y)

    def conv_back(self, name, cache, dy):
        W, b, ks, relu = self.p[name]
        x, col, y = cache
        n, h, w, _ = x.shape
        dz = dy.reshape(n * h * w, -1)
        if relu:
            dz = dz * (y > 0)
        dz = self.r(dz)
        dW = col.T @ dz
        db = dz.sum(axis=0)
        dcol = dz @ W.T
        dx = col2im(dcol, x.shape) if ks == 3 else dcol.reshape(x.shape)
        return dx, dW, db

    def forward(self, x):
        c, r = {}, self.r
        a, c["E0A"] = self.conv("E0A", r(x))
        a0, c["E0B"] = self.conv("E0B", a)
        a, c["E1A"] = self.conv("E1A", r(pool(a0)))
        a1, c["E1B"] = self.conv("E1B", a)
        a, c["E2A"] = self.conv("E2A", r(pool(a1)))
        a2, c["E2B"] = self.conv("E2B", a)
        a, c["E3A"] = self.conv("E3A", r(pool(a2)))
        a3, c["E3B"] = self.conv("E3B", a)
        a, c["D2A"] = self.conv("D2A", np.concatenate([up(a3), a2], axis=3))
        b2, c["D2B"] = self.conv("D2B", a)
        a, c["D1A"] = self.conv("D1A", np.concatenate([up(b2), a1], axis=3))
        b1, c["D1B"] = self.conv("D1B", a)
        a, c["D0A"] = self.conv("D0A", np.concatenate([up(b1), a0], axis=3))
        b0, c["D0B"] = self.conv("D0B", a)
        out, c["OUT"] = self.conv("OUT", b0)
        self.cache = c
        return out

    def backward(self, dout):
        """Gradients (flat vector, same layout as theta) and d/d(input)."""
        c, g = self.cache, {}
        B_ = self.base
        d, *g["OUT"] = self.conv_back("OUT", c["OUT"], dout)
        d, *g["D0B"] = self.conv_back("D0B", c["D0B"], d)
        d, *g["D0A"] = self.conv_back("D0A", c["D0A"], d)
        d_b1, d_a0 = up_back(d[..., :2 * B_]), d[..., 2 * B_:]
        d, *g["D1B"] = self.conv_back("D1B", c["D1B"], d_b1)
        d, *g["D1A"] = self.conv_back("D1A", c["D1A"], d)
        d_b2, d_a1 = up_back(d[..., :4 * B_]), d[..., 4 * B_:]
        d, *g["D2B"] = self.conv_back("D2B", c["D2B"], d_b2)
        d, *g["D2A"] = self.conv_back("D2A", c["D2A"], d)
        d_a3, d_a2 = up_back(d[..., :8 * B_]), d[..., 8 * B_:]
        d, *g["E3B"] = self.conv_back("E3B", c["E3B"], d_a3)
        d, *g["E3A"] = self.conv_back("E3A", c["E3A"], d)
        d_a2 = d_a2 + pool_back(d)
        d, *g["E2B"] = self.conv_back("E2B", c["E2B"], d_a2)
        d, *g["E2A"] = self.conv_back("E2A", c["E2A"], d)
        d_a1 = d_a1 + pool_back(d)
        d, *g["E1B"] = self.conv_back("E1B", c["E1B"], d_a1)
        d, *g["E1A"] = self.conv_back("E1A", c["E1A"], d)
        d_a0 = d_a0 + pool_back(d)
        d, *g["E0B"] = self.conv_back("E0B", c["E0B"], d_a0)
        dx, *g["E0A"] = self.conv_back("E0A", c["E0A"], d)
        flat = []
        for name, *_ in layers(self.base):
            dW, db = g[name]
            flat += [dW.ravel(), db.ravel()]
        return np.concatenate(flat), dx


def mse_loss_and_grad(out, target):
    d = out - target
    n = d.size
    return float(np.mean(d * d)), 2.0 * d / n


def to_bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
