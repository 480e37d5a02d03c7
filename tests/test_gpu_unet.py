"""The learned-guess network's GPU kernels against the CPU oracle.

* the MFMA GEMM (bf16 in, fp32 accumulate) against numpy, all operand
  layouts the network uses, with and without split-K;
* forward and parameter/input gradients of the U-Net against the numpy
  oracle (oracle/unet_ref.py) run on the same bf16 weights and with bf16
  rounding at the points where the GPU stores bf16 (tolerance 1e-2 relative
  per layer: remaining differences are fp32 vs fp64 accumulation and
  round-to-nearest ties); the plain float64 network is within 5e-2;
* a few Adam steps reduce the training loss.
"""
import numpy as np
import pytest

from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod(gpu_ready):
    import torch
    return torch


def bf16_bits(torch, a):
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16)
    return t.view(torch.int16).to("cuda")


@pytest.mark.parametrize("M,N,K,ta,tb,split", [
    (64, 64, 32, 0, 0, 1), (100, 37, 75, 0, 0, 1), (130, 96, 288, 0, 0, 1),
    (77, 45, 999, 1, 0, 4), (288, 32, 5000, 1, 0, 16), (300, 288, 64, 0, 1, 1)])
def test_mfma_gemm_matches_numpy(torch_mod, M, N, K, ta, tb, split):
    torch = torch_mod
    from trans_bl_amd import guess
    rng = np.random.default_rng(M * N + K)
    A = U.to_bf16(rng.normal(size=(M, K)))
    B = U.to_bf16(rng.normal(size=(K, N)))
    a = bf16_bits(torch, A.T if ta else A)
    b = bf16_bits(torch, B.T if tb else B)
    c = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    guess.mfma_gemm(a.data_ptr(), b.data_ptr(), c.data_ptr(), M, N, K, ta, tb, split)
    ref = A @ B
    got = c.cpu().numpy()
    assert np.max(np.abs(got - ref)) < 1e-4 * np.max(np.abs(ref)) + 1e-4


def _problem(rng, n=2, H=16, W=24):
    x = rng.normal(size=(n, H, W, 4)).astype(np.float32)
    t = rng.normal(scale=0.5, size=(n, H, W, 2)).astype(np.float32)
    return x, t


def test_forward_and_gradients_match_oracle(torch_mod):
    torch = torch_mod
    from trans_bl_amd import guess
    rng = np.random.default_rng(5)
    x, t = _problem(rng)
    net = guess.Net(16, 24, 2, seed=11)
    theta = U.to_bf16(net.get_params())           # compute weights are bf16
    net.set_params(theta.astype(np.float32))
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(t).cuda()
    yd = torch.zeros((2, 16, 24, 2), dtype=torch.float32, device="cuda")
    gx = torch.zeros_like(xd)
    torch.cuda.synchronize()
    net.forward(xd.data_ptr(), yd.data_ptr(), 2)
    loss, g = net.gradients(xd.data_ptr(), td.data_ptr(), 2, gx.data_ptr())
    ref = U.UNet(theta, bf16=True)
    out = ref.forward(x)
    rloss, dout = U.mse_loss_and_grad(out, t)
    rg, rdx = ref.backward(dout)
    y = yd.cpu().numpy()
    assert np.linalg.norm(y - out) / np.linalg.norm(out) < 1e-2
    assert abs(loss - rloss) / rloss < 1e-2
    exact = U.UNet(theta).forward(x)
    assert np.linalg.norm(y - exact) / np.linalg.norm(exact) < 5e-2
    # per-layer relative error of the parameter gradient
    off, errs = 0, {}
    for name, cin, cout, ks, lvl, relu in U.layers():
        n = ks * ks * cin * cout + cout
        a, b = g[off:off + n], rg[off:off + n]
        errs[name] = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        off += n
    assert max(errs.values()) < 1e-2, errs
    gxv = gx.cpu().numpy()
    assert np.linalg.norm(gxv - rdx) / np.linalg.norm(rdx) < 1e-2


def test_forward_wide_image_matches_oracle(torch_mod):
    """Forward at the width of the 1025x193 grid's coarsest level (image
    32x136): the convolution kernels' LDS tiles then span more than one image
    row and reach across rows at both ends."""
    torch = torch_mod
    from trans_bl_amd import guess
    rng = np.random.default_rng(9)
    x, _ = _problem(rng, n=1, H=32, W=136)
    net = guess.Net(32, 136, 1, seed=13)
    theta = U.to_bf16(net.get_params())
    net.set_params(theta.astype(np.float32))
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros((1, 32, 136, 2), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    net.forward(xd.data_ptr(), yd.data_ptr(), 1)
    out = U.UNet(theta, bf16=True).forward(x)
    y = yd.cpu().numpy()
    assert np.linalg.norm(y - out) / np.linalg.norm(out) < 1e-2


def test_training_reduces_loss(torch_mod):
    torch = torch_mod
    from trans_bl_amd import guess
    rng = np.random.default_rng(9)
    x = rng.normal(size=(4, 16, 24, 4)).astype(np.float32)
    # a smooth target the network can fit: a blurred copy of input channel 0
    t = np.stack([np.roll(x[..., 0], 1, axis=2) * 0.5, x[..., 1] * 0.25], axis=-1).astype(np.float32)
    net = guess.Net(16, 24, 4, seed=3)
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(t).cuda()
    torch.cuda.synchronize()
    first = net.train_step(xd.data_ptr(), td.data_ptr(), 4, 1e-3)
    for _ in range(60):
        last = net.train_step(xd.data_ptr(), td.data_ptr(), 4, 1e-3)
    assert last < 0.3 * first, (first, last)


def test_set_params_after_forward_repacks_weights(torch_mod):
    """New weights given after a forward reach every layer, the 3x3
    convolutions' packed copies included."""
    torch = torch_mod
    from trans_bl_amd import guess
    rng = np.random.default_rng(21)
    x, _ = _problem(rng, n=1)
    net = guess.Net(16, 24, 1, seed=5)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros((1, 16, 24, 2), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    net.forward(xd.data_ptr(), yd.data_ptr(), 1)            # packs the first weights
    theta = U.to_bf16(guess.Net(16, 24, 1, seed=77).get_params())
    net.set_params(theta.astype(np.float32))
    net.forward(xd.data_ptr(), yd.data_ptr(), 1)
    out = U.UNet(theta, bf16=True).forward(x)
    y = yd.cpu().numpy()
    assert np.linalg.norm(y - out) / np.linalg.norm(out) < 1e-2
