"""Pin of the U-Net oracle's hand-written backward pass by central finite
differences (float64), on a narrow copy of the topology (base 4 channels)."""
import numpy as np

from oracle import unet_ref as U


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    base = 4
    theta = rng.normal(scale=0.3, size=U.num_params(base))
    x = rng.normal(size=(2, 8, 16, 4))
    t = rng.normal(size=(2, 8, 16, 2))
    net = U.UNet(theta, base)
    out = net.forward(x)
    loss, dout = U.mse_loss_and_grad(out, t)
    g, dx = net.backward(dout)

    def L(th, xx):
        return U.mse_loss_and_grad(U.UNet(th, base).forward(xx), t)[0]

    eps = 1e-6
    for _ in range(3):
        u = rng.normal(size=theta.shape)
        fd = (L(theta + eps * u, x) - L(theta - eps * u, x)) / (2 * eps)
        assert abs(fd - g @ u) < 1e-6 * max(1.0, abs(fd)), (fd, g @ u)
    u = rng.normal(size=x.shape)
    fd = (L(theta, x + eps * u) - L(theta, x - eps * u)) / (2 * eps)
    assert abs(fd - np.sum(dx * u)) < 1e-6 * max(1.0, abs(fd))


def test_parameter_layout_count():
    # 15 convolutions: 1 947 202 parameters at base 32
    assert U.num_params(32) == sum(ks * ks * ci * co + co for _, ci, co, ks, _, _ in U.layers(32))
    assert U.num_params(32) == 1947202


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.0 + 2 ** -9, -3.14159, 0.0])
    r = U.to_bf16(v)
    assert r[0] == 1.0 and r[4] == 0.0
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8 + 1e-300)
