"""Pins the U-Net oracle: its BF16 rounding against torch's bfloat16
conversion; its forward pass (convolution, pooling, upsampling and the whole
network in float64) against torch's conv2d, avg_pool2d and nearest
interpolation; its backward pass against central finite differences of its
own forward pass in float64 with rounding off."""
import numpy as np
import pytest

from oracle import unet as U


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 20000),
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, -0.0, 65504.0],
                                 np.float32)])
    mine = U.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert mine[-6] == 1.0 and mine[-5] == np.float32(1.0 + 2.0 ** -6)


def _flat_loss(net, P, x, t, m):
    net.P = P
    y = net.forward(x)
    return U.masked_mse(y, t, m)[0]


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_backward_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    cin, cout, B = 3, 2, 2
    P = U.init_params(cin, cout, B, seed)
    for k in P:  # non-zero biases so that ReLUs are not degenerate
        P[k] = P[k].astype(np.float64)
        if k.endswith(".b"):
            P[k] = rng.standard_normal(P[k].shape) * 0.1
    N, H, W = 1, 8, 16
    x = rng.standard_normal((N, H, W, cin))
    t = rng.standard_normal((N, H, W, cout))
    m = rng.random((N, H, W)) > 0.2
    net = U.UNet(P, cin, cout, B, exact=True)
    y = net.forward(x)
    _, dy = U.masked_mse(y, t, m)
    G = net.backward(dy)
    eps = 1e-6
    checked = 0
    for k in sorted(P):
        flat = P[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(3, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp = _flat_loss(net, P, x, t, m)
            flat[idx] = old - eps
            lm = _flat_loss(net, P, x, t, m)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            g = G[k].reshape(-1)[idx]
            assert abs(fd - g) <= 1e-6 + 1e-4 * abs(fd), f"{k}[{idx}]: finite difference {fd}, backward {g}"
            checked += 1
    assert checked >= 40


def test_oracle_adam_first_step_is_signed_lr():
    """With zero moments, Adam's first step is -lr * sign(g) (for |g| >> eps)."""
    P = {"w": np.ones(4, np.float32)}
    G = {"w": np.array([3.0, -2.0, 0.5, -1e-3], np.float32)}
    U.adam(P, G, {}, lr=0.01)
    assert np.allclose(P["w"], 1 - 0.01 * np.sign(G["w"]), rtol=0, atol=1e-6)


def _torch_weight(w, k):
    """Oracle weights W[k*k][ci][co] (taps row-major over (dy, dx)) as torch's
    conv2d weight [co][ci][ky][kx]."""
    torch = pytest.importorskip("torch")
    kk, ci, co = w.shape
    return torch.from_numpy(np.ascontiguousarray(w.reshape(k, k, ci, co).transpose(3, 2, 0, 1)))


def _nchw(x):
    torch = pytest.importorskip("torch")
    return torch.from_numpy(np.ascontiguousarray(x.transpose(0, 3, 1, 2)))


def _nhwc(t):
    return t.numpy().transpose(0, 2, 3, 1)


@pytest.mark.parametrize("k", [3, 1])
def test_oracle_conv_matches_torch(k):
    """The oracle's convolution (zero padding, bias) against
    torch.nn.functional.conv2d in float64: a transposed tap order, a flipped
    kernel or a padding off by one would each fail."""
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(k)
    x = rng.standard_normal((2, 8, 12, 5))
    w = rng.standard_normal((k * k, 5, 7))
    b = rng.standard_normal(7)
    mine = U.conv(x, w, b, k)
    ref = _nhwc(F.conv2d(_nchw(x), _torch_weight(w, k), torch.from_numpy(b), padding=k // 2))
    assert np.allclose(mine, ref, rtol=1e-12, atol=1e-12)


def test_oracle_pool_and_upsample_match_torch():
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 8, 12, 3))
    assert np.allclose(U.pool(x), _nhwc(F.avg_pool2d(_nchw(x), 2)), rtol=1e-15, atol=0)
    assert np.array_equal(U.up(x), _nhwc(F.interpolate(_nchw(x), scale_factor=2, mode="nearest")))


def test_oracle_forward_matches_torch_network():
    """The whole forward pass (exact float64, no BF16 rounding) against the
    same network written with torch's conv2d, avg_pool2d, nearest
    interpolation and channel concatenation."""
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(5)
    cin, cout, B = 5, 2, 4
    P = {k: v.astype(np.float64) for k, v in U.init_params(cin, cout, B, 5).items()}
    for k in P:
        if k.endswith(".b"):
            P[k] = rng.standard_normal(P[k].shape) * 0.1
    x = rng.standard_normal((2, 16, 24, cin))
    mine = U.UNet(P, cin, cout, B, exact=True).forward(x)

    def cr(name, t):
        k = 1 if name == "head" else 3
        y = F.conv2d(t, _torch_weight(P[name + ".w"], k), torch.from_numpy(P[name + ".b"]), padding=k // 2)
        return y if name == "head" else F.relu(y)

    t = _nchw(x)
    e1 = cr("e1b", cr("e1a", t))
    e2 = cr("e2b", cr("e2a", F.avg_pool2d(e1, 2)))
    e3 = cr("e3b", cr("e3a", F.avg_pool2d(e2, 2)))
    bt = cr("b_b", cr("b_a", F.avg_pool2d(e3, 2)))
    up = lambda z: F.interpolate(z, scale_factor=2, mode="nearest")  # noqa: E731
    d3 = cr("d3b", cr("d3a", torch.cat([up(bt), e3], dim=1)))
    d2 = cr("d2b", cr("d2a", torch.cat([up(d3), e2], dim=1)))
    d1 = cr("d1b", cr("d1a", torch.cat([up(d2), e1], dim=1)))
    ref = _nhwc(cr("head", d1))
    assert np.allclose(mine, ref, rtol=1e-10, atol=1e-12)
