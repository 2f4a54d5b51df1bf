"""Pins the U-Net oracle: its BF16 rounding against torch's bfloat16
conversion, and its backward pass against central finite differences of its
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
