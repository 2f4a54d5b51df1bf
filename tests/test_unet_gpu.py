"""The U-Net's HIP forward pass, backward pass (every parameter gradient) and
Adam step against the CPU oracle (oracle/unet.py) on the same parameters and
inputs, and a short training run that must lower the loss.

Tolerances: the GPU sums in FP32 in another order than numpy, so a BF16
rounding of an intermediate activation or gradient occasionally lands on
the other side (one BF16 ulp, 2^-8 relative); through 15 layers this leaves
relative L2 differences of order 1e-3 (3.7e-3 measured at base 32, 0 at
base 8).  Forward outputs must agree to 5e-3, each gradient tensor to 3e-2
relative L2 (cosine > 0.999), Adam to 1e-6."""
import importlib

import numpy as np
import pytest

from oracle import unet as U

pytestmark = pytest.mark.gpu

CIN, COUT = 5, 2


def flat(P, base):
    return np.concatenate([np.concatenate([P[n + ".w"].ravel(), P[n + ".b"].ravel()])
                           for n, _, _, _ in U.conv_specs(CIN, COUT, base)]).astype(np.float32)


def unflat(v, base):
    P, o = {}, 0
    for n, ci, co, k in U.conv_specs(CIN, COUT, base):
        P[n + ".w"] = v[o:o + k * k * ci * co].reshape(k * k, ci, co)
        o += k * k * ci * co
        P[n + ".b"] = v[o:o + co]
        o += co
    assert o == v.size
    return P


def problem(base, N, H, W, seed):
    rng = np.random.default_rng(seed)
    P = U.init_params(CIN, COUT, base, seed)
    for k in P:
        if k.endswith(".b"):
            P[k] = (rng.standard_normal(P[k].shape) * 0.05).astype(np.float32)
    x = rng.standard_normal((N, H, W, CIN)).astype(np.float32)
    x[..., 4] = (rng.random((N, H, W)) > 0.1)  # mask channel
    t = rng.standard_normal((N, H, W, COUT)).astype(np.float32)
    m = x[..., 4] > 0
    return P, x, t, m


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def net_for(base, N, H, W, P):
    pkg = importlib.import_module("simple-ts_amd")
    net = pkg.UNet(CIN, COUT, base, N, H, W)
    assert net.n_params == flat(P, base).size
    net.set_params(flat(P, base))
    return net


@pytest.mark.parametrize("base", [8, 32])
def test_forward_matches_oracle(base):
    N, H, W = 2, 16, 64
    P, x, _, _ = problem(base, N, H, W, 1)
    print(f"U-Net forward base={base} N={N} H={H} W={W}", flush=True)
    net = net_for(base, N, H, W, P)
    y = net.forward(x)
    yr = U.UNet(P, CIN, COUT, base).forward(x)
    e = rel_l2(y, yr)
    print(f"forward rel L2 {e:.2e}", flush=True)
    assert np.all(np.isfinite(y)) and e <= 5e-3


@pytest.mark.parametrize("base", [8, 32])
def test_inference_equals_forward(base):
    """The inference schedule (concatenations read in place) gives the
    forward pass's outputs bit for bit."""
    N, H, W = 2, 16, 64
    P, x, _, _ = problem(base, N, H, W, 7)
    net = net_for(base, N, H, W, P)
    assert np.array_equal(net.infer(x), net.forward(x))


@pytest.mark.parametrize("base", [8, 32])
def test_backward_matches_oracle(base):
    N, H, W = 2, 16, 64
    P, x, t, m = problem(base, N, H, W, 2)
    print(f"U-Net backward base={base}", flush=True)
    net = net_for(base, N, H, W, P)
    net.forward(x)
    loss = net.backward_mse(t, m)
    G = unflat(net.grads(), base)
    ref = U.UNet(P, CIN, COUT, base)
    lr_, dy = U.masked_mse(ref.forward(x), t, m)
    Gr = ref.backward(dy)
    assert abs(loss - lr_) <= 1e-2 * abs(lr_)
    worst = 0.0
    for k in Gr:
        e = rel_l2(G[k], Gr[k])
        cos = float(np.dot(G[k].ravel(), Gr[k].ravel()) /
                    max(np.linalg.norm(G[k]) * np.linalg.norm(Gr[k]), 1e-30))
        worst = max(worst, e)
        assert e <= 3e-2 and cos > 0.999, f"{k}: rel L2 {e:.2e}, cosine {cos:.6f}"
    print(f"worst gradient rel L2 {worst:.2e}", flush=True)


def test_backward_given_dy_matches_oracle():
    base, N, H, W = 8, 1, 8, 64
    P, x, _, _ = problem(base, N, H, W, 3)
    rng = np.random.default_rng(5)
    dy = rng.standard_normal((N, H, W, COUT)).astype(np.float32) * 1e-2
    net = net_for(base, N, H, W, P)
    net.forward(x)
    net.backward(dy)
    G = unflat(net.grads(), base)
    ref = U.UNet(P, CIN, COUT, base)
    ref.forward(x)
    Gr = ref.backward(dy)
    for k in Gr:
        assert rel_l2(G[k], Gr[k]) <= 3e-2, k


def test_adam_matches_oracle():
    base, N, H, W = 8, 2, 16, 64
    P, x, t, m = problem(base, N, H, W, 4)
    net = net_for(base, N, H, W, P)
    net.forward(x)
    net.backward_mse(t, m)
    p0, g = net.params(), net.grads()
    net.adam(1e-3)
    p1 = net.params()
    Pr, st = unflat(p0.copy(), base), {}
    U.adam(Pr, unflat(g, base), st, lr=1e-3)
    assert np.max(np.abs(flat(Pr, base) - p1)) <= 1e-6


def test_training_lowers_loss():
    """Fits a fixed batch whose target is a smooth function of the input."""
    base, N, H, W = 8, 2, 16, 64
    P, x, _, m = problem(base, N, H, W, 6)
    t = np.stack([0.5 * x[..., 0] - 0.2 * x[..., 2], 0.3 * x[..., 1]], axis=-1).astype(np.float32)
    net = net_for(base, N, H, W, P)
    losses = []
    for it in range(60):
        net.forward(x)
        losses.append(net.backward_mse(t, m))
        net.adam(3e-3)
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
