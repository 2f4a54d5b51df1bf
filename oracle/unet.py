"""CPU oracle of the warm-start U-Net (test infrastructure only).

The network has no counterpart in the reference: it predicts, from the state
of time step n, the change of pressure and temperature over the step, used as
the starting guess of the first Jacobi sweep.  This module is the numpy
statement of exactly what the HIP kernels compute, used by the tests to check
the GPU forward pass, the backward pass (every parameter gradient and the
input-side gradients) and the optimiser step.  Its own backward pass is pinned
by finite differences in float64 with the BF16 rounding switched off
(tests/test_unet_cpu.py).

Architecture (depth 4, base width B):
  enc1: conv3x3(Cin->B), ReLU, conv3x3(B->B), ReLU            -> s1
  avg-pool 2x2
  enc2: conv(B->2B), ReLU, conv(2B->2B), ReLU                  -> s2
  avg-pool 2x2
  enc3: conv(2B->4B), ReLU, conv(4B->4B), ReLU                 -> s3
  avg-pool 2x2
  bott: conv(4B->8B), ReLU, conv(8B->8B), ReLU
  nearest 2x up, concat [up, s3]: conv(12B->4B), ReLU, conv(4B->4B), ReLU
  nearest 2x up, concat [up, s2]: conv(6B->2B), ReLU, conv(2B->2B), ReLU
  nearest 2x up, concat [up, s1]: conv(3B->B), ReLU, conv(B->B), ReLU
  head: conv1x1(B->Cout), linear
All 3x3 convolutions are "same" with zero padding and a bias.  Tensors are
NHWC (channels fastest), H = rows (j), W = columns (i).

Precision points (the GPU kernels round at the same places): weights are used
as BF16 (master copies FP32); every product sum accumulates in FP32; each
conv's output (bias added, ReLU applied) is stored as BF16; the pooled tensor
is stored as BF16; the head's output is FP32.  Backward: the loss gradient is
FP32; the gradient reaching each conv's output is masked by its ReLU and
rounded to BF16 before the weight, bias and input gradients are formed (FP32
accumulation); input-side gradients are stored as BF16.
"""
import numpy as np

DEPTH = 4


def bf16(x, on=True):
    """Round float32 values to BF16 (round to nearest even), kept as float32."""
    if not on:
        return x
    a = np.ascontiguousarray(x, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def conv_specs(cin, cout, base):
    """(name, cin, cout, k) of every convolution, in forward order."""
    B = base
    return [("e1a", cin, B, 3), ("e1b", B, B, 3), ("e2a", B, 2 * B, 3), ("e2b", 2 * B, 2 * B, 3),
            ("e3a", 2 * B, 4 * B, 3), ("e3b", 4 * B, 4 * B, 3), ("b_a", 4 * B, 8 * B, 3), ("b_b", 8 * B, 8 * B, 3),
            ("d3a", 12 * B, 4 * B, 3), ("d3b", 4 * B, 4 * B, 3), ("d2a", 6 * B, 2 * B, 3), ("d2b", 2 * B, 2 * B, 3),
            ("d1a", 3 * B, B, 3), ("d1b", B, B, 3), ("head", B, cout, 1)]


def init_params(cin, cout, base, seed):
    """He-normal weights W[k*k][ci][co], zero biases (float32)."""
    rng = np.random.default_rng(seed)
    P = {}
    for name, ci, co, k in conv_specs(cin, cout, base):
        std = np.sqrt(2.0 / (k * k * ci)) if name != "head" else np.sqrt(1.0 / ci)
        P[name + ".w"] = (rng.standard_normal((k * k, ci, co)) * std).astype(np.float32)
        P[name + ".b"] = np.zeros(co, np.float32)
    return P


def _taps(k):
    r = k // 2
    return [(dy, dx) for dy in range(-r, r + 1) for dx in range(-r, r + 1)]


def conv(x, w, b, k):
    """x [N,H,W,Ci], w [k*k,Ci,Co] -> FP32 sum + bias, zero padding."""
    N, H, W, _ = x.shape
    acc = x.dtype.type
    out = np.zeros((N, H, W, w.shape[2]), x.dtype)
    r = k // 2
    xp = np.pad(x, ((0, 0), (r, r), (r, r), (0, 0)))
    for t, (dy, dx) in enumerate(_taps(k)):
        out += xp[:, r + dy:r + dy + H, r + dx:r + dx + W, :] @ w[t]
    return out + b.astype(acc)


def conv_back(x, w, dz, k, need_dx=True):
    """Gradients of conv: dW [k*k,Ci,Co], db [Co], dX [N,H,W,Ci]."""
    N, H, W, Ci = x.shape
    r = k // 2
    xp = np.pad(x, ((0, 0), (r, r), (r, r), (0, 0)))
    dzf = dz.reshape(-1, dz.shape[-1])
    dw = np.zeros_like(w)
    for t, (dy, dx) in enumerate(_taps(k)):
        dw[t] = xp[:, r + dy:r + dy + H, r + dx:r + dx + W, :].reshape(-1, Ci).T @ dzf
    db = dzf.sum(axis=0)
    dxo = None
    if need_dx:
        dxp = np.zeros_like(xp)
        for t, (dy, dx) in enumerate(_taps(k)):
            dxp[:, r + dy:r + dy + H, r + dx:r + dx + W, :] += dz @ w[t].T
        dxo = dxp[:, r:r + H, r:r + W, :]
    return dw, db, dxo


def pool(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4), dtype=x.dtype)


def pool_back(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) * d.dtype.type(0.25)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_back(d):
    N, H, W, C = d.shape
    return d.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4), dtype=d.dtype)


class UNet:
    """Forward/backward of the network.  exact=True computes in float64 with
    no BF16 rounding (used to pin the backward pass by finite differences)."""

    def __init__(self, params, cin, cout, base, exact=False):
        self.P, self.cin, self.cout, self.base, self.exact = params, cin, cout, base, exact
        self.dt = np.float64 if exact else np.float32
        self.specs = {s[0]: s for s in conv_specs(cin, cout, base)}

    def _r(self, x):
        return x if self.exact else bf16(x)

    def _w(self, name):
        return self._r(self.P[name + ".w"].astype(self.dt))

    def _conv_relu(self, name, x):
        k = self.specs[name][3]
        z = conv(x, self._w(name), self.P[name + ".b"].astype(self.dt), k)
        return self._r(np.maximum(z, 0))

    def forward(self, x):
        """x [N,H,W,Cin] (H, W divisible by 8) -> y [N,H,W,Cout]; keeps the
        tensors the backward pass needs."""
        T = {}
        T["x"] = self._r(x.astype(self.dt))
        T["e1a"] = self._conv_relu("e1a", T["x"])
        T["e1b"] = self._conv_relu("e1b", T["e1a"])
        T["p1"] = self._r(pool(T["e1b"]))
        T["e2a"] = self._conv_relu("e2a", T["p1"])
        T["e2b"] = self._conv_relu("e2b", T["e2a"])
        T["p2"] = self._r(pool(T["e2b"]))
        T["e3a"] = self._conv_relu("e3a", T["p2"])
        T["e3b"] = self._conv_relu("e3b", T["e3a"])
        T["p3"] = self._r(pool(T["e3b"]))
        T["b_a"] = self._conv_relu("b_a", T["p3"])
        T["b_b"] = self._conv_relu("b_b", T["b_a"])
        T["c3"] = np.concatenate([up(T["b_b"]), T["e3b"]], axis=-1)
        T["d3a"] = self._conv_relu("d3a", T["c3"])
        T["d3b"] = self._conv_relu("d3b", T["d3a"])
        T["c2"] = np.concatenate([up(T["d3b"]), T["e2b"]], axis=-1)
        T["d2a"] = self._conv_relu("d2a", T["c2"])
        T["d2b"] = self._conv_relu("d2b", T["d2a"])
        T["c1"] = np.concatenate([up(T["d2b"]), T["e1b"]], axis=-1)
        T["d1a"] = self._conv_relu("d1a", T["c1"])
        T["d1b"] = self._conv_relu("d1b", T["d1a"])
        y = conv(T["d1b"], self._w("head"), self.P["head.b"].astype(self.dt), 1)
        self.T = T
        return y

    def backward(self, dy):
        """dy = dL/dy [N,H,W,Cout] -> dict of parameter gradients (same
        shapes as the parameters)."""
        T, G = self.T, {}
        r = self._r
        dw, db, d = conv_back(T["d1b"], self._w("head"), dy.astype(self.dt), 1)
        G["head.w"], G["head.b"] = dw, db
        d = r(d)

        def back(name, x_in, d_out, need_dx=True):
            dz = r(d_out * (T[name] > 0))
            k = self.specs[name][3]
            dw, db, dx = conv_back(x_in, self._w(name), dz, k, need_dx)
            G[name + ".w"], G[name + ".b"] = dw, db
            return None if dx is None else r(dx)

        B = self.base
        d = back("d1b", T["d1a"], d)
        dc1 = back("d1a", T["c1"], d)
        d_e1b = dc1[..., 2 * B:]
        d = r(up_back(dc1[..., :2 * B]))
        d = back("d2b", T["d2a"], d)
        dc2 = back("d2a", T["c2"], d)
        d_e2b = dc2[..., 4 * B:]
        d = r(up_back(dc2[..., :4 * B]))
        d = back("d3b", T["d3a"], d)
        dc3 = back("d3a", T["c3"], d)
        d_e3b = dc3[..., 8 * B:]
        d = r(up_back(dc3[..., :8 * B]))
        d = back("b_b", T["b_a"], d)
        d = back("b_a", T["p3"], d)
        d = r(r(pool_back(d)) + d_e3b)
        d = back("e3b", T["e3a"], d)
        d = back("e3a", T["p2"], d)
        d = r(r(pool_back(d)) + d_e2b)
        d = back("e2b", T["e2a"], d)
        d = back("e2a", T["p1"], d)
        d = r(r(pool_back(d)) + d_e1b)
        d = back("e1b", T["e1a"], d)
        back("e1a", T["x"], d, need_dx=False)
        return G


def masked_mse(y, t, mask):
    """L = sum over cells with mask of (y - t)^2 / (2 * count * Cout); returns
    (L, dL/dy)."""
    m = mask[..., None].astype(y.dtype)
    n = max(1.0, float(mask.sum())) * y.shape[-1]
    e = (y - t) * m
    return float((e * e).sum() / (2 * n)), (e / y.dtype.type(n)).astype(y.dtype)


def adam(P, G, state, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step on the FP32 master parameters (in place)."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k in P:
        m = state.setdefault("m." + k, np.zeros_like(P[k]))
        v = state.setdefault("v." + k, np.zeros_like(P[k]))
        g = G[k].astype(np.float32)
        m[...] = np.float32(b1) * m + np.float32(1 - b1) * g
        v[...] = np.float32(b2) * v + np.float32(1 - b2) * g * g
        mh = m / np.float32(1 - b1 ** t)
        vh = v / np.float32(1 - b2 ** t)
        P[k] -= (np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))).astype(np.float32)
