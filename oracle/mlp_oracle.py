"""NumPy restatement of the warm-start network (forward pass, mean-squared-error loss, backward
pass, Adam) that the GPU kernels implement.

TEST INFRASTRUCTURE ONLY: the checker of the network kernels in tests/. The reference simulator
has no learned component, so there is no reference file to follow; the restatement is pinned by
the mathematics instead: its gradients are checked against central finite differences of its own
FP64 loss (tests/test_oracle_warm.py), and its BF16 rounding points are the ones the kernels use:
inputs after normalization, weights, hidden activations after bias + ReLU, and the backward
signals dY and dZ are rounded to BF16 (round to nearest even); every sum is exact (FP64) here and
FP32 on the device.
"""
import numpy as np

IN, H, OUT, MAX_IN, MAX_OUT = 48, 64, 32, 40, 12
W1, B1 = 0, IN * H
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H * H
W4 = B3 + H
B4 = W4 + H * OUT
NPARAMS = B4 + OUT


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(IN, H), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3].reshape(H, H), p[B3:W4],
            p[W4:B4].reshape(H, OUT), p[B4:NPARAMS])


def _round(a, rounding):
    return bf16(a) if rounding else np.asarray(a, dtype=np.float64)


def normalize(x, mean, inv_std, rounding=True):
    """(x - mean) * inv_std in FP32 (as the kernels compute it), padded to 48 inputs."""
    x = np.asarray(x, dtype=np.float32)
    n, k = x.shape
    z = np.zeros((n, IN), dtype=np.float32)
    z[:, :k] = (x - np.asarray(mean, np.float32)[:k]) * np.asarray(inv_std, np.float32)[:k]
    return _round(z, rounding)


def forward(p, xn, rounding=True):
    """Network outputs (all 32) and the cached activations for normalized inputs xn [n][48]."""
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    if rounding:
        w1, w2, w3, w4 = bf16(w1), bf16(w2), bf16(w3), bf16(w4)
    a0 = xn
    z1 = a0 @ w1 + b1
    a1 = _round(np.maximum(z1, 0.0).astype(np.float32) if rounding else np.maximum(z1, 0.0), rounding)
    z2 = a1 @ w2 + b2
    a2 = _round(np.maximum(z2, 0.0).astype(np.float32) if rounding else np.maximum(z2, 0.0), rounding)
    z3 = a2 @ w3 + b3
    a3 = _round(np.maximum(z3, 0.0).astype(np.float32) if rounding else np.maximum(z3, 0.0), rounding)
    y = a3 @ w4 + b4
    return y, (a0, a1, a2, a3, (w1, w2, w3, w4))


def loss_grad(p, xn, t, nout, rounding=True):
    """MSE over the first nout outputs against t [n][nout] (already divided by out_scale), and the
    gradient with respect to the flat parameter vector."""
    n = len(xn)
    y, (a0, a1, a2, a3, (w1, w2, w3, w4)) = forward(p, xn, rounding)
    d = np.zeros_like(y)
    d[:, :nout] = y[:, :nout] - t
    loss = float(np.sum(d[:, :nout] ** 2) / (n * nout))
    dy = _round((2.0 / (n * nout) * d).astype(np.float32) if rounding else 2.0 / (n * nout) * d, rounding)
    g = np.zeros(NPARAMS)
    g[W4:B4] = (a3.T @ dy).reshape(-1)
    g[B4:NPARAMS] = dy.sum(0)
    dz3 = _round(((dy @ w4.T) * (a3 > 0)).astype(np.float32) if rounding else (dy @ w4.T) * (a3 > 0), rounding)
    g[W3:B3] = (a2.T @ dz3).reshape(-1)
    g[B3:W4] = dz3.sum(0)
    dz2 = _round(((dz3 @ w3.T) * (a2 > 0)).astype(np.float32) if rounding else (dz3 @ w3.T) * (a2 > 0), rounding)
    g[W2:B2] = (a1.T @ dz2).reshape(-1)
    g[B2:W3] = dz2.sum(0)
    dz1 = _round(((dz2 @ w2.T) * (a1 > 0)).astype(np.float32) if rounding else (dz2 @ w2.T) * (a1 > 0), rounding)
    g[W1:B1] = (a0.T @ dz1).reshape(-1)
    g[B1:W2] = dz1.sum(0)
    return loss, g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected), returning the new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    return p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps), m, v
