"""CPU restatement of the per-tile rate network (test infrastructure only).

A 32 -> 64 -> 64 -> 1 perceptron with ReLU hidden layers; parameter block
(6337 floats): W1 [32][64], b1 [64], W2 [64][64], b2 [64], W3 [64], b3 [1].
With quant=True the operands are rounded to bfloat16 (round to nearest even)
at the points where the GPU kernels store bfloat16 (inputs, weights, hidden
activations, hidden-layer gradients), and everything else runs in FP64; with
quant=False the whole network runs in FP64 (for finite-difference checks).
The network has no counterpart in the reference.
"""
import numpy as np

N_IN, N_HID = 32, 64
OFF_W1 = 0
OFF_B1 = N_IN * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID
N_PARAMS = OFF_B3 + 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, np.float64)
    return (P[OFF_W1:OFF_B1].reshape(N_IN, N_HID), P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(N_HID, N_HID),
            P[OFF_B2:OFF_W3], P[OFF_W3:OFF_B3], P[OFF_B3])


def _q(x, quant):
    return bf16(x) if quant else np.asarray(x, np.float64)


def forward(X, P, quant=True):
    """Returns (y, cache)."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    Xb = _q(X, quant)
    W1b, W2b, W3b = _q(W1, quant), _q(W2, quant), _q(W3, quant)
    H1 = _q(np.maximum(Xb @ W1b + b1, 0.0), quant)
    H2 = _q(np.maximum(H1 @ W2b + b2, 0.0), quant)
    y = H2 @ W3b + b3
    return y, (Xb, W1b, W2b, W3b, H1, H2)


def loss_grad(X, P, target, weight=None, quant=True):
    """Weighted mean-square loss sum w (y - t)^2 / sum w and its parameter gradient."""
    y, (Xb, W1b, W2b, W3b, H1, H2) = forward(X, P, quant)
    w = np.ones(len(y)) if weight is None else np.asarray(weight, np.float64)
    e = y - np.asarray(target, np.float64)
    loss = (w * e * e).sum() / w.sum()
    dy = 2.0 * w * e / w.sum()
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = H2.T @ dy
    g[OFF_B3] = dy.sum()
    dZ2 = np.outer(dy, W3b) * (H2 > 0)
    dZ2b = _q(dZ2, quant)
    g[OFF_W2:OFF_B2] = (H1.T @ dZ2b).ravel()
    g[OFF_B2:OFF_W3] = dZ2.sum(0)
    dZ1 = (dZ2b @ W2b.T) * (H1 > 0)
    dZ1b = _q(dZ1, quant)
    g[OFF_W1:OFF_B1] = (Xb.T @ dZ1b).ravel()
    g[OFF_B1:OFF_W2] = dZ1.sum(0)
    return loss, g


def adam_step(P, g, m, v, lr, b1, b2, eps, step):
    """Bias-corrected Adam (Kingma & Ba 2015); returns new (P, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return P - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
