"""Pins the oracle's rate network (oracle/mlp.py): bfloat16 rounding against known
bit patterns, and the FP64 backward against central differences."""
import numpy as np

from oracle import mlp as M


def test_bf16_rounding():
    x = np.array([1.0, -2.5, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 3.4e38, np.inf, 0.0], np.float32)
    r = M.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins;
    # 1 + 3 * 2^-9 rounds up to 1 + 2^-7
    assert list(r[:4]) == [1.0, -2.5, 1.0, 1.0 + 2 ** -7]
    assert r[5] == np.inf and r[6] == 0.0
    assert np.isnan(M.bf16(np.array([np.nan], np.float32)))[0]
    # relative error of a random sample: at most half a unit in the 8-bit significand
    y = np.random.default_rng(0).normal(size=10000).astype(np.float32)
    assert np.max(np.abs(M.bf16(y) - y) / np.abs(y)) <= 2.0 ** -8


def test_backward_matches_central_differences():
    rng = np.random.default_rng(1)
    T = 40
    X = rng.normal(size=(T, M.N_IN))
    P = rng.normal(size=M.N_PARAMS) * 0.2
    t = rng.normal(size=T)
    w = rng.uniform(0.5, 2.0, size=T)
    loss, g = M.loss_grad(X, P, t, w, quant=False)
    eps = 1e-6
    for k in rng.choice(M.N_PARAMS, size=60, replace=False):
        Pp = P.copy(); Pp[k] += eps
        Pm = P.copy(); Pm[k] -= eps
        fd = (M.loss_grad(X, Pp, t, w, quant=False)[0] - M.loss_grad(X, Pm, t, w, quant=False)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_adam_reduces_loss():
    rng = np.random.default_rng(2)
    X = rng.normal(size=(200, M.N_IN))
    true = rng.normal(size=M.N_PARAMS) * 0.2
    t, _ = M.forward(X, true, quant=False)
    P = rng.normal(size=M.N_PARAMS) * 0.2
    m = np.zeros_like(P); v = np.zeros_like(P)
    l0 = M.loss_grad(X, P, t, quant=False)[0]
    for s in range(1, 301):
        l, g = M.loss_grad(X, P, t, quant=False)
        P, m, v = M.adam_step(P, g, m, v, 1e-2, 0.9, 0.999, 1e-8, s)
    assert M.loss_grad(X, P, t, quant=False)[0] < 0.1 * l0
