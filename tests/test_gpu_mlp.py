"""GPU rate network (bfloat16 matrix cores, FP32 accumulation) against the FP64
oracle with the same bfloat16 rounding points: forward, loss gradient, Adam,
and a training run that fits a known network."""
import numpy as np
import pytest
import torch

from oracle import mlp as M

pytestmark = pytest.mark.gpu


def t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)


def rel(a, b):
    return np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("T", [1, 17, 1024, 5000])
def test_mlp_forward_and_grad_match_oracle(dev, T):
    from splatting_amd.learned import RateNet
    rng = np.random.default_rng(T)
    X = rng.normal(size=(T, M.N_IN)).astype(np.float32)
    P = (rng.normal(size=M.N_PARAMS) * 0.2).astype(np.float32)
    tgt = rng.normal(size=T).astype(np.float32)
    w = rng.uniform(0.5, 2.0, size=T).astype(np.float32)
    net = RateNet(t(P, dev), t(np.zeros(M.N_IN), dev), t(np.ones(M.N_IN), dev))
    y = net.forward(t(X, dev)).cpu().numpy()
    yr, _ = M.forward(X, P, quant=True)
    # FP32 accumulation vs FP64, then occasional one-ulp bfloat16 flips of hidden units
    assert rel(y, yr) < 5e-3
    loss, g = net.loss_grad(t(X, dev), t(tgt, dev), t(w, dev))
    lr, gr = M.loss_grad(X, P, tgt, w, quant=True)
    assert abs(float(loss.item()) - lr) <= 1e-2 * lr
    g = g.cpu().numpy()
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAMS)):
        assert rel(g[a:b], gr[a:b]) < 2e-2, (a, b, rel(g[a:b], gr[a:b]))


def test_adam_step_matches_oracle(dev):
    from cudaExtensions import _lib
    rng = np.random.default_rng(3)
    n = 6337
    P, g, m, v = (rng.normal(size=n).astype(np.float32) for _ in range(4))
    v = np.abs(v)
    Pt, gt, mt, vt = t(P, dev), t(g, dev), t(m, dev), t(v, dev)
    rc = _lib.lib().ude_adam_step(n, _lib.ptr(Pt), _lib.ptr(gt), _lib.ptr(mt), _lib.ptr(vt), 1e-3, 0.9, 0.999, 1e-8,
                                  5, _lib.stream(dev))
    assert rc == 0
    Pr, mr, vr = M.adam_step(P.astype(np.float64), g, m, v, 1e-3, 0.9, 0.999, 1e-8, 5)
    assert rel(Pt.cpu().numpy(), Pr) < 1e-6 and rel(mt.cpu().numpy(), mr) < 1e-6 and rel(vt.cpu().numpy(), vr) < 1e-6


def test_training_fits_a_known_network(dev):
    from splatting_amd.learned import RateNet
    rng = np.random.default_rng(4)
    T = 4096
    X = rng.normal(size=(T, M.N_IN)).astype(np.float32)
    true = (rng.normal(size=M.N_PARAMS) * 0.15).astype(np.float32)
    y, _ = M.forward(X, true, quant=False)
    net = RateNet.init(dev, seed=1)
    hist = net.fit(t(X, dev), t(y, dev), steps=1500, lr=3e-3, log_every=500)
    assert hist[-1][1] < 0.05 * hist[0][1], hist
