"""nearestsamples.nearest through the C ABI (FMMKdTreeSampling/nearestSamples.cpp:30-75,
nearestSamplesCuda.cu:290-564) on the regularizer's inputs -- indices of the
reference sampler (k = 16) and the compositing weights of the 4,096-splat
sphere at 128x128, two views -- equal to the oracle bit for bit (float32 and
float64), independent of the order the pixels are inserted in (a permuted
pixel order gives the same rows), and empty inputs."""
import numpy as np
import pytest
import torch

from helpers import scene_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def inputs(dev):
    from cudaExtensions import filtering, sampling, splatting
    sc, p, n, nb, sd = scene_inputs(0, sd_scale=2.0)
    W, H, f = sc.width, sc.height, sc.focal
    rot = np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 0.0], [-1.0, 0.0, 0.0]], np.float32)
    p1 = ((p + np.float32([0, 0, 2.5])) @ rot.T - np.float32([0, 0, 2.5])).astype(np.float32)
    n1 = (n @ rot.T).astype(np.float32)
    P, Nn = t(np.stack([p, p1]), dev), t(np.stack([n, n1]), dev)
    NB = torch.where(Nn[..., 2:3] > 0, -Nn, Nn).contiguous()
    S = t(sd[None], dev)
    idx = sampling.sample(W, H, f, P, NB, None, S, 16, 0.0, 0.01, seed=2)
    w = filtering.forward(idx, splatting.forward(W, H, f, idx, P, Nn, S, 0.0), W, H, f, P, Nn)
    return sc, p, idx, w


@pytest.mark.parametrize("K", [1, 10])
def test_nearest_equals_oracle(dev, K):
    from cudaExtensions import nearestsamples
    sc, p, idx, w = inputs(dev)
    got = nearestsamples.nearest(idx, w, t(p[None], dev), K)
    ref = O.nearest_samples(idx.cpu().numpy(), w.cpu().numpy(), p[None], K)
    assert (ref >= 0).sum() > 1000
    assert np.array_equal(got.cpu().numpy(), ref)
    got64 = nearestsamples.nearest(idx, w.double(), t(p[None], dev).double(), K)
    assert np.array_equal(got64.cpu().numpy(), ref)


def test_nearest_independent_of_pixel_order(dev):
    from cudaExtensions import nearestsamples
    sc, p, idx, w = inputs(dev)
    perm = torch.randperm(idx.shape[1] * idx.shape[2], generator=torch.Generator().manual_seed(0)).to(dev)
    pi = idx.reshape(2, -1, 16)[:, perm].reshape(idx.shape).contiguous()
    pw = w.reshape(2, -1, 16)[:, perm].reshape(w.shape).contiguous()
    a = nearestsamples.nearest(idx, w, t(p[None], dev), 8)
    b = nearestsamples.nearest(pi, pw, t(p[None], dev), 8)
    assert torch.equal(a, b)


def test_nearest_empty(dev):
    from cudaExtensions import nearestsamples
    idx = torch.full((1, 4, 4, 5), -1, dtype=torch.int32, device=dev)
    out = nearestsamples.nearest(idx, torch.zeros(1, 4, 4, 5, device=dev), torch.zeros(1, 7, 3, device=dev), 3)
    assert out.shape == (1, 7, 3) and bool((out == -1).all())
    with pytest.raises(ValueError):
        nearestsamples.nearest(idx, torch.zeros(1, 4, 4, 5, device=dev), torch.zeros(2, 7, 3, device=dev), 3)
