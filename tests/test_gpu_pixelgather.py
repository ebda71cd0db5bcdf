"""pixelgather.scatter / gather (splatting/extensions/pixelGather.cpp:10-23,
pixelGatherCuda.cu:7-152) through the C ABI: scatter equals the oracle bit for
bit in float32, float64 and float16 (a copy); gather equals the oracle's FP64
sums elementwise within 1e-5 relative plus 64 float32 roundoffs of each output's
magnitude (float32 atomics) and to FP64 rounding in float64; the autograd
wrapper's backward is the adjoint (torch gradcheck in float64); sampled index
tensors with empty slots and an all-empty tensor."""
import numpy as np
import pytest
import torch

from helpers import EPS64, assert_close, magnitude, random_slots
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def t(x, dev, dt=None):
    x = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    return x if dt is None else x.to(dt)


@pytest.mark.parametrize("C", [1, 3, 16])
def test_scatter_gather_against_oracle(dev, C):
    from cudaExtensions import pixelgather
    rng = np.random.default_rng(C)
    bn, W, H, k, N = 2, 33, 21, 12, 500
    idx = random_slots(rng, bn, W, H, k, N)
    v = rng.normal(size=(bn, N, C))
    for dt in (torch.float32, torch.float64, torch.float16):
        vd = t(v, dev, dt)
        out = pixelgather.scatter(t(idx, dev), vd)
        assert out.dtype == dt and out.shape == (bn, W, H, k, C)
        ref = O.pixel_scatter(idx, vd.cpu().numpy())
        assert np.array_equal(out.cpu().numpy(), ref), dt
    u = rng.normal(size=(bn, W, H, k, C))
    ref = O.pixel_gather(idx, u, N)
    mag = O.pixel_gather(idx, np.abs(u), N)
    g32 = pixelgather.gather(t(idx, dev), t(u, dev, torch.float32), N)
    assert_close(g32.cpu().numpy(), O.pixel_gather(idx, u.astype(np.float32), N), 1e-5, "gather f32", mag=mag)
    g64 = pixelgather.gather(t(idx, dev), t(u, dev, torch.float64), N)
    assert_close(g64.cpu().numpy(), ref, 1e-12, "gather f64", mag=mag, ops=256, eps=EPS64)


def test_autograd_adjoint(dev):
    from splatting_amd.extensions import PixelScatterGather
    rng = np.random.default_rng(5)
    idx = t(random_slots(rng, 1, 6, 5, 4, 20), dev)
    v = t(rng.normal(size=(1, 20, 3)), dev, torch.float64).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: PixelScatterGather.apply(idx, x), (v,))


def test_empty_slots_and_shapes(dev):
    from cudaExtensions import pixelgather
    idx = torch.full((1, 4, 4, 3), -1, dtype=torch.int32, device=dev)
    v = torch.randn(1, 10, 2, device=dev)
    assert float(pixelgather.scatter(idx, v).abs().sum()) == 0.0
    assert float(pixelgather.gather(idx, torch.ones(1, 4, 4, 3, 2, device=dev), 10).abs().sum()) == 0.0
    with pytest.raises(ValueError):
        pixelgather.scatter(idx, torch.randn(2, 10, 2, device=dev))
    with pytest.raises(ValueError):
        pixelgather.gather(idx, torch.ones(1, 4, 4, 2, 2, device=dev), 10)
