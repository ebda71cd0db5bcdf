"""Pins the oracle's pixelgather restatement (oracle.pixel_scatter / pixel_gather,
pixelGatherCuda.cu:7-152): a hand-checked tiny case, gather as the exact adjoint
of scatter (<scatter(v), u> = <v, gather(u)>), and gather of ones counting how
often each splat is sampled."""
import numpy as np

from oracle import oracle as O


def test_tiny_known_answer():
    idx = np.array([[[[2, -1], [0, 2]]]], np.int32)            # (1, 1, 2, 2)
    v = np.array([[[1.0, 10.0], [2.0, 20.0], [3.0, 30.0]]])    # (1, 3, 2)
    out = O.pixel_scatter(idx, v)
    assert np.array_equal(out, np.array([[[[[3, 30], [0, 0]], [[1, 10], [3, 30]]]]], np.float64))
    g = O.pixel_gather(idx, out, 3)
    assert np.array_equal(g, np.array([[[1, 10], [0, 0], [6, 60]]], np.float64))


def test_adjoint_and_counts():
    rng = np.random.default_rng(0)
    bn, W, H, k, N, C = 2, 7, 5, 6, 40, 3
    idx = rng.integers(-1, N, size=(bn, W, H, k)).astype(np.int32)
    v = rng.normal(size=(bn, N, C))
    u = rng.normal(size=(bn, W, H, k, C))
    lhs = (O.pixel_scatter(idx, v) * u).sum()
    rhs = (v * O.pixel_gather(idx, u, N)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    cnt = O.pixel_gather(idx, np.ones((bn, W, H, k, 1)), N)[..., 0]
    for b in range(bn):
        ref = np.bincount(idx[b][idx[b] >= 0], minlength=N)
        assert np.array_equal(cnt[b], ref)
