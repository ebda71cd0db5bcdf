"""GPU parity of sampling.approximateFMMWeights: the density image equals the
oracle's bit for bit (same tree, FP64 moments rounded once, FP32 evaluation in
the same order)."""
import numpy as np
import pytest
import torch

from helpers import scene_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def run_gpu(dev, W, H, f, p, nb, sd, prec=0.01, smoothing=0.0):
    from cudaExtensions import sampling
    out = sampling.approximateFMMWeights(W, H, f, t(p[None], dev), t(nb[None], dev), None, t(sd[None], dev),
                                         smoothing, prec)
    return out[0].cpu().numpy()


@pytest.mark.parametrize("cfg,sds,prec", [(2, 8.0, 0.01), (0, 2.0, 0.01), (0, 4.0, 0.001), (0, 1.0, 1e-6)])
def test_fmm_bit_exact(dev, cfg, sds, prec):
    sc, p, n, nb, sd = scene_inputs(cfg, sd_scale=sds)
    got = run_gpu(dev, sc.width, sc.height, sc.focal, p, nb, sd, prec)
    ref = O.fmm_weights_view(sc.width, sc.height, sc.focal, p, nb, sd, 0.0, prec)
    assert ref.max() > 0
    bad = np.argwhere(got != ref)
    assert bad.shape[0] == 0, "%d pixels differ, first %s: gpu %r oracle %r" % (
        bad.shape[0], bad[0], got[tuple(bad[0])], ref[tuple(bad[0])])


def test_fmm_non_square_smoothing_and_batch(dev):
    from cudaExtensions import sampling
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = 48, 32, sc.focal * 48 / sc.width
    got = run_gpu(dev, W, H, f, p, nb, sd, 0.01, smoothing=0.3)
    ref = O.fmm_weights_view(W, H, f, p, nb, sd, 0.3, 0.01)
    assert np.array_equal(got, ref)
    P = np.stack([p, p * np.float32([1.0, -1.0, 1.0])])
    NB = np.stack([nb, nb])
    out = sampling.approximateFMMWeights(W, H, f, t(P, dev), t(NB, dev), None, t(sd[None], dev), 0.0, 0.01)
    ref = O.fmm_weights(W, H, f, P, NB, sd[None], 0.0, 0.01)
    assert np.array_equal(out.cpu().numpy(), ref)


def test_fmm_edge_cases(dev):
    from cudaExtensions import sampling
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    assert np.all(run_gpu(dev, sc.width, sc.height, sc.focal, -p, nb, sd) == 0)
    p1 = np.array([[0.1, -0.05, -2.0]], np.float32)
    got = run_gpu(dev, sc.width, sc.height, sc.focal, p1, nb[:1], sd[:1])
    assert np.array_equal(got, O.fmm_weights_view(sc.width, sc.height, sc.focal, p1, nb[:1], sd[:1]))
    z = np.zeros((1, 0, 3), np.float32)
    out = sampling.approximateFMMWeights(8, 8, 9.6, t(z, dev), t(z, dev), None,
                                         t(np.zeros((1, 0, 2), np.float32), dev), 0.0, 0.01)
    assert out.shape == (1, 8, 8) and bool((out == 0).all())


def test_fmm_large_workload(dev):
    """1M surfels on 64 ellipsoids at 512x512: bit-exact against the oracle."""
    sc, p, n, nb, sd = scene_inputs(1)
    got = run_gpu(dev, sc.width, sc.height, sc.focal, p, nb, sd, 0.01)
    ref = O.fmm_weights_view(sc.width, sc.height, sc.focal, p, nb, sd, 0.0, 0.01)
    assert np.array_equal(got, ref)


def test_fmm_large_workload_against_direct_sum(dev):
    """1M surfels at 512x512: the GPU density at 4096 random pixels against the
    FP64 direct Gaussian sum c sum_i exp(-|q - s_i|^2 / 2h^2) over every culled
    splat (evaluated on the GPU in FP64): between the direct sum minus the mass of
    the leaves beyond the query radius and the direct sum, up to the degree-6
    expansion error of the accepted nodes (1e-4 of the largest density)."""
    import math
    sc, p, n, nb, sd = scene_inputs(1)
    W, H = sc.width, sc.height
    got = run_gpu(dev, W, H, sc.focal, p, nb, sd, 0.01)
    T = O.Tree(W, H, sc.focal, p, nb, sd, 0.0, 0.01)
    h, rad = float(T.h), float(T.radius)
    c = 1.0 / (2.0 * math.pi * math.sqrt(h))
    lv = torch.from_numpy(T.cen[T.nc - 1:].astype(np.float64)).to(dev)
    rng = np.random.default_rng(4096)
    q = np.stack([rng.integers(0, W, 4096), rng.integers(0, H, 4096)], 1)
    qt = torch.from_numpy(q.astype(np.float64)).to(dev)
    dens = torch.zeros(4096, dtype=torch.float64, device=dev)
    far = torch.zeros_like(dens)
    for s in range(0, lv.shape[0], 1 << 16):
        d2 = ((qt[:, None, :] - lv[None, s:s + (1 << 16), :]) ** 2).sum(-1)
        g = c * torch.exp(-d2 / (2 * h * h))
        dens += g.sum(1)
        far += (g * (d2 >= rad * rad)).sum(1)
    dens, far = dens.cpu().numpy(), far.cpu().numpy()
    img = got[q[:, 0], q[:, 1]].astype(np.float64)
    tol = 1e-4 * dens.max()
    print("direct-sum check: max over %.3g, max under-minus-far %.3g, max density %.3g, max far %.3g"
          % ((img - dens).max(), (dens - far - img).max(), dens.max(), far.max()))
    assert dens.max() > 1.0
    assert np.all(img <= dens + tol), (img - dens).max()
    assert np.all(img >= dens - far - tol), (dens - far - img).max()


@pytest.mark.parametrize("p", [4, 5, 7, 8])
def test_fmm_other_orders_bit_exact(dev, p):
    from cudaExtensions import sampling
    sc, p_, n, nb, sd = scene_inputs(0, sd_scale=2.0)
    sampling.set_fixed_degree(p)
    O.set_fixed_degree(p)
    try:
        got = run_gpu(dev, sc.width, sc.height, sc.focal, p_, nb, sd, 0.01)
        ref = O.fmm_weights_view(sc.width, sc.height, sc.focal, p_, nb, sd, 0.0, 0.01)
    finally:
        sampling.set_fixed_degree(6)
        O.set_fixed_degree(6)
    assert sampling.get_fixed_degree() == 6
    assert np.array_equal(got, ref)
