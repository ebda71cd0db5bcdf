"""Pins the oracle's approximateFMMWeights (fixed-order Hermite FGT density image,
sampling.cpp:136-175, samplingCuda.cu:398-493, transformRadixTree.cuh:15-69, 152-222)
against the FP64 direct density c sum_i exp(-|q - s_i|^2 / 2h^2), c = 1 / (2 pi sqrt(h)):

  * with a vanishing precision no internal node is accepted and the query radius
    covers every leaf that matters: the image is the direct sum up to FP32 rounding;
  * at working precision the image lies between the direct sum minus the mass of
    the leaves beyond the query radius (the only ones the traversal may prune) and
    the direct sum, up to the degree-6 expansion error of the accepted nodes;
  * the image is x-major (pixel (i, j) at j + H i) also when W != H.
"""
import math

import numpy as np
import pytest

from helpers import scene_inputs
from oracle import oracle as O


def direct(sc, p, nb, sd, prec, W=None, H=None, f=None):
    W = W or sc.width
    H = H or sc.height
    f = f or sc.focal
    T = O.Tree(W, H, f, p, nb, sd, 0.0, prec)
    h = float(T.h)
    c = 1.0 / (2.0 * math.pi * math.sqrt(h))
    lv = T.cen[T.nc - 1:].astype(np.float64)
    ii, jj = np.meshgrid(np.arange(W), np.arange(H), indexing="ij")
    q = np.stack([ii.ravel(), jj.ravel()], 1).astype(np.float64)
    d2 = ((q[:, None, :] - lv[None, :, :]) ** 2).sum(-1)
    g = c * np.exp(-d2 / (2 * h * h))
    far = d2 >= float(T.radius) ** 2
    return g.sum(1).reshape(W, H), (g * far).sum(1).reshape(W, H), T


@pytest.mark.parametrize("cfg,sds", [(2, 8.0), (0, 2.0)])
def test_fmm_exact_limit(cfg, sds):
    sc, p, n, nb, sd = scene_inputs(cfg, sd_scale=sds)
    img = O.fmm_weights_view(sc.width, sc.height, sc.focal, p, nb, sd, 0.0, 1e-12)
    dens, far, T = direct(sc, p, nb, sd, 1e-12)
    assert np.abs(img - dens).max() <= 2e-6 * dens.max()


@pytest.mark.parametrize("cfg,sds,prec", [(2, 8.0, 0.01), (0, 2.0, 0.01), (0, 4.0, 0.001)])
def test_fmm_density_bracket(cfg, sds, prec):
    sc, p, n, nb, sd = scene_inputs(cfg, sd_scale=sds)
    img = O.fmm_weights_view(sc.width, sc.height, sc.focal, p, nb, sd, 0.0, prec)
    dens, far, T = direct(sc, p, nb, sd, prec)
    # expansion error of the accepted nodes (bound <= precision each, relative to c n);
    # observed below 6e-6 on these workloads
    tol = 5e-5
    assert np.all(img <= dens + tol)
    assert np.all(img >= dens - far - tol)
    assert far.max() > 0 and img.max() > 0.1


def test_fmm_layout_non_square():
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = 48, 32, sc.focal * 48 / sc.width
    img = O.fmm_weights_view(W, H, f, p, nb, sd, 0.0, 1e-12)
    dens, far, T = direct(sc, p, nb, sd, 1e-12, W, H, f)
    assert img.shape == (W, H)
    assert np.abs(img - dens).max() <= 2e-6 * dens.max()


def test_fmm_single_splat_and_empty():
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    p1 = np.array([[0.1, -0.05, -2.0]], np.float32)
    img = O.fmm_weights_view(sc.width, sc.height, sc.focal, p1, nb[:1], sd[:1], 0.0, 0.01)
    dens, far, T = direct(sc, p1, nb[:1], sd[:1], 0.01)
    assert T.nc == 1
    assert np.abs(img - dens).max() <= 2e-6 * dens.max()
    z = np.zeros((0, 3), np.float32)
    assert np.all(O.fmm_weights_view(8, 8, 9.6, z, z, np.zeros((0, 2), np.float32)) == 0)
    assert np.all(O.fmm_weights_view(sc.width, sc.height, sc.focal, -p, nb, sd) == 0)


@pytest.mark.parametrize("p", [4, 5, 7, 8])
def test_fmm_density_bracket_other_orders(p):
    """Every expansion order keeps the bracket: a node expands only where its
    order-p truncation bound meets the precision."""
    sc, p_, n, nb, sd = scene_inputs(0, sd_scale=2.0)
    O.set_fixed_degree(p)
    try:
        img = O.fmm_weights_view(sc.width, sc.height, sc.focal, p_, nb, sd, 0.0, 0.01)
    finally:
        O.set_fixed_degree(6)
    dens, far, T = direct(sc, p_, nb, sd, 0.01)
    tol = 5e-5
    assert np.all(img <= dens + tol)
    assert np.all(img >= dens - far - tol)
    with pytest.raises(ValueError):
        O.set_fixed_degree(3)


def test_direct_sums_closed_form():
    """o_direct_sums against numpy's FP64 sum of the same Gaussians."""
    import math
    rng = np.random.default_rng(3)
    s = rng.uniform(0, 64, size=(3000, 2)).astype(np.float32)
    q = rng.uniform(0, 64, size=(50, 2))
    h, c = 1.7, 1.0 / (2.0 * math.pi * math.sqrt(1.7))
    got = O.direct_sums(s, h, c, q)
    d2 = ((q[:, None, :] - s[None, :, :].astype(np.float64)) ** 2).sum(-1)
    ref = c * np.exp(-d2 / (2 * h * h)).sum(1)
    assert np.allclose(got, ref, rtol=1e-13, atol=0)
