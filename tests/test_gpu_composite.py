"""Composited Horvitz-Thompson estimate (estimator.forward / backward with
composite=True) through the C ABI:

  * parity with the oracle's FP64 restatement (oracle.htc): image and gradients
    of points, normals, std-devs and attribute elementwise within 1e-4 relative
    plus a floor of 64 float32 roundoffs of each output's magnitude
    sum_t |term_t|, for one and two views, per-view and shared attributes and
    std-devs, vectorised and scalar slot reads, full pixels and tiles whose
    distinct splats overflow the backward's LDS table;
  * the image equals the unfused ops (splatting.forward, times 1/pi,
    filtering.forward, interpolation.forward) to float32 rounding;
  * unbiasedness: over 4096 Philox seeds the mean image and the mean gradient
    match the image and gradient composited over every splat of the sampling
    support (the oracle, FP64) within 3 standard errors, z-scores of mean 0 and
    variance 1.
"""
import numpy as np
import pytest
import torch

from helpers import assert_close, magnitude, near_slots, scene_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def check(dev, W, H, f, ids, ipi, P, Nc, SD, A, g, what):
    from cudaExtensions import estimator
    img = estimator.forward(W, H, f, ids, ipi, P, Nc, SD, A, composite=True)
    grads = estimator.backward(W, H, f, g, ids, ipi, P, Nc, SD, A, composite=True)
    args = [x.cpu().numpy() for x in (ids, ipi, P, Nc, SD, A)]
    args[1] = args[1].astype(np.float64)
    gn = g.cpu().numpy().astype(np.float64)
    rimg, rgrads = O.htc(W, H, f, *args, gimg=gn)
    mimg, mgrads = magnitude(O.htc, W, H, f, *args, gimg=np.abs(gn))
    assert np.abs(rimg).max() > 0
    assert_close(img.cpu().numpy(), rimg, RTOL, what + " composited image", mag=mimg)
    for name, a, r, m in zip(("points", "normals", "std-devs", "attribute"), grads, rgrads, mgrads):
        assert_close(a.cpu().numpy(), r, RTOL, "%s composited grad %s" % (what, name), mag=m)
    # the backward reusing the forward's records
    ws = estimator.workspace(ids, SD, composite=True)
    assert torch.equal(estimator.forward(W, H, f, ids, ipi, P, Nc, SD, A, workspace=ws, composite=True), img)
    pg = estimator.backward(W, H, f, g, ids, ipi, P, Nc, SD, A, workspace=ws, prepared=True, composite=True)
    for name, a, b, m in zip(("points", "normals", "std-devs", "attribute"), pg, grads, mgrads):
        assert_close(a.cpu().numpy(), b.cpu().numpy(), 1e-5, "%s prepared grad %s" % (what, name), mag=m)
    return img


def poisson_views(dev, views, kcap=40, rate=8.0, seed=5):
    from cudaExtensions import sampling
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    flips = [np.float32([1.0, 1.0, 1.0]), np.float32([1.0, -1.0, 1.0])]
    P = t(np.stack([p * flips[v] for v in range(views)]), dev)
    Nc = t(np.stack([n * flips[v] for v in range(views)]), dev)
    NB = t(np.stack([nb * flips[v] for v in range(views)]), dev)
    SD = t(sd[None], dev)
    ids, ipi = sampling.samplePoisson(sc.width, sc.height, sc.focal, P, NB, None, SD, kcap, 0.0, 0.01, seed=seed,
                                      rate=rate)
    sampling.check_errors()
    assert int((ids >= 0).sum()) > 1000
    return sc, P, Nc, SD, ids, ipi


@pytest.mark.parametrize("views,per_view_attr,per_view_sd", [(1, False, False), (2, False, False),
                                                             (2, True, True)])
def test_composite_matches_oracle(dev, views, per_view_attr, per_view_sd):
    sc, P, Nc, SD, ids, ipi = poisson_views(dev, views)
    W, H, f = sc.width, sc.height, sc.focal
    rng = np.random.default_rng(views)
    A = sc.albedo[None].astype(np.float32)
    if per_view_attr:
        A = np.concatenate([A, rng.uniform(0, 1, size=A.shape).astype(np.float32)])
    if per_view_sd:
        SD = torch.cat([SD, SD * 1.25]).contiguous()
    g = t(rng.normal(size=(views, W, H, 3)).astype(np.float32), dev)
    check(dev, W, H, f, ids, ipi, P, Nc, SD, t(A, dev), g, "views %d" % views)


@pytest.mark.parametrize("k", [5, 8])
def test_composite_full_pixels_and_scalar_slots(dev, k):
    sc, P, Nc, SD, ids, ipi = poisson_views(dev, 1)
    W, H, f = sc.width, sc.height, sc.focal
    ids, ipi = ids[..., :k].contiguous(), ipi[..., :k].contiguous()
    assert int((ids[..., -1] >= 0).sum()) > 10
    g = t(np.random.default_rng(k).normal(size=(1, W, H, 3)).astype(np.float32), dev)
    check(dev, W, H, f, ids, ipi, P, Nc, SD, t(sc.albedo[None].astype(np.float32), dev), g, "k %d" % k)


def test_composite_table_overflow_and_dense_slots(dev):
    """Every slot filled (40 per pixel, about 4000 distinct splats per 16 x 16 tile)."""
    sc, p, n, nb, sd = scene_inputs(0, sd_scale=4.0)
    W, H, f = sc.width, sc.height, sc.focal
    rng = np.random.default_rng(11)
    ids = near_slots(rng, p, W, H, f, 40, p_empty=0.0)
    ids = -np.sort(-ids, axis=-1)
    assert (ids >= 0).all() and np.unique(ids[0, :16, :16]).size > 2048
    ipi = rng.uniform(1.0, 4.0, size=ids.shape).astype(np.float32)
    g = t(rng.normal(size=(1, W, H, 3)).astype(np.float32), dev)
    check(dev, W, H, f, t(ids, dev), t(ipi, dev), t(p[None], dev), t(n[None], dev), t(sd[None], dev),
          t(sc.albedo[None].astype(np.float32), dev), g, "dense")


def test_composite_equals_unfused_ops(dev):
    from cudaExtensions import estimator, filtering, interpolation, splatting
    sc, P, Nc, SD, ids, ipi = poisson_views(dev, 2)
    W, H, f = sc.width, sc.height, sc.focal
    A = t(sc.albedo[None].astype(np.float32), dev)
    img = estimator.forward(W, H, f, ids, ipi, P, Nc, SD, A, composite=True)
    w = splatting.forward(W, H, f, ids, P, Nc, SD, 0.0)
    alpha = (w * ipi).contiguous()
    ref = interpolation.forward(ids, filtering.forward(ids, alpha, W, H, f, P, Nc), A)
    args = [x.cpu().numpy() for x in (ids, ipi, P, Nc, SD, A)]
    args[1] = args[1].astype(np.float64)
    mimg = magnitude(O.htc, W, H, f, *args)
    assert_close(img.cpu().numpy(), ref.cpu().numpy(), RTOL, "composited vs unfused", mag=mimg)


def test_composite_rejects_wide_slots(dev):
    from cudaExtensions import estimator
    sc, P, Nc, SD, ids, ipi = poisson_views(dev, 1, kcap=120)
    A = t(sc.albedo[None].astype(np.float32), dev)
    with pytest.raises(ValueError):
        estimator.forward(sc.width, sc.height, sc.focal, ids, ipi, P, Nc, SD, A, composite=True)


def test_composite_unbiased(dev):
    """Functional L = sum_q <g(q), image(q)>; the mean over 4096 seeds of the sampled
    image and gradient against the FP64 composite over the full support (every
    (pixel, splat) pair the proposal can draw, taken with certainty at rate 1e30)."""
    from cudaExtensions import estimator, sampling
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = sc.width, sc.height, sc.focal
    P, Nc, NB, SD = t(p[None], dev), t(n[None], dev), t(nb[None], dev), t(sd[None], dev)
    A = t(sc.albedo[None].astype(np.float32), dev)
    rng = np.random.default_rng(1)
    gn = rng.normal(size=(1, W, H, 3))
    g = t(gn.astype(np.float32), dev)
    ids_d, ipi_d = sampling.samplePoisson(W, H, f, P, NB, None, SD, 256, 0.0, 0.01, seed=0, rate=1e30)
    sampling.check_errors()
    ids_d = ids_d.cpu().numpy()
    dimg, dgr = O.htc(W, H, f, ids_d, (ids_d >= 0).astype(np.float64), p[None], n[None], sd[None],
                      sc.albedo[None], gimg=gn)
    dense = np.concatenate([x.reshape(-1) for x in dgr])
    S = 4096
    ws = estimator.workspace(torch.empty((1, W, H, 40), dtype=torch.int32, device=dev), SD, composite=True)
    s1 = s2 = i1 = i2 = None
    for s in range(S):
        ids, ipi = sampling.samplePoisson(W, H, f, P, NB, None, SD, 40, 0.0, 0.01, seed=20_000 + s, rate=8.0)
        img = estimator.forward(W, H, f, ids, ipi, P, Nc, SD, A, workspace=ws, composite=True).double()
        gr = estimator.backward(W, H, f, g, ids, ipi, P, Nc, SD, A, workspace=ws, prepared=True, composite=True)
        x = torch.cat([v.reshape(-1) for v in gr]).double()
        s1 = x if s1 is None else s1 + x
        s2 = x * x if s2 is None else s2 + x * x
        i1 = img if i1 is None else i1 + img
        i2 = img * img if i2 is None else i2 + img * img
    sampling.check_errors()

    def zscores(a1, a2, ref):
        mean = (a1 / S).cpu().numpy().reshape(-1)
        var = ((a2 - a1 * a1 / S) / (S - 1)).clamp(min=0).cpu().numpy().reshape(-1)
        se = np.sqrt(var / S)
        ref = ref.reshape(-1)
        m = se > 1e-12 * np.abs(ref).max()
        return (mean[m] - ref[m]) / se[m], mean[~m], ref[~m]

    for what, (z, fixed, rfixed) in (("image", zscores(i1, i2, dimg)), ("gradient", zscores(s1, s2, dense))):
        print("%s: %d values, max|z| %.2f, mean z %.3f, mean z^2 %.3f" % (what, z.size, np.abs(z).max(), z.mean(),
                                                                         (z ** 2).mean()))
        assert z.size > 1000
        assert abs(z.mean()) < 0.1 and 0.8 < (z ** 2).mean() < 1.25
        assert np.mean(np.abs(z) > 3.0) < 0.01
        assert np.allclose(fixed, rfixed, rtol=1e-4, atol=1e-6)
