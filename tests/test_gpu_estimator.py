"""Fused Horvitz-Thompson image estimate (cudaExtensions.estimator) through the
C ABI: the image equals the unfused ops (splatting.forward, times 1/pi,
interpolation.forward) bit for bit, and the gradients equal the unfused
backward ops' within 1e-5 (they differ only in the order of atomic sums) and
the oracle's FP64 composition within 1e-4, elementwise with a floor of 64
float32 roundoffs of the magnitude sum_t |term_t| of each output, for one and two
views, per-view and shared attributes and std-devs, vectorised (k % 4 == 0)
and scalar slot reads, full pixels (no -1), tiles whose distinct splats overflow
the per-tile LDS table, and empty inputs."""
import numpy as np
import pytest
import torch

from helpers import assert_close, magnitude, near_slots, scene_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL_ORACLE = 1e-4      # FP32 atomics and EWA in FP32 against the FP64 oracle
RTOL_UNFUSED = 1e-5     # same FP32 slot values, different atomic order


def t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def unfused(W, H, f, ids, ipi, P, NB, SD, A, g):
    from cudaExtensions import interpolation, splatting
    w = splatting.forward(W, H, f, ids, P, NB, SD, 0.0)
    wt = (w * ipi).contiguous()
    img = interpolation.forward(ids, wt, A)
    gW, gA = interpolation.backward(g, ids, wt, A)
    gw = (gW * ipi).contiguous()
    gP, gN, gS = splatting.backward(W, H, f, gw, ids, P, NB, SD, 0.0)
    return img, (gP, gN, gS, gA)


def oracle(W, H, f, ids, ipi, P, NB, SD, A, g):
    ids, ipi = ids.cpu().numpy(), ipi.cpu().numpy().astype(np.float64)
    P, NB, SD, A, g = (x.cpu().numpy() for x in (P, NB, SD, A, g))
    w = O.splat_forward(W, H, f, ids, P, NB, SD)
    wt = w * ipi
    img = O.interp_forward(ids, wt, A)
    gW, gA = O.interp_backward(g, ids, wt, A)
    gP, gN, gS = O.splat_backward(W, H, f, gW * ipi, ids, P, NB, SD)
    # magnitudes sum_t |term_t| of the same chain (the floors of the elementwise checks),
    # the weights scaled by their float32 evaluation condition
    mwt = magnitude(O.splat_forward, W, H, f, ids, P, NB, SD) * ipi
    mimg = magnitude(O.interp_forward, ids, mwt, A)
    mW, mA = magnitude(O.interp_backward, g, ids, mwt, A)
    mP, mN, mS = magnitude(O.splat_backward, W, H, f, mW * ipi, ids, P, NB, SD)
    return img, (gP, gN, gS, gA), mimg, (mP, mN, mS, mA)


def check(dev, W, H, f, ids, ipi, P, NB, SD, A, g, what):
    from cudaExtensions import estimator
    img = estimator.forward(W, H, f, ids, ipi, P, NB, SD, A)
    grads = estimator.backward(W, H, f, g, ids, ipi, P, NB, SD, A)
    uimg, ugrads = unfused(W, H, f, ids, ipi, P, NB, SD, A, g)
    assert torch.equal(img, uimg), what
    rimg, rgrads, mimg, mgrads = oracle(W, H, f, ids, ipi, P, NB, SD, A, g)
    assert_close(img.cpu().numpy(), rimg, RTOL_ORACLE, what + " image vs oracle", mag=mimg)
    for name, a, b, r, m in zip(("points", "normals", "std-devs", "attribute"), grads, ugrads, rgrads, mgrads):
        assert a.shape == b.shape, (what, name)
        assert_close(a.cpu().numpy(), b.cpu().numpy(), RTOL_UNFUSED, "%s grad %s vs unfused" % (what, name), mag=m)
        assert_close(a.cpu().numpy(), r, RTOL_ORACLE, "%s grad %s vs oracle" % (what, name), mag=m)
    # the backward reusing the forward's per-splat records
    ws = estimator.workspace(ids, SD)
    assert torch.equal(estimator.forward(W, H, f, ids, ipi, P, NB, SD, A, workspace=ws), img), what
    pgrads = estimator.backward(W, H, f, g, ids, ipi, P, NB, SD, A, workspace=ws, prepared=True)
    for name, a, b, m in zip(("points", "normals", "std-devs", "attribute"), pgrads, grads, mgrads):
        assert_close(a.cpu().numpy(), b.cpu().numpy(), RTOL_UNFUSED, "%s grad %s, prepared" % (what, name), mag=m)


def poisson_views(dev, views, kcap=40, rate=8.0):
    from cudaExtensions import sampling
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    flips = [np.float32([1.0, 1.0, 1.0]), np.float32([1.0, -1.0, 1.0])]
    P = t(np.stack([p * flips[v] for v in range(views)]), dev)
    NB = t(np.stack([nb * flips[v] for v in range(views)]), dev)
    SD = t(sd[None], dev)
    ids, ipi = sampling.samplePoisson(sc.width, sc.height, sc.focal, P, NB, None, SD, kcap, 0.0, 0.01, seed=5,
                                      rate=rate)
    sampling.check_errors()
    assert int((ids >= 0).sum()) > 1000
    return sc, P, NB, SD, ids, ipi


@pytest.mark.parametrize("views,per_view_attr,per_view_sd", [(1, False, False), (2, False, False),
                                                             (2, True, True)])
def test_estimate_matches_unfused_and_oracle(dev, views, per_view_attr, per_view_sd):
    sc, P, NB, SD, ids, ipi = poisson_views(dev, views)
    W, H, f = sc.width, sc.height, sc.focal
    rng = np.random.default_rng(views)
    A = sc.albedo[None].astype(np.float32)
    if per_view_attr:
        A = np.concatenate([A, rng.uniform(0, 1, size=A.shape).astype(np.float32)])
    if per_view_sd:
        SD = torch.cat([SD, SD * 1.25]).contiguous()
    g = t(rng.normal(size=(views, W, H, 3)).astype(np.float32), dev)
    check(dev, W, H, f, ids, ipi, P, NB, SD, t(A, dev), g, "views %d" % views)


@pytest.mark.parametrize("k", [5, 8])
def test_estimate_full_pixels_and_scalar_slots(dev, k):
    """Pixels whose k slots are all filled (no -1 stops the read) and k not a
    multiple of 4 (scalar slot reads)."""
    sc, P, NB, SD, ids, ipi = poisson_views(dev, 1, kcap=40)
    W, H, f = sc.width, sc.height, sc.focal
    ids, ipi = ids[..., :k].contiguous(), ipi[..., :k].contiguous()
    full = int((ids[..., -1] >= 0).sum())
    assert full > 10, full
    g = t(np.random.default_rng(k).normal(size=(1, W, H, 3)).astype(np.float32), dev)
    check(dev, W, H, f, ids, ipi, P, NB, SD, t(sc.albedo[None].astype(np.float32), dev), g, "k %d" % k)


def test_estimate_table_overflow(dev):
    """Dense slots (every slot filled, about 4000 distinct splats per 16 x 16 tile):
    samples beyond the tile's LDS table add to the records directly."""
    sc, p, n, nb, sd = scene_inputs(0, sd_scale=4.0)
    W, H, f = sc.width, sc.height, sc.focal
    rng = np.random.default_rng(11)
    ids = near_slots(rng, p, W, H, f, 40, p_empty=0.0)
    ids = -np.sort(-ids, axis=-1)                    # compact: samples first (all filled here)
    assert (ids >= 0).all()
    tile_ids = ids[0, :16, :16].reshape(-1)
    assert np.unique(tile_ids).size > 2048
    ipi = rng.uniform(1.0, 4.0, size=ids.shape).astype(np.float32)
    g = t(rng.normal(size=(1, W, H, 3)).astype(np.float32), dev)
    check(dev, W, H, f, t(ids, dev), t(ipi, dev), t(p[None], dev), t(nb[None], dev), t(sd[None], dev),
          t(sc.albedo[None].astype(np.float32), dev), g, "table overflow")


def test_estimate_empty_and_bad_arguments(dev):
    from cudaExtensions import estimator
    sc, P, NB, SD, ids, ipi = poisson_views(dev, 1)
    W, H, f = sc.width, sc.height, sc.focal
    A = t(sc.albedo[None].astype(np.float32), dev)
    empty = torch.full_like(ids, -1)
    zero = torch.zeros_like(ipi)
    g = torch.ones((1, W, H, 3), dtype=torch.float32, device=dev)
    assert float(estimator.forward(W, H, f, empty, zero, P, NB, SD, A).abs().sum()) == 0.0
    for x in estimator.backward(W, H, f, g, empty, zero, P, NB, SD, A):
        assert float(x.abs().sum()) == 0.0
    z3 = torch.zeros((1, 0, 3), dtype=torch.float32, device=dev)
    z2 = torch.zeros((1, 0, 2), dtype=torch.float32, device=dev)
    none = torch.full((1, 4, 4, 8), -1, dtype=torch.int32, device=dev)
    assert float(estimator.forward(4, 4, f, none, torch.zeros((1, 4, 4, 8), device=dev), z3, z3, z2, z3)
                 .abs().sum()) == 0.0
    with pytest.raises(ValueError):
        estimator.forward(W, H, f, ids, ipi, P, NB, SD, torch.zeros((1, A.shape[1], 4), device=dev))
    with pytest.raises(ValueError):
        estimator.forward(W, H, f, ids, ipi[..., :8].contiguous(), P, NB, SD, A)
    with pytest.raises(ValueError):
        estimator.forward(W, H, f, ids, ipi, P, NB, SD, torch.cat([A, A, A]).contiguous())
    with pytest.raises(ValueError):
        estimator.backward(W, H, f, g, ids, ipi, P, NB, SD, A, prepared=True)
    with pytest.raises(ValueError):
        estimator.forward(W, H, f, ids, ipi, P, NB, SD, A, workspace=torch.empty(16, dtype=torch.uint8, device=dev))


def test_l2_loss_and_gradient(dev):
    """estimator.l2: the gradient equals 2 (image - target) bit for bit, the loss the
    FP64 sum of squares (also for an empty image)."""
    from cudaExtensions import estimator
    rng = np.random.default_rng(3)
    for shape in [(1, 64, 48, 3), (2, 129, 7, 3), (1, 0, 5, 3)]:
        a = t(rng.normal(size=shape).astype(np.float32), dev)
        b = t(rng.normal(size=shape).astype(np.float32), dev)
        loss, g = estimator.l2(a, b)
        assert torch.equal(g, 2.0 * (a - b))
        ref = float(((a - b).double() ** 2).sum())
        assert abs(float(loss) - ref) <= 1e-12 * max(ref, 1.0), (shape, float(loss), ref)
