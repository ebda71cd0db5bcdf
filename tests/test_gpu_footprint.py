"""GPU Poisson-rate sampling with the footprint proposal (the wide band plus
each splat's own EWA footprint): bit-exact against the oracle (ids and 1/pi) for a
given normaliser, single views, batches and tile multipliers; with the
normaliser computed on the device, the same draws as the oracle's up to the
last bits of that sum; and the Horvitz-Thompson gradients it feeds are
unbiased (same check as the band proposal's)."""
import numpy as np
import pytest
import torch

from helpers import scene_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def gpu_fp(dev, W, H, f, P, NB, SD, kcap, rate, prec=0.01, seed=0, ts=None, scale=None):
    from cudaExtensions import sampling
    ids, ipi = sampling.samplePoisson(W, H, f, t(P, dev), t(NB, dev), None, t(SD, dev), kcap, 0.0, prec,
                                      seed=seed, rate=rate, tileScale=None if ts is None else t(ts, dev),
                                      proposal="footprint", footprintScale=scale)
    return ids.cpu().numpy(), ipi.cpu().numpy()


def view_scale(sc, p, nb, sd):
    """The normaliser kw = mean footprint mass / h_1^1.5 of one view, as the kernels form it."""
    leaf, ids, par, mass, a1 = O.foot_view(sc.width, sc.height, sc.focal, p, nb, sd)
    h = float(O.poisson_wide_h(sc.width, sc.height, sc.focal, p, nb, sd, 0.0, 0.01))
    return float(np.float32((mass / leaf.shape[0]) / (h * np.sqrt(h))))


@pytest.mark.parametrize("cfg,sds,kcap,rate,prec", [(2, 8.0, 48, 6.0, 0.01), (0, 2.0, 40, 8.0, 0.01),
                                                    (0, 4.0, 64, 12.0, 0.001)])
def test_footprint_bit_exact(dev, cfg, sds, kcap, rate, prec):
    sc, p, n, nb, sd = scene_inputs(cfg, sd_scale=sds)
    for scale in (view_scale(sc, p, nb, sd), 0.37):
        ids, ipi = gpu_fp(dev, sc.width, sc.height, sc.focal, p[None], nb[None], sd[None], kcap, rate, prec,
                          seed=77, scale=scale)
        rid, ripi, dens, over = O.poisson_view(sc.width, sc.height, sc.focal, p, nb, sd, kcap, rate, 0.0, prec,
                                               seed=77, proposal="footprint", footprint_scale=scale)
        assert over == 0 and (rid >= 0).sum() > 1000
        bad = np.argwhere((ids[0] != rid).any(-1))
        assert bad.shape[0] == 0, "scale %g: %d pixels differ, first %s: gpu %s oracle %s" % (
            scale, bad.shape[0], bad[0], ids[0][tuple(bad[0])], rid[tuple(bad[0])])
        assert np.array_equal(ipi[0], ripi)


def test_footprint_batch_and_tile_scale(dev):
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = sc.width, sc.height, sc.focal
    P = np.stack([p, p * np.float32([1.0, -1.0, 1.0])])
    NB = np.stack([nb, nb])
    ts = np.random.default_rng(0).uniform(0.2, 3.0, size=(2, (W + 15) // 16, (H + 15) // 16)).astype(np.float32)
    ids, ipi = gpu_fp(dev, W, H, f, P, NB, sd[None], 48, 5.0, seed=3, ts=ts, scale=0.5)
    rid, ripi = O.poisson(W, H, f, P, NB, sd[None], 48, 5.0, seed=3, tile_scale=ts, proposal="footprint",
                          footprint_scale=0.5)
    assert np.array_equal(ids, rid) and np.array_equal(ipi, ripi)
    # nothing visible
    ids, ipi = gpu_fp(dev, W, H, f, -p[None], nb[None], sd[None], 8, 4.0, scale=0.5)
    assert np.all(ids == -1) and np.all(ipi == 0)


def test_footprint_device_normaliser(dev):
    """The device sums the footprint masses with atomics (order-dependent last
    bits): its draws equal the oracle's at the oracle's own normaliser in almost
    every pixel, and its densities output carries D0 and D1."""
    from cudaExtensions import sampling
    sc, p, n, nb, sd = scene_inputs(0, sd_scale=2.0)
    W, H, f = sc.width, sc.height, sc.focal
    dens = torch.zeros((1, 2, W, H), dtype=torch.float32, device=dev)
    ids, ipi = sampling.samplePoisson(W, H, f, t(p[None], dev), t(nb[None], dev), None, t(sd[None], dev), 40, 0.0,
                                      0.01, seed=21, rate=8.0, densities=dens, proposal="footprint")
    ids = ids.cpu().numpy()[0]
    rid, ripi, rdens, over = O.poisson_view(W, H, f, p, nb, sd, 40, 8.0, 0.0, 0.01, seed=21, proposal="footprint")
    same = (ids == rid).all(-1).mean()
    assert same > 0.99, same
    d = dens.cpu().numpy()[0]
    assert np.array_equal(d, rdens)
    # the two halves expect about rate / 2 each where the image is covered
    cov = rdens[0] >= 0.01
    assert 0.2 * 8.0 < (rid >= 0).sum(-1)[cov].mean() < 2.0 * 8.0


def test_footprint_gradients_unbiased(dev):
    """Linear image I(q) = sum_s w_s(q) a_s / pi_s; functional L = sum_q <g(q), I(q)>;
    the mean sampled gradient over 4096 seeds against the support's dense gradient."""
    from cudaExtensions import splatting, interpolation, sampling
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = sc.width, sc.height, sc.focal
    P, NB, SD = t(p[None], dev), t(nb[None], dev), t(sd[None], dev)
    A = t(sc.albedo[None].astype(np.float32), dev)
    g = t(np.random.default_rng(1).normal(size=(1, W, H, 3)).astype(np.float32), dev)

    def grads(ids, ipi):
        w = splatting.forward(W, H, f, ids, P, NB, SD, 0.0)
        wt = (w * ipi).contiguous()
        gW, gA = interpolation.backward(g, ids, wt, A)
        gw = (gW * ipi).contiguous()
        gP, gN, gS = splatting.backward(W, H, f, gw, ids, P, NB, SD, 0.0)
        return torch.cat([gP.reshape(-1), gN.reshape(-1), gS.reshape(-1), gA.reshape(-1)]).double()

    ids_d, ipi_d = sampling.samplePoisson(W, H, f, P, NB, None, SD, 512, 0.0, 0.01, seed=0, rate=1e30,
                                          proposal="footprint")
    sampling.check_errors()
    assert bool(((ipi_d == 1.0) | ((ids_d < 0) & (ipi_d == 0.0))).all())
    dense = grads(ids_d, ipi_d)
    S = 4096
    s1 = torch.zeros_like(dense)
    s2 = torch.zeros_like(dense)
    for s in range(S):
        ids, ipi = sampling.samplePoisson(W, H, f, P, NB, None, SD, 64, 0.0, 0.01, seed=10_000 + s, rate=8.0,
                                          proposal="footprint")
        x = grads(ids, ipi)
        s1 += x
        s2 += x * x
    sampling.check_errors()
    mean = s1 / S
    var = (s2 - S * mean * mean) / (S - 1)
    se = torch.sqrt(torch.clamp(var, min=0) / S)
    m = se > 1e-12 * dense.abs().max()
    z = ((mean - dense)[m] / se[m]).cpu().numpy()
    print("params %d, max|z| %.2f, mean z %.3f, mean z^2 %.3f" % (z.size, np.abs(z).max(), z.mean(), (z ** 2).mean()))
    assert z.size > 1000
    assert abs(z.mean()) < 0.1 and 0.8 < (z ** 2).mean() < 1.25
    assert np.mean(np.abs(z) > 3.0) < 0.01
    dm = (~m).cpu().numpy()
    assert np.allclose(mean.cpu().numpy()[dm], dense.cpu().numpy()[dm], rtol=1e-4, atol=1e-6)


def test_footprint_proposal_at_full_scale(dev, record_property):
    """The footprint proposal against the band proposal on the benchmark's
    Poisson-rate frame at full size (1M surfels on 64 ellipsoids, 512x512,
    Horvitz-Thompson estimate without occlusion) at the same expected sample
    count, 128 seeds in 16 groups per measurement.  A variance gain is asserted
    only when the whole 95 % interval of kappa lies below 1; otherwise the test
    records the interval and is skipped, naming it."""
    from splatting_amd import learned as L, scenes
    sc = scenes.config(1)
    target = L.make_target(sc, dev, 10.0, 40, seeds=range(10_000, 10_004))
    fr = L.PoissonFrame(sc, dev, rate=10.0, kcap=40, target=target)
    r = L.compare_proposals(fr, list(range(900, 1028)), groups=16)
    for k, v in r.items():
        record_property(k, v)
    print("footprint vs bands at 1M splats:", r)
    fr.check_errors()
    assert abs(r["samples_per_pixel_footprint"] / r["samples_per_pixel_bands"] - 1.0) < 0.02, r
    lo, hi = r["kappa_ci95"]
    assert 0.0 < lo <= hi
    if hi >= 1.0:
        pytest.skip("no significant gain from the footprint proposal at 1M splats, 512x512: kappa = %.3f, "
                    "95%% interval [%.3f, %.3f]" % (r["kappa"], lo, hi))
    assert hi < 1.0
