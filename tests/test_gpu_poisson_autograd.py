"""Autograd layer of the Poisson-rate renderer (splatting_amd.poisson): the
backward that draws the samples again from the Philox key gives the gradients
of the backward that keeps the index and 1/pi tensors, keeps no per-slot tensor
between the passes, and matches the hand-chained frame (PoissonFrame)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _scene(dev, n=40_000, V=2):
    from splatting_amd import scenes
    sc = scenes.Scene(*scenes.ellipsoids(16, n // 16, seed=9), 96, 80, 110.0,
                      np.stack([scenes.camera_matrix(2.5), scenes.camera_matrix(3.0)])[:V])
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
    return sc, t(sc.view), t(sc.positions)[None], t(sc.normals)[None], t(sc.stddevs)[None], t(sc.albedo)[None]


def _grads(dev, M, P, N, S, A, sc, recompute, target, seed=21, composite=False):
    from splatting_amd import poisson
    leaves = [x.clone().requires_grad_(True) for x in (P, N, S, A)]
    img = poisson.render(M, *leaves, sc.width, sc.height, sc.focal, numSamples=32, rate=6.0, seed=seed,
                         recompute=recompute, composite=composite)
    ((img - target) ** 2).sum().backward()
    return img.detach(), [x.grad for x in leaves]


def close(a, b, tol=1e-5):
    return float((a - b).abs().max()) <= tol * max(float(b.abs().max()), 1e-30)


@pytest.mark.parametrize("composite", [False, True])
def test_recompute_equals_stored_and_frame(dev, composite):
    from splatting_amd import learned as L
    sc, M, P, N, S, A = _scene(dev)
    target = torch.full((2, sc.width, sc.height, 3), 0.2, dtype=torch.float32, device=dev)
    img_r, g_r = _grads(dev, M, P, N, S, A, sc, True, target, composite=composite)
    img_s, g_s = _grads(dev, M, P, N, S, A, sc, False, target, composite=composite)
    assert torch.equal(img_r, img_s) and float(img_r.abs().sum()) > 0
    # gradients sum with atomics, so two backward passes agree to rounding
    for a, b in zip(g_r, g_s):
        assert close(a, b), float((a - b).abs().max())
    fr = L.PoissonFrame(sc, dev, rate=6.0, kcap=32, precision=0.01, target=target, composite=composite)
    _, img_f, (gPw, gNw, gS, gA), _ = fr.step(21)
    assert torch.equal(img_f, img_r)
    for a, b in zip((gPw, gNw, gS, gA), g_r):
        assert close(a.reshape(b.shape), b), float((a.reshape(b.shape) - b).abs().max())


def test_recompute_keeps_no_slot_tensors(dev):
    from splatting_amd import poisson
    sc, M, P, N, S, A = _scene(dev, V=1)
    sizes = {}
    for recompute in (True, False):
        saved = []

        def pack(x):
            saved.append(tuple(x.shape))
            return x

        with torch.autograd.graph.saved_tensors_hooks(pack, lambda x: x):
            leaves = [x.clone().requires_grad_(True) for x in (P, N, S, A)]
            img = poisson.render(M, *leaves, sc.width, sc.height, sc.focal, numSamples=32, rate=6.0, seed=3,
                                 recompute=recompute)
        img.sum().backward()
        sizes[recompute] = saved
    slot = (1, sc.width, sc.height, 32)
    assert slot not in sizes[True]
    assert slot in sizes[False]


def test_default_seed_advances_and_recompute_redraws_it(dev):
    """Without a seed every forward takes the next key of the per-process counter
    (two calls draw different sets), and the recompute backward redraws the set
    of its own forward: the same gradients as the stored-sample backward on the
    same key."""
    from cudaExtensions import sampling
    sc, M, P, N, S, A = _scene(dev, V=1)
    target = torch.full((1, sc.width, sc.height, 3), 0.2, dtype=torch.float32, device=dev)
    sampling.manual_seed(1234)
    img1, g1 = _grads(dev, M, P, N, S, A, sc, True, target, seed=None)
    img2, _ = _grads(dev, M, P, N, S, A, sc, True, target, seed=None)
    assert not torch.equal(img1, img2)
    sampling.manual_seed(1234)
    img3, g3 = _grads(dev, M, P, N, S, A, sc, False, target, seed=None)
    assert torch.equal(img1, img3)
    for a, b in zip(g1, g3):
        assert close(a, b), float((a - b).abs().max())


def test_capacity_overflow_raises(dev):
    """More inclusions than numSamples at a pixel drop samples, so the forward raises."""
    from splatting_amd import poisson
    sc, M, P, N, S, A = _scene(dev, V=1)
    with pytest.raises(RuntimeError, match="more than numSamples"):
        poisson.render(M, P, N, S, A, sc.width, sc.height, sc.focal, numSamples=2, rate=50.0, seed=1)
    assert int(poisson.PoissonRender.last_errors.item()) & 4
