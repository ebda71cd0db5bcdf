"""The batched-view workload at its real size: the 8 x 6 x 3 m box room with 2M
surfels on the walls and 200 boxes, 16 cameras at 640x480 with uneven culled
counts, Poisson-rate sampling at rate 10 (kcap 40):

  * the batched sampler (every stage one launch over all views) equals the
    view-by-view sampler bit for bit on all 16 views;
  * two whole views equal the oracle bit for bit (indices and 1/pi; the Philox
    key carries the view index);
  * size-independent properties of all views: indices in range and in front of
    the camera, 1/pi >= 1 exactly where a sample is, a pixel's samples first,
    no capacity overflow, the same key draws the same sets and another key
    different ones;
  * the batched fwd+bwd frame (Horvitz-Thompson estimate without occlusion;
    the composited estimate's importance weights min(w, 1) / pi reach the
    hundreds where hundreds of opaque splats overlap a pixel, as here, and its
    products overflow) is finite and its gradients reach every view's splats.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def room(dev):
    from splatting_amd import learned as L, scenes
    sc = scenes.room_scene(2_000_000, 16)
    fr = L.PoissonFrame(sc, dev, rate=10.0, kcap=40)
    return sc, fr


def draw(fr, seed, batched=True, errors=None):
    from cudaExtensions import sampling
    P, Nc, nb = fr.camera()
    return sampling.samplePoisson(fr.W, fr.H, fr.f, P, nb, None, fr.S, 40, 0.0, 0.01, seed=seed, rate=10.0,
                                  batched=batched, errors=errors)


def test_room_batched_equals_view_by_view(dev, room):
    sc, fr = room
    err = torch.zeros(1, dtype=torch.int32, device=dev)
    a_ids, a_ipi = draw(fr, 7, True, err)
    b_ids, b_ipi = draw(fr, 7, False, err)
    assert int(err.item()) == 0
    assert torch.equal(a_ids, b_ids) and torch.equal(a_ipi, b_ipi)
    counts = (a_ids >= 0).sum((1, 2, 3)).tolist()
    assert min(counts) > 0 and len(set(counts)) == 16, counts


@pytest.mark.parametrize("view", [2, 9])
def test_room_views_equal_oracle(dev, room, view):
    sc, fr = room
    ids, ipi = draw(fr, 11)
    P, Nc, nb = fr.camera()
    p = P[view].cpu().numpy()
    n = nb[view].cpu().numpy()
    rid, ripi, _, over = O.poisson_view(fr.W, fr.H, fr.f, p, n, sc.stddevs, 40, 10.0, 0.0, 0.01, seed=11, view=view)
    assert over == 0 and (rid >= 0).sum() > 100_000
    got = ids[view].cpu().numpy()
    bad = np.argwhere((got != rid).any(-1))
    assert bad.shape[0] == 0, "%d pixels differ, first %s" % (bad.shape[0], bad[0])
    assert np.array_equal(ipi[view].cpu().numpy(), ripi)


def test_room_properties_and_frame(dev, room):
    sc, fr = room
    ids, ipi = draw(fr, 13)
    again, _ = draw(fr, 13)
    other, _ = draw(fr, 14)
    assert torch.equal(ids, again) and not torch.equal(ids, other)
    P, Nc, nb = fr.camera()
    valid = ids >= 0
    assert int(ids.max()) < sc.n and int(ids.min()) >= -1
    assert bool(((ipi >= 1.0) == valid).all())
    # samples first: no sample after a pixel's first empty slot
    first_empty = (~valid).int().cumsum(-1) > 0
    assert not bool((valid & first_empty).any())
    # every sampled splat is in front of its view's camera
    for v in range(16):
        sel = ids[v][valid[v]].long()
        assert bool((P[v, sel, 2] < 0).all()), v
    fr.target = torch.zeros((16, fr.W, fr.H, 3), dtype=torch.float32, device=dev)
    loss, img, grads, _ = fr.step(21)
    fr.check_errors()
    assert torch.isfinite(img).all() and float(img.abs().sum()) > 0
    for g in grads:
        assert torch.isfinite(g).all()
    touched = (grads[0][0].abs().sum(-1) > 0)
    for v in range(16):
        sel = ids[v][valid[v]].long().unique()
        assert float(touched[sel].float().mean()) > 0.5, v
