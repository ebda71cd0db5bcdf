"""Launch shapes, view batching and HIP graphs of Poisson-rate sampling:
every pixel-tile shape / workgroup size / block order of the per-pixel passes
gives the same bits as the oracle; a batch of views in one launch per pass
equals the view-by-view path; a Philox key read from device memory equals the
same key passed by value; and a captured graph of the sampler, replayed with
eager work and synchronisations in between, reproduces the eager result of
every replay's key."""
import itertools

import numpy as np
import pytest
import torch

from helpers import scene_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.fixture
def restore_tiles():
    from cudaExtensions import sampling
    saved = {k: sampling.get_tile_config(k) for k in ("fmm_eval", "poisson")}
    yield
    for k, c in saved.items():
        sampling.set_tile_config(k, *c)


SHAPES = [(8, 8), (16, 4), (4, 16), (32, 2), (2, 32), (64, 1), (1, 64)]


def test_tile_shapes_bit_exact(dev, restore_tiles):
    from cudaExtensions import sampling
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = 52, 36, sc.focal * 52 / sc.width          # ragged tiles in both directions
    rid, ripi, rdens, _ = O.poisson_view(W, H, f, p, nb, sd, 48, 6.0, 0.0, 0.01, seed=5)
    rfmm = O.fmm_weights_view(W, H, f, p, nb, sd, 0.0, 0.01)
    P, NB, SD = t(p[None], dev), t(nb[None], dev), t(sd[None], dev)
    for (tw, th), wpb, xcd in itertools.product(SHAPES, (1, 2, 4, 8), (0, 1)):
        sampling.set_tile_config("fmm_eval", tw, th, wpb, xcd)
        sampling.set_tile_config("poisson", th, tw, 8 // wpb if wpb > 1 else 8, 1 - xcd)
        dens = torch.zeros((1, 2, W, H), dtype=torch.float32, device=dev)
        ids, ipi = sampling.samplePoisson(W, H, f, P, NB, None, SD, 48, 0.0, 0.01, seed=5, rate=6.0, densities=dens)
        fmm = sampling.approximateFMMWeights(W, H, f, P, NB, None, SD, 0.0, 0.01)
        tag = (tw, th, wpb, xcd)
        assert np.array_equal(dens[0].cpu().numpy(), rdens), tag
        assert np.array_equal(ids[0].cpu().numpy(), rid), tag
        assert np.array_equal(ipi[0].cpu().numpy(), ripi), tag
        assert np.array_equal(fmm[0].cpu().numpy(), rfmm), tag


def test_tile_config_rejects_bad_shapes():
    from cudaExtensions import sampling
    for args in [("poisson", 8, 4, 4), ("poisson", 3, 21, 4), ("fmm_eval", 8, 8, 3), ("fmm_eval", 8, 8, 16)]:
        with pytest.raises(RuntimeError):
            sampling.set_tile_config(*args)


def test_batched_views_equal_view_by_view(dev):
    """Uneven culled counts: each view sees a different part of the scene."""
    from cudaExtensions import sampling
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = sc.width, sc.height, sc.focal
    views = []
    for v in range(5):
        q = p.copy()
        q[:, 0] += 0.15 * (v - 2)
        q[:, 1] *= 1.0 if v % 2 == 0 else -1.0
        views.append(q)
    P = t(np.stack(views), dev)
    NB = t(np.stack([nb] * 5), dev)
    SD = t(sd[None], dev)
    ts = t(np.random.default_rng(1).uniform(0.5, 2.0, size=(5, (W + 15) // 16, (H + 15) // 16)).astype(np.float32),
           dev)
    outs = {}
    for batched in (True, False):
        dens = torch.zeros((5, 2, W, H), dtype=torch.float32, device=dev)
        ids, ipi = sampling.samplePoisson(W, H, f, P, NB, None, SD, 48, 0.0, 0.01, seed=9, rate=5.0, tileScale=ts,
                                          densities=dens, batched=batched)
        outs[batched] = (ids.cpu().numpy(), ipi.cpu().numpy(), dens.cpu().numpy())
    counts = (outs[True][0] >= 0).sum((1, 2, 3))
    assert len(set(counts.tolist())) == 5 and counts.min() > 0
    for a, b in zip(outs[True], outs[False]):
        assert np.array_equal(a, b)
    rid, ripi = O.poisson(W, H, f, np.stack(views), np.stack([nb] * 5), sd[None], 48, 5.0, seed=9,
                          tile_scale=ts.cpu().numpy())
    assert np.array_equal(outs[True][0], rid) and np.array_equal(outs[True][1], ripi)


def test_device_seed_equals_seed(dev):
    from cudaExtensions import sampling
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    args = (sc.width, sc.height, sc.focal, t(p[None], dev), t(nb[None], dev), None, t(sd[None], dev), 32, 0.0, 0.01)
    key = (3 << 40) + 12345
    a = sampling.samplePoisson(*args, seed=key, rate=4.0)
    st = torch.tensor([key], dtype=torch.int64, device=dev)
    b = sampling.samplePoisson(*args, seedTensor=st, rate=4.0)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def test_graph_replays_match_eager(dev):
    from cudaExtensions import sampling
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    P = t(np.stack([p, p * np.float32([1.0, -1.0, 1.0])]), dev)
    NB = t(np.stack([nb, nb]), dev)
    SD = t(sd[None], dev)
    W, H, f = sc.width, sc.height, sc.focal
    key = torch.tensor([1000], dtype=torch.int64, device=dev)
    dens = torch.zeros((2, 2, W, H), dtype=torch.float32, device=dev)

    def frame():
        ids, ipi = sampling.samplePoisson(W, H, f, P, NB, None, SD, 48, 0.0, 0.01, seedTensor=key, rate=6.0,
                                          densities=dens)
        return ids, ipi

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        frame()                                            # warm the allocator outside the capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        gids, gipi = frame()
    for r in range(4):
        key.fill_(2000 + 17 * r)
        g.replay()
        torch.cuda.synchronize()
        got_ids, got_ipi = gids.clone(), gipi.clone()
        eids, eipi = sampling.samplePoisson(W, H, f, P, NB, None, SD, 48, 0.0, 0.01, seed=2000 + 17 * r, rate=6.0)
        torch.cuda.synchronize()
        assert torch.equal(got_ids, eids), r
        assert torch.equal(got_ipi, eipi), r
        assert (got_ids >= 0).sum() > 1000


def test_multiview_frame_graph(dev):
    """A whole fwd+bwd frame over 4 room views (uneven culled counts) captured as
    one graph that advances its own Philox key: every replay equals the eager
    frame of the same key (images bit for bit; gradients, which sum over views
    with atomics, within 1e-4 relative)."""
    from splatting_amd import learned as L, scenes
    sc = scenes.Scene(*scenes.box_room(60_000, n_boxes=20), 96, 72, 60.0,
                      scenes.room_cameras(4, boxes=scenes.room_boxes(20)))
    fr = L.PoissonFrame(sc, dev, rate=6.0, kcap=32)
    fr.target = torch.full((4, 96, 72, 3), 0.25, dtype=torch.float32, device=dev)
    key = torch.tensor([500], dtype=torch.int64, device=dev)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fr.step(0, seed_tensor=key, with_loss=False)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        key.add_(1)
        _, gimg, ggrads, _ = fr.step(0, seed_tensor=key, with_loss=False)
    for r in range(3):
        graph.replay()
        torch.cuda.synchronize()
        k = int(key.item())
        img, grads = gimg.clone(), [x.clone() for x in ggrads]
        _, eimg, egrads, _ = fr.step(k, with_loss=False)
        torch.cuda.synchronize()
        assert k == 501 + r
        assert torch.equal(img, eimg), r
        for a, b in zip(grads, egrads):
            scale = b.abs().max().clamp(min=1e-30)
            assert ((a - b).abs().max() / scale) < 1e-4, r
        counts = (fr.last_ids >= 0).sum((1, 2, 3)).tolist()
        assert min(counts) > 0 and len(set(counts)) > 1, counts
