"""Pins the oracle's sequential Poisson sampler (sampling.sample,
sequentialPoissonDynamicTree.cuh:16-196, sampleRadixTree.cuh:13-99,
sequentialPoissonRadixTree.cuh:15-41) with a brute restatement built from the
already-pinned pieces (tree weights, Philox uniforms):

  * per launch: stop when the root weight is below the precision; descend while
    the node holds more than k splats, to the child with the larger
    capacity-scaled normalized weight (ties to the left); then lower the
    capacities along the path by the path probability, bottom-up;
  * every leaf under the reached node draws the key u / p, with u the Philox
    uniform of (seed, view, pixel, launch, leaf ordinal);
  * the output is the k smallest finite keys over all launches, in key order,
    ties in draw order (a stable sort replaces the insertion list), padded -1.
"""
import numpy as np
import pytest

from helpers import scene_inputs
from oracle import oracle as O

f32 = np.float32


def brute_pixel(T, i, j, H, k, seed, view=0):
    L0 = T.nc - 1
    qx, qy = f32(i), f32(j)

    def W(nodes):
        nodes = np.asarray(nodes, np.int32)
        w, _ = T.weight(nodes, np.full(nodes.size, qx, f32), np.full(nodes.size, qy, f32))
        return np.maximum(f32(0), w)

    cap = {}
    draws = []
    pixel = j + H * i
    for la in range(O.num_launches(k)):
        if W([0])[0] < f32(T.precision):
            break
        nd, path = 0, [(0, f32(1))]
        while T.count[nd] > k:
            l, r = T.left[nd], T.right[nd]
            wl, wr = W([l, r])
            s = f32(wl + wr)
            with np.errstate(invalid="ignore"):        # 0 / 0: NaN compares false -> left
                wl, wr = f32(wl / s), f32(wr / s)
            if f32(wl * f32(cap.get(l, 1.0))) < f32(wr * f32(cap.get(r, 1.0))):
                nd = r
                path.append((r, wr))
            else:
                nd = l
                path.append((l, wl))
        prob = f32(1)
        for node, pr in reversed(path):
            cap[node] = max(f32(0), f32(f32(cap.get(node, 1.0)) - prob))
            prob = f32(prob * pr)
        leaves = L0 + np.arange(T.first[nd], T.last[nd] + 1)
        pw = W(leaves)
        for kk, (lf, p) in enumerate(zip(leaves, pw)):
            u = f32(O.uniform(seed, view, pixel, la, kk))
            with np.errstate(divide="ignore", over="ignore"):
                key = f32(u / p)
            if np.isfinite(key):
                draws.append((float(key), int(T.ids_sorted[lf - L0])))
    order = sorted(range(len(draws)), key=lambda t: draws[t][0])     # stable
    out = [draws[t][1] for t in order[:k]]
    return np.array(out + [-1] * (k - len(out)), np.int32)


CASES = [(2, 8.0, 8, 0.01), (2, 8.0, 40, 0.01), (0, 4.0, 16, 0.01), (0, 2.0, 5, 0.001)]


@pytest.mark.parametrize("cfg,sds,k,prec", CASES)
def test_sampler_matches_brute(cfg, sds, k, prec):
    sc, p, n, nb, sd = scene_inputs(cfg, sd_scale=sds)
    seed = 4242
    got = O.sample_view(sc.width, sc.height, sc.focal, p, nb, sd, k, 0.0, prec, seed=seed)
    T = O.Tree(sc.width, sc.height, sc.focal, p, nb, sd, 0.0, prec)
    rng = np.random.default_rng(cfg * 100 + k)
    pix = rng.integers(0, [sc.width, sc.height], size=(120, 2))
    nonempty = 0
    for i, j in pix:
        ref = brute_pixel(T, int(i), int(j), sc.height, k, seed)
        assert np.array_equal(got[i, j], ref), (i, j, got[i, j], ref)
        nonempty += int(ref[0] >= 0)
    assert nonempty >= 30


def test_sampler_invariants():
    """Valid ids, -1 only as trailing padding, every sampled splat has a positive
    leaf weight at the pixel, deterministic per seed, different across seeds."""
    sc, p, n, nb, sd = scene_inputs(0, sd_scale=4.0)
    k = 24
    a = O.sample_view(sc.width, sc.height, sc.focal, p, nb, sd, k, 0.0, 0.01, seed=1)
    b = O.sample_view(sc.width, sc.height, sc.focal, p, nb, sd, k, 0.0, 0.01, seed=1)
    c = O.sample_view(sc.width, sc.height, sc.focal, p, nb, sd, k, 0.0, 0.01, seed=2)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.max() < p.shape[0] and a.min() >= -1
    filled = a >= 0
    assert np.all(filled[..., :-1] | ~filled[..., 1:]), "-1 must only trail"
    assert filled.any(axis=-1).mean() > 0.2
    T = O.Tree(sc.width, sc.height, sc.focal, p, nb, sd, 0.0, 0.01)
    leaf_of = {int(s): T.nc - 1 + q for q, s in enumerate(T.ids_sorted)}
    rng = np.random.default_rng(0)
    for i, j in rng.integers(0, [sc.width, sc.height], size=(200, 2)):
        ids = a[i, j][a[i, j] >= 0]
        if ids.size == 0:
            continue
        leaves = np.array([leaf_of[int(s)] for s in ids], np.int32)
        w, _ = T.weight(leaves, np.full(leaves.size, i, f32), np.full(leaves.size, j, f32))
        assert np.all(w > 0)


def test_sampler_views_use_view_counter():
    """Batched views draw from Philox streams keyed by the view index."""
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    P = np.stack([p, p])
    NB = np.stack([nb, nb])
    out = O.sample(sc.width, sc.height, sc.focal, P, NB, sd[None], 8, 0.0, 0.01, seed=3)
    v0 = O.sample_view(sc.width, sc.height, sc.focal, p, nb, sd, 8, 0.0, 0.01, seed=3, view=0)
    v1 = O.sample_view(sc.width, sc.height, sc.focal, p, nb, sd, 8, 0.0, 0.01, seed=3, view=1)
    assert np.array_equal(out[0], v0) and np.array_equal(out[1], v1)
    assert not np.array_equal(v0, v1)


def test_sampler_empty_and_culled():
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    out = O.sample_view(sc.width, sc.height, sc.focal, -p, nb, sd, 8, 0.0, 0.01, seed=0)
    assert np.all(out == -1)
    z = np.zeros((0, 3), np.float32)
    out = O.sample_view(8, 8, 9.6, z, z, np.zeros((0, 2), np.float32), 4, 0.0, 0.01, seed=0)
    assert out.shape == (8, 8, 4) and np.all(out == -1)
