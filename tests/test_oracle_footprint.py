"""The oracle's footprint proposal of Poisson-rate sampling, pinned on the CPU:
every node bound dominates the footprint term of each of its leaves at any pixel
(the thinning's U >= pi needs it); the footprint term equals the splat's EWA
weight min(w, 1) of the FP64 splatting restatement; and the footprint mass of a
leaf, the closed form 2 pi sqrt(det V) (amp if amp <= 1, else 1 + ln amp), equals
a numerical integral of min(w, 1) over the plane."""
import numpy as np
import pytest

from helpers import scene_inputs
from oracle import oracle as O


def boxes_and_chains(leaf, par):
    """Per node the box of its leaves' points, and per leaf its ancestors (itself first)."""
    nc = leaf.shape[0]
    L0 = nc - 1
    lo = np.full((2 * nc - 1, 2), np.inf, np.float32)
    hi = np.full((2 * nc - 1, 2), -np.inf, np.float32)
    chains = []
    for pos in range(nc):
        nd = L0 + pos
        ch = [nd]
        while nd != 0:
            nd = int(par[nd])
            ch.append(nd)
        for a in ch:
            lo[a] = np.minimum(lo[a], leaf[pos, :2])
            hi[a] = np.maximum(hi[a], leaf[pos, :2])
        chains.append(ch)
    return lo, hi, chains


@pytest.mark.parametrize("cfg,sds", [(0, 2.0), (2, 8.0)])
def test_node_bounds_dominate_leaf_terms(cfg, sds):
    sc, p, n, nb, sd = scene_inputs(cfg, sd_scale=sds)
    leaf, ids, par, mass, a1 = O.foot_view(sc.width, sc.height, sc.focal, p, nb, sd)
    assert leaf.shape[0] > 100 and (leaf[:, 5] > 0).mean() > 0.9 and a1 > 0
    lo, hi, chains = boxes_and_chains(leaf, par)
    rng = np.random.default_rng(cfg)
    pos = rng.integers(0, leaf.shape[0], 4000)
    # pixels near the leaf (where its term is large) and anywhere in the image
    near = leaf[pos, :2] + rng.normal(0.0, 3.0, size=(pos.size, 2))
    far = rng.uniform(0, [sc.width, sc.height], size=(pos.size, 2))
    q = np.rint(np.concatenate([near, far])).astype(np.float32)
    pos = np.concatenate([pos, pos])
    fvals = O.foot_terms(leaf[pos], a1, q)
    assert (fvals > 1e-3).sum() > 500
    rows, d2s, fs = [], [], []
    for k, ps in enumerate(pos):
        for a in chains[ps]:
            c = np.minimum(np.maximum(q[k], lo[a]), hi[a])
            d = (q[k] - c).astype(np.float32)
            d2 = np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])
            rows.append(a); d2s.append(d2); fs.append(fvals[k])
    b = O.foot_bounds(a1, np.array(d2s, np.float32))
    fs = np.array(fs, np.float32)
    assert np.all(b >= fs), "a node bound below a leaf term: %d of %d" % ((b < fs).sum(), b.size)


@pytest.mark.parametrize("cfg,sds", [(0, 2.0), (2, 8.0)])
def test_footprint_term_is_the_ewa_weight(cfg, sds):
    sc, p, n, nb, sd = scene_inputs(cfg, sd_scale=sds)
    W, H, f = sc.width, sc.height, sc.focal
    leaf, ids, par, mass, a1 = O.foot_view(W, H, f, p, nb, sd)
    rng = np.random.default_rng(7 + cfg)
    idx = np.full((1, W, H, 1), -1, np.int32)
    qs, ls = [], []
    for pos in rng.permutation(leaf.shape[0])[:3000]:
        qi = int(np.rint(leaf[pos, 0] + rng.normal(0, 1.5)))
        qj = int(np.rint(leaf[pos, 1] + rng.normal(0, 1.5)))
        if 0 <= qi < W and 0 <= qj < H and idx[0, qi, qj, 0] < 0:
            idx[0, qi, qj, 0] = ids[pos]
            qs.append((qi, qj)); ls.append(pos)
    qs = np.array(qs, np.float32); ls = np.array(ls)
    wref = O.splat_forward(W, H, f, idx, p[None].astype(np.float64), nb[None].astype(np.float64),
                           sd[None].astype(np.float64))[0, qs[:, 0].astype(int), qs[:, 1].astype(int), 0]
    fv = O.foot_terms(leaf[ls], a1, qs)
    # the cap E(d) = exp(-max(0, d - 2)^2 a1), in FP64 from its definition
    d = np.hypot(qs[:, 0] - leaf[ls, 0].astype(np.float64), qs[:, 1] - leaf[ls, 1].astype(np.float64))
    cap = np.exp(-np.maximum(0.0, d - 2.0) ** 2 * float(a1))
    ref = np.minimum(np.minimum(wref, 1.0), cap)
    ok = (leaf[ls, 5] > 0) & (ref > 1e-4)
    assert ok.sum() > 150 and (cap < wref).sum() > 10
    rel = np.abs(fv[ok] - ref[ok]) / ref[ok]
    # FP32 covariance against the FP64 chain: a few hundred ulp for ordinary splats,
    # more for surfels seen nearly edge-on; the median pins the formula
    assert np.median(rel) < 2e-5 and np.mean(rel < 1e-3) > 0.98, (np.median(rel), np.max(rel))


def test_footprint_mass_closed_form():
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    leaf, ids, par, mass, a1 = O.foot_view(sc.width, sc.height, sc.focal, p, nb, sd)
    L = leaf.astype(np.float64)
    good = L[:, 5] > 0
    area = 2.0 * np.pi / np.where(good, L[:, 2] * L[:, 4], 1.0)   # 2 pi sqrt(det V), det V^-1 = (t00 t11)^2
    closed = np.where(L[:, 5] <= 1.0, area * L[:, 5], area * (1.0 + np.log(np.maximum(L[:, 5], 1e-30))))
    closed = np.where(good, closed, 0.0)
    # the oracle's total (FP32 per leaf, summed in FP64) against the closed form in FP64
    assert abs(mass / closed.sum() - 1.0) < 1e-4
    # the closed form against a numerical integral of min(w, 1) for a few leaves
    rng = np.random.default_rng(3)
    for pos in rng.choice(np.flatnonzero(good), 6, replace=False):
        t00, t01, t11, amp = L[pos, 2:6]
        # integrate in whitened coordinates (u, v) = T d: dq = du dv / (t00 t11)
        r = np.sqrt(2.0 * max(np.log(max(amp, 1.0)), 0.0) + 2.0 * 40.0)
        g = np.linspace(-r, r, 1601)
        du = g[1] - g[0]
        U, V = np.meshgrid(g, g, indexing="ij")
        val = np.minimum(amp * np.exp(-0.5 * (U * U + V * V)), 1.0).sum() * du * du / (t00 * t11)
        assert abs(val / closed[pos] - 1.0) < 2e-3, (pos, val, closed[pos])
