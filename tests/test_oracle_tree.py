"""Pins the oracle's LBVH + Hermite FGT (sampling.sample's tree stages) against
brute force and the mathematics:

  cull / Morton / stable sort   samplingCuda.cu:23-55, 251-269; buildRadixTree.cuh:18-54, 212-259
  Karras hierarchy              buildRadixTree.cuh:66-167 (vs a top-down split at the highest
                                differing bit of the 64-bit keys)
  bottom-up boxes / degrees     buildDynamicRadixTree.cuh:126-223, hermitefmm.cuh:184-195
  coefficient offsets           buildDynamicRadixTree.cuh:225-252, 399-410
  Hermite moments               buildDynamicRadixTree.cuh:254-314 (vs the Hermite generating
                                function; evaluation vs numpy's physicists' Hermite series)
  getWeight                     evaluateDynamicRadixTree.cuh:118-242 (vs the FP64 direct
                                Gaussian sum, within the Hermite truncation bound)
  bandwidth h / radius          samplingCuda.cu:106-195, 244-258 (vs the FP64 covariance)
"""
import math

import numpy as np
import pytest
from numpy.polynomial import hermite as HP

from helpers import scene_inputs
from oracle import oracle as O

CASES = [(2, 8.0, 0.01), (0, 2.0, 0.01), (0, 4.0, 0.001)]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: "cfg%d-sd%g-prec%g" % c)
def tree(request):
    cfg, sds, prec = request.param
    sc, p, n, nb, sd = scene_inputs(cfg, sd_scale=sds)
    T = O.Tree(sc.width, sc.height, sc.focal, p, nb, sd, 0.0, prec)
    assert T.nc > 1
    return sc, p, nb, sd, T


def screen32(sc, p):
    f32 = np.float32
    W, H, f = f32(sc.width), f32(sc.height), f32(sc.focal)
    sx = (-(p[:, 0] * f)) / p[:, 2] + f32(0.5) * W
    sy = (p[:, 1] * f) / p[:, 2] + f32(0.5) * H
    return sx.astype(np.float32), sy.astype(np.float32)


def interleave(v):
    out = np.zeros_like(v, dtype=np.uint64)
    for bit in range(10):
        out |= ((v.astype(np.uint64) >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit)
    return out


def test_cull_morton_stable_sort(tree):
    sc, p, nb, sd, T = tree
    sx, sy = screen32(sc, p)
    lo, hix, hiy = np.float32(T.lo), np.float32(T.hix), np.float32(T.hiy)
    vis = (p[:, 2] <= 0) & (lo <= sx) & (hix >= sx) & (lo <= sy) & (hiy >= sy)
    ids = np.nonzero(vis)[0]
    assert T.nc == ids.size
    ux = ((sx[ids] - lo) / (hix - lo)).astype(np.float32)
    uy = ((sy[ids] - lo) / (hiy - lo)).astype(np.float32)
    cx = np.clip(ux * np.float32(1024), 0, 1023).astype(np.uint32)
    cy = np.clip(uy * np.float32(1024), 0, 1023).astype(np.uint32)
    key = (2 * interleave(cx) + interleave(cy)).astype(np.uint32)
    order = np.argsort(key, kind="stable")
    assert np.array_equal(T.morton_sorted, key[order])
    assert np.array_equal(T.ids_sorted, ids[order])


def _clz64(x):
    return 64 - int(x).bit_length()


def test_karras_matches_topdown_split(tree):
    """Every internal node is the top-down split of its range at the highest
    differing bit of the (morton << 32 | position) keys, and every range of the
    top-down construction appears exactly once."""
    sc, p, nb, sd, T = tree
    nc, L0 = T.nc, T.nc - 1
    keys = [(int(m) << 32) | i for i, m in enumerate(T.morton_sorted)]
    rng_of = {(int(T.first[i]), int(T.last[i])): i for i in range(L0)}
    assert len(rng_of) == L0
    assert rng_of[(0, nc - 1)] == 0
    stack, seen = [(0, nc - 1)], 0
    while stack:
        i, j = stack.pop()
        if i == j:
            continue
        d = _clz64(keys[i] ^ keys[j])
        m = max(x for x in range(i, j) if _clz64(keys[i] ^ keys[x]) > d)
        nd = rng_of[(i, j)]
        seen += 1
        l, r = T.left[nd], T.right[nd]
        assert (T.first[l], T.last[l]) == (i, m)
        assert (T.first[r], T.last[r]) == (m + 1, j)
        assert (l >= L0) == (m == i) and (r >= L0) == (m + 1 == j)
        stack += [(i, m), (m + 1, j)]
    assert seen == L0


def test_tree_invariants(tree):
    sc, p, nb, sd, T = tree
    nc, L0 = T.nc, T.nc - 1
    nn = 2 * nc - 1
    assert T.parent[0] == -1
    for i in range(L0):
        l, r = T.left[i], T.right[i]
        assert T.parent[l] == i and T.parent[r] == i
        assert T.count[i] == T.count[l] + T.count[r] == T.last[i] - T.first[i] + 1
    assert np.all(T.left[L0:nn] == -1) and np.all(T.count[L0:nn] == 1)
    assert np.array_equal(T.first[L0:nn], np.arange(nc))
    # every leaf reached exactly once from the root
    hits = np.zeros(nc, int)
    stack = [0]
    while stack:
        x = stack.pop()
        if x >= L0:
            hits[x - L0] += 1
        else:
            stack += [T.left[x], T.right[x]]
    assert np.all(hits == 1)


def test_boxes_and_centres(tree):
    sc, p, nb, sd, T = tree
    nc, L0 = T.nc, T.nc - 1
    sx, sy = screen32(sc, p)
    lx, ly = sx[T.ids_sorted], sy[T.ids_sorted]
    assert np.array_equal(T.cen[L0:, 0], lx) and np.array_equal(T.cen[L0:, 1], ly)
    for i in range(L0):
        a, b = T.first[i], T.last[i] + 1
        assert T.box[i, 0] == lx[a:b].min() and T.box[i, 2] == lx[a:b].max()
        assert T.box[i, 1] == ly[a:b].min() and T.box[i, 3] == ly[a:b].max()
        mean = np.array([lx[a:b].astype(np.float64).mean(), ly[a:b].astype(np.float64).mean()])
        assert np.abs(T.cen[i] - mean).max() <= 2e-5 * max(1.0, np.abs(mean).max())


def _radius_ratio(T, i):
    c, b = T.cen[i], T.box[i]
    f32 = np.float32
    mn = f32(0)
    for d in range(2):
        mn = max(mn, max(f32(c[d] - b[d]), f32(b[2 + d] - c[d])))
    return float(f32(mn) / f32(T.h))


def test_degree_rule_and_offsets(tree):
    """Accepted degree = the smallest p >= both children's degrees whose Hermite
    bound is <= precision (p <= 16, r < 1); -1 = no expansion."""
    sc, p, nb, sd, T = tree
    nc, L0 = T.nc, T.nc - 1
    assert np.all(T.deg[L0:] == 0)
    for i in range(L0):
        dl, dr, d = T.deg[T.left[i]], T.deg[T.right[i]], T.deg[i]
        r = _radius_ratio(T, i)
        if d >= 0:
            assert d >= max(dl, dr) and d <= 16 and r < 1.0
            assert O.hermite_error(r, float(T.count[i]), d) <= T.precision
            if d > max(dl, dr):
                assert O.hermite_error(r, float(T.count[i]), d - 1) > T.precision
        else:
            assert (dl < 0 and dr < 0) or r >= 1.0 or \
                O.hermite_error(r, float(T.count[i]), 16) > T.precision
    sizes = np.where(T.deg >= 0, (T.deg + 1) ** 2, 0)
    assert np.array_equal(T.off, np.concatenate([[0], np.cumsum(sizes)[:-1]]))
    assert T.ncoef == sizes.sum()


def _expanded_nodes(T, limit=None):
    L0 = T.nc - 1
    nodes = [i for i in range(L0) if T.deg[i] >= 1]
    return nodes[:limit] if limit else nodes


def test_moments_are_hermite_source_moments(tree):
    """coef[a, b] = sum over the node's leaves of s_x^a s_y^b / (a! b!),
    s = (leaf - centre) / sqrt(2 h^2): the generating-function coefficients of
    exp(-|t - s|^2) = sum_ab s_x^a s_y^b / (a! b!) H_a(t_x) H_b(t_y) exp(-|t|^2)."""
    sc, p, nb, sd, T = tree
    L0 = T.nc - 1
    s = 1.0 / math.sqrt(2.0 * float(T.h) ** 2)
    for i in _expanded_nodes(T, 400):
        d = T.deg[i]
        lv = T.cen[L0 + T.first[i]: L0 + T.last[i] + 1].astype(np.float64)
        sxy = (lv - T.cen[i].astype(np.float64)) * s
        fa = np.array([math.factorial(a) for a in range(d + 1)], np.float64)
        mx = sxy[:, 0:1] ** np.arange(d + 1) / fa
        my = sxy[:, 1:2] ** np.arange(d + 1) / fa
        ref = mx.T @ my
        got = T.coef[T.off[i]: T.off[i] + (d + 1) ** 2].reshape(d + 1, d + 1)
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-5 * scale, i


def test_node_value_is_hermite_series(tree):
    """W(node, q) equals the FP64 Hermite series of the node's moments, evaluated
    with numpy's physicists' Hermite polynomials."""
    sc, p, nb, sd, T = tree
    s = 1.0 / math.sqrt(2.0 * float(T.h) ** 2)
    rng = np.random.default_rng(5)
    for i in _expanded_nodes(T, 200):
        d = T.deg[i]
        C = T.coef[T.off[i]: T.off[i] + (d + 1) ** 2].reshape(d + 1, d + 1).astype(np.float64)
        q = T.cen[i][None, :] + rng.uniform(-2.0 * T.h, 2.0 * T.h, size=(8, 2)).astype(np.float32)
        w, _ = T.weight(np.full(8, i), q[:, 0], q[:, 1])
        t = (q.astype(np.float64) - T.cen[i].astype(np.float64)) * s
        ref = HP.hermval2d(t[:, 0], t[:, 1], C) * np.exp(-(t ** 2).sum(1))
        assert np.abs(w - ref).max() <= 1e-5 * max(np.abs(C).sum(), 1.0), i


def test_fgt_error_within_hermite_bound(tree):
    """|W(node, q) - sum_leaves exp(-|q - y|^2 / 2h^2)| <= hermiteError(r, n, p)
    at every node with an accepted expansion (plus FP32 rounding)."""
    sc, p, nb, sd, T = tree
    rng = np.random.default_rng(6)
    L0 = T.nc - 1
    for i in [x for x in range(L0) if T.deg[x] >= 0]:
        bound = O.hermite_error(_radius_ratio(T, i), float(T.count[i]), int(T.deg[i]))
        q = T.cen[i][None, :] + rng.uniform(-3.0 * T.h, 3.0 * T.h, size=(8, 2)).astype(np.float32)
        w, direct = T.weight(np.full(8, i), q[:, 0], q[:, 1])
        assert np.abs(w - direct).max() <= bound + 2e-6 * T.count[i], i


def test_leaf_weight_is_gaussian(tree):
    """Degree-0 leaves evaluate exp(-|q - y|^2 / 2h^2) exactly up to FP32 rounding."""
    sc, p, nb, sd, T = tree
    L0 = T.nc - 1
    rng = np.random.default_rng(8)
    leaves = L0 + rng.integers(0, T.nc, size=256)
    q = T.cen[leaves] + rng.uniform(-2, 2, size=(256, 2)).astype(np.float32)
    w, direct = T.weight(leaves, q[:, 0], q[:, 1])
    assert np.abs(w - direct).max() <= 2e-6


def _truncated(T, nd, qx, qy):
    f32 = np.float32
    b = T.box[nd]
    cx = min(max(f32(qx), b[0]), b[2])
    cy = min(max(f32(qy), b[1]), b[3])
    dx, dy = f32(qx) - cx, f32(qy) - cy
    d2 = f32(f32(dx * dx) + f32(dy * dy))
    return not (d2 < T.r2[nd])


def test_tree_weight_frontier(tree):
    """getWeight(root, q) = sum of the node expansions on a frontier that covers
    every leaf exactly once; subtrees without an expansion are pruned only when
    the query is outside their truncation radius, where their whole Gaussian mass
    is <= precision (n exp(-R^2 / 2h^2) = precision)."""
    sc, p, nb, sd, T = tree
    L0 = T.nc - 1
    rng = np.random.default_rng(9)
    for qx, qy in rng.uniform(0, [sc.width, sc.height], size=(12, 2)).astype(np.float32):
        ev, tr, stack = [], [], [0]
        while stack:
            x = stack.pop()
            if T.deg[x] >= 0:
                ev.append(x)
            elif _truncated(T, x, qx, qy):
                tr.append(x)
            else:
                stack += [T.right[x], T.left[x]]
        cover = np.zeros(T.nc, int)
        for x in ev + tr:
            cover[T.first[x]: T.last[x] + 1] += 1
        assert np.all(cover == 1)
        w_root, d_root = T.weight(np.array([0]), np.array([qx]), np.array([qy]))
        w_ev, d_ev = T.weight(np.array(ev), np.full(len(ev), qx), np.full(len(ev), qy))
        assert abs(w_root[0] - w_ev.astype(np.float64).sum()) <= 1e-5 * max(np.abs(w_ev).sum(), 1e-3)
        if tr:
            _, d_tr = T.weight(np.array(tr), np.full(len(tr), qx), np.full(len(tr), qy))
            assert np.all(d_tr <= T.precision * (1 + 1e-4))
        bounds = sum(O.hermite_error(_radius_ratio(T, x), float(T.count[x]), int(T.deg[x]))
                     for x in ev if x < L0)
        assert abs(w_root[0] - d_root[0]) <= bounds + T.precision * len(tr) + 1e-5 * T.nc


def test_bandwidth_radius_truncation(tree):
    """h_k = min(1, 1/sqrt(lambda_min(V^-1))) = min(1, sqrt(lambda_max(V))) with V the
    FP64 EWA covariance; h = mean h_k; the cull radius satisfies
    exp(-radius^2 / 2h^2) / (2 pi sqrt(h)) = precision and the node truncation
    radius n exp(-R^2 / 2h^2) = precision."""
    sc, p, nb, sd, T = tree
    V = O.covariance(sc.width, sc.height, sc.focal, p[None].astype(np.float64), nb[None].astype(np.float64),
                     sd[None].astype(np.float64))[0]
    ev = np.linalg.eigvalsh(V)
    ref = np.minimum(1.0, np.sqrt(ev[:, 1]))
    kappa = ev[:, 1] / ev[:, 0]
    hk = np.minimum(np.float32(1.0), T.hk)          # exported before the one-pixel cap
    ok = np.isfinite(hk)
    # the FP32 eigenvalue of V^-1 (as the reference computes it) carries a relative
    # error of about eps_f32 * cond(V): near edge-on splats have cond up to 1e7
    err = np.abs(hk[ok] - ref[ok])
    assert np.all(err <= 1e-4 + 1.2e-7 * kappa[ok])
    assert err[kappa[ok] < 1e3].max() <= 1e-4
    assert abs(T.h - hk[ok].astype(np.float64).mean()) <= 1e-6
    h = float(T.h)
    rel = math.exp(-T.radius ** 2 / (2 * h * h)) / (2 * math.pi * math.sqrt(h)) / T.precision
    assert abs(rel - 1.0) < 1e-4
    for i in [0, T.nc - 1, 2 * T.nc - 2]:
        v = T.count[i] * math.exp(-float(T.r2[i]) / (2 * h * h)) / T.precision
        assert abs(v - 1.0) < 1e-3
