"""Pins the oracle's composited Horvitz-Thompson estimate (oracle.htc), the
Poisson-rate counterpart of splatting -> filtering -> interpolation:

  * with every inclusion certain (1/pi = 1) it is the reference's composited
    image, i.e. the oracle's filter_forward + interp_forward chain (which the
    other oracle tests pin against closed forms and central differences);
  * its adjoint equals FP64 central differences of its forward;
  * unbiasedness, exactly: at one pixel with 10 candidate splats every one of
    the 2^10 inclusion patterns is a view of its own; their probability-weighted
    sum of images (and of gradients) equals the image (gradients) composited
    over all 10 -- the expectation taken by enumeration, not sampling.
"""
import itertools

import numpy as np

from oracle import oracle as O


def _pixel_scene(n=10, seed=0):
    """n splats facing the camera near the ray through pixel (2, 2) of a 6x6 view
    at distinct depths, each covering the pixel with a weight in (0, 1)."""
    rng = np.random.default_rng(seed)
    W = H = 6
    f = 6.0
    z = -rng.uniform(1.5, 3.0, size=n)
    # screen (2, 2) +- 0.4 px: x = -(sx - W/2) z / f, y = (sy - H/2) z / f
    sx = 2.0 + rng.uniform(-0.4, 0.4, size=n)
    sy = 2.0 + rng.uniform(-0.4, 0.4, size=n)
    p = np.stack([-(sx - W / 2) * z / f, (sy - H / 2) * z / f, z], 1)
    nrm = rng.normal(size=(n, 3)) * 0.2 + np.array([0.0, 0.0, 1.0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    sd = rng.uniform(0.25, 0.5, size=(n, 2)) * (-z[:, None]) / f
    attr = rng.uniform(size=(n, 3))
    return W, H, f, p.astype(np.float32), nrm.astype(np.float32), sd, attr


def test_certain_inclusions_equal_filtering_chain():
    W, H, f, p, n, sd, a = _pixel_scene()
    rng = np.random.default_rng(1)
    k = 12
    idx = np.full((1, W, H, k), -1, np.int32)
    for i in range(W):
        for j in range(H):
            m = rng.integers(0, 11)
            idx[0, i, j, :m] = rng.permutation(10)[:m]
    ipi = np.where(idx >= 0, 1.0, 0.0)
    img = O.htc(W, H, f, idx, ipi, p[None], n[None], sd[None], a[None])
    w = O.splat_forward(W, H, f, idx, p[None].astype(np.float64), n[None].astype(np.float64), sd[None])
    dw = O.filter_forward(idx, w, W, H, f, p[None], n[None])
    ref = O.interp_forward(idx, dw, a[None])
    assert np.abs(ref).max() > 0.1
    assert np.allclose(img, ref, rtol=1e-12, atol=1e-14)


def test_adjoint_matches_central_differences():
    W, H, f, p, n, sd, a = _pixel_scene(seed=3)
    rng = np.random.default_rng(4)
    k = 12
    idx = np.full((1, W, H, k), -1, np.int32)
    for i in range(W):
        for j in range(H):
            m = rng.integers(3, 11)
            idx[0, i, j, :m] = rng.permutation(10)[:m]
    ipi = np.where(idx >= 0, rng.uniform(1.0, 3.0, size=idx.shape), 0.0)
    g = rng.normal(size=(1, W, H, 3))
    _, (gP, gN, gS, gA) = O.htc(W, H, f, idx, ipi, p[None], n[None], sd[None], a[None], gimg=g)
    L = lambda pp, nn, ss, aa: float((O.htc(W, H, f, idx, ipi, pp, nn, ss, aa) * g).sum())
    h = 2.0 ** -12           # exact in float32 for |x| < 2^11
    for s in range(10):
        for c in range(3):
            for arr, grad in ((p, gP), (n, gN)):
                up = arr[None].copy(); up[0, s, c] += h
                dn = arr[None].copy(); dn[0, s, c] -= h
                args_u = (up, n[None], sd[None], a[None]) if arr is p else (p[None], up, sd[None], a[None])
                args_d = (dn, n[None], sd[None], a[None]) if arr is p else (p[None], dn, sd[None], a[None])
                fd = (L(*args_u) - L(*args_d)) / (2 * h)
                assert abs(fd - grad[0, s, c]) <= 1e-4 * abs(fd) + 1e-6 * np.abs(grad).max(), (s, c, fd, grad[0, s, c])
            ua = a[None].copy(); ua[0, s, c] += 1e-6
            da = a[None].copy(); da[0, s, c] -= 1e-6
            fd = (L(p[None], n[None], sd[None], ua) - L(p[None], n[None], sd[None], da)) / 2e-6
            assert abs(fd - gA[0, s, c]) <= 1e-8 * abs(fd) + 1e-9 * np.abs(gA).max()
        for c in range(2):
            us = sd[None].copy(); us[0, s, c] += 1e-7
            ds = sd[None].copy(); ds[0, s, c] -= 1e-7
            fd = (L(p[None], n[None], us, a[None]) - L(p[None], n[None], ds, a[None])) / 2e-7
            assert abs(fd - gS[0, s, c]) <= 1e-5 * abs(fd) + 1e-7 * np.abs(gS).max(), (s, c, fd, gS[0, s, c])


def test_expectation_over_all_inclusion_patterns_is_the_dense_composite():
    W, H, f, p, n, sd, a = _pixel_scene(seed=5)
    rng = np.random.default_rng(6)
    N = p.shape[0]
    pi = rng.uniform(0.15, 0.95, size=N)
    pats = np.array(list(itertools.product([0, 1], repeat=N)), bool)      # (1024, N)
    V = pats.shape[0]
    k = N
    idx = np.full((V, W, H, k), -1, np.int32)
    ipi = np.zeros((V, W, H, k))
    prob = np.ones(V)
    for v in range(V):
        ids = np.flatnonzero(pats[v])
        idx[v, 2, 2, :ids.size] = rng.permutation(ids)           # slot order must not matter
        ipi[v, 2, 2, :ids.size] = 1.0 / pi[idx[v, 2, 2, :ids.size]]
        prob[v] = np.prod(np.where(pats[v], pi, 1.0 - pi))
    assert abs(prob.sum() - 1.0) < 1e-12
    P = np.repeat(p[None], V, 0)
    Nn = np.repeat(n[None], V, 0)
    G = np.zeros((V, W, H, 3))
    G[:, 2, 2] = np.array([0.3, -1.1, 0.7]) * prob[:, None]
    img, (gP, gN, gS, gA) = O.htc(W, H, f, idx, ipi, P, Nn, sd[None], a[None], gimg=G)
    mean_img = (img[:, 2, 2] * prob[:, None]).sum(0)
    # dense: all N splats, 1/pi = 1
    didx = np.full((1, W, H, k), -1, np.int32)
    didx[0, 2, 2] = np.arange(N)
    dipi = np.where(didx >= 0, 1.0, 0.0)
    Gd = np.zeros((1, W, H, 3))
    Gd[0, 2, 2] = [0.3, -1.1, 0.7]
    dimg, (dP, dN, dS, dA) = O.htc(W, H, f, didx, dipi, p[None], n[None], sd[None], a[None], gimg=Gd)
    assert np.abs(dimg[0, 2, 2]).max() > 0.05
    assert np.allclose(mean_img, dimg[0, 2, 2], rtol=1e-10, atol=1e-13)
    # gradients: views carry their own copies of points / normals, shared std-devs and attributes
    assert np.allclose(gP.sum(0), dP[0], rtol=1e-9, atol=1e-12)
    assert np.allclose(gN.sum(0), dN[0], rtol=1e-9, atol=1e-12)
    assert np.allclose(gS, dS, rtol=1e-9, atol=1e-12)
    assert np.allclose(gA, dA, rtol=1e-9, atol=1e-12)
    # while a single draw is not the dense image
    assert np.abs(img[:, 2, 2] - dimg[0, 2, 2]).max() > 1e-3
