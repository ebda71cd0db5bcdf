"""Pins the oracle's Poisson-rate sampler (independent inclusion of every culled
splat i at pixel q with the defensive-mixture probability
pi_i = min(1, lambda/2 (g0_i / max(D0, c0) + g1_i / max(D1, c1))), g_b the Gaussian of
the tree bandwidth (b = 0) or the wide bandwidth (b = 1) times c_b = 1 / (2 pi sqrt(h_b)),
D_b the fixed-order density of bandwidth b) by what the design fixes:

  * the empirical inclusion frequency of every (pixel, splat) pair over many
    Philox seeds matches pi_i (z-scores of mean 0 and variance 1, and every
    count within its exact binomial tails, Bonferroni over the pairs);
  * the returned 1/pi equals the inclusion probability computed independently
    from the splat's screen position, h and D;
  * the expected sample count is sum_i pi_i, so the Horvitz-Thompson sum
    sum_{i in S} x_i / pi_i is unbiased for sum_i x_i;
  * no samples where D < precision; deterministic per seed.
"""
import math

import numpy as np
import pytest

from helpers import scene_inputs
from oracle import oracle as O

W = H = 16


def small_case():
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    f = sc.focal * W / sc.width
    return p, nb, sd, f


def proposal(p, nb, sd, f, prec):
    """Per (pixel, sorted leaf): the two mixture components c_b exp(-|q - s_i|^2 / 2 h_b^2)
    for the tree bandwidth h_0 = h and the wide bandwidth h_1."""
    T = O.Tree(W, H, f, p, nb, sd, 0.0, prec)
    hs = [np.float32(T.h), np.float32(O.poisson_wide_h(W, H, f, p, nb, sd, 0.0, prec))]
    lv = T.cen[T.nc - 1:].astype(np.float64)
    ii, jj = np.meshgrid(np.arange(W), np.arange(H), indexing="ij")
    q = np.stack([ii.ravel(), jj.ravel()], 1).astype(np.float64)
    out = []
    for h in hs:
        c = 1.0 / (2.0 * math.pi * math.sqrt(h))
        # one exponential of |q - s|^2 / 2h^2; the FP32 exponential flushes arguments below -87 to 0
        t2 = (((q[:, None, :] - lv[None, :, :]) / (math.sqrt(2.0) * float(h))) ** 2).sum(-1)
        e = np.where(t2 <= 87.0, np.exp(-t2), 0.0)
        out.append((c * e, c))
    return out, T


def inclusion_probs(p, nb, sd, f, lam, prec, dens, raw=False):
    """pi_i = min(1, lam p_i), p_i = (g0_i / max(D0, c0) + g1_i / max(D1, c1)) / 2; 0 where
    D0 < prec or p_i < 1e-4 (outside the support)."""
    (g0, c0), (g1, c1) = proposal(p, nb, sd, f, prec)[0]
    T = proposal(p, nb, sd, f, prec)[1]
    D0 = dens[0].reshape(-1, 1).astype(np.float64)
    D1 = dens[1].reshape(-1, 1).astype(np.float64)
    pr = 0.5 * (g0 / np.maximum(D0, c0) + g1 / np.maximum(D1, c1))
    ratio = np.where((D0 >= prec) & (pr >= 1e-4), lam * pr, 0.0)
    if raw:
        return ratio, T.ids_sorted
    return np.minimum(1.0, ratio), T.ids_sorted


def test_inclusion_frequencies_match_pi():
    p, nb, sd, f = small_case()
    lam, prec, kcap, S = 6.0, 0.01, 48, 3000
    counts = None
    for s in range(S):
        ids, ipi, dens, over = O.poisson_view(W, H, f, p, nb, sd, kcap, lam, 0.0, prec, seed=1000 + s)
        assert over == 0
        if counts is None:
            counts = np.zeros((W * H, p.shape[0]))
        flat = ids.reshape(W * H, kcap)
        for pix in range(W * H):
            sel = flat[pix][flat[pix] >= 0]
            counts[pix, sel] += 1
    pi_sorted, ids_sorted = inclusion_probs(p, nb, sd, f, lam, prec, dens)
    pi = np.zeros_like(counts)
    pi[:, ids_sorted] = pi_sorted
    freq = counts / S
    m = (pi > 1e-3) & (pi < 1 - 1e-6)
    z = (freq[m] - pi[m]) / np.sqrt(pi[m] * (1 - pi[m]) / S)
    assert m.sum() > 500
    assert abs(np.mean(z)) < 0.1 and 0.8 < np.mean(z ** 2) < 1.2, (np.mean(z), np.mean(z ** 2))
    # every pair's count against its exact binomial tails (the normal approximation
    # understates the upper tail for pi near 1e-3), Bonferroni over all pairs
    from scipy.stats import binom
    k, q = counts[m], pi[m]
    pval = 2.0 * np.minimum(binom.sf(k - 1, S, q), binom.cdf(k, S, q))
    assert pval.min() > 1e-3 / m.sum(), (pval.min(), k[np.argmin(pval)], q[np.argmin(pval)])
    # certain inclusions always happen, impossible ones never
    raw, _ = inclusion_probs(p, nb, sd, f, lam, prec, dens, raw=True)
    rawf = np.zeros_like(counts)
    rawf[:, ids_sorted] = raw
    assert np.all(counts[rawf >= 1.01] == S)
    assert np.all(counts[pi == 0] == 0)


def test_inverse_probabilities_and_expected_count():
    p, nb, sd, f = small_case()
    lam, prec, kcap = 6.0, 0.01, 48
    ids, ipi, dens, over = O.poisson_view(W, H, f, p, nb, sd, kcap, lam, 0.0, prec, seed=7)
    pi_sorted, ids_sorted = inclusion_probs(p, nb, sd, f, lam, prec, dens)
    col = {int(s): q for q, s in enumerate(ids_sorted)}
    flat_ids, flat_ipi = ids.reshape(W * H, kcap), ipi.reshape(W * H, kcap)
    for pix in range(W * H):
        for s in range(kcap):
            if flat_ids[pix, s] < 0:
                assert flat_ipi[pix, s] == 0 and np.all(flat_ids[pix, s:] < 0)
                break
            ref = pi_sorted[pix, col[int(flat_ids[pix, s])]]
            assert abs(1.0 / flat_ipi[pix, s] - ref) <= 2e-5 * ref + 1e-7
    # expected count per pixel: sum of pi (mean over seeds)
    tot = np.zeros(W * H)
    S = 400
    for s in range(S):
        ids, _, _, _ = O.poisson_view(W, H, f, p, nb, sd, kcap, lam, 0.0, prec, seed=50 + s)
        tot += (ids.reshape(W * H, kcap) >= 0).sum(1)
    exp = pi_sorted.sum(1)
    var = (pi_sorted * (1 - pi_sorted)).sum(1)
    z = (tot / S - exp) / np.sqrt(np.maximum(var, 1e-12) / S)
    assert np.abs(z[var > 0]).max() < 5.5


def test_horvitz_thompson_unbiased():
    """E[sum_{i in S(q)} x_i(q) / pi_i(q)] = sum_i x_i(q) for an attribute field that
    decays like the wide component, x_i(q) = a_i exp(-|q - s_i|^2 / 2 h_1^2) (so x / pi
    is bounded, as for EWA weights of splats not wider than h_1); pixels with
    D0 < precision have no samples and a zero estimate by construction."""
    p, nb, sd, f = small_case()
    lam, prec, kcap, S = 4.0, 0.01, 48, 2000
    a = np.random.default_rng(0).uniform(0.5, 1.5, size=p.shape[0])
    pi_sorted, ids_sorted, xg = None, None, None
    est = []
    for s in range(S):
        ids, ipi, dens, _ = O.poisson_view(W, H, f, p, nb, sd, kcap, lam, 0.0, prec, seed=9000 + s)
        if xg is None:
            pi_sorted, ids_sorted = inclusion_probs(p, nb, sd, f, lam, prec, dens)
            (g0, c0), (g1, c1) = proposal(p, nb, sd, f, prec)[0]
            xg = np.zeros((W * H, p.shape[0]))
            xg[:, ids_sorted] = (g1 / c1) * a[ids_sorted][None, :]
        flat = ids.reshape(W * H, kcap)
        x = np.take_along_axis(xg, np.maximum(flat, 0), 1) * (flat >= 0)
        est.append((x * ipi.reshape(W * H, kcap)).sum(-1))
    est = np.array(est)
    pif = np.zeros_like(xg)
    pif[:, ids_sorted] = pi_sorted
    dense = (xg * (pif > 0)).sum(1)
    mean, se = est.mean(0), est.std(0, ddof=1) / math.sqrt(S)
    m = se > 0
    z = (mean[m] - dense[m]) / se[m]
    assert m.sum() > 100
    assert np.abs(z).max() < 5.0 and abs(z.mean()) < 0.2 and 0.7 < np.mean(z ** 2) < 1.3, \
        (np.abs(z).max(), z.mean(), np.mean(z ** 2))
    assert np.allclose(mean[~m], dense[~m], rtol=1e-5, atol=1e-12)


def test_poisson_edge_cases():
    p, nb, sd, f = small_case()
    a = O.poisson_view(W, H, f, p, nb, sd, 16, 4.0, seed=3)
    b = O.poisson_view(W, H, f, p, nb, sd, 16, 4.0, seed=3)
    c = O.poisson_view(W, H, f, p, nb, sd, 16, 4.0, seed=4)
    assert np.array_equal(a[0], b[0]) and not np.array_equal(a[0], c[0])
    ids, ipi, dens, over = O.poisson_view(W, H, f, -p, nb, sd, 16, 4.0)
    assert np.all(ids == -1) and np.all(ipi == 0)
    z = np.zeros((0, 3), np.float32)
    ids, ipi, dens, over = O.poisson_view(8, 8, 9.6, z, z, np.zeros((0, 2), np.float32), 4, 2.0)
    assert np.all(ids == -1)
    # capacity overflow is reported
    ids, ipi, dens, over = O.poisson_view(W, H, f, p, nb, sd, 2, 50.0, seed=1)
    assert over == 1

