"""Shared test inputs: seeded scenes, random slot tensors and tolerance checks."""
import numpy as np

from splatting_amd import scenes


def scene_inputs(cfg, n_override=None, sd_scale=1.0):
    """Camera-space float32 inputs of a workload (single view)."""
    sc = scenes.config(cfg, n_override)
    p, n = sc.camera_space()
    nb = np.where(n[:, 2:3] > 0, -n, n).astype(np.float32)
    sd = (sc.stddevs * sd_scale).astype(np.float32)
    return sc, p, n, nb, sd


def random_slots(rng, bn, W, H, k, N, p_empty=0.2):
    idx = rng.integers(0, N, size=(bn, W, H, k)).astype(np.int32)
    idx[rng.uniform(size=idx.shape) < p_empty] = -1
    return idx


def near_slots(rng, p, W, H, f, k, radius=3.0, p_empty=0.1):
    """Slots holding splats whose screen position is near the pixel (non-trivial weights)."""
    sx = -(p[:, 0] * f) / p[:, 2] + W / 2.0
    sy = (p[:, 1] * f) / p[:, 2] + H / 2.0
    idx = np.full((1, W, H, k), -1, np.int32)
    ii, jj = np.meshgrid(np.arange(W), np.arange(H), indexing="ij")
    for s in range(k):
        cand = rng.integers(0, p.shape[0], size=(W, H))
        idx[0, :, :, s] = cand
    # bias: for each pixel replace half the slots by the nearest splats on a coarse grid
    order = np.argsort(sx)
    for i in range(W):
        lo, hi = np.searchsorted(sx[order], [i - radius, i + radius])
        sel = order[lo:hi]
        if sel.size == 0:
            continue
        for j in range(H):
            m = sel[np.abs(sy[sel] - j) < radius]
            if m.size:
                t = min(k // 2, m.size)
                idx[0, i, j, :t] = rng.choice(m, size=t, replace=False)
    idx[rng.uniform(size=idx.shape) < p_empty] = -1
    return idx


EPS32 = 2.0 ** -24     # unit roundoff of float32
EPS64 = 2.0 ** -53     # unit roundoff of float64


def magnitude(fn, *args, **kw):
    """Per-element sum of |terms| of an oracle operator (its abs-terms mode)."""
    from oracle import oracle as O
    with O.abs_terms():
        return fn(*args, **kw)


def assert_close(got, ref, rtol, what, mag=None, ops=64, atol=1e-30, eps=EPS32):
    """Elementwise check |got - ref| <= rtol |ref| + ops * eps * mag + atol (eps: the
    unit roundoff of the kernel's arithmetic, float32 by default).

    ``mag`` is the per-element sum of |terms| the kernel accumulates (from the
    oracle's abs-terms mode), each term scaled by its own float32 evaluation
    condition where it carries an EWA weight (a tail weight exp(-e / 2) evaluated
    from float32 whitening coefficients is exact only to eps * kappa relative,
    kappa growing with |e|; the oracle states it): a float32 sum of m terms loses
    at most about m * EPS32 * sum |terms|, so the floor ``ops * EPS32 * mag``
    admits what float32 evaluation and accumulation can lose where terms cancel
    or weights are far in their tails, and nothing more.  With ``mag`` None the
    bound is relative only (plus ``atol``, below float32's normal range).
    Returns the largest ratio |got - ref| / bound."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    bound = rtol * np.abs(ref) + atol
    if mag is not None:
        mag = np.asarray(mag, np.float64)
        assert mag.shape == ref.shape, (what, mag.shape, ref.shape)
        bound = bound + ops * eps * mag
    err = np.abs(got - ref)
    ratio = err / bound
    worst = int(np.argmax(ratio))
    r = float(ratio.flat[worst])
    assert np.all(np.isfinite(got)), "%s: non-finite output" % what
    print("tolerance %-40s worst |diff| / bound = %.3g" % (what, r))
    assert r <= 1.0, ("%s: element %s: got %.9e, ref %.9e, |diff| %.3e > bound %.3e (rtol %.1e, mag %s)"
                      % (what, np.unravel_index(worst, ref.shape), got.flat[worst], ref.flat[worst], err.flat[worst],
                         bound.flat[worst], rtol, None if mag is None else "%.3e" % mag.flat[worst]))
    return r
