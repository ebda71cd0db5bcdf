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


def assert_close(got, ref, rtol, what, atol_frac=None):
    """max |got - ref| <= rtol * max(|ref|) (normwise, infinity norm)."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    scale = max(np.abs(ref).max(), 1e-30)
    err = np.abs(got - ref).max()
    assert err <= rtol * scale, "%s: max err %.3e > %.1e * %.3e" % (what, err, rtol, scale)
    return err / scale
